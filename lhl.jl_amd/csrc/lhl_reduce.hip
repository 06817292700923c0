// The reduction J = Z H Z^-1 on one GPU (SciML/LHL.jl lhl_reduce!, src:355-390).
//
//  * balancing (src:297-340): Gauss-Seidel sweeps over i in blocks of 64.  A
//    grid of workgroups sums each block's row/column norms over the columns
//    outside the block with the scales current at the block start; the last
//    workgroup to finish walks the block's 64 indices in order with one wave,
//    adding the in-block terms with the scales as they change.  J itself is
//    never scaled during the sweeps: |A[j,i]| = |J[j,i]| d_i / d_j exactly,
//    because every d is a power of two.
//  * the ELMHES step loop (src:466-501 with the fused trailing update of
//    src:547-678): three kernels per step --
//      pivot    adds the previous step's GEMV partials into column k, scans
//               |A[k+1:n, k]| for the first maximum (strict >, NaN never wins)
//               and, in the last workgroup to finish, publishes the pivot;
//      swap     row/column interchange k+1 <-> p, scaling of the multipliers
//               and the left update of column k+1, all row-local;
//      trailing the rank-1 left update of the trailing block fused with the
//               right update's GEMV, column chunks x row tiles, coalesced
//               down the columns, one partial per chunk.
//  * Ht = H^T (src:2024-2039) by an LDS-tiled transpose.
#include "lhl_common.h"

#include <climits>

namespace lhl {
namespace {

constexpr int kBalB = 64;        // balancing block of indices
constexpr int kRedThreads = 256;
constexpr int kRedMaxBlocks = 64;
constexpr int kChunk = 64;       // trailing columns per GEMV partial

template <typename T> __device__ __forceinline__ T dabs(T x) { return x < 0 ? -x : x; }

template <typename T> __device__ __forceinline__ T ld_coherent(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T> __device__ __forceinline__ void st_coherent(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T> __device__ __forceinline__ T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ---------------------------------------------------------------- balancing

// One block of indices [i0, i0+B) of one sweep.  part holds 2*kBalB values
// per workgroup (column sums, then row sums).
template <typename TJ>
__global__ void __launch_bounds__(256) k_bal_block(const TJ *__restrict__ J, int64_t ldj, int n, int i0,
                                                   double *__restrict__ d, double *__restrict__ isc,
                                                   TJ *__restrict__ part, unsigned *counter, int *changed) {
    const int B = min(kBalB, n - i0);
    const int G = gridDim.x;
    const int g = blockIdx.x;
    const int S = (n + G - 1) / G;
    const int ja = g * S, jb = min(n, ja + S);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __shared__ TJ red_r[4][kBalB];
    __shared__ int last;

    // column sums: wave wv takes block columns u = wv, wv+4, ...; lanes run down j.
    for (int u = wv; u < B; u += 4) {
        const int i = i0 + u;
        TJ acc = 0;
        for (int j = ja + lane; j < jb; j += 64) {
            if (j >= i0 && j < i0 + B) continue;
            acc += dabs(J[j + (int64_t)i * ldj]) * (TJ)isc[j];
        }
        acc = wave_sum(acc);
        if (lane == 0) part[(int64_t)g * 2 * kBalB + u] = acc;
    }
    // row sums: lane = block row, waves split the slice's columns.
    {
        TJ acc = 0;
        if (lane < B) {
            for (int j = ja + wv; j < jb; j += 4) {
                if (j >= i0 && j < i0 + B) continue;
                acc += dabs(J[(i0 + lane) + (int64_t)j * ldj]) * (TJ)d[j];
            }
        }
        red_r[wv][lane] = acc;
    }
    __syncthreads();
    if (threadIdx.x < kBalB)
        part[(int64_t)g * 2 * kBalB + kBalB + threadIdx.x] =
            (red_r[0][threadIdx.x] + red_r[1][threadIdx.x]) + (red_r[2][threadIdx.x] + red_r[3][threadIdx.x]);
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = (atomicAdd(counter, 1u) == (unsigned)(G - 1));
    __syncthreads();
    if (!last) return;
    __threadfence();

    // The last workgroup: in-block terms in index order, one wave.
    __shared__ TJ blk[kBalB][kBalB + 1];   // blk[r][c] = J[i0+r, i0+c]
    __shared__ double sd[kBalB], si[kBalB];
    for (int t = threadIdx.x; t < B * B; t += blockDim.x) {
        int r = t % B, c = t / B;
        blk[r][c] = J[(i0 + r) + (int64_t)(i0 + c) * ldj];
    }
    if (threadIdx.x < B) { sd[threadIdx.x] = d[i0 + threadIdx.x]; si[threadIdx.x] = isc[i0 + threadIdx.x]; }
    __syncthreads();
    if (wv != 0) return;
    TJ cp = 0, rp = 0;
    if (lane < B) {
        for (int q = 0; q < G; q++) {
            cp += ld_coherent(&part[(int64_t)q * 2 * kBalB + lane]);
            rp += ld_coherent(&part[(int64_t)q * 2 * kBalB + kBalB + lane]);
        }
    }
    int any = 0;
    for (int u = 0; u < B; u++) {
        TJ tc = 0, tr = 0;
        if (lane < B && lane != u) {
            tc = dabs(blk[lane][u]) * (TJ)si[lane];
            tr = dabs(blk[u][lane]) * (TJ)sd[lane];
        }
        tc = wave_sum(tc);
        tr = wave_sum(tr);
        const TJ cpu = __shfl(cp, u, 64), rpu = __shfl(rp, u, 64);
        TJ c = (TJ)sd[u] * (cpu + tc);
        TJ r = (TJ)si[u] * (rpu + tr);
        if (c == 0 || r == 0) continue;
        TJ f = 1, s = c + r;
        while (c < r / 2) { c *= 2; r /= 2; f *= 2; }
        while (c >= 2 * r) { c /= 2; r *= 2; f /= 2; }
        if (c + r < (TJ)0.95 * s) {
            any = 1;
            if (lane == 0) { sd[u] = sd[u] * (double)f; si[u] = si[u] / (double)f; }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (lane < B) { d[i0 + lane] = sd[lane]; isc[i0 + lane] = si[lane]; }
    if (lane == 0) {
        if (any) *changed = 1;
        *counter = 0;
    }
}

// A = D^-1 J D (exact: power-of-two products) converted to T.
template <typename TJ, typename T>
__global__ void k_copy_scale(const TJ *__restrict__ J, int64_t ldj, T *__restrict__ A, int64_t lda, int n,
                             const double *__restrict__ d, const double *__restrict__ isc, int balanced) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= n) return;
    TJ v = J[i + (int64_t)j * ldj];
    if (balanced) v = v * (TJ)(isc[i] * d[j]);
    A[i + (int64_t)j * lda] = (T)v;
}

// ---------------------------------------------------------------- step loop

// Column k += sum of the previous step's partials, then the pivot scan of
// step k (unless final_only).  The last workgroup publishes ctl[k & 1],
// ipiv[k] and the interchange of perm (src:381-384).
template <typename T>
__global__ void __launch_bounds__(kRedThreads) k_pivot(T *__restrict__ A, int64_t lda, int n, int k,
                                                       const T *__restrict__ part, int64_t ldp,
                                                       const RedCtl<T> *ctl, T *cand_v, int *cand_i, T *cand_x,
                                                       T *cand_d, T *vk1, int final_only) {
    const RedCtl<T> prev = ctl[(k + 1) & 1];
    const int add = (k >= 1) && !prev.skip;
    const int nch = prev.nchunk;
    T bv = (T)-1, bx = 0;
    int bi = INT_MAX;
    T *col = A + (int64_t)k * lda;
    for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += gridDim.x * kRedThreads) {
        T v = col[i];
        if (add) {
            // partials summed in chunk order; loads issued eight at a time
            T s = 0;
            int c = 0;
            for (; c + 8 <= nch; c += 8) {
                T pv[8];
#pragma unroll
                for (int u = 0; u < 8; u++) pv[u] = part[i + (int64_t)(c + u) * ldp];
#pragma unroll
                for (int u = 0; u < 8; u++) s += pv[u];
            }
            for (; c < nch; c++) s += part[i + (int64_t)c * ldp];
            v = v + s;
            col[i] = v;
        }
        if (!final_only) {
            if (i == k + 1) *vk1 = v;   // A[k+1, k] before k_swap touches it
            if (i >= k + 1) {
                T a = dabs(v);
                if (a == a && a > bv) { bv = a; bi = i; bx = v; }   // rows ascend per thread
            }
        }
    }
    if (final_only) return;
    // block argmax: larger value wins, ties to the smaller row
    __shared__ T sv[kRedThreads], sx[kRedThreads];
    __shared__ int si[kRedThreads];
    sv[threadIdx.x] = bv; si[threadIdx.x] = bi; sx[threadIdx.x] = bx;
    __syncthreads();
    for (int o = kRedThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            T v2 = sv[threadIdx.x + o];
            int i2 = si[threadIdx.x + o];
            if (v2 > sv[threadIdx.x] || (v2 == sv[threadIdx.x] && i2 < si[threadIdx.x])) {
                sv[threadIdx.x] = v2; si[threadIdx.x] = i2; sx[threadIdx.x] = sx[threadIdx.x + o];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        // the block's candidate and its row's diagonal (k_swap overwrites
        // A[p, p] while it decides, so the decision reads this copy)
        cand_v[blockIdx.x] = sv[0]; cand_i[blockIdx.x] = si[0]; cand_x[blockIdx.x] = sx[0];
        const int ci = si[0] == INT_MAX ? k + 1 : si[0];
        cand_d[blockIdx.x] = A[ci + (int64_t)ci * lda];
    }
}

// Interchange rows and columns k+1 <-> p, scale the multipliers by 1/piv and
// apply the left update to column k+1 (src:479-497).  Thread t swaps rows k+1
// and p of column t (t not in {k, k+1, p}) and owns row t of columns k, k+1
// and p; the thread of row k+1 also does row p, so no two threads touch one
// element.
template <typename T>
__global__ void __launch_bounds__(256) k_swap(T *__restrict__ A, int64_t lda, int n, int k, RedCtl<T> *__restrict__ ctl,
                                              const T *__restrict__ cand_v, const int *__restrict__ cand_i,
                                              const T *__restrict__ cand_x, const T *__restrict__ cand_d,
                                              const T *__restrict__ vk1, int ncand, int *__restrict__ ipiv,
                                              int *__restrict__ perm) {
#pragma clang fp contract(off)
    // The pivot decision, in every workgroup from k_pivot's candidates: larger
    // |value| wins, ties to the smaller row; a NaN at row k+1 keeps p = k+1
    // (src:470-477).  Workgroup 0 publishes it for the trailing update.
    __shared__ RedCtl<T> sc;
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        T gv = (T)-1, gx = 0, gd = 0;
        int gi = INT_MAX;
        if (lane < ncand) {
            gv = cand_v[lane];
            gi = cand_i[lane];
            gx = cand_x[lane];
            gd = cand_d[lane];
        }
        for (int o = 32; o > 0; o >>= 1) {
            const T v2 = __shfl_xor(gv, o, 64);
            const int i2 = __shfl_xor(gi, o, 64);
            const T x2 = __shfl_xor(gx, o, 64);
            const T d2 = __shfl_xor(gd, o, 64);
            if (v2 > gv || (v2 == gv && i2 < gi)) { gv = v2; gi = i2; gx = x2; gd = d2; }
        }
        if (lane == 0) {
            const T a1 = *vk1;
            int p = k + 1;
            T piv = a1;
            if (a1 == a1 && gi != INT_MAX) {
                p = gi;
                piv = gx;
            }
            RedCtl<T> c;
            c.p = p;
            c.piv = piv;
            c.pk = (p != k + 1) ? gd : A[(k + 1) + (int64_t)(k + 1) * lda];   // no swap: A[k+1, k+1] is untouched
            c.skip = (piv == (T)0);
            c.nchunk = (n - k - 2 + kChunk - 1) / kChunk;
            c.pad = 0;
            sc = c;
            if (blockIdx.x == 0) {
                ctl[k & 1] = c;
                ipiv[k] = p;
                const int t = perm[k + 1]; perm[k + 1] = perm[p]; perm[p] = t;
            }
        }
    }
    __syncthreads();
    const RedCtl<T> c = sc;
    const int p = c.p;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    T *ck = A + (int64_t)k * lda;
    T *ck1 = A + (int64_t)(k + 1) * lda;
    T *cp = A + (int64_t)p * lda;
    if (p != k + 1 && t != k && t != k + 1 && t != p) {
        T *ct = A + (int64_t)t * lda;
        T u = ct[k + 1];
        ct[k + 1] = ct[p];
        ct[p] = u;
    }
    const int i = t;
    if (p == k + 1) {
        if (!c.skip && i >= k + 2) {
            T l = ck[i] / c.piv;
            ck[i] = l;
            if (c.pk != (T)0) ck1[i] = ck1[i] - l * c.pk;
        }
        return;
    }
    if (i == p) return;
    if (i == k + 1) {
        const T a_k = ck[k + 1], a_k1 = ck1[k + 1], a_p = cp[k + 1];
        const T b_k = ck[p], b_k1 = ck1[p], b_p = cp[p];
        ck[k + 1] = b_k; ck1[k + 1] = b_p; cp[k + 1] = b_k1;
        T lp = a_k, np1 = a_p;
        if (!c.skip) {
            lp = a_k / c.piv;
            if (c.pk != (T)0) np1 = a_p - lp * c.pk;
        }
        ck[p] = lp; ck1[p] = np1; cp[p] = a_k1;
        return;
    }
    const T a = ck1[i], b = cp[i];
    T nb = b;
    if (!c.skip && i >= k + 2) {
        T l = ck[i] / c.piv;
        ck[i] = l;
        if (c.pk != (T)0) nb = b - l * c.pk;
    }
    ck1[i] = nb;
    cp[i] = a;
}

// Trailing update of step k (src:583-648): rows i >= k+2 take
// a_ij = fma(l_i, -A[k+1,j], a_ij); every row accumulates sum_j fma(A[j,k], a_ij)
// over its chunk of kChunk columns into part[i, chunk].
template <typename T>
__global__ void __launch_bounds__(kRedThreads) k_trailing(T *__restrict__ A, int64_t lda, int n, int k,
                                                         const RedCtl<T> *__restrict__ ctl,
                                                         T *__restrict__ part, int64_t ldp) {
    if (ctl[k & 1].skip) return;
    const int chunk = blockIdx.y;
    const int j0 = k + 2 + chunk * kChunk;
    const int j1 = min(n, j0 + kChunk);
    __shared__ T rj[kChunk], vj[kChunk];
    const T *ck = A + (int64_t)k * lda;
    for (int t = threadIdx.x; t < j1 - j0; t += blockDim.x) {
        rj[t] = -A[(k + 1) + (int64_t)(j0 + t) * lda];
        vj[t] = ck[j0 + t];
    }
    __syncthreads();
    const int i = blockIdx.x * kRedThreads + threadIdx.x;
    if (i >= n) return;
    const bool upd = i >= k + 2;
    const T li = upd ? ck[i] : (T)0;
    T acc = 0;
    T *a = A + i + (int64_t)j0 * lda;
    const int m = j1 - j0;
    int t = 0;
    // columns in batches of 16 loads issued together; the FMA order is the
    // column order either way
    if (upd) {
        for (; t + 16 <= m; t += 16) {
            T av[16];
#pragma unroll
            for (int u = 0; u < 16; u++) av[u] = a[(int64_t)(t + u) * lda];
#pragma unroll
            for (int u = 0; u < 16; u++) {
                av[u] = fma_t(li, rj[t + u], av[u]);
                a[(int64_t)(t + u) * lda] = av[u];
                acc = fma_t(vj[t + u], av[u], acc);
            }
        }
        for (; t < m; t++) {
            T a0 = fma_t(li, rj[t], a[(int64_t)t * lda]);
            a[(int64_t)t * lda] = a0;
            acc = fma_t(vj[t], a0, acc);
        }
    } else {
        for (; t + 16 <= m; t += 16) {
            T av[16];
#pragma unroll
            for (int u = 0; u < 16; u++) av[u] = a[(int64_t)(t + u) * lda];
#pragma unroll
            for (int u = 0; u < 16; u++) acc = fma_t(vj[t + u], av[u], acc);
        }
        for (; t < m; t++) acc = fma_t(vj[t], a[(int64_t)t * lda], acc);
    }
    part[i + (int64_t)chunk * ldp] = acc;
}

// Ht[j, i] = A[i, j] for i <= j+1, zero elsewhere (src:2024-2039).
template <typename T>
__global__ void k_ht_fill(const T *__restrict__ A, int64_t lda, T *__restrict__ Ht, int64_t ldh, int n) {
    __shared__ T tile[32][33];
    const int bi = blockIdx.y * 32, bj = blockIdx.x * 32;   // tile rows i of A, cols j of A
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        int i = bi + threadIdx.x, j = bj + r;
        T v = 0;
        if (i < n && j < n && i <= j + 1) v = A[i + (int64_t)j * lda];
        tile[r][threadIdx.x] = v;
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        int j = bj + threadIdx.x, i = bi + r;   // Ht[j, i]
        if (i < n && j < n) Ht[j + (int64_t)i * ldh] = tile[threadIdx.x][r];
    }
}

__global__ void k_iperm(const int *__restrict__ perm, int *__restrict__ iperm, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) iperm[perm[i]] = i;
}

__global__ void k_iota(int *__restrict__ v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}

template <typename T>
__global__ void k_cast_scale(const double *__restrict__ s, const double *__restrict__ is, T *__restrict__ st,
                             T *__restrict__ ist, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { st[i] = (T)s[i]; ist[i] = (T)is[i]; }
}

__global__ void k_fill1(double *__restrict__ a, double *__restrict__ b, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { a[i] = 1.0; b[i] = 1.0; }
}

template <typename TJ>
Status balance_typed(lhl_handle *h) {
    const int n = (int)h->n;
    const TJ *J = h->J.as<TJ>();
    double *d = h->scale64.as<double>(), *isc = h->iscale64.as<double>();
    const int G = std::max(1, std::min(64, (n + 63) / 64));
    LHL_TRY(h->part.ensure(sizeof(TJ) * 2 * kBalB * (size_t)G));
    LHL_TRY(h->counter.ensure(64));
    LHL_TRY(h->flag.ensure(64));
    LHL_HIP_TRY(hipMemsetAsync(h->counter.p, 0, 64, h->stream));
    for (int sweep = 0; sweep < 20; sweep++) {
        LHL_HIP_TRY(hipMemsetAsync(h->flag.p, 0, sizeof(int), h->stream));
        for (int i0 = 0; i0 < n; i0 += kBalB)
            LHL_LAUNCH(h, "balance", k_bal_block<TJ>, dim3(G), dim3(256), 0, J, h->ldj, n, i0, d, isc,
                               h->part.as<TJ>(), h->counter.as<unsigned>(), h->flag.as<int>());
        LHL_HIP_TRY(hipGetLastError());
        int changed = 0;
        LHL_HIP_TRY(hipMemcpyAsync(&changed, h->flag.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        LHL_HIP_TRY(hipStreamSynchronize(h->stream));
        if (!changed) break;
    }
    return {};
}

template <typename T>
Status steps_typed(lhl_handle *h) {
    const int n = (int)h->n;
    T *A = h->A.as<T>();
    const int64_t lda = h->lda;
    const int64_t ldp = pad_ld(n, h->dtype);
    const int maxchunk = std::max(1, (n + kChunk - 1) / kChunk);
    LHL_TRY(h->part.ensure(sizeof(T) * (size_t)ldp * maxchunk));
    LHL_TRY(h->cand.ensure((sizeof(T) * 3 + sizeof(int)) * kRedMaxBlocks + 64));
    LHL_TRY(h->ctl.ensure(sizeof(RedCtl<T>) * 2 + 64));
    LHL_HIP_TRY(hipMemsetAsync(h->ctl.p, 0, sizeof(RedCtl<T>) * 2 + 64, h->stream));
    char *cb = h->cand.as<char>();
    T *cand_v = reinterpret_cast<T *>(cb);
    T *cand_x = cand_v + kRedMaxBlocks;
    T *cand_d = cand_x + kRedMaxBlocks;
    int *cand_i = reinterpret_cast<int *>(cand_d + kRedMaxBlocks);
    RedCtl<T> *ctl = h->ctl.as<RedCtl<T>>();
    T *vk1 = reinterpret_cast<T *>(ctl + 2);
    int *ipiv = h->ipiv.as<int>(), *perm = h->perm.as<int>();
    const int pblocks = std::max(1, std::min(kRedMaxBlocks, (n + kRedThreads - 1) / kRedThreads));
    const int rowblocks = (n + kRedThreads - 1) / kRedThreads;
    for (int k = 0; k + 2 < n; k++) {
        LHL_LAUNCH(h, "reduce_pivot", k_pivot<T>, dim3(pblocks), dim3(kRedThreads), 0, A, lda, n, k,
                   (const T *)h->part.as<T>(), ldp, (const RedCtl<T> *)ctl, cand_v, cand_i, cand_x, cand_d, vk1, 0);
        LHL_LAUNCH(h, "reduce_swap", k_swap<T>, dim3((n + 255) / 256), dim3(256), 0, A, lda, n, k, ctl,
                   (const T *)cand_v, (const int *)cand_i, (const T *)cand_x, (const T *)cand_d, (const T *)vk1,
                   pblocks, ipiv, perm);
        const int nch = (n - k - 2 + kChunk - 1) / kChunk;
        LHL_LAUNCH(h, "reduce_trailing", k_trailing<T>, dim3(rowblocks, nch), dim3(kRedThreads), 0, A, lda, n,
                           k, (const RedCtl<T> *)ctl, h->part.as<T>(), ldp);
    }
    if (n >= 3)
        LHL_LAUNCH(h, "reduce_pivot", k_pivot<T>, dim3(pblocks), dim3(kRedThreads), 0, A, lda, n, n - 2,
                   (const T *)h->part.as<T>(), ldp, (const RedCtl<T> *)ctl, cand_v, cand_i, cand_x, cand_d, vk1, 1);
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

template <typename T>
Status finish_typed(lhl_handle *h) {
    const int n = (int)h->n;
    dim3 tb(32, 8);
    dim3 gr((n + 31) / 32, (n + 31) / 32);
    LHL_LAUNCH(h, "reduce_finish", k_ht_fill<T>, gr, tb, 0, (const T *)h->A.as<T>(), h->lda, h->Ht.as<T>(),
                       h->lda, n);
    LHL_LAUNCH(h, "reduce_finish", k_iperm, dim3((n + 255) / 256), dim3(256), 0, (const int *)h->perm.as<int>(),
                       h->iperm.as<int>(), n);
    LHL_LAUNCH(h, "reduce_finish", k_cast_scale<T>, dim3((n + 255) / 256), dim3(256), 0,
                       (const double *)h->scale64.as<double>(), (const double *)h->iscale64.as<double>(),
                       h->scaleT.as<T>(), h->iscaleT.as<T>(), n);
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

template <typename TJ, typename T>
Status copy_scale_typed(lhl_handle *h, int balanced) {
    const int n = (int)h->n;
    LHL_LAUNCH(h, "reduce_copy", (k_copy_scale<TJ, T>), dim3((n + 255) / 256, n), dim3(256), 0,
                       (const TJ *)h->J.as<TJ>(), h->ldj, h->A.as<T>(), h->lda, n,
                       (const double *)h->scale64.as<double>(), (const double *)h->iscale64.as<double>(), balanced);
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

}  // namespace

Status balance_device(lhl_handle *h) {
    return h->j_dtype == LHL_F64 ? balance_typed<double>(h) : balance_typed<float>(h);
}

// The resident J (h->J, element type h->j_dtype) is already on the device.
Status reduce_device(lhl_handle *h, int balance) {
    const int n = (int)h->n;
    if (n == 0) return {};
    LHL_LAUNCH(h, "reduce_copy", k_fill1, dim3((n + 255) / 256), dim3(256), 0, h->scale64.as<double>(),
                       h->iscale64.as<double>(), n);
    LHL_LAUNCH(h, "reduce_copy", k_iota, dim3((n + 255) / 256), dim3(256), 0, h->perm.as<int>(), n);
    LHL_HIP_TRY(hipGetLastError());
    if (balance) LHL_TRY(balance_device(h));
    if (h->j_dtype == LHL_F64) {
        if (h->dtype == LHL_F64) LHL_TRY((copy_scale_typed<double, double>(h, balance)));
        else LHL_TRY((copy_scale_typed<double, float>(h, balance)));
    } else {
        if (h->dtype == LHL_F64) LHL_TRY((copy_scale_typed<float, double>(h, balance)));
        else LHL_TRY((copy_scale_typed<float, float>(h, balance)));
    }
    if (h->dtype == LHL_F64) {
        LHL_TRY(steps_typed<double>(h));
        LHL_TRY(finish_typed<double>(h));
    } else {
        LHL_TRY(steps_typed<float>(h));
        LHL_TRY(finish_typed<float>(h));
    }
    return linv_device(h);
}

}  // namespace lhl
