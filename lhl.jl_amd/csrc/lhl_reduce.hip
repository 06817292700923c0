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

#include <algorithm>
#include <cstring>
#include <vector>

#include <climits>
#include <cstdint>
#include <cstdlib>

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

    // column sums: wave wv takes block columns u = wv + 4m (m < 16); lanes run
    // down j.  All 16 columns' loads of a 64-row chunk are issued together;
    // in-block rows add an exact zero.
    {
        TJ acc[16];
#pragma unroll
        for (int m = 0; m < 16; m++) acc[m] = 0;
        for (int jc = ja; jc < jb; jc += 64) {
            const int j = jc + lane;
            const bool use = j < jb && !(j >= i0 && j < i0 + B);
            const int jj = min(j, n - 1);
            const TJ w = use ? (TJ)isc[jj] : (TJ)0;
            TJ a[16];
#pragma unroll
            for (int m = 0; m < 16; m++) a[m] = J[jj + (int64_t)(i0 + min(wv + 4 * m, B - 1)) * ldj];
#pragma unroll
            for (int m = 0; m < 16; m++) acc[m] += use ? dabs(a[m]) * w : (TJ)0;
        }
#pragma unroll
        for (int m = 0; m < 16; m++) {
            const TJ t = wave_sum_fast(acc[m]);
            if (lane == 0 && wv + 4 * m < B) part[(int64_t)g * 2 * kBalB + wv + 4 * m] = t;
        }
    }
    // row sums: lane = block row, waves split the slice's columns (j = ja + wv
    // + 4t, in order), eight loads in flight.
    {
        TJ acc = 0;
        const int r = i0 + min(lane, B - 1);
        for (int jb8 = ja + wv; jb8 < jb; jb8 += 32) {
            TJ a[8], w[8];
            bool use[8];
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const int j = jb8 + 4 * t;
                use[t] = j < jb && !(j >= i0 && j < i0 + B);
                const int jj = min(j, n - 1);
                a[t] = J[r + (int64_t)jj * ldj];
                w[t] = (TJ)d[jj];
            }
#pragma unroll
            for (int t = 0; t < 8; t++) acc += use[t] ? dabs(a[t]) * w[t] : (TJ)0;
        }
        red_r[wv][lane] = lane < B ? acc : (TJ)0;
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
    __shared__ TJ sp[2 * kBalB];           // the partials summed over the workgroups
    for (int t = threadIdx.x; t < B * B; t += blockDim.x) {
        int r = t % B, c = t / B;
        blk[r][c] = J[(i0 + r) + (int64_t)(i0 + c) * ldj];
    }
    if (threadIdx.x < B) { sd[threadIdx.x] = d[i0 + threadIdx.x]; si[threadIdx.x] = isc[i0 + threadIdx.x]; }
    if (threadIdx.x < 2 * kBalB) {
        // summed in workgroup order, 16 coherent loads in flight
        const int t = threadIdx.x;
        TJ acc = 0;
        int q = 0;
        for (; q + 16 <= G; q += 16) {
            TJ v[16];
#pragma unroll
            for (int u = 0; u < 16; u++) v[u] = ld_coherent(&part[(int64_t)(q + u) * 2 * kBalB + t]);
#pragma unroll
            for (int u = 0; u < 16; u++) acc += v[u];
        }
        for (; q < G; q++) acc += ld_coherent(&part[(int64_t)q * 2 * kBalB + t]);
        sp[t] = acc;
    }
    __syncthreads();
    if (wv != 0) return;
    TJ cp = 0, rp = 0;
    if (lane < B) { cp = sp[lane]; rp = sp[kBalB + lane]; }
    int any = 0;
    for (int u = 0; u < B; u++) {
        TJ tc = 0, tr = 0;
        if (lane < B && lane != u) {
            tc = dabs(blk[lane][u]) * (TJ)si[lane];
            tr = dabs(blk[u][lane]) * (TJ)sd[lane];
        }
        tc = wave_sum_fast(tc);
        tr = wave_sum_fast(tr);
        const TJ cpu = wsum_detail::lane_read(cp, u), rpu = wsum_detail::lane_read(rp, u);
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

// ------------------------------------------------- row-resident step loop
// One launch per step.  A is kept row-major (rowA[r * ldb + c] = A[r, c]) and
// rows never move: physical row r sits at logical position ilog[r], so the
// reference's row interchange k+1 <-> p (src:479-487) is a relabelling and
// only the column interchange touches memory (two elements per row).  A
// workgroup owns a block of physical rows across all columns, so the rank-1
// left update and the right update's GEMV (src:583-648) finish every row of
// column k+1 inside the workgroup: no partials, no second pass.  What crosses
// workgroups goes through small records written by step k-1 and read by step
// k (ping-pong by step parity):
//   * each workgroup's pivot candidate for column k (largest |A[i, k]| over
//     its rows with logical index >= k+1, ties to the smaller index) with the
//     two entries of that row the interchange rewrites, A[r, k+1] and A[r, p];
//   * the row at logical position k+1: its physical index, A[k+1, k] (the
//     NaN rule of the pivot decision, src:470-477) and A[k+1, k+1];
//   * column k in logical order (gcol), the multipliers' numerators.
// Every workgroup makes the pivot decision from the candidates itself.
template <typename T>
struct RowCand {
    T v, x;        // |A[r, k]| (-1: none) and A[r, k]
    T ak1, ap;     // A[r, k+1] and A[r, i] (i = the row's logical index)
    int i, r;      // logical index (INT_MAX: none), physical row
};
template <typename T>
struct RowNext {
    T vk1, a11;    // A[k+1, k], A[k+1, k+1] of the row at logical position k+1
    int r1, pad;   // its physical row
};

constexpr int kRowWaves = 8;   // waves per workgroup of the step kernel

template <typename T>
__device__ __forceinline__ bool cand_better(T v2, int i2, T v, int i) {
    return v2 > v || (v2 == v && i2 < i);
}

// rowA = (balanced) J transposed, through a 32x32 LDS tile.
template <typename TJ, typename T>
__global__ void k_rows_load(const TJ *__restrict__ J, int64_t ldj, T *__restrict__ B, int64_t ldb, int n,
                            const double *__restrict__ d, const double *__restrict__ isc, int balanced) {
    __shared__ T tile[32][33];
    const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
    for (int q = threadIdx.y; q < 32; q += blockDim.y) {
        const int r = r0 + threadIdx.x, c = c0 + q;   // read J down column c
        T v = 0;
        if (r < n && c < n) {
            TJ w = J[r + (int64_t)c * ldj];
            if (balanced) w = w * (TJ)(isc[r] * d[c]);
            v = (T)w;
        }
        tile[threadIdx.x][q] = v;
    }
    __syncthreads();
    for (int q = threadIdx.y; q < 32; q += blockDim.y) {
        const int r = r0 + q, c = c0 + threadIdx.x;   // write along row r
        if (r < n && c < n) B[(int64_t)r * ldb + c] = tile[q][threadIdx.x];
    }
}

// Records for step 0: identity labels, column 0, candidates over rows >= 1.
template <typename T>
__global__ void __launch_bounds__(256) k_rows_init(const T *__restrict__ B, int64_t ldb, int n, int rows_per_wg,
                                                   RowCand<T> *__restrict__ cout, RowNext<T> *__restrict__ rout,
                                                   T *__restrict__ gout, int *__restrict__ ilog) {
    T bv = (T)-1, bx = 0, bk1 = 0, bp = 0;
    int bi = INT_MAX, br = 0;
    const int r0 = blockIdx.x * rows_per_wg, r1 = min(n, r0 + rows_per_wg);
    for (int r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const T *row = B + (int64_t)r * ldb;
        const T x = row[0];
        ilog[r] = r;
        gout[r] = x;
        if (r == 1) { RowNext<T> q; q.vk1 = x; q.a11 = row[1]; q.r1 = 1; q.pad = 0; *rout = q; }
        const T a = dabs(x);
        if (r >= 1 && a == a && a > bv) { bv = a; bi = r; bx = x; bk1 = row[1]; bp = row[r]; br = r; }
    }
    __shared__ RowCand<T> sc[256];
    RowCand<T> c;
    c.v = bv; c.x = bx; c.ak1 = bk1; c.ap = bp; c.i = bi; c.r = br;
    sc[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        RowCand<T> b = sc[0];
        for (int t = 1; t < (int)blockDim.x; t++)
            if (cand_better(sc[t].v, sc[t].i, b.v, b.i)) b = sc[t];
        cout[blockIdx.x] = b;
    }
}

// The records for step k of the row-resident loop from the rows as they
// stand after k steps of another schedule (the delayed-update panels hand
// over here): candidates over logical rows >= k+1 of column k, the row at
// logical k+1.  Column k in logical order is already in place.
template <typename T>
__global__ void __launch_bounds__(256) k_rows_bridge(const T *__restrict__ B, int64_t ldb, int n, int k,
                                                     int rows_per_wg, const int *__restrict__ ilog,
                                                     RowCand<T> *__restrict__ cout, RowNext<T> *__restrict__ rout) {
    T bv = (T)-1, bx = 0, bk1 = 0, bp = 0;
    int bi = INT_MAX, br = 0;
    const int r0 = blockIdx.x * rows_per_wg, r1 = min(n, r0 + rows_per_wg);
    for (int r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const T *row = B + (int64_t)r * ldb;
        const int il = ilog[r];
        if (il < k + 1) continue;
        const T x = row[k];
        if (il == k + 1) { RowNext<T> q; q.vk1 = x; q.a11 = row[k + 1]; q.r1 = r; q.pad = 0; *rout = q; }
        const T a = dabs(x);
        if (a == a && cand_better(a, il, bv, bi)) { bv = a; bi = il; bx = x; bk1 = row[k + 1]; bp = row[il]; br = r; }
    }
    __shared__ RowCand<T> sc[256];
    RowCand<T> c;
    c.v = bv; c.x = bx; c.ak1 = bk1; c.ap = bp; c.i = bi; c.r = br;
    sc[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int bt = 0;
        for (int t = 1; t < (int)blockDim.x; t++)
            if (cand_better(sc[t].v, sc[t].i, sc[bt].v, sc[bt].i)) bt = t;
        cout[blockIdx.x] = sc[bt];
    }
}

// Step k.  Wave w owns rows blockIdx.x*R + w*RPW*PASSES + i, taken in
// PASSES passes of RPW rows; lanes run along the rows, U column groups of 64
// per row in flight, the next group (or the next pass's first group) loaded
// while the current one is processed.  Dynamic LDS holds the pivot row
// (post-interchange row k+1) and the multipliers l_j, j >= k+2.  The loads
// that do not depend on the pivot (the first column group, the labels and
// columns k, k+1 of the rows, column k in logical order) are issued before the
// decision, so one step costs about three memory round trips plus its
// streaming.
// (The body of the step; `bx` is the workgroup's row block.)
template <typename T, int RPW, int U, int PASSES>
__device__ __forceinline__ void rows_step_body(
    T *__restrict__ B, int64_t ldb, int n, int k, const RowCand<T> *__restrict__ cin, RowCand<T> *__restrict__ cout,
    int ncand, const RowNext<T> *__restrict__ rin, RowNext<T> *__restrict__ rout, const T *__restrict__ gin,
    T *__restrict__ gout, int *__restrict__ ilog, int *__restrict__ ipiv, const int bx) {
#pragma clang fp contract(off)
    constexpr int RW = RPW * PASSES;            // rows per wave
    constexpr int R = kRowWaves * RW;
    constexpr int NT = kRowWaves * 64;
    constexpr int FILL = 8;          // pivot-row / multiplier loads in flight per thread
    constexpr int CPL = 4;           // candidates per lane (up to 256 workgroups)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    extern __shared__ __align__(16) unsigned char lds_dyn[];
    T *prow = reinterpret_cast<T *>(lds_dyn);   // prow[j], j in [k+1, n)
    T *lv = prow + n;                           // lv[j], j in [k+2, n)
    __shared__ T s_piv, s_pk, s_pak1, s_a1;
    __shared__ int s_p, s_rp, s_skip;
    __shared__ RowCand<T> wc[kRowWaves];

    // the rows of this wave and what they need before the decision
    int il[RW], rr[RW];
    int64_t ro[RW];
    bool ok[RW];
    T ck[RW], k1[RW];
#pragma unroll
    for (int q = 0; q < RW; q++) {
        const int r = bx * R + wv * RW + q;
        ok[q] = r < n;
        rr[q] = min(r, n - 1);
        ro[q] = (int64_t)rr[q] * ldb;
        il[q] = ilog[rr[q]];
        ck[q] = B[ro[q] + k];
        k1[q] = B[ro[q] + k + 1];
    }
    const int j0 = k + 2;
    T v[RPW][U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int jc = min(j0 + 64 * u + lane, n - 1);
#pragma unroll
        for (int q = 0; q < RPW; q++) v[q][u] = B[ro[q] + jc];
    }
    // the multipliers' numerators (column k in logical order) of the first fill pass
    T gpre[FILL];
#pragma unroll
    for (int u = 0; u < FILL; u++) gpre[u] = gin[min(k + 1 + NT * u + (int)threadIdx.x, n - 1)];

    // the pivot decision (src:470-477): larger |value| wins, ties to the
    // smaller logical row; a NaN at row k+1 keeps p = k+1
    if (wv == 0) {
        RowCand<T> e[CPL];
#pragma unroll
        for (int u = 0; u < CPL; u++) e[u] = cin[min(lane + 64 * u, ncand - 1)];
        const RowNext<T> nx = *rin;
        // each lane's best, then (value, index) alone through the butterfly
        // with the owning lane; the winner's fields are read from that lane
        int bu = 0;
        T gv = (T)-1;
        int gi = INT_MAX;
#pragma unroll
        for (int u = 0; u < CPL; u++)
            if (lane + 64 * u < ncand && cand_better(e[u].v, e[u].i, gv, gi)) { gv = e[u].v; gi = e[u].i; bu = u; }
        T gx = e[0].x, gk1 = e[0].ak1, gp = e[0].ap;
        int gr = e[0].r;
#pragma unroll
        for (int u = 1; u < CPL; u++)
            if (bu == u) { gx = e[u].x; gk1 = e[u].ak1; gp = e[u].ap; gr = e[u].r; }
        int gl = lane;
        for (int o = 32; o > 0; o >>= 1) {
            const T v2 = __shfl_xor(gv, o, 64);
            const int i2 = __shfl_xor(gi, o, 64), l2 = __shfl_xor(gl, o, 64);
            if (cand_better(v2, i2, gv, gi)) { gv = v2; gi = i2; gl = l2; }
        }
        gx = __shfl(gx, gl, 64); gk1 = __shfl(gk1, gl, 64); gp = __shfl(gp, gl, 64); gr = __shfl(gr, gl, 64);
        if (lane == 0) {
            const T a1 = nx.vk1;
            int p = k + 1, rp = nx.r1;
            T piv = a1, pk = nx.a11, pak1 = nx.a11;
            if (a1 == a1 && gi != INT_MAX) {
                p = gi;
                piv = gx;
                if (p != k + 1) { rp = gr; pk = gp; pak1 = gk1; }
            }
            s_p = p; s_rp = rp; s_piv = piv; s_pk = pk; s_pak1 = pak1; s_a1 = a1;
            s_skip = (piv == (T)0);
            if (bx == 0) ipiv[k] = p;
        }
    }
    __syncthreads();
    const int p = s_p, rp = s_rp, skip = s_skip;
    const T piv = s_piv, pk = s_pk;
    // the element moving into column k+1 (the old column p)
    T t0[RW];
#pragma unroll
    for (int q = 0; q < RW; q++) t0[q] = B[ro[q] + p];
    {
        // post-interchange row k+1 and the multipliers (src:489-491)
        const T pak1 = s_pak1, a1 = s_a1;
        const T *prw = B + (int64_t)rp * ldb;
        auto fill = [&](int jb, const T(&g)[FILL], const T(&c)[FILL]) __attribute__((always_inline)) {
#pragma unroll
            for (int u = 0; u < FILL; u++) {
                const int j = jb + NT * u + (int)threadIdx.x;
                if (j < n) {
                    prow[j] = (j == k + 1) ? pk : (j == p ? pak1 : g[u]);
                    const T num = (j == p) ? a1 : c[u];
                    if (j >= k + 2) lv[j] = skip ? (T)0 : num / piv;
                }
            }
        };
        {
            T g[FILL];
#pragma unroll
            for (int u = 0; u < FILL; u++) g[u] = prw[min(k + 1 + NT * u + (int)threadIdx.x, n - 1)];
            fill(k + 1, g, gpre);
        }
        for (int jb = k + 1 + NT * FILL; jb < n; jb += NT * FILL) {
            T g[FILL], c[FILL];
#pragma unroll
            for (int u = 0; u < FILL; u++) {
                const int jc = min(jb + NT * u + (int)threadIdx.x, n - 1);
                g[u] = prw[jc];
                c[u] = gin[jc];
            }
            fill(jb, g, c);
        }
    }

    // per row: new label and multiplier
    int nl[RW];
    bool act[RW];
    T lq[RW];
#pragma unroll
    for (int q = 0; q < RW; q++) {
        nl[q] = (rr[q] == rp) ? k + 1 : (il[q] == k + 1 ? p : il[q]);
        act[q] = ok[q] && !skip && nl[q] >= k + 2;
        lq[q] = act[q] ? ck[q] / piv : (T)0;
        if (act[q] && lane == 0) B[ro[q] + k] = lq[q];
    }
    __syncthreads();

    RowCand<T> wb;
    wb.v = (T)-1; wb.x = 0; wb.ak1 = 0; wb.ap = 0; wb.i = INT_MAX; wb.r = 0;
#pragma unroll
    for (int ps = 0; ps < PASSES; ps++) {
        // the sweep over columns j >= k+2: a_rj = fma(l_r, -prow_j, a_rj) on
        // the rows below the pivot, and every row accumulates sum_j fma(l_j, a_rj)
        T acc[RPW], c2[RPW], cn[RPW];
#pragma unroll
        for (int q = 0; q < RPW; q++) { acc[q] = 0; c2[q] = 0; cn[q] = 0; }
        for (int jb = j0; jb < n; jb += 64 * U) {
            // next group: this pass's, or past the end the next pass's first
            const bool last = jb + 64 * U >= n;
            const bool jump = last && ps + 1 < PASSES;
            T vn[RPW][U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int jc = min((jump ? j0 : jb + 64 * U) + 64 * u + lane, n - 1);
#pragma unroll
                for (int q = 0; q < RPW; q++) {
                    const int64_t base = (ps + 1 < PASSES && jump) ? ro[min(ps + 1, PASSES - 1) * RPW + q]
                                                                   : ro[ps * RPW + q];
                    vn[q][u] = B[base + jc];
                }
            }
            T pr[U], lj[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int jc = min(jb + 64 * u + lane, n - 1);
                pr[u] = prow[jc];
                lj[u] = lv[jc];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int j = jb + 64 * u + lane;
                const bool in = j < n;
#pragma unroll
                for (int q = 0; q < RPW; q++) {
                    const int w = ps * RPW + q;
                    const T x = (j == p) ? k1[w] : v[q][u];    // column p takes the old column k+1
                    const T y = act[w] ? fma_t(lq[w], -pr[u], x) : x;
                    if (in && ok[w] && (act[w] || j == p)) B[ro[w] + j] = y;
                    acc[q] = in ? fma_t(lj[u], y, acc[q]) : acc[q];
                    c2[q] = (j == j0) ? y : c2[q];
                    cn[q] = (j == nl[w]) ? y : cn[q];
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int q = 0; q < RPW; q++) v[q][u] = vn[q][u];
        }

        // column k+1: left update (src:493-497), then the GEMV sum;
        // candidates and records for step k+1
#pragma unroll
        for (int q = 0; q < RPW; q++) {
            const int w = ps * RPW + q;
            const T s = wave_sum_fast(acc[q]);
            const int lnl = (nl[w] - j0) & 63;
            const T a_nl = wsum_detail::lane_read(cn[q], lnl);
            T t = t0[w];
            if (act[w] && pk != (T)0) t = t - lq[w] * pk;
            if (!skip) t = t + s;
            RowCand<T> c;
            c.v = (T)-1; c.x = t; c.ak1 = c2[q]; c.ap = a_nl; c.i = INT_MAX; c.r = rr[w];
            if (lane == 0 && ok[w]) {
                B[ro[w] + k + 1] = t;
                gout[nl[w]] = t;
                if (nl[w] != il[w]) ilog[rr[w]] = nl[w];
                if (nl[w] == k + 2) { RowNext<T> x; x.vk1 = t; x.a11 = c2[q]; x.r1 = rr[w]; x.pad = 0; *rout = x; }
                const T a = dabs(t);
                if (nl[w] >= k + 2 && a == a) { c.v = a; c.i = nl[w]; }
            }
            if (cand_better(c.v, c.i, wb.v, wb.i)) wb = c;
        }
    }
    if (lane == 0) wc[wv] = wb;
    __syncthreads();
    if (threadIdx.x == 0) {
        int bt = 0;
        T bv = wc[0].v;
        int bi = wc[0].i;
#pragma unroll
        for (int t = 1; t < kRowWaves; t++) {
            const T v2 = wc[t].v;
            const int i2 = wc[t].i;
            if (cand_better(v2, i2, bv, bi)) { bv = v2; bi = i2; bt = t; }
        }
        cout[bx] = wc[bt];
    }
}

template <typename T, int RPW, int U, int PASSES>
__global__ void __launch_bounds__(kRowWaves * 64) k_rows_step(
    T *__restrict__ B, int64_t ldb, int n, int k, const RowCand<T> *__restrict__ cin, RowCand<T> *__restrict__ cout,
    int ncand, const RowNext<T> *__restrict__ rin, RowNext<T> *__restrict__ rout, const T *__restrict__ gin,
    T *__restrict__ gout, int *__restrict__ ilog, int *__restrict__ ipiv) {
    rows_step_body<T, RPW, U, PASSES>(B, ldb, n, k, cin, cout, ncand, rin, rout, gin, gout, ilog, ipiv,
                                      (int)blockIdx.x);
}

// One matrix of a grouped reduction: its row-major working copy, its step
// loop's candidate / next-row / column buffers (ping-pong by step parity) and
// its labels and pivots.
template <typename T>
struct RowGroupDesc {
    T *B;
    int64_t ldb;
    int n, G;                 // order, workgroups of its steps
    RowCand<T> *cand;         // 2 G
    RowNext<T> *nxt;          // 2
    T *g;                     // 2 n
    int *ilog, *ipiv;
};

// Step k of every matrix of a group (blockIdx.y = matrix): one launch per step
// for matrices of one size class.  A matrix whose steps are over (k + 2 >= n)
// or a row block past its rows returns at once; the others run the same step
// body as a lone reduction, so the results are those of lhl_reduce.
template <typename T, int RPW, int U, int PASSES>
__global__ void __launch_bounds__(kRowWaves * 64) k_rows_step_grp(const RowGroupDesc<T> *__restrict__ desc, int k) {
    const RowGroupDesc<T> d = desc[blockIdx.y];
    if (k + 2 >= d.n || (int)blockIdx.x >= d.G) return;
    const int a = k & 1, b = a ^ 1;
    rows_step_body<T, RPW, U, PASSES>(d.B, d.ldb, d.n, k, d.cand + a * d.G, d.cand + b * d.G, d.G, d.nxt + a,
                                      d.nxt + b, d.g + (int64_t)a * d.n, d.g + (int64_t)b * d.n, d.ilog, d.ipiv,
                                      (int)blockIdx.x);
}

// perm[ilog[r]] = r (the row now at each logical position), iperm = ilog.
__global__ void k_rows_perm(const int *__restrict__ ilog, int *__restrict__ perm, int *__restrict__ iperm, int n) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) { const int i = ilog[r]; perm[i] = r; iperm[r] = i; }
}

// factors A[i, c] = rowA[perm[i], c] (column-major, through a 32x32 tile) and
// Ht[c, i] = A[i, c] for i <= c+1 (src:2024-2039), both from the rows in
// logical order.
template <typename T>
__global__ void k_rows_out(const T *__restrict__ B, int64_t ldb, const int *__restrict__ perm, T *__restrict__ A,
                           int64_t lda, T *__restrict__ Ht, int64_t ldh, int n) {
    __shared__ T tile[32][33];
    const int i0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
    for (int q = threadIdx.y; q < 32; q += blockDim.y) {
        const int i = i0 + q, c = c0 + threadIdx.x;
        T v = 0;
        if (i < n && c < n) {
            v = B[(int64_t)perm[i] * ldb + c];
            Ht[c + (int64_t)i * ldh] = (i <= c + 1) ? v : (T)0;
        }
        tile[q][threadIdx.x] = v;
    }
    __syncthreads();
    for (int q = threadIdx.y; q < 32; q += blockDim.y) {
        const int i = i0 + threadIdx.x, c = c0 + q;
        if (i < n && c < n) A[i + (int64_t)c * lda] = tile[threadIdx.x][q];
    }
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

// The row-resident step loop at one (rows per wave, unroll) shape.
template <typename T, int RPW, int U, int PASSES>
Status rows_steps(lhl_handle *h, int64_t ldb, size_t lds, int kstart) {
    const int n = (int)h->n;
    constexpr int R = kRowWaves * RPW * PASSES;
    const int G = (n + R - 1) / R;
    T *B = h->rowA.as<T>();
    LHL_TRY(h->cand.ensure(sizeof(RowCand<T>) * 2 * (size_t)G + sizeof(RowNext<T>) * 2 + 64));
    RowCand<T> *cand = h->cand.as<RowCand<T>>();
    RowNext<T> *nxt = reinterpret_cast<RowNext<T> *>(cand + 2 * G);
    T *g = h->gcol.as<T>();
    int *ilog = h->ilog.as<int>();
    const int a0 = kstart & 1;
    if (kstart == 0)
        LHL_LAUNCH(h, "reduce_copy", k_rows_init<T>, dim3(G), dim3(256), 0, (const T *)B, ldb, n, R, cand, nxt, g, ilog);
    else
        LHL_LAUNCH(h, "reduce_copy", k_rows_bridge<T>, dim3(G), dim3(256), 0, (const T *)B, ldb, n, kstart, R,
                   (const int *)ilog, cand + a0 * G, nxt + a0);
    auto fn = k_rows_step<T, RPW, U, PASSES>;
    LHL_TRY(allow_dynamic_lds(fn));
    for (int k = kstart; k + 2 < n; k++) {
        const int a = k & 1, b = a ^ 1;
        LHL_LAUNCH(h, "reduce_step", fn, dim3(G), dim3(kRowWaves * 64), lds, B, ldb, n, k,
                   (const RowCand<T> *)(cand + a * G), cand + b * G, G, (const RowNext<T> *)(nxt + a), nxt + b,
                   (const T *)(g + (int64_t)a * n), g + (int64_t)b * n, ilog, h->ipiv.as<int>());
    }
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

// Whether the row-resident loop takes this reduction: the pivot row and the
// multipliers (2 n elements) must fit the LDS of one workgroup.
inline bool rows_fit(int64_t n, int dtype) {
    // at most 256 workgroups (the decision reads 4 candidates per lane)
    return n >= 3 && n <= 16384 && (size_t)(2 * n) * elt_size(dtype) + 8 * 1024 <= lds_limit();
}

// Whether the delayed-update (blocked) step loop takes this reduction, and
// its panel width and fewest rows per step workgroup.  The handle's tuning
// knobs decide (the environment variables LHL_BLOCKED_MIN, LHL_NB and
// LHL_RMIN override them, for sweeps); the defaults follow measurement.
inline bool use_blocked(const lhl_handle *h, int n, int *nb, int *rmin, int *kswitch) {
    auto env = [](const char *k, int64_t d) -> int64_t {
        const char *e = std::getenv(k);
        return e ? std::strtoll(e, nullptr, 10) : d;
    };
    // defaults from measurement (one MI355X): at n = 4096 FP64 runs the panels
    // over the first 0.4 n steps (96.9 ms against 102.8 ms for the row-resident
    // loop alone) while FP32 gains nothing from them (67.7 against 67.3 ms); at
    // n = 1024 and 2048 the panels lose in FP64 too (13.3 / 30.3 ms against
    // 8.8 / 22.0 ms), so they start above 3072
    const int64_t bdef = h->dtype == LHL_F64 ? 3073 : INT64_MAX;
    const int64_t bmin = env("LHL_BLOCKED_MIN", h->tune_blocked_min > 0 ? h->tune_blocked_min : bdef);
    *nb = (int)env("LHL_NB", h->tune_nb > 0 ? h->tune_nb : 16);
    *rmin = (int)env("LHL_RMIN", h->tune_rmin > 0 ? h->tune_rmin : 16);
    if (*nb != 8 && *nb != 16) *nb = 16;
    if (*rmin < 1) *rmin = 1;
    // the step from which the row-resident loop takes over (rounded up to a
    // panel boundary), as a fraction of n in thousandths
    const int64_t sw = env("LHL_SWITCH", h->tune_switch > 0 ? h->tune_switch : (h->tune_switch < 0 ? 0 : 400));
    *kswitch = (int)std::min<int64_t>(n, (int64_t)n * std::max<int64_t>(sw, 0) / 1000);
    // the step kernel's LDS: l (n elements) beside up to ~45 KB of fixed tables
    if ((size_t)n * elt_size(h->dtype) + 48 * 1024 > lds_limit()) return false;
    return n >= bmin && n >= 3;
}

template <typename TJ, typename T>
Status rows_typed(lhl_handle *h, int balanced) {
    const int n = (int)h->n;
    const int64_t ldb = pad_ld(n, h->dtype);
    const int maxG = (n + kRowWaves - 1) / kRowWaves;
    LHL_TRY(h->rowA.ensure(sizeof(T) * (size_t)ldb * n));
    LHL_TRY(h->gcol.ensure(sizeof(T) * 2 * (size_t)n));
    LHL_TRY(h->ilog.ensure(sizeof(int) * (size_t)n));
    LHL_TRY(h->cand.ensure(sizeof(RowCand<T>) * 2 * (size_t)maxG + sizeof(RowNext<T>) * 2 + 64));
    T *B = h->rowA.as<T>();
    dim3 tb(32, 8), gr((n + 31) / 32, (n + 31) / 32);
    LHL_LAUNCH(h, "reduce_copy", (k_rows_load<TJ, T>), gr, tb, 0, (const TJ *)h->J.as<TJ>(), h->ldj, B, ldb, n,
               (const double *)h->scale64.as<double>(), (const double *)h->iscale64.as<double>(), balanced);
    const size_t lds = sizeof(T) * 2 * (size_t)n;
    int nb = 0, rmin = 0, kswitch = 0;
    int kstart = 0;
    if (use_blocked(h, n, &nb, &rmin, &kswitch)) {
        // delayed-update panels while the trailing block is large; the
        // row-resident loop takes over from step kstart (n - 2: none left)
        LHL_TRY(blocked_device(h, ldb, nb, rmin, kswitch, &kstart));
    }
    if (kstart + 2 < n) {
        // rows per wave so that about 256 workgroups cover the rows; loads in
        // flight per lane = rows per wave x unroll
        constexpr int W = sizeof(T) == 8 ? 1 : 2;
        if (n <= 2048) LHL_TRY((rows_steps<T, 1, 8 * W, 1>(h, ldb, lds, kstart)));
        else if (n <= 4096) LHL_TRY((rows_steps<T, 2, 8 * W, 1>(h, ldb, lds, kstart)));
        else if (n <= 8192) LHL_TRY((rows_steps<T, 2, 8 * W, 2>(h, ldb, lds, kstart)));
        else LHL_TRY((rows_steps<T, 2, 8 * W, 4>(h, ldb, lds, kstart)));
    }
    LHL_LAUNCH(h, "reduce_finish", k_rows_perm, dim3((n + 255) / 256), dim3(256), 0, (const int *)h->ilog.as<int>(),
               h->perm.as<int>(), h->iperm.as<int>(), n);
    LHL_LAUNCH(h, "reduce_finish", k_rows_out<T>, gr, tb, 0, (const T *)B, ldb, (const int *)h->perm.as<int>(),
               h->A.as<T>(), h->lda, h->Ht.as<T>(), h->lda, n);
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

template <typename T>
Status finish_typed(lhl_handle *h, int rows) {
    const int n = (int)h->n;
    dim3 tb(32, 8);
    dim3 gr((n + 31) / 32, (n + 31) / 32);
    if (!rows) {
        LHL_LAUNCH(h, "reduce_finish", k_ht_fill<T>, gr, tb, 0, (const T *)h->A.as<T>(), h->lda, h->Ht.as<T>(),
                   h->lda, n);
        LHL_LAUNCH(h, "reduce_finish", k_iperm, dim3((n + 255) / 256), dim3(256), 0, (const int *)h->perm.as<int>(),
                   h->iperm.as<int>(), n);
    }
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

// A group of matrices of one element type and one step-kernel shape: each
// matrix's prologue (scales, labels, balancing, the row-major copy, the loop's
// start state), then one launch per step for all of them, then each matrix's
// epilogue.  Every launch goes to the handles' streams, which the caller has
// pointed at one stream.
template <typename TJ, typename T, int RPW, int U, int PASSES>
Status group_rows(const std::vector<lhl_handle *> &v, int balance) {
    constexpr int R = kRowWaves * RPW * PASSES;
    std::vector<RowGroupDesc<T>> d(v.size());
    int maxn = 0, maxG = 0;
    for (size_t m = 0; m < v.size(); m++) {
        lhl_handle *h = v[m];
        const int n = (int)h->n;
        LHL_LAUNCH(h, "reduce_copy", k_fill1, dim3((n + 255) / 256), dim3(256), 0, h->scale64.as<double>(),
                   h->iscale64.as<double>(), n);
        LHL_LAUNCH(h, "reduce_copy", k_iota, dim3((n + 255) / 256), dim3(256), 0, h->perm.as<int>(), n);
        if (balance) LHL_TRY(balance_device(h));
        const int64_t ldb = pad_ld(n, h->dtype);
        const int maxGh = (n + kRowWaves - 1) / kRowWaves;
        LHL_TRY(h->rowA.ensure(sizeof(T) * (size_t)ldb * n));
        LHL_TRY(h->gcol.ensure(sizeof(T) * 2 * (size_t)n));
        LHL_TRY(h->ilog.ensure(sizeof(int) * (size_t)n));
        LHL_TRY(h->cand.ensure(sizeof(RowCand<T>) * 2 * (size_t)maxGh + sizeof(RowNext<T>) * 2 + 64));
        T *B = h->rowA.as<T>();
        const int G = (n + R - 1) / R;
        RowCand<T> *cand = h->cand.as<RowCand<T>>();
        RowNext<T> *nxt = reinterpret_cast<RowNext<T> *>(cand + 2 * G);
        dim3 tb(32, 8), gr((n + 31) / 32, (n + 31) / 32);
        LHL_LAUNCH(h, "reduce_copy", (k_rows_load<TJ, T>), gr, tb, 0, (const TJ *)h->J.as<TJ>(), h->ldj, B, ldb, n,
                   (const double *)h->scale64.as<double>(), (const double *)h->iscale64.as<double>(), balance);
        LHL_LAUNCH(h, "reduce_copy", k_rows_init<T>, dim3(G), dim3(256), 0, (const T *)B, ldb, n, R, cand, nxt,
                   h->gcol.as<T>(), h->ilog.as<int>());
        d[m] = RowGroupDesc<T>{B, ldb, n, G, cand, nxt, h->gcol.as<T>(), h->ilog.as<int>(), h->ipiv.as<int>()};
        maxn = std::max(maxn, n);
        maxG = std::max(maxG, G);
    }
    lhl_handle *h0 = v[0];
    LHL_TRY(h0->grp.ensure(sizeof(RowGroupDesc<T>) * d.size()));
    unsigned char *keep = host_keep(h0, sizeof(RowGroupDesc<T>) * d.size());
    std::memcpy(keep, d.data(), sizeof(RowGroupDesc<T>) * d.size());
    LHL_HIP_TRY(hipMemcpyAsync(h0->grp.p, keep, sizeof(RowGroupDesc<T>) * d.size(), hipMemcpyHostToDevice, h0->stream));
    auto fn = k_rows_step_grp<T, RPW, U, PASSES>;
    LHL_TRY(allow_dynamic_lds(fn));
    const size_t lds = sizeof(T) * 2 * (size_t)maxn;
    const RowGroupDesc<T> *dd = h0->grp.as<RowGroupDesc<T>>();
    for (int k = 0; k + 2 < maxn; k++)
        LHL_LAUNCH(h0, "reduce_step", fn, dim3((unsigned)maxG, (unsigned)v.size()), dim3(kRowWaves * 64), lds, dd, k);
    for (lhl_handle *h : v) {
        const int n = (int)h->n;
        dim3 tb(32, 8), gr((n + 31) / 32, (n + 31) / 32);
        LHL_LAUNCH(h, "reduce_finish", k_rows_perm, dim3((n + 255) / 256), dim3(256), 0,
                   (const int *)h->ilog.as<int>(), h->perm.as<int>(), h->iperm.as<int>(), n);
        LHL_LAUNCH(h, "reduce_finish", k_rows_out<T>, gr, tb, 0, (const T *)h->rowA.as<T>(), pad_ld(n, h->dtype),
                   (const int *)h->perm.as<int>(), h->A.as<T>(), h->lda, h->Ht.as<T>(), h->lda, n);
        LHL_TRY(finish_typed<T>(h, 1));
        LHL_TRY(linv_device(h));
    }
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

template <typename T>
Status group_class(const std::vector<lhl_handle *> &v, int cls, int balance) {
    constexpr int W = sizeof(T) == 8 ? 1 : 2;
    const bool j64 = v[0]->j_dtype == LHL_F64;
    switch (cls) {
    case 0: return j64 ? group_rows<double, T, 1, 8 * W, 1>(v, balance) : group_rows<float, T, 1, 8 * W, 1>(v, balance);
    case 1: return j64 ? group_rows<double, T, 2, 8 * W, 1>(v, balance) : group_rows<float, T, 2, 8 * W, 1>(v, balance);
    case 2: return j64 ? group_rows<double, T, 2, 8 * W, 2>(v, balance) : group_rows<float, T, 2, 8 * W, 2>(v, balance);
    default: return j64 ? group_rows<double, T, 2, 8 * W, 4>(v, balance) : group_rows<float, T, 2, 8 * W, 4>(v, balance);
    }
}

}  // namespace

// Reductions of many matrices (each resident J already on its handle's
// device): the matrices the row-resident loop takes alone (not blocked) are
// grouped by element type, resident-J type and step-kernel shape (the size
// classes rows_typed uses); each group runs on the stream of its first handle
// with one launch per step for all of its matrices.  The others reduce one by
// one on their own streams.  The caller synchronizes every handle's stream.
Status reduce_group_device(lhl_handle *const *hs, int64_t count, int balance) {
    std::vector<lhl_handle *> cls[2][2][4];   // [factor type][J type][shape]
    for (int64_t m = 0; m < count; m++) {
        lhl_handle *h = hs[m];
        const int n = (int)h->n;
        int nb = 0, rmin = 0, ksw = 0;
        if (h->cplx || n < 3 || !rows_fit(n, h->dtype) || use_blocked(h, n, &nb, &rmin, &ksw)) {
            LHL_TRY(reduce_device(h, balance));
            continue;
        }
        const int c = n <= 2048 ? 0 : n <= 4096 ? 1 : n <= 8192 ? 2 : 3;
        cls[h->dtype == LHL_F64 ? 0 : 1][h->j_dtype == LHL_F64 ? 0 : 1][c].push_back(h);
    }
    for (int t = 0; t < 2; t++)
        for (int jt = 0; jt < 2; jt++)
            for (int c = 0; c < 4; c++) {
                std::vector<lhl_handle *> &v = cls[t][jt][c];
                if (v.empty()) continue;
                // every member on the group's stream while it is issued
                std::vector<hipStream_t> keep(v.size());
                hipStream_t gs = v[0]->stream;
                for (size_t m = 0; m < v.size(); m++) {
                    keep[m] = v[m]->stream;
                    if (v[m]->stream != gs) {
                        // the member's earlier work (its J upload) comes first
                        LHL_HIP_TRY(hipStreamSynchronize(v[m]->stream));
                        v[m]->stream = gs;
                    }
                }
                // sub-groups whose row-major copies together stay within a
                // budget (the knob group_bytes), run one after another.  Measured
                // on 64 stiff J with n in [128, 3000] (one MI355X): 16 MiB
                // 1455 / 1099 ms (FP64 / FP32), 64 MiB 1325 / 892, 128 MiB
                // 1179 / 742, 256 MiB 990 / 680, all in one group 1074 / 690,
                // one lhl() after another 1465 / 1152: the default is 256 MiB
                const int64_t budget = v[0]->tune_group_bytes > 0 ? v[0]->tune_group_bytes : (int64_t)256 << 20;
                std::vector<lhl_handle *> order = v;
                std::stable_sort(order.begin(), order.end(),
                                 [](const lhl_handle *x, const lhl_handle *y) { return x->n > y->n; });
                Status st;
                std::vector<lhl_handle *> sub;
                int64_t bytes = 0;
                for (size_t m = 0; m <= order.size() && st.ok(); m++) {
                    const int64_t b = m < order.size() ? order[m]->n * order[m]->n * (int64_t)elt_size(order[m]->dtype) : 0;
                    if (!sub.empty() && (m == order.size() || bytes + b > budget)) {
                        st = t == 0 ? group_class<double>(sub, c, balance) : group_class<float>(sub, c, balance);
                        sub.clear();
                        bytes = 0;
                    }
                    if (m < order.size()) {
                        sub.push_back(order[m]);
                        bytes += b;
                    }
                }
                for (size_t m = 0; m < v.size(); m++) v[m]->stream = keep[m];
                LHL_TRY(st);
            }
    return {};
}

Status balance_device(lhl_handle *h) {
    return h->j_dtype == LHL_F64 ? balance_typed<double>(h) : balance_typed<float>(h);
}

// The resident J (h->J, element type h->j_dtype) is already on the device.
Status reduce_device(lhl_handle *h, int balance) {
    const int n = (int)h->n;
    if (n == 0) return {};
    if (h->cplx) return zreduce_device(h, balance);
    LHL_LAUNCH(h, "reduce_copy", k_fill1, dim3((n + 255) / 256), dim3(256), 0, h->scale64.as<double>(),
                       h->iscale64.as<double>(), n);
    LHL_LAUNCH(h, "reduce_copy", k_iota, dim3((n + 255) / 256), dim3(256), 0, h->perm.as<int>(), n);
    LHL_HIP_TRY(hipGetLastError());
    if (balance) LHL_TRY(balance_device(h));
    const int rows = rows_fit(n, h->dtype) ? 1 : 0;
    if (rows) {
        if (h->j_dtype == LHL_F64) {
            if (h->dtype == LHL_F64) LHL_TRY((rows_typed<double, double>(h, balance)));
            else LHL_TRY((rows_typed<double, float>(h, balance)));
        } else {
            if (h->dtype == LHL_F64) LHL_TRY((rows_typed<float, double>(h, balance)));
            else LHL_TRY((rows_typed<float, float>(h, balance)));
        }
    } else {
        if (h->j_dtype == LHL_F64) {
            if (h->dtype == LHL_F64) LHL_TRY((copy_scale_typed<double, double>(h, balance)));
            else LHL_TRY((copy_scale_typed<double, float>(h, balance)));
        } else {
            if (h->dtype == LHL_F64) LHL_TRY((copy_scale_typed<float, double>(h, balance)));
            else LHL_TRY((copy_scale_typed<float, float>(h, balance)));
        }
        if (h->dtype == LHL_F64) LHL_TRY(steps_typed<double>(h));
        else LHL_TRY(steps_typed<float>(h));
    }
    if (h->dtype == LHL_F64) LHL_TRY(finish_typed<double>(h, rows));
    else LHL_TRY(finish_typed<float>(h, rows));
    return linv_device(h);
}

}  // namespace lhl
