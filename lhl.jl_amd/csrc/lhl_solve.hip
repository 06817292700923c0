// The per-shift Hessenberg solves of the batched solve (SciML/LHL.jl
// _hessenberg_solve!, src:2746-2816, and its planar complex form
// _hessenberg_solve_planar!, src:2823-2884), applyZ!/applyZinv! on a few
// vectors (src:2069-2126), and the planar layout conversions.
//
// k_hess: one workgroup owns NR right-hand sides (PL planes each: 2 for
// complex values) of one shift with the columns resident in LDS (the batched
// solve runs it for the forward chain only; the back substitution of a batch
// runs as blocked MFMA steps over all shifts at once):
//   forward      the interchange/elimination chain, one lane per column, with
//                the flags and multipliers staged through LDS;
//   back subst.  right-looking over 32-row blocks from the bottom: the
//                block's triangle of U is staged in LDS and solved by one wave
//                with lane broadcasts, then every row above subtracts its
//                contiguous 32-element U segment times the finished values.
// The multipliers and U come from the packed per-shift storage.
#include "lhl_common.h"

#include <cstring>

namespace lhl {
namespace {

constexpr int kSolveThreads = 256;
constexpr int kBlk = 64;
constexpr int kHB = 32;    // back-substitution block rows
constexpr int kCH = 512;   // forward-chain staging chunk

template <typename T> __device__ __forceinline__ T shfl(T v, int src) { return __shfl(v, src, 64); }
template <typename T> __device__ __forceinline__ T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

struct Group {
    int shift;     // shift index
    int col0;      // first column
    int ncol;      // columns in this group (<= NR)
    int pad;
};

template <typename T, int NR, int PL, int UC, bool BACK>
__global__ void __launch_bounds__(kSolveThreads) k_hess(
    int n, const T *__restrict__ Ubase, int64_t ustride, const T *__restrict__ lbase, const T *__restrict__ rdbase,
    int64_t lstride, const uint8_t *__restrict__ swbase, int64_t sstride, const Group *__restrict__ groups,
    T *__restrict__ Yg, int64_t ldy) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *Y = reinterpret_cast<T *>(smem_raw);   // Y[v*n + i], v = r*PL + pl
    constexpr int NV = NR * PL;
    constexpr int UP = UC ? 2 : 1;            // planes of U, l, rdiag
    __shared__ T tri[UP][kHB][kHB + 1];        // tri[.][r][c] = U[j0+r, j0+c], c >= r
    __shared__ T trd[UP][kHB];
    __shared__ T stl[UP][kCH];                 // staged multipliers of the forward chain
    __shared__ uint8_t sts[kCH];               // staged interchange flags
    const Group g = groups[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nc = g.ncol;
    const T *U = Ubase + (int64_t)g.shift * ustride;
    const T *lv = lbase + (int64_t)g.shift * lstride;
    const T *rd = rdbase + (int64_t)g.shift * lstride;
    const uint8_t *sw = swbase + (int64_t)g.shift * sstride;
    const int64_t up = ustride / 2, lp = lstride / 2;   // imaginary planes (UC)
    auto y = [&](int v, int i) -> T & { return Y[(int64_t)v * n + i]; };
    auto gcol = [&](int v) -> T * { return Yg + (int64_t)(g.col0 * PL + v) * ldy; };

    for (int i = tid; i < n; i += kSolveThreads)
#pragma unroll
        for (int v = 0; v < NV; v++) y(v, i) = (v / PL < nc) ? gcol(v)[i] : (T)0;
    __syncthreads();

    // ---- forward chain (src:2806-2816): one lane per right-hand side walks the
    // interchange flags and multipliers, staged through LDS kCH at a time.
    const bool chain = wv == 0 && lane < NR;
    const int r = chain ? lane : 0;
    T ca[PL];
#pragma unroll
    for (int pl = 0; pl < PL; pl++) ca[pl] = chain && n > 0 ? y(r * PL + pl, 0) : (T)0;
    for (int c0 = 0; c0 < n - 1; c0 += kCH) {
        const int c1 = min(c0 + kCH, n - 1);
        for (int t = tid; t < c1 - c0; t += kSolveThreads) {
            sts[t] = sw[c0 + t];
            stl[0][t] = lv[c0 + t];
            if constexpr (UC) stl[UP - 1][t] = lv[c0 + t + lp];
        }
        __syncthreads();
        if (chain) {
            for (int k = c0; k < c1; k++) {
                const bool s = sts[k - c0] != 0;
                if constexpr (UC) {
                    const T br = y(r * PL, k + 1), bi = y(r * PL + 1, k + 1);
                    const T xr = s ? br : ca[0], xi = s ? bi : ca[1];
                    y(r * PL, k) = xr;
                    y(r * PL + 1, k) = xi;
                    const T lr = stl[0][k - c0], li = stl[UP - 1][k - c0];
                    const T orr = s ? ca[0] : br, oi = s ? ca[1] : bi;
                    ca[0] = orr - (lr * xr - li * xi);
                    ca[1] = oi - (lr * xi + li * xr);
                } else {
                    const T l = stl[0][k - c0];
#pragma unroll
                    for (int pl = 0; pl < PL; pl++) {
                        const T b = y(r * PL + pl, k + 1);
                        const T xk = s ? b : ca[pl];
                        y(r * PL + pl, k) = xk;
                        ca[pl] = (s ? ca[pl] : b) - l * xk;
                    }
                }
            }
        }
        __syncthreads();
    }
    if (chain && n > 0)
#pragma unroll
        for (int pl = 0; pl < PL; pl++) y(r * PL + pl, n - 1) = ca[pl];
    __syncthreads();

    if constexpr (!BACK) {
        for (int i = tid; i < n; i += kSolveThreads)
#pragma unroll
            for (int v = 0; v < NV; v++)
                if (v / PL < nc) gcol(v)[i] = y(v, i);
        return;
    }
    // ---- back substitution, right-looking over kHB-row blocks from the bottom:
    // solve the block's triangle from LDS, then every row above subtracts its
    // contiguous U segment times the block's finished values.
    for (int j1 = n; j1 > 0; j1 -= kHB) {
        const int j0 = max(0, j1 - kHB);
        const int nb = j1 - j0;
        for (int t = tid; t < kHB * kHB; t += kSolveThreads) {
            const int c = t % kHB, rr = t / kHB;
            T v0 = 0, v1 = 0;
            if (rr < nb && c < nb && c >= rr) {
                const T *Ur = U + upack_off(n, j0 + rr) - (j0 + rr);
                v0 = Ur[j0 + c];
                if constexpr (UC) v1 = Ur[j0 + c + up];
            }
            tri[0][rr][c] = v0;
            if constexpr (UC) tri[UP - 1][rr][c] = v1;
        }
        if (tid < nb) {
            trd[0][tid] = rd[j0 + tid];
            if constexpr (UC) trd[UP - 1][tid] = rd[j0 + tid + lp];
        }
        __syncthreads();
        if (wv == 0) {
            const bool own = lane < nb;
            T v[NV];
#pragma unroll
            for (int q = 0; q < NV; q++) v[q] = own ? y(q, j0 + lane) : (T)0;
            const T dr = own ? trd[0][lane] : (T)0;
            T di = 0;
            if constexpr (UC) di = own ? trd[UP - 1][lane] : (T)0;
            for (int c = nb - 1; c >= 0; c--) {
                if (lane == c) {
                    if constexpr (UC) {
#pragma unroll
                        for (int rr = 0; rr < NR; rr++) {
                            const T a = v[rr * PL], b = v[rr * PL + 1];
                            v[rr * PL] = a * dr - b * di;
                            v[rr * PL + 1] = a * di + b * dr;
                        }
                    } else {
#pragma unroll
                        for (int q = 0; q < NV; q++) v[q] = v[q] * dr;
                    }
                }
                const bool take = own && lane < c;
                const T u = take ? tri[0][lane][c] : (T)0;
                if constexpr (UC) {
                    const T ui = take ? tri[UP - 1][lane][c] : (T)0;
#pragma unroll
                    for (int rr = 0; rr < NR; rr++) {
                        const T xr = shfl(v[rr * PL], c), xi = shfl(v[rr * PL + 1], c);
                        v[rr * PL] = v[rr * PL] - (u * xr - ui * xi);
                        v[rr * PL + 1] = v[rr * PL + 1] - (u * xi + ui * xr);
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < NV; q++) {
                        const T xc = shfl(v[q], c);
                        v[q] = fma_t(-u, xc, v[q]);
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < NV; q++)
                if (own) y(q, j0 + lane) = v[q];
        }
        __syncthreads();
        if (j0 > 0) {
            // here nb == kHB: only the top block is partial, and it has no rows above.
            // RU rows per thread at a time, all their U segments loaded before use.
            constexpr int RU = (UC ? 1 : 2) * (sizeof(T) == 4 ? 2 : 1);
            for (int i0 = tid; i0 < j0; i0 += RU * kSolveThreads) {
                T uu[RU][kHB], ui[RU][kHB];
#pragma unroll
                for (int ru = 0; ru < RU; ru++) {
                    const int i = i0 + ru * kSolveThreads;
                    const int ii = i < j0 ? i : 0;
                    const T *Ur = U + upack_off(n, ii) - ii + j0;   // Ur[c] = U[ii, j0+c]
#pragma unroll
                    for (int c = 0; c < kHB; c++) {
                        uu[ru][c] = Ur[c];
                        if constexpr (UC) ui[ru][c] = Ur[c + up];
                    }
                }
#pragma unroll
                for (int ru = 0; ru < RU; ru++) {
                    const int i = i0 + ru * kSolveThreads;
                    T acc[NV];
#pragma unroll
                    for (int q = 0; q < NV; q++) acc[q] = 0;
#pragma unroll
                    for (int c = 0; c < kHB; c++) {
                        if constexpr (UC) {
#pragma unroll
                            for (int rr = 0; rr < NR; rr++) {
                                const T xr = y(rr * PL, j0 + c), xi = y(rr * PL + 1, j0 + c);
                                acc[rr * PL] = fma_t(uu[ru][c], xr, acc[rr * PL]);
                                acc[rr * PL] = fma_t(-ui[ru][c], xi, acc[rr * PL]);
                                acc[rr * PL + 1] = fma_t(uu[ru][c], xi, acc[rr * PL + 1]);
                                acc[rr * PL + 1] = fma_t(ui[ru][c], xr, acc[rr * PL + 1]);
                            }
                        } else {
#pragma unroll
                            for (int q = 0; q < NV; q++) acc[q] = fma_t(uu[ru][c], y(q, j0 + c), acc[q]);
                        }
                    }
                    if (i < j0)
#pragma unroll
                        for (int q = 0; q < NV; q++) y(q, i) -= acc[q];
                }
            }
            __syncthreads();
        }
    }
    for (int i = tid; i < n; i += kSolveThreads)
#pragma unroll
        for (int v = 0; v < NV; v++)
            if (v / PL < nc) gcol(v)[i] = y(v, i);
}

// applyZ / applyZinv on NR columns (planes independent).
template <typename T, typename TX, int NV>
__global__ void __launch_bounds__(kSolveThreads) k_applyz(int n, int inverse, const T *__restrict__ A, int64_t lda,
                                                          const int *__restrict__ perm, const int *__restrict__ iperm,
                                                          const double *__restrict__ scale,
                                                          const double *__restrict__ iscale, TX *__restrict__ x,
                                                          int64_t ldx, int64_t xplane, int planes, int ncol) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *Y = reinterpret_cast<T *>(smem_raw);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int v0 = blockIdx.x * NV;   // first flattened (column, plane) of this group
    const int nv = min(NV, ncol * planes - v0);
    auto y = [&](int v, int i) -> T & { return Y[(int64_t)v * n + i]; };
    auto xp = [&](int v) -> TX * { int c = (v0 + v) / planes, pl = (v0 + v) % planes; return x + (int64_t)c * ldx + pl * xplane; };
    if (inverse) {
        for (int i = tid; i < n; i += kSolveThreads) {
            const int p = perm[i];
            for (int q = 0; q < NV; q++) y(q, i) = (q < nv) ? (T)(xp(q)[p] * (TX)iscale[p]) : (T)0;
        }
        __syncthreads();
        for (int kb0 = 0; kb0 < n - 2; kb0 += kBlk) {
            const int kb1 = min(kb0 + kBlk, n - 2);
            const int nb = kb1 - kb0;
            if (wv == 0) {
                const int row = kb0 + 1 + lane;
                T v[NV];
                for (int q = 0; q < NV; q++) v[q] = (lane < nb) ? y(q, row) : (T)0;
                for (int kk = 0; kk < nb; kk++) {
                    const T lik = (lane > kk && lane < nb) ? A[row + (int64_t)(kb0 + kk) * lda] : (T)0;
                    for (int q = 0; q < NV; q++) v[q] = fma_t(-lik, shfl(v[q], kk), v[q]);
                }
                for (int q = 0; q < NV; q++) if (lane < nb) y(q, row) = v[q];
            }
            __syncthreads();
            for (int i = kb1 + 1 + tid; i < n; i += kSolveThreads) {
                T acc[NV];
                for (int q = 0; q < NV; q++) acc[q] = 0;
                for (int kk = 0; kk < nb; kk++) {
                    const T l = A[i + (int64_t)(kb0 + kk) * lda];
                    for (int q = 0; q < NV; q++) acc[q] = fma_t(l, y(q, kb0 + 1 + kk), acc[q]);
                }
                for (int q = 0; q < NV; q++) y(q, i) -= acc[q];
            }
            __syncthreads();
        }
        for (int i = tid; i < n; i += kSolveThreads)
            for (int q = 0; q < nv; q++) xp(q)[i] = (TX)y(q, i);
    } else {
        for (int i = tid; i < n; i += kSolveThreads)
            for (int q = 0; q < NV; q++) y(q, i) = (q < nv) ? (T)xp(q)[i] : (T)0;
        __syncthreads();
        const int chunk = kSolveThreads * 4;
        for (int r1 = n; r1 > 2; r1 -= chunk) {
            const int r0 = max(2, r1 - chunk);
            T acc[4][NV];
            for (int m = 0; m < 4; m++) for (int q = 0; q < NV; q++) acc[m][q] = 0;
            for (int k = 0; k < r1 - 2; k++) {
                for (int m = 0; m < 4; m++) {
                    const int i = r0 + tid + m * kSolveThreads;
                    if (i < r1 && i >= k + 2) {
                        const T l = A[i + (int64_t)k * lda];
                        for (int q = 0; q < NV; q++) acc[m][q] = fma_t(l, y(q, k + 1), acc[m][q]);
                    }
                }
            }
            __syncthreads();
            for (int m = 0; m < 4; m++) {
                const int i = r0 + tid + m * kSolveThreads;
                if (i < r1) for (int q = 0; q < NV; q++) y(q, i) += acc[m][q];
            }
            __syncthreads();
        }
        for (int i = tid; i < n; i += kSolveThreads) {
            const int p = iperm[i];
            const TX s = (TX)scale[i];
            for (int q = 0; q < nv; q++) xp(q)[i] = (TX)y(q, p) * s;
        }
    }
}

// planar <-> interleaved / precision conversion
template <typename TS, typename TD>
__global__ void k_to_planar(const TS *__restrict__ src, int64_t lds, int src_complex, TD *__restrict__ dst,
                            int64_t ldd, int64_t plane, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n) return;
    if (src_complex) {
        dst[i + (int64_t)c * ldd] = (TD)src[2 * (i + (int64_t)c * lds)];
        dst[i + (int64_t)c * ldd + plane] = (TD)src[2 * (i + (int64_t)c * lds) + 1];
    } else {
        dst[i + (int64_t)c * ldd] = (TD)src[i + (int64_t)c * lds];
    }
}
template <typename TS, typename TD>
__global__ void k_from_planar(const TS *__restrict__ src, int64_t lds, int64_t plane, TD *__restrict__ dst,
                              int64_t ldd, int dst_complex, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n) return;
    if (dst_complex) {
        dst[2 * (i + (int64_t)c * ldd)] = (TD)src[i + (int64_t)c * lds];
        dst[2 * (i + (int64_t)c * ldd) + 1] = (TD)src[i + (int64_t)c * lds + plane];
    } else {
        dst[i + (int64_t)c * ldd] = (TD)src[i + (int64_t)c * lds];
    }
}

// ---------------------------------------------------------------- launchers


// Right-hand sides per workgroup: as many as fit in LDS, at most 4.
inline int pick_nr(int64_t n, int planes, size_t es) {
    const size_t budget = lds_limit() - 28 * 1024;   // static LDS of k_hess
    const size_t per = (size_t)planes * (size_t)n * es;
    int nr = per ? (int)(budget / per) : 4;
    return nr >= 4 ? 4 : nr >= 2 ? 2 : nr;
}

template <typename T, int NR, int PL, int UC>
Status launch_hess_t(lhl_handle *h, const SolveArgs &a, const Group *groups, int ngroups, void *Y, int64_t ldy) {
    const int n = (int)h->n;
    const size_t lds = sizeof(T) * (size_t)NR * PL * n;
    auto fn = k_hess<T, NR, PL, UC, false>;   // the forward chain only; U is solved by the blocked kernels
    LHL_HIP_TRY(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    LHL_LAUNCH(h, "fwd_chain", fn, dim3(ngroups), dim3(kSolveThreads), lds, n, (const T *)a.U, a.ustride, (const T *)a.l,
               (const T *)a.rdiag, a.lstride, a.swp, a.sstride, groups, (T *)Y, ldy);
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

template <typename T>
Status hess_typed(lhl_handle *h, const SolveArgs &a, void *Y, int64_t ldy) {
    const int PLn = a.x_complex ? 2 : 1;
    const int nr = pick_nr(h->n, PLn, sizeof(T));
    if (nr < 1) return {LHL_ERR_ARG, "lhl_ldiv: n too large for the LDS-resident Hessenberg solve"};
    std::vector<Group> hg;
    std::vector<int64_t> cb((size_t)a.nshift + 1);
    if (a.col_begin) {
        for (int64_t s = 0; s <= a.nshift; s++) cb[s] = a.col_begin[s];
    } else {
        cb[0] = 0;
        cb[1] = a.ncol;
    }
    for (int64_t s = 0; s < a.nshift; s++)
        for (int64_t c = cb[s]; c < cb[s + 1]; c += nr)
            hg.push_back(Group{(int)s, (int)c, (int)std::min<int64_t>(nr, cb[s + 1] - c), 0});
    if (hg.empty()) return {};
    LHL_TRY(h->nbuf.ensure(sizeof(Group) * hg.size()));
    h->host_hg.resize(sizeof(Group) * hg.size());
    std::memcpy(h->host_hg.data(), hg.data(), h->host_hg.size());
    LHL_HIP_TRY(hipMemcpyAsync(h->nbuf.p, h->host_hg.data(), h->host_hg.size(), hipMemcpyHostToDevice, h->stream));
    const Group *dg = h->nbuf.as<Group>();
    const int ng = (int)hg.size();
    Status st;
    if (a.shift_complex) {
        if (nr >= 8) st = launch_hess_t<T, 8, 2, 1>(h, a, dg, ng, Y, ldy);
        else if (nr >= 4) st = launch_hess_t<T, 4, 2, 1>(h, a, dg, ng, Y, ldy);
        else if (nr >= 2) st = launch_hess_t<T, 2, 2, 1>(h, a, dg, ng, Y, ldy);
        else st = launch_hess_t<T, 1, 2, 1>(h, a, dg, ng, Y, ldy);
    } else if (a.x_complex) {
        if (nr >= 8) st = launch_hess_t<T, 8, 2, 0>(h, a, dg, ng, Y, ldy);
        else if (nr >= 4) st = launch_hess_t<T, 4, 2, 0>(h, a, dg, ng, Y, ldy);
        else if (nr >= 2) st = launch_hess_t<T, 2, 2, 0>(h, a, dg, ng, Y, ldy);
        else st = launch_hess_t<T, 1, 2, 0>(h, a, dg, ng, Y, ldy);
    } else {
        if (nr >= 8) st = launch_hess_t<T, 8, 1, 0>(h, a, dg, ng, Y, ldy);
        else if (nr >= 4) st = launch_hess_t<T, 4, 1, 0>(h, a, dg, ng, Y, ldy);
        else if (nr >= 2) st = launch_hess_t<T, 2, 1, 0>(h, a, dg, ng, Y, ldy);
        else st = launch_hess_t<T, 1, 1, 0>(h, a, dg, ng, Y, ldy);
    }
    return st;
}

template <typename T, typename TX>
Status applyz_typed(lhl_handle *h, int inverse, void *x, int64_t ncol, int64_t ldx, int64_t xplane, int planes) {
    const int n = (int)h->n;
    constexpr int NV = 4;
    const size_t lds = sizeof(T) * (size_t)NV * n;
    if (lds > lds_limit() - 4096) {
        // fall back to one vector per group
        const size_t lds1 = sizeof(T) * (size_t)n;
        if (lds1 > lds_limit() - 4096) return {LHL_ERR_ARG, "applyZ: n too large for the LDS-resident sweep"};
        auto fn = k_applyz<T, TX, 1>;
        LHL_HIP_TRY(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds1));
        const int groups = (int)(ncol * planes);
        LHL_LAUNCH(h, "applyz", fn, dim3(groups), dim3(kSolveThreads), lds1, n, inverse,
                           (const T *)h->A.as<T>(), h->lda, (const int *)h->perm.as<int>(),
                           (const int *)h->iperm.as<int>(), (const double *)h->scale64.as<double>(),
                           (const double *)h->iscale64.as<double>(), (TX *)x, ldx, xplane, planes, (int)ncol);
    } else {
        auto fn = k_applyz<T, TX, NV>;
        LHL_HIP_TRY(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        const int groups = (int)((ncol * planes + NV - 1) / NV);
        LHL_LAUNCH(h, "applyz", fn, dim3(groups), dim3(kSolveThreads), lds, n, inverse,
                           (const T *)h->A.as<T>(), h->lda, (const int *)h->perm.as<int>(),
                           (const int *)h->iperm.as<int>(), (const double *)h->scale64.as<double>(),
                           (const double *)h->iscale64.as<double>(), (TX *)x, ldx, xplane, planes, (int)ncol);
    }
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

}  // namespace

Status hess_device(lhl_handle *h, const SolveArgs &a, void *Y, int64_t ldy) {
    if (h->n == 0 || a.ncol == 0) return {};
    return h->dtype == LHL_F64 ? hess_typed<double>(h, a, Y, ldy) : hess_typed<float>(h, a, Y, ldy);
}

Status applyz_device(lhl_handle *h, int inverse, void *x, int64_t ncol, int64_t ldx, int64_t xplane, int xdt,
                     int x_complex) {
    if (h->n == 0 || ncol == 0) return {};
    const int planes = x_complex ? 2 : 1;
    if (h->dtype == LHL_F64) return applyz_typed<double, double>(h, inverse, x, ncol, ldx, xplane, planes);
    if (xdt == LHL_F64) return applyz_typed<float, double>(h, inverse, x, ncol, ldx, xplane, planes);
    return applyz_typed<float, float>(h, inverse, x, ncol, ldx, xplane, planes);
}

Status to_planar(lhl_handle *h, const void *src, int64_t ld_src, int src_dt, int src_complex, void *dst,
                 int64_t ld_dst, int64_t plane_dst, int dst_dt, int64_t n, int64_t ncol) {
    if (n == 0 || ncol == 0) return {};
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)ncol);
#define TP(TS, TD)                                                                                        \
    LHL_LAUNCH(h, "convert", (k_to_planar<TS, TD>), grid, dim3(256), 0, (const TS *)src, ld_src,     \
                       src_complex, (TD *)dst, ld_dst, plane_dst, (int)n)
    if (src_dt == LHL_F64 && dst_dt == LHL_F64) TP(double, double);
    else if (src_dt == LHL_F64) TP(double, float);
    else if (dst_dt == LHL_F64) TP(float, double);
    else TP(float, float);
#undef TP
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

Status from_planar(lhl_handle *h, const void *src, int64_t ld_src, int64_t plane_src, int src_dt, void *dst,
                   int64_t ld_dst, int dst_dt, int dst_complex, int64_t n, int64_t ncol) {
    if (n == 0 || ncol == 0) return {};
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)ncol);
#define FP(TS, TD)                                                                                        \
    LHL_LAUNCH(h, "convert", (k_from_planar<TS, TD>), grid, dim3(256), 0, (const TS *)src, ld_src,   \
                       plane_src, (TD *)dst, ld_dst, dst_complex, (int)n)
    if (src_dt == LHL_F64 && dst_dt == LHL_F64) FP(double, double);
    else if (src_dt == LHL_F64) FP(double, float);
    else if (dst_dt == LHL_F64) FP(float, double);
    else FP(float, float);
#undef FP
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

}  // namespace lhl
