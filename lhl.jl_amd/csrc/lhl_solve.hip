// The solves x <- (sigma I + tau J)^-1 x = Z (sigma I + tau H)^-1 Z^-1 x
// (SciML/LHL.jl lhl_ldiv!, src:3052-3123) and the refinement residual.
//
// One workgroup owns NR right-hand sides (PL planes each: 2 for complex values)
// for the whole solve, with the vectors resident in LDS:
//   gather       y = P^-1 D^-1 x                                    (src:2130-2140)
//   L^-1 sweep   blocks of 64 steps: one wave solves the 64x64 triangle with
//                lane broadcasts, then every wave updates the rows below with the
//                block's 64 multiplier columns (coalesced down the columns).
//   Hessenberg   forward interchange/elimination chain (src:2806-2816), then the
//                back substitution in 64-row blocks: all waves form the dot
//                products against the finished rows (U rows are contiguous),
//                one wave finishes the 64x64 triangle column by column.
//   L sweep      y <- L y is a triangular product with no recurrence; rows are
//                done bottom-up in chunks so no row is read after it is written.
//   scatter      x = D P y                                            (src:3064-3068)
// The multipliers are read from the factors A itself (column k below row k+1).
#include "lhl_common.h"

namespace lhl {
namespace {

constexpr int kSolveThreads = 256;
constexpr int kBlk = 64;

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

template <typename T, typename TX, int NR, int PL, int UC>
__global__ void __launch_bounds__(kSolveThreads) k_ldiv(
    int n, const T *__restrict__ A, int64_t lda, const int *__restrict__ perm, const int *__restrict__ iperm,
    const double *__restrict__ scale, const double *__restrict__ iscale, const T *__restrict__ Ubase,
    int64_t ustride, const T *__restrict__ lbase, const T *__restrict__ rdbase, int64_t lstride,
    const uint8_t *__restrict__ swbase, int64_t sstride, const Group *__restrict__ groups,
    const TX *__restrict__ xin, TX *__restrict__ xout, int64_t ldx, int64_t xplane, int accumulate) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *Y = reinterpret_cast<T *>(smem_raw);   // Y[(r*PL + pl)*n + i]
    __shared__ T tdot[NR * PL][kBlk];
    const Group g = groups[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    constexpr int NW = kSolveThreads / 64;
    constexpr int NV = NR * PL;
    const int nc = g.ncol;
    const T *U = Ubase + (int64_t)g.shift * ustride;
    const T *lv = lbase + (int64_t)g.shift * lstride;
    const T *rd = rdbase + (int64_t)g.shift * lstride;
    const uint8_t *sw = swbase + (int64_t)g.shift * sstride;
    const int64_t up = ustride / 2, lp = lstride / 2;   // imaginary planes (UC)
    auto y = [&](int v, int i) -> T & { return Y[(int64_t)v * n + i]; };

    // ---- gather
    for (int i = tid; i < n; i += kSolveThreads) {
        const int p = perm[i];
        const double s = iscale[p];
        for (int r = 0; r < NR; r++)
            for (int pl = 0; pl < PL; pl++) {
                T v = 0;
                if (r < nc) v = (T)(xin[p + (int64_t)(g.col0 + r) * ldx + pl * xplane] * (TX)s);
                y(r * PL + pl, i) = v;
            }
    }
    __syncthreads();

    // ---- L^-1 sweep: steps k = 0..n-3, step k: y[i] -= A[i,k] * y[k+1], i >= k+2
    for (int kb0 = 0; kb0 < n - 2; kb0 += kBlk) {
        const int kb1 = min(kb0 + kBlk, n - 2);
        const int nb = kb1 - kb0;
        if (wv == 0) {
            // rows kb0+1+t, t < nb
            const int row = kb0 + 1 + lane;
            T v[NV];
#pragma unroll
            for (int q = 0; q < NV; q++) v[q] = (lane < nb) ? y(q, row) : (T)0;
            for (int kk = 0; kk < nb; kk++) {
                const T lik = (lane > kk && lane < nb) ? A[row + (int64_t)(kb0 + kk) * lda] : (T)0;
#pragma unroll
                for (int q = 0; q < NV; q++) {
                    const T yk = shfl(v[q], kk);
                    v[q] = __builtin_fma(-lik, yk, v[q]);
                }
            }
#pragma unroll
            for (int q = 0; q < NV; q++) if (lane < nb) y(q, row) = v[q];
        }
        __syncthreads();
        for (int i = kb1 + 1 + tid; i < n; i += kSolveThreads) {
            T acc[NV];
#pragma unroll
            for (int q = 0; q < NV; q++) acc[q] = 0;
            const T *Li = A + i + (int64_t)kb0 * lda;
            int kk = 0;
            for (; kk + 4 <= nb; kk += 4) {
                const T l0 = Li[(int64_t)kk * lda], l1 = Li[(int64_t)(kk + 1) * lda];
                const T l2 = Li[(int64_t)(kk + 2) * lda], l3 = Li[(int64_t)(kk + 3) * lda];
#pragma unroll
                for (int q = 0; q < NV; q++) {
                    acc[q] = __builtin_fma(l0, y(q, kb0 + 1 + kk), acc[q]);
                    acc[q] = __builtin_fma(l1, y(q, kb0 + 2 + kk), acc[q]);
                    acc[q] = __builtin_fma(l2, y(q, kb0 + 3 + kk), acc[q]);
                    acc[q] = __builtin_fma(l3, y(q, kb0 + 4 + kk), acc[q]);
                }
            }
            for (; kk < nb; kk++) {
                const T l0 = Li[(int64_t)kk * lda];
#pragma unroll
                for (int q = 0; q < NV; q++) acc[q] = __builtin_fma(l0, y(q, kb0 + 1 + kk), acc[q]);
            }
#pragma unroll
            for (int q = 0; q < NV; q++) y(q, i) -= acc[q];
        }
        __syncthreads();
    }

    // ---- Hessenberg forward chain (one lane per right-hand side)
    if (wv == 0 && lane < NR) {
        const int r = lane;
        if constexpr (UC) {
            T ar = y(r * PL, 0), ai = y(r * PL + 1, 0);
            for (int k = 0; k < n - 1; k++) {
                const bool s = sw[k] != 0;
                const T br = y(r * PL, k + 1), bi = y(r * PL + 1, k + 1);
                const T xr = s ? br : ar, xi = s ? bi : ai;
                y(r * PL, k) = xr; y(r * PL + 1, k) = xi;
                const T lr = lv[k], li = lv[k + lp];
                const T orr = s ? ar : br, oi = s ? ai : bi;
                ar = orr - (lr * xr - li * xi);
                ai = oi - (lr * xi + li * xr);
            }
            if (n >= 1) { y(r * PL, n - 1) = ar; y(r * PL + 1, n - 1) = ai; }
        } else {
            T a[PL];
#pragma unroll
            for (int pl = 0; pl < PL; pl++) a[pl] = y(r * PL + pl, 0);
            for (int k = 0; k < n - 1; k++) {
                const bool s = sw[k] != 0;
                const T l = lv[k];
#pragma unroll
                for (int pl = 0; pl < PL; pl++) {
                    const T b = y(r * PL + pl, k + 1);
                    const T xk = s ? b : a[pl];
                    y(r * PL + pl, k) = xk;
                    a[pl] = (s ? a[pl] : b) - l * xk;
                }
            }
#pragma unroll
            for (int pl = 0; pl < PL; pl++) if (n >= 1) y(r * PL + pl, n - 1) = a[pl];
        }
    }
    __syncthreads();

    // ---- back substitution: x_j = (y_j - sum_{i>j} U[j,i] x_i) * rdiag[j]
    for (int j1 = n; j1 > 0; j1 -= kBlk) {
        const int j0 = max(0, j1 - kBlk);
        const int nb = j1 - j0;
        // far part: rows j0..j1-1 against finished rows i >= j1; wave wv takes rows j0+wv+NW*q
        for (int jr = wv; jr < nb; jr += NW) {
            const int j = j0 + jr;
            const T *Ur = U + upack_off(n, j) - j;   // Ur[i] = U[j, i]
            T acc[NV];
#pragma unroll
            for (int q = 0; q < NV; q++) acc[q] = 0;
            for (int i = j1 + lane; i < n; i += 64) {
                const T u = Ur[i];
                if constexpr (UC) {
                    const T ui = Ur[i + up];
#pragma unroll
                    for (int r = 0; r < NR; r++) {
                        const T xr = y(r * PL, i), xi = y(r * PL + 1, i);
                        acc[r * PL] = __builtin_fma(u, xr, acc[r * PL]);
                        acc[r * PL] = __builtin_fma(-ui, xi, acc[r * PL]);
                        acc[r * PL + 1] = __builtin_fma(u, xi, acc[r * PL + 1]);
                        acc[r * PL + 1] = __builtin_fma(ui, xr, acc[r * PL + 1]);
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < NV; q++) acc[q] = __builtin_fma(u, y(q, i), acc[q]);
                }
            }
#pragma unroll
            for (int q = 0; q < NV; q++) {
                const T s = wave_sum(acc[q]);
                if (lane == 0) tdot[q][jr] = s;
            }
        }
        __syncthreads();
        if (wv == 0) {
            // lane t owns row j0+t; columns i = j1-1 .. j0 finish in turn
            const int j = j0 + lane;
            const bool own = lane < nb;
            const T *Ur = U + upack_off(n, own ? j : 0) - (own ? j : 0);
            T v[NV];
#pragma unroll
            for (int q = 0; q < NV; q++) v[q] = own ? (y(q, j) - tdot[q][lane]) : (T)0;
            T dr = 0, di = 0;
            if (own) { dr = rd[j]; if constexpr (UC) di = rd[j + lp]; }
            for (int c = nb - 1; c >= 0; c--) {
                const int i = j0 + c;
                // lane c finishes x_i
                if (lane == c) {
                    if constexpr (UC) {
#pragma unroll
                        for (int r = 0; r < NR; r++) {
                            const T a = v[r * PL], b = v[r * PL + 1];
                            v[r * PL] = a * dr - b * di;
                            v[r * PL + 1] = a * di + b * dr;
                        }
                    } else {
#pragma unroll
                        for (int q = 0; q < NV; q++) v[q] = v[q] * dr;
                    }
                }
                const bool take = own && lane < c;
                const T u = take ? Ur[i] : (T)0;
                T ui = 0;
                if constexpr (UC) ui = take ? Ur[i + up] : (T)0;
                if constexpr (UC) {
#pragma unroll
                    for (int r = 0; r < NR; r++) {
                        const T xr = shfl(v[r * PL], c), xi = shfl(v[r * PL + 1], c);
                        v[r * PL] = v[r * PL] - (u * xr - ui * xi);
                        v[r * PL + 1] = v[r * PL + 1] - (u * xi + ui * xr);
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < NV; q++) {
                        const T xi = shfl(v[q], c);
                        v[q] = __builtin_fma(-u, xi, v[q]);
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < NV; q++) if (own) y(q, j) = v[q];
        }
        __syncthreads();
    }

    // ---- L sweep: y[i] += sum_{k <= i-2} A[i,k] * y[k+1], bottom-up row chunks
    constexpr int RPT = 4;   // rows per thread per chunk
    const int chunk = kSolveThreads * RPT;
    for (int r1 = n; r1 > 2; r1 -= chunk) {
        const int r0 = max(2, r1 - chunk);
        T acc[RPT][NV];
#pragma unroll
        for (int m = 0; m < RPT; m++)
#pragma unroll
            for (int q = 0; q < NV; q++) acc[m][q] = 0;
        const int kmax = r1 - 2;   // steps k with k+2 <= i <= r1-1
        for (int k = 0; k < kmax; k++) {
            T yk[NV];
#pragma unroll
            for (int q = 0; q < NV; q++) yk[q] = y(q, k + 1);
            const T *Lk = A + (int64_t)k * lda;
#pragma unroll
            for (int m = 0; m < RPT; m++) {
                const int i = r0 + tid + m * kSolveThreads;
                if (i < r1 && i >= k + 2) {
                    const T l = Lk[i];
#pragma unroll
                    for (int q = 0; q < NV; q++) acc[m][q] = __builtin_fma(l, yk[q], acc[m][q]);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < RPT; m++) {
            const int i = r0 + tid + m * kSolveThreads;
            if (i < r1)
#pragma unroll
                for (int q = 0; q < NV; q++) y(q, i) += acc[m][q];
        }
        __syncthreads();
    }

    // ---- scatter x = D P y
    for (int i = tid; i < n; i += kSolveThreads) {
        const int p = iperm[i];
        const TX s = (TX)scale[i];
        for (int r = 0; r < nc; r++)
            for (int pl = 0; pl < PL; pl++) {
                TX v = (TX)y(r * PL + pl, p) * s;
                TX *dst = xout + i + (int64_t)(g.col0 + r) * ldx + pl * xplane;
                if (accumulate) *dst += v;
                else *dst = v;
            }
    }
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
                    for (int q = 0; q < NV; q++) v[q] = __builtin_fma(-lik, shfl(v[q], kk), v[q]);
                }
                for (int q = 0; q < NV; q++) if (lane < nb) y(q, row) = v[q];
            }
            __syncthreads();
            for (int i = kb1 + 1 + tid; i < n; i += kSolveThreads) {
                T acc[NV];
                for (int q = 0; q < NV; q++) acc[q] = 0;
                for (int kk = 0; kk < nb; kk++) {
                    const T l = A[i + (int64_t)(kb0 + kk) * lda];
                    for (int q = 0; q < NV; q++) acc[q] = __builtin_fma(l, y(q, kb0 + 1 + kk), acc[q]);
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
                        for (int q = 0; q < NV; q++) acc[m][q] = __builtin_fma(l, y(q, k + 1), acc[m][q]);
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

// ---- residual r = b - W x (W dense) or b - (sigma x + tau J x), planar columns
template <typename TW, typename TX, int NC>
__global__ void __launch_bounds__(256) k_residual(int n, const TW *__restrict__ W, int64_t ldw, int w_complex,
                                                  int64_t wplane, const double *__restrict__ colshift,
                                                  const TX *__restrict__ b, const TX *__restrict__ x,
                                                  TX *__restrict__ r, int64_t ld, int64_t plane, int planes,
                                                  int ncol) {
    // block: 256 rows x NC columns (planes folded into the columns of a complex x)
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int c0 = blockIdx.y * NC;
    __shared__ TX xs[2][NC][256];
    TX acc[2][NC];
    for (int c = 0; c < NC; c++) { acc[0][c] = 0; acc[1][c] = 0; }
    for (int j0 = 0; j0 < n; j0 += 256) {
        __syncthreads();
        for (int c = 0; c < NC; c++)
            for (int pl = 0; pl < planes; pl++) {
                const int j = j0 + threadIdx.x;
                xs[pl][c][threadIdx.x] = (j < n && c0 + c < ncol) ? x[j + (int64_t)(c0 + c) * ld + pl * plane] : (TX)0;
            }
        __syncthreads();
        if (i < n) {
            const int jn = min(256, n - j0);
            for (int jj = 0; jj < jn; jj++) {
                const int j = j0 + jj;
                const TX wr = (TX)W[i + (int64_t)j * ldw];
                if (w_complex) {
                    const TX wi = (TX)W[i + (int64_t)j * ldw + wplane];
                    for (int c = 0; c < NC; c++) {
                        const TX xr = xs[0][c][jj], xi = xs[1][c][jj];
                        acc[0][c] = __builtin_fma(wr, xr, acc[0][c]);
                        acc[0][c] = __builtin_fma(-wi, xi, acc[0][c]);
                        acc[1][c] = __builtin_fma(wr, xi, acc[1][c]);
                        acc[1][c] = __builtin_fma(wi, xr, acc[1][c]);
                    }
                } else {
                    for (int c = 0; c < NC; c++)
                        for (int pl = 0; pl < planes; pl++) acc[pl][c] = __builtin_fma(wr, xs[pl][c][jj], acc[pl][c]);
                }
            }
        }
    }
    if (i >= n) return;
    for (int c = 0; c < NC; c++) {
        const int col = c0 + c;
        if (col >= ncol) break;
        const int64_t o = i + (int64_t)col * ld;
        if (colshift) {
            // W = sigma I + tau J: acc holds J x
            const TX sr = (TX)colshift[4 * col], si = (TX)colshift[4 * col + 1];
            const TX tr = (TX)colshift[4 * col + 2], ti = (TX)colshift[4 * col + 3];
            const TX xr = x[o], xi = planes > 1 ? x[o + plane] : (TX)0;
            const TX jr = acc[0][c], ji = planes > 1 ? acc[1][c] : (TX)0;
            const TX wr = (sr * xr - si * xi) + (tr * jr - ti * ji);
            const TX wi = (sr * xi + si * xr) + (tr * ji + ti * jr);
            r[o] = b[o] - wr;
            if (planes > 1) r[o + plane] = b[o + plane] - wi;
        } else {
            for (int pl = 0; pl < planes; pl++) r[o + pl * plane] = b[o + pl * plane] - acc[pl][c];
        }
    }
}

// norms[c] = max_i |r_c(i)| (complex: |re| + |im|... the modulus), norms[ncol + c] = same for b
template <typename TX>
__global__ void k_colnorms(int n, const TX *__restrict__ r, const TX *__restrict__ b, int64_t ld, int64_t plane,
                           int planes, int ncol, double *__restrict__ norms) {
    const int c = blockIdx.x;
    double mr = 0, mb = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int64_t o = i + (int64_t)c * ld;
        double vr = 0, vb = 0;
        if (planes > 1) {
            vr = hypot((double)r[o], (double)r[o + plane]);
            vb = hypot((double)b[o], (double)b[o + plane]);
        } else {
            vr = fabs((double)r[o]);
            vb = fabs((double)b[o]);
        }
        mr = fmax(mr, vr);
        mb = fmax(mb, vb);
    }
    __shared__ double sr[256], sb[256];
    sr[threadIdx.x] = mr; sb[threadIdx.x] = mb;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sr[threadIdx.x] = fmax(sr[threadIdx.x], sr[threadIdx.x + o]);
            sb[threadIdx.x] = fmax(sb[threadIdx.x], sb[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { norms[c] = sr[0]; norms[ncol + c] = sb[0]; }
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

template <typename T, typename TX, int NR, int PL, int UC>
Status launch_ldiv_t(lhl_handle *h, const SolveArgs &a, const Group *groups, int ngroups) {
    const int n = (int)h->n;
    const size_t lds = sizeof(T) * (size_t)NR * PL * n;
    auto fn = k_ldiv<T, TX, NR, PL, UC>;
    LHL_HIP_TRY(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    LHL_LAUNCH(h, "solve", fn, dim3(ngroups), dim3(kSolveThreads), lds, n, (const T *)h->A.as<T>(), h->lda,
                       (const int *)h->perm.as<int>(), (const int *)h->iperm.as<int>(),
                       (const double *)h->scale64.as<double>(), (const double *)h->iscale64.as<double>(),
                       (const T *)a.U, a.ustride, (const T *)a.l, (const T *)a.rdiag, a.lstride, a.swp, a.sstride,
                       groups, (const TX *)a.xin, (TX *)a.xout, a.ldx, a.xplane, a.accumulate);
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

// LDS one workgroup may use on the current device (160 KB on gfx950).
inline size_t lds_limit() {
    static int cached = -1;
    if (cached < 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = 64 * 1024;
        cached = v < kMaxLds ? v : kMaxLds;
    }
    return (size_t)cached;
}

// Right-hand sides per workgroup: as many as fit in LDS, at most 4.
inline int pick_nr(int64_t n, int planes, size_t es) {
    const size_t budget = lds_limit() - 8 * 1024;
    const size_t per = (size_t)planes * (size_t)n * es;
    int nr = per ? (int)(budget / per) : 4;
    return nr >= 4 ? 4 : nr >= 2 ? 2 : nr;
}

template <typename T, typename TX>
Status solve_typed(lhl_handle *h, const SolveArgs &a) {
    const int PLn = a.x_complex ? 2 : 1;
    const int nr = pick_nr(h->n, PLn, sizeof(T));
    if (nr < 1) return {LHL_ERR_ARG, "lhl_ldiv: n too large for the LDS-resident solve"};
    // groups: (shift, first column, count)
    std::vector<Group> hg;
    std::vector<int64_t> cb((size_t)a.nshift + 1);
    if (a.col_begin) {
        // host copy of the column offsets
        LHL_HIP_TRY(hipMemcpy(cb.data(), a.col_begin, sizeof(int64_t) * (a.nshift + 1), hipMemcpyDeviceToHost));
    } else {
        cb[0] = 0;
        cb[1] = a.ncol;
    }
    for (int64_t s = 0; s < a.nshift; s++)
        for (int64_t c = cb[s]; c < cb[s + 1]; c += nr)
            hg.push_back(Group{(int)s, (int)c, (int)std::min<int64_t>(nr, cb[s + 1] - c), 0});
    if (hg.empty()) return {};
    LHL_TRY(h->nbuf.ensure(sizeof(Group) * hg.size()));
    LHL_HIP_TRY(hipMemcpyAsync(h->nbuf.p, hg.data(), sizeof(Group) * hg.size(), hipMemcpyHostToDevice, h->stream));
    const Group *dg = h->nbuf.as<Group>();
    const int ng = (int)hg.size();
    Status st;
    if (a.shift_complex) {
        if (nr >= 4) st = launch_ldiv_t<T, TX, 4, 2, 1>(h, a, dg, ng);
        else if (nr >= 2) st = launch_ldiv_t<T, TX, 2, 2, 1>(h, a, dg, ng);
        else st = launch_ldiv_t<T, TX, 1, 2, 1>(h, a, dg, ng);
    } else if (a.x_complex) {
        if (nr >= 4) st = launch_ldiv_t<T, TX, 4, 2, 0>(h, a, dg, ng);
        else if (nr >= 2) st = launch_ldiv_t<T, TX, 2, 2, 0>(h, a, dg, ng);
        else st = launch_ldiv_t<T, TX, 1, 2, 0>(h, a, dg, ng);
    } else {
        if (nr >= 4) st = launch_ldiv_t<T, TX, 4, 1, 0>(h, a, dg, ng);
        else if (nr >= 2) st = launch_ldiv_t<T, TX, 2, 1, 0>(h, a, dg, ng);
        else st = launch_ldiv_t<T, TX, 1, 1, 0>(h, a, dg, ng);
    }
    if (!st.ok()) return st;
    // the group table must stay valid until the kernel has read it
    LHL_HIP_TRY(hipStreamSynchronize(h->stream));
    return {};
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

template <typename TW, typename TX>
Status residual_typed(lhl_handle *h, const ResidArgs &a) {
    constexpr int NC = 4;
    const int n = (int)h->n;
    const int planes = a.x_complex ? 2 : 1;
    dim3 grid((n + 255) / 256, (unsigned)((a.ncol + NC - 1) / NC));
    const TW *W;
    int64_t ldw, wplane = 0;
    int wc;
    if (a.W) {
        W = (const TW *)a.W;
        ldw = a.ldw;
        wc = a.w_complex;
        wplane = a.w_complex ? a.ldw * n : 0;
    } else {
        W = (const TW *)h->J.p;
        ldw = h->ldj;
        wc = 0;
    }
    LHL_LAUNCH(h, "residual", (k_residual<TW, TX, NC>), grid, dim3(256), 0, n, W, ldw, wc, wplane,
                       a.W ? nullptr : a.colshift, (const TX *)a.b, (const TX *)a.x, (TX *)a.r, a.ld, a.plane,
                       planes, (int)a.ncol);
    LHL_LAUNCH(h, "residual", k_colnorms<TX>, dim3((unsigned)a.ncol), dim3(256), 0, n, (const TX *)a.r,
                       (const TX *)a.b, a.ld, a.plane, planes, (int)a.ncol, a.norms);
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

}  // namespace

Status solve_device(lhl_handle *h, const SolveArgs &a) {
    if (h->n == 0 || a.ncol == 0) return {};
    if (h->dtype == LHL_F64) {
        if (a.xdt != LHL_F64) return {LHL_ERR_ARG, "lhl_ldiv: a Float32 vector cannot take a Float64 solve"};
        return solve_typed<double, double>(h, a);
    }
    if (a.xdt == LHL_F64) return solve_typed<float, double>(h, a);
    return solve_typed<float, float>(h, a);
}

Status applyz_device(lhl_handle *h, int inverse, void *x, int64_t ncol, int64_t ldx, int64_t xplane, int xdt,
                     int x_complex) {
    if (h->n == 0 || ncol == 0) return {};
    const int planes = x_complex ? 2 : 1;
    if (h->dtype == LHL_F64) return applyz_typed<double, double>(h, inverse, x, ncol, ldx, xplane, planes);
    if (xdt == LHL_F64) return applyz_typed<float, double>(h, inverse, x, ncol, ldx, xplane, planes);
    return applyz_typed<float, float>(h, inverse, x, ncol, ldx, xplane, planes);
}

Status residual_device(lhl_handle *h, const ResidArgs &a) {
    if (h->n == 0 || a.ncol == 0) return {};
    const int wdt = a.W ? a.xdt : h->j_dtype;
    if (a.xdt == LHL_F64) {
        if (wdt == LHL_F64) return residual_typed<double, double>(h, a);
        return residual_typed<float, double>(h, a);
    }
    if (wdt == LHL_F64) return residual_typed<double, float>(h, a);
    return residual_typed<float, float>(h, a);
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
