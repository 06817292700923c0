// The solves for many right-hand sides at once (SciML/LHL.jl lhl_ldiv!,
// src:3052-3123, applied to a batch), and the refinement residual.
//
// x <- Z (sigma I + tau H)^-1 Z^-1 x with Z = D P L runs as four launches over
// all columns of the batch:
//   gather   Y = P^-1 D^-1 X into the working precision           (src:2130-2140)
//   L^-1     a blocked triangular solve over 64-index blocks: one launch per
//            block, whose workgroups each take a (row tile, 64 columns) pair;
//            every workgroup re-solves the block's 64x64 triangle for its
//            columns (cheap, from LDS) and the workgroups below the block
//            subtract the block's rank-64 product from their rows.  The block's
//            own tile writes its finished rows to a second buffer, so no
//            workgroup reads a row another one is rewriting.
//   H solve  per shift: the Hessenberg forward chain and the blocked back
//            substitution with the columns resident in LDS (lhl_solve.hip).
//   L, D P   y <- L y is a triangular matrix product without recurrence: one
//            launch of (row tile, column tile) workgroups, whose epilogue
//            scatters x[perm[p]] = y[p] * d[perm[p]] (src:3064-3068).
// In y' = y[1:n] the multipliers form the strictly lower triangle
// L'[i, m] = A[i+1, m] (i > m): step m of the reference's sweeps.
#include "lhl_common.h"
#include "lhl_mfma.h"

namespace lhl {

constexpr int kT = 64;   // tile edge

// Y[i + q*ldy] = x[perm[i], col] * iscale[perm[i]] for internal column q = col*planes + pl;
// row 0 is also copied to Y2 (the second buffer of the triangular solve).
template <typename T, typename TX>
__global__ void k_gather(int n, const int *__restrict__ perm, const double *__restrict__ iscale,
                         const TX *__restrict__ x, int64_t ldx, int64_t xplane, int planes, T *__restrict__ Y,
                         T *__restrict__ Y2, int64_t ldy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (i >= n) return;
    const int col = q / planes, pl = q % planes;
    const int p = perm[i];
    const T v = (T)(x[p + (int64_t)col * ldx + pl * xplane] * (TX)iscale[p]);
    Y[i + (int64_t)q * ldy] = v;
    if (i == 0) Y2[(int64_t)q * ldy] = v;
}

// Linv[b] = (L'[64b:64b+64, 64b:64b+64])^-1, column-major 64x64 per block: the
// explicit inverse of each unit lower triangular diagonal block (the
// reference forms its nb x nb panel inverse the same way, src:1942-1954).
template <typename T>
__global__ void __launch_bounds__(64) k_linv(int n, const T *__restrict__ A, int64_t lda, T *__restrict__ Linv) {
    const int N1 = n - 1;
    const int b = blockIdx.x;
    const int r0 = b * kT;
    const int nb = min(kT, N1 - r0);
    __shared__ T Ls[kT][kT + 1];
    for (int t = threadIdx.x; t < kT * kT; t += 64) {
        const int i = t % kT, m = t / kT;
        Ls[i][m] = (i > m && i < nb) ? A[(r0 + i + 1) + (int64_t)(r0 + m) * lda] : (T)0;
    }
    __syncthreads();
    const int c = threadIdx.x;
    T x[kT];
#pragma unroll
    for (int i = 0; i < kT; i++) x[i] = (i == c) ? (T)1 : (T)0;
#pragma unroll
    for (int m = 0; m < kT - 1; m++) {
        const T xm = x[m];
#pragma unroll
        for (int i = m + 1; i < kT; i++) x[i] = __builtin_fma(-Ls[i][m], xm, x[i]);
    }
    T *dst = Linv + (int64_t)b * kT * kT + (int64_t)c * kT;
#pragma unroll
    for (int i = 0; i < kT; i++) dst[i] = x[i];
}

// One block step b of y' <- L'^-1 y': the block's rows become Linv[b] * y'_b
// (redundantly in every workgroup of the step: a 64x64x64 MFMA product); the
// block's own tile stores them to Yo, the tiles below subtract L'[I, b] * y'_b.
template <typename T>
__global__ void __launch_bounds__(256) k_trsm_step(int n, const T *__restrict__ A, int64_t lda,
                                                   const T *__restrict__ Linv, T *__restrict__ Yw,
                                                   T *__restrict__ Yo, int64_t ldy, int ncols, int b) {
    const int N1 = n - 1;
    const int I = b + blockIdx.x;
    const int c0 = blockIdx.y * kT;
    const int r0 = b * kT;
    const int nb = min(kT, N1 - r0);
    const int tid = threadIdx.x;
    __shared__ T As[kT][kT + kTilePad];   // As[m][i]: A operand (i, m)
    __shared__ T Bs[kT][kT + kTilePad];   // Bs[m][c]: B operand (m, c)
    const T *Lb = Linv + (int64_t)b * kT * kT;
    for (int t = tid; t < kT * kT; t += 256) {
        const int x = t % kT, y = t / kT;
        As[y][x] = Lb[x + (int64_t)y * kT];
        Bs[x][y] = (x < nb && c0 + y < ncols) ? Yw[(r0 + x + 1) + (int64_t)(c0 + y) * ldy] : (T)0;
    }
    __syncthreads();
    typename Mfma<T>::acc_t acc[2][2];
    mfma_tile64_zero<T>(acc);
    mfma_tile64<T>(As, Bs, kT, acc);
    __syncthreads();
    mfma_tile64_each<T>(acc, [&](int i, int c, T v) { Bs[i][c] = v; });   // y'_b, the next B operand
    const int ri = I * kT;
    const int ni = min(kT, N1 - ri);
    if (I != b)
        for (int t = tid; t < kT * kT; t += 256) {
            const int x = t % kT, y = t / kT;
            As[y][x] = (x < ni && y < nb) ? A[(ri + x + 1) + (int64_t)(r0 + y) * lda] : (T)0;
        }
    __syncthreads();
    if (I == b) {
        for (int t = tid; t < kT * kT; t += 256) {
            const int x = t % kT, y = t / kT;
            if (x < nb && c0 + y < ncols) Yo[(r0 + x + 1) + (int64_t)(c0 + y) * ldy] = Bs[x][y];
        }
        return;
    }
    mfma_tile64_zero<T>(acc);
    mfma_tile64<T>(As, Bs, kT, acc);
    __syncthreads();
    mfma_tile64_each<T>(acc, [&](int i, int c, T v) { As[i][c] = v; });
    __syncthreads();
    for (int t = tid; t < kT * kT; t += 256) {
        const int x = t % kT, y = t / kT;
        if (x < ni && c0 + y < ncols) {
            T *dst = Yw + (ri + x + 1) + (int64_t)(c0 + y) * ldy;
            *dst = *dst - As[x][y];
        }
    }
}

// y_new' = y' + L' y' for one (row tile, column tile), scattered to
// x[perm[p], col] (+)= y_new[p] * d[perm[p]], p = i' + 1; the row-0 tile also
// scatters p = 0.
template <typename T, typename TX>
__global__ void __launch_bounds__(256) k_trmm_scatter(int n, const T *__restrict__ A, int64_t lda,
                                                      const T *__restrict__ Y, int64_t ldy, int ncols, int planes,
                                                      const int *__restrict__ perm, const double *__restrict__ scale,
                                                      TX *__restrict__ x, int64_t ldx, int64_t xplane,
                                                      int accumulate) {
    const int N1 = n - 1;
    const int I = blockIdx.x;
    const int c0 = blockIdx.y * kT;
    const int ri = I * kT;
    const int ni = min(kT, N1 - ri);
    const int tid = threadIdx.x;
    __shared__ T As[kT][kT + kTilePad];   // As[m][i] = L'(ri+i, rk+m)
    __shared__ T Bs[kT][kT + kTilePad];   // Bs[m][c] = y'(rk+m, c0+c)
    typename Mfma<T>::acc_t acc[2][2];
    mfma_tile64_zero<T>(acc);
    for (int K = 0; K <= I; K++) {
        const int rk = K * kT;
        const int nk = min(kT, N1 - rk);
        __syncthreads();
        for (int t = tid; t < kT * kT; t += 256) {
            const int xx = t % kT, yy = t / kT;
            const bool ok = xx < ni && yy < nk && (K < I || yy < xx);
            As[yy][xx] = ok ? A[(ri + xx + 1) + (int64_t)(rk + yy) * lda] : (T)0;
            Bs[xx][yy] = (xx < nk && c0 + yy < ncols) ? Y[(rk + xx + 1) + (int64_t)(c0 + yy) * ldy] : (T)0;
        }
        __syncthreads();
        mfma_tile64<T>(As, Bs, kT, acc);
    }
    __syncthreads();
    mfma_tile64_each<T>(acc, [&](int i, int c, T v) { As[i][c] = v; });
    __syncthreads();
    for (int t = tid; t < kT * kT; t += 256) {
        const int xx = t % kT, yy = t / kT;
        const int q = c0 + yy;
        if (xx >= ni || q >= ncols) continue;
        const int col = q / planes, pl = q % planes;
        const int p = ri + xx + 1;
        const T yn = Y[p + (int64_t)q * ldy] + As[xx][yy];
        const int dst_row = perm[p];
        TX *dst = x + dst_row + (int64_t)col * ldx + pl * xplane;
        const TX v = (TX)yn * (TX)scale[dst_row];
        if (accumulate) *dst += v;
        else *dst = v;
    }
    if (I == 0 && tid < kT && c0 + tid < ncols) {
        const int q = c0 + tid;
        const int col = q / planes, pl = q % planes;
        const int dst_row = perm[0];
        TX *dst = x + dst_row + (int64_t)col * ldx + pl * xplane;
        const TX v = (TX)Y[(int64_t)q * ldy] * (TX)scale[dst_row];
        if (accumulate) *dst += v;
        else *dst = v;
    }
}


// ---- U: the shifted Hessenberg's triangular factor, per shift -------------
//
// Uinv[s][b] = (U_s[64b:64b+64, 64b:64b+64])^-1, column-major, real plane then
// imaginary plane (complex shifts): back substitution on the unit columns with
// the reciprocal pivots, one column per thread.
template <typename T, int UC>
__global__ void __launch_bounds__(64) k_uinv(int n, const T *__restrict__ Ubase, int64_t ustride,
                                             const T *__restrict__ rdbase, int64_t lstride, T *__restrict__ Uinv,
                                             int64_t uistride) {
#pragma clang fp contract(off)
    constexpr int UP = UC ? 2 : 1;
    const int b = blockIdx.x, s = blockIdx.y;
    const int r0 = b * kT, nb = min(kT, n - r0);
    const T *U = Ubase + (int64_t)s * ustride;
    const T *rd = rdbase + (int64_t)s * lstride;
    const int64_t up = ustride / 2, lp = lstride / 2;
    __shared__ T Us[UP][kT][kT + 1];
    __shared__ T ds[UP][kT];
    for (int t = threadIdx.x; t < kT * kT; t += 64) {
        const int c = t % kT, r = t / kT;
        const bool ok = r < nb && c < nb && c > r;
        const int64_t o = upack_off(n, ok ? r0 + r : 0) - (ok ? r0 + r : 0) + r0 + c;
        Us[0][r][c] = ok ? U[o] : (T)0;
        if constexpr (UC) Us[UP - 1][r][c] = ok ? U[o + up] : (T)0;
    }
    {
        const int t = threadIdx.x;
        ds[0][t] = t < nb ? rd[r0 + t] : (T)1;
        if constexpr (UC) ds[UP - 1][t] = t < nb ? rd[r0 + t + lp] : (T)0;
    }
    __syncthreads();
    const int c = threadIdx.x;
    T xr[kT], xi[kT];
#pragma unroll
    for (int i = kT - 1; i >= 0; i--) {
        T ar = (i == c) ? (T)1 : (T)0, ai = 0;
#pragma unroll
        for (int j = i + 1; j < kT; j++) {
            ar = __builtin_fma(-Us[0][i][j], xr[j], ar);
            if constexpr (UC) {
                ar = __builtin_fma(Us[UP - 1][i][j], xi[j], ar);
                ai = __builtin_fma(-Us[0][i][j], xi[j], ai);
                ai = __builtin_fma(-Us[UP - 1][i][j], xr[j], ai);
            }
        }
        if constexpr (UC) {
            const T dr = ds[0][i], di = ds[UP - 1][i];
            xr[i] = (i <= c) ? ar * dr - ai * di : (T)0;
            xi[i] = (i <= c) ? ar * di + ai * dr : (T)0;
        } else {
            xr[i] = (i <= c) ? ar * ds[0][i] : (T)0;
            xi[i] = 0;
        }
    }
    T *dst = Uinv + (int64_t)s * uistride + (int64_t)b * kT * kT * UP + (int64_t)c * kT;
#pragma unroll
    for (int i = 0; i < kT; i++) {
        dst[i] = xr[i];
        if constexpr (UC) dst[i + kT * kT] = xi[i];
    }
}

struct UGroup {
    int shift, q0, nq, pad;   // internal columns q0 .. q0+nq-1 (nq <= 16), all of one shift
};

// One block step b of the back substitution on every group: the block's rows
// become Uinv[b] * y_b (redundantly per workgroup); the block's own tile
// stores them to Yf, the tiles above subtract U[I, b] * x_b from Yw.
// Output tile: 64 rows x 16 internal columns; wave w owns rows 16w..16w+15.
template <typename T, int UC>
__global__ void __launch_bounds__(256) k_utrsm_step(int n, const T *__restrict__ Ubase, int64_t ustride,
                                                    const T *__restrict__ Uinv, int64_t uistride,
                                                    const UGroup *__restrict__ groups, T *__restrict__ Yw,
                                                    T *__restrict__ Yf, int64_t ldy, int b) {
    constexpr int UP = UC ? 2 : 1;
    constexpr int NC = 16;
    const int I = blockIdx.x;
    const UGroup g = groups[blockIdx.y];
    const int rb = b * kT, nbk = min(kT, n - rb);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    __shared__ T As[UP][kT][kT + kTilePad];   // As[.][k][i]: A operand (i, k)
    __shared__ T Bs[kT][NC + 1];              // Bs[k][c]:   B operand (k, c)
    const T *Ui = Uinv + (int64_t)g.shift * uistride + (int64_t)b * kT * kT * UP;
    for (int t = tid; t < kT * kT; t += 256) {
        const int x = t % kT, y = t / kT;
        As[0][y][x] = Ui[x + (int64_t)y * kT];
        if constexpr (UC) As[UP - 1][y][x] = Ui[kT * kT + x + (int64_t)y * kT];
    }
    for (int t = tid; t < kT * NC; t += 256) {
        const int x = t % kT, y = t / kT;
        Bs[x][y] = (x < nbk && y < g.nq) ? Yw[(rb + x) + (int64_t)(g.q0 + y) * ldy] : (T)0;
    }
    __syncthreads();
    const int li = lane & 15, lk = lane >> 4;
    auto product = [&]() {
        typename Mfma<T>::acc_t p = Mfma<T>::zero(), q = Mfma<T>::zero();
#pragma unroll 4
        for (int k0 = 0; k0 < kT; k0 += 4) {
            const int k = k0 + lk;
            const T bv = Bs[k][li];
            p = Mfma<T>::mac(As[0][k][16 * w + li], bv, p);
            if constexpr (UC) q = Mfma<T>::mac(As[UP - 1][k][16 * w + li], bv, q);
        }
        if constexpr (UC) {
            // internal columns alternate re, im: re = Pr - Qi, im = Pi + Qr (neighbour lane holds col ^ 1)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const T qn = __shfl_xor(q[r], 1, 64);
                p[r] = (li & 1) ? p[r] + qn : p[r] - qn;
            }
        }
        return p;
    };
    typename Mfma<T>::acc_t xb = product();
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; r++) Bs[16 * w + Mfma<T>::row(lane, r)][li] = xb[r];
    const int ri = I * kT;
    if (I != b) {
        for (int t = tid; t < kT * kT; t += 256) {
            const int k = t % kT, i = t / kT;
            const int row = ri + i;
            const int64_t o = upack_off(n, row) - row + rb + k;
            const bool ok = k < nbk;
            As[0][k][i] = ok ? Ubase[(int64_t)g.shift * ustride + o] : (T)0;
            if constexpr (UC) As[UP - 1][k][i] = ok ? Ubase[(int64_t)g.shift * ustride + ustride / 2 + o] : (T)0;
        }
    }
    __syncthreads();
    if (I == b) {
        for (int t = tid; t < kT * NC; t += 256) {
            const int x = t % kT, y = t / kT;
            if (x < nbk && y < g.nq) Yf[(rb + x) + (int64_t)(g.q0 + y) * ldy] = Bs[x][y];
        }
        return;
    }
    typename Mfma<T>::acc_t up = product();
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; r++) As[0][li][16 * w + Mfma<T>::row(lane, r)] = up[r];   // As[0][c][i] staging
    __syncthreads();
    for (int t = tid; t < kT * NC; t += 256) {
        const int x = t % kT, y = t / kT;
        if (y < g.nq) {
            T *dst = Yw + (ri + x) + (int64_t)(g.q0 + y) * ldy;
            *dst = *dst - As[0][y][x];
        }
    }
}


// ---- column-owner triangular solves ---------------------------------------
// One workgroup owns kCN internal columns and walks every 64-row block in
// order, left-looking: Y_b <- inv(T_bb) (Y_b - sum_K T[b,K] X_K).  Columns are
// independent, so no workgroup waits on another and one launch covers the
// whole solve.  kCW waves split the K sum (partial tiles reduced through LDS);
// four of them then apply the precomputed 64x64 diagonal-block inverse.
// Within each 16-wide K slab lane (i, q) takes k = 16j + 4q + t, t = 0..3, so
// both operands are read as runs of four consecutive elements.
constexpr int kCW = 8;
constexpr int kCN = 16;

// Four consecutive elements through one (float) or two (double) vector loads
// that need only element alignment.
template <typename T> struct V4T;
template <> struct V4T<float> { typedef float type __attribute__((ext_vector_type(4), aligned(4))); };
template <> struct V4T<double> { typedef double type __attribute__((ext_vector_type(4), aligned(8))); };
template <typename T> using V4 = typename V4T<T>::type;
template <typename T> __device__ __forceinline__ V4<T> ld4u(const T *p) { return *(const V4<T> *)p; }

// y' <- L'^-1 y': Yw holds the gathered columns (rows 1..n-1), Yo receives the
// solution (row 0 is already there).
template <typename T>
__global__ void __launch_bounds__(64 * kCW) k_ltrsm_cols(int n, const T *__restrict__ A, int64_t lda,
                                                         const T *__restrict__ Linv, const T *__restrict__ Yw,
                                                         T *__restrict__ Yo, int64_t ldy, int Q) {
    const int N1 = n - 1, nbt = (N1 + kT - 1) / kT;
    const int c0 = blockIdx.x * kCN;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const T *Xc = Yo + (int64_t)min(c0 + li, Q - 1) * ldy + 1;   // solved rows m at Xc[m]
    __shared__ T red[kCW][kT][kCN + 1];
    __shared__ T Ys[kT][kCN + 1];
    for (int b = 0; b < nbt; b++) {
        const int r0 = b * kT, nb = min(kT, N1 - r0);
        // MFMA row i of accumulator rr is block row 4i + rr, so each lane reads
        // four consecutive rows of a column of L' as one run.
        typename Mfma<T>::acc_t acc[4];
#pragma unroll
        for (int rr = 0; rr < 4; rr++) acc[rr] = Mfma<T>::zero();
        const T *Ar = A + r0 + 4 * li + 1;   // L'(i, m) = A(i + 1, m)
        const bool full = nb == kT;
        for (int j = w; j < 4 * b; j += kCW) {
            const int m0 = 16 * j + 4 * lk;
            T av[4][4];
            const V4<T> xv = ld4u(Xc + m0);
            if (full) {
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const V4<T> a = ld4u(Ar + (int64_t)(m0 + t) * lda);
#pragma unroll
                    for (int rr = 0; rr < 4; rr++) av[rr][t] = a[rr];
                }
            } else {
#pragma unroll
                for (int t = 0; t < 4; t++)
#pragma unroll
                    for (int rr = 0; rr < 4; rr++)
                        av[rr][t] = A[min(r0 + 4 * li + rr, N1 - 1) + 1 + (int64_t)(m0 + t) * lda];
            }
#pragma unroll
            for (int t = 0; t < 4; t++)
#pragma unroll
                for (int rr = 0; rr < 4; rr++) acc[rr] = Mfma<T>::mac(av[rr][t], xv[t], acc[rr]);
        }
#pragma unroll
        for (int rr = 0; rr < 4; rr++)
#pragma unroll
            for (int r = 0; r < 4; r++) red[w][4 * Mfma<T>::row(lane, r) + rr][li] = acc[rr][r];
        __syncthreads();
        for (int t = threadIdx.x; t < kT * kCN; t += 64 * kCW) {
            const int i = t % kT, c = t / kT;
            T v = 0;
            if (i < nb && c0 + c < Q) {
                T sum = 0;
#pragma unroll
                for (int ww = 0; ww < kCW; ww++) sum += red[ww][i][c];
                v = Yw[(r0 + i + 1) + (int64_t)(c0 + c) * ldy] - sum;
            }
            Ys[i][c] = v;
        }
        __syncthreads();
        if (w < 4) {
            const T *Lb = Linv + (int64_t)b * kT * kT + 16 * w + li;
            typename Mfma<T>::acc_t x = Mfma<T>::zero();
#pragma unroll
            for (int sl = 0; sl < 4; sl++)
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const int k = 16 * sl + 4 * lk + t;
                    x = Mfma<T>::mac(Lb[(int64_t)k * kT], Ys[k][li], x);
                }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = 16 * w + Mfma<T>::row(lane, r);
                if (i < nb && c0 + li < Q) Yo[(r0 + i + 1) + (int64_t)(c0 + li) * ldy] = x[r];
            }
        }
        __syncthreads();   // the block's solved rows are visible to every wave
    }
}

// x <- U_s^-1 y for one group of internal columns of shift s: Yin holds the
// forward-chain output, Yf receives the solution.  Complex shifts: the
// columns alternate (re, im) and U is planar (imaginary plane ustride/2 on).
template <typename T, int UC>
__global__ void __launch_bounds__(64 * kCW) k_utrsm_cols(int n, const T *__restrict__ Ubase, int64_t ustride,
                                                         const T *__restrict__ Uinv, int64_t uistride,
                                                         const UGroup *__restrict__ groups, const T *__restrict__ Yin,
                                                         T *__restrict__ Yf, int64_t ldy) {
    constexpr int UP = UC ? 2 : 1;
    const UGroup g = groups[blockIdx.x];
    const int nbu = (n + kT - 1) / kT;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const T *Xc = Yf + (int64_t)(g.q0 + min(li, g.nq - 1)) * ldy;
    const T *Us = Ubase + (int64_t)g.shift * ustride;
    const int64_t up = ustride / 2;
    __shared__ T red[kCW][kT][kCN + 1];
    __shared__ T Ys[kT][kCN + 1];
    for (int b = nbu - 1; b >= 0; b--) {
        const int rb = b * kT, nbk = min(kT, n - rb);
        typename Mfma<T>::acc_t acc[4], qc[4];
        const T *Ur[4];
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
            acc[rr] = Mfma<T>::zero();
            qc[rr] = Mfma<T>::zero();
            const int row = min(rb + 16 * rr + li, n - 1);
            Ur[rr] = Us + upack_off(n, row) - row;   // U(row, m) = Ur[m], m >= row
        }
        const int kbeg = rb + kT;
        const int nsl = kbeg < n ? (n - kbeg + 15) / 16 : 0;
        for (int j = w; j < nsl; j += kCW) {
            const int m0 = kbeg + 16 * j + 4 * lk;
            T bv[4], av[4][4], ai[4][4];
            if (m0 + 3 < n) {
                // one 4-element run per lane and row (dword aligned vector loads)
                const V4<T> xv = ld4u(Xc + m0);
#pragma unroll
                for (int t = 0; t < 4; t++) bv[t] = xv[t];
#pragma unroll
                for (int rr = 0; rr < 4; rr++) {
                    const V4<T> u = ld4u(Ur[rr] + m0);
#pragma unroll
                    for (int t = 0; t < 4; t++) av[rr][t] = u[t];
                    if constexpr (UC) {
                        const V4<T> v = ld4u(Ur[rr] + m0 + up);
#pragma unroll
                        for (int t = 0; t < 4; t++) ai[rr][t] = v[t];
                    }
                }
            } else {
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const int m = m0 + t;
                    const bool ok = m < n;
                    const int mc = ok ? m : n - 1;
                    bv[t] = ok ? Xc[mc] : (T)0;
#pragma unroll
                    for (int rr = 0; rr < 4; rr++) {
                        av[rr][t] = Ur[rr][mc];
                        if constexpr (UC) ai[rr][t] = Ur[rr][mc + up];
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < 4; t++)
#pragma unroll
                for (int rr = 0; rr < 4; rr++) {
                    acc[rr] = Mfma<T>::mac(av[rr][t], bv[t], acc[rr]);
                    if constexpr (UC) qc[rr] = Mfma<T>::mac(ai[rr][t], bv[t], qc[rr]);
                }
        }
#pragma unroll
        for (int rr = 0; rr < 4; rr++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                T v = acc[rr][r];
                if constexpr (UC) {
                    // re = Ur*xr - Ui*xi, im = Ur*xi + Ui*xr (the neighbour lane holds col ^ 1)
                    const T qn = __shfl_xor(qc[rr][r], 1, 64);
                    v = (li & 1) ? v + qn : v - qn;
                }
                red[w][16 * rr + Mfma<T>::row(lane, r)][li] = v;
            }
        __syncthreads();
        for (int t = threadIdx.x; t < kT * kCN; t += 64 * kCW) {
            const int i = t % kT, c = t / kT;
            T v = 0;
            if (i < nbk && c < g.nq) {
                T sum = 0;
#pragma unroll
                for (int ww = 0; ww < kCW; ww++) sum += red[ww][i][c];
                v = Yin[(rb + i) + (int64_t)(g.q0 + c) * ldy] - sum;
            }
            Ys[i][c] = v;
        }
        __syncthreads();
        if (w < 4) {
            const T *Ui = Uinv + (int64_t)g.shift * uistride + (int64_t)b * kT * kT * UP + 16 * w + li;
            typename Mfma<T>::acc_t x = Mfma<T>::zero(), xq = Mfma<T>::zero();
#pragma unroll
            for (int sl = 0; sl < 4; sl++)
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const int k = 16 * sl + 4 * lk + t;
                    const T yv = Ys[k][li];
                    x = Mfma<T>::mac(Ui[(int64_t)k * kT], yv, x);
                    if constexpr (UC) xq = Mfma<T>::mac(Ui[kT * kT + (int64_t)k * kT], yv, xq);
                }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                T v = x[r];
                if constexpr (UC) {
                    const T qn = __shfl_xor(xq[r], 1, 64);
                    v = (li & 1) ? v + qn : v - qn;
                }
                const int i = 16 * w + Mfma<T>::row(lane, r);
                if (i < nbk && li < g.nq) Yf[(rb + i) + (int64_t)(g.q0 + li) * ldy] = v;
            }
        }
        __syncthreads();
    }
}

// ---- the refinement residual: JX = J X (FP64 tiled GEMM), then
// r = b - (sigma x + tau JX) per column with its shift, and column norms.
// X, B, R are in the staged planar layout (ld, plane); JX uses the internal
// column order q = col*planes + pl with leading dimension ldq.
template <typename TW, typename TX>
__global__ void __launch_bounds__(256) k_gemm_jx(int n, const TW *__restrict__ J, int64_t ldj,
                                                 const TX *__restrict__ X, int64_t ld, int64_t plane, int planes,
                                                 int ncols, TX *__restrict__ JX, int64_t ldq) {
    const int r0 = blockIdx.x * kT, c0 = blockIdx.y * kT;
    const int tid = threadIdx.x;
    __shared__ TX As[kT][kT + kTilePad];   // As[k][i] = J[r0+i, k0+k]
    __shared__ TX Bs[kT][kT + kTilePad];   // Bs[k][c] = X[k0+k, q = c0+c]
    typename Mfma<TX>::acc_t acc[2][2];
    mfma_tile64_zero<TX>(acc);
    for (int k0 = 0; k0 < n; k0 += kT) {
        __syncthreads();
        for (int t = tid; t < kT * kT; t += 256) {
            const int xx = t % kT, yy = t / kT;
            As[yy][xx] = (r0 + xx < n && k0 + yy < n) ? (TX)J[(r0 + xx) + (int64_t)(k0 + yy) * ldj] : (TX)0;
            const int q = c0 + yy;
            TX v = 0;
            if (k0 + xx < n && q < ncols) {
                const int col = q / planes, pl = q % planes;
                v = X[(k0 + xx) + (int64_t)col * ld + pl * plane];
            }
            Bs[xx][yy] = v;
        }
        __syncthreads();
        mfma_tile64<TX>(As, Bs, kT, acc);
    }
    __syncthreads();
    mfma_tile64_each<TX>(acc, [&](int i, int c, TX v) { As[i][c] = v; });
    __syncthreads();
    for (int t = tid; t < kT * kT; t += 256) {
        const int xx = t % kT, yy = t / kT;
        const int q = c0 + yy;
        if (r0 + xx < n && q < ncols) JX[(r0 + xx) + (int64_t)q * ldq] = As[xx][yy];
    }
}

// dense W x for a user-supplied W (planar complex when w_complex): same GEMM with
// the complex combination done by k_resid_combine.
template <typename TX>
__global__ void k_resid_combine(int n, const TX *__restrict__ JX, int64_t ldq, const TX *__restrict__ JXi,
                                const TX *__restrict__ X, const TX *__restrict__ B, TX *__restrict__ R, int64_t ld,
                                int64_t plane, int planes, const double *__restrict__ colshift, int w_complex) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int col = blockIdx.y;
    if (i >= n) return;
    const int64_t o = i + (int64_t)col * ld;
    if (colshift) {
        const TX sr = (TX)colshift[4 * col], si = (TX)colshift[4 * col + 1];
        const TX tr = (TX)colshift[4 * col + 2], ti = (TX)colshift[4 * col + 3];
        const TX xr = X[o], xi = planes > 1 ? X[o + plane] : (TX)0;
        const TX jr = JX[i + (int64_t)(col * planes) * ldq];
        const TX ji = planes > 1 ? JX[i + (int64_t)(col * planes + 1) * ldq] : (TX)0;
        const TX wr = (sr * xr - si * xi) + (tr * jr - ti * ji);
        const TX wi = (sr * xi + si * xr) + (tr * ji + ti * jr);
        R[o] = B[o] - wr;
        if (planes > 1) R[o + plane] = B[o + plane] - wi;
    } else if (w_complex) {
        // JX = Wr X (both planes), JXi = Wi X (both planes): W x = (Wr xr - Wi xi) + i (Wr xi + Wi xr)
        const TX a = JX[i + (int64_t)(col * 2) * ldq], b = JX[i + (int64_t)(col * 2 + 1) * ldq];
        const TX c = JXi[i + (int64_t)(col * 2) * ldq], d = JXi[i + (int64_t)(col * 2 + 1) * ldq];
        R[o] = B[o] - (a - d);
        R[o + plane] = B[o + plane] - (b + c);
    } else {
        for (int pl = 0; pl < planes; pl++)
            R[o + pl * plane] = B[o + pl * plane] - JX[i + (int64_t)(col * planes + pl) * ldq];
    }
}


// norms[c] = ||r_c||_inf, norms[ncol + c] = ||b_c||_inf (complex: the modulus)
template <typename TX>
__global__ void k_colnorms(int n, const TX *__restrict__ r, const TX *__restrict__ b, int64_t ld, int64_t plane,
                           int planes, int ncol, double *__restrict__ norms) {
    const int c = blockIdx.x;
    double mr = 0, mb = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int64_t o = i + (int64_t)c * ld;
        double vr, vb;
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
    sr[threadIdx.x] = mr;
    sb[threadIdx.x] = mb;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sr[threadIdx.x] = fmax(sr[threadIdx.x], sr[threadIdx.x + o]);
            sb[threadIdx.x] = fmax(sb[threadIdx.x], sb[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        norms[c] = sr[0];
        norms[ncol + c] = sb[0];
    }
}

namespace {

inline int64_t ldy_for(int64_t n) { return std::max<int64_t>(64, ((n + 63) / 64) * 64); }

template <typename T, typename TX>
Status solve_batch_typed(lhl_handle *h, const SolveArgs &a) {
    const int n = (int)h->n;
    const int planes = a.x_complex ? 2 : 1;
    const int Q = (int)(a.ncol * planes);
    const int64_t ldy = ldy_for(n);
    LHL_TRY(h->ybuf.ensure(sizeof(T) * (size_t)ldy * Q));
    LHL_TRY(h->rbuf.ensure(sizeof(T) * (size_t)ldy * Q));
    T *Yw = h->ybuf.as<T>(), *Yo = h->rbuf.as<T>();
    const T *A = h->A.as<T>();
    // the back substitution's column groups (within one shift)
    std::vector<int64_t> cb((size_t)a.nshift + 1);
    if (a.col_begin) {
        LHL_HIP_TRY(hipMemcpy(cb.data(), a.col_begin, sizeof(int64_t) * (a.nshift + 1), hipMemcpyDeviceToHost));
    } else {
        cb[0] = 0;
        cb[1] = a.ncol;
    }
    std::vector<UGroup> ug;
    for (int64_t s = 0; s < a.nshift; s++) {
        for (int64_t q = cb[s] * planes; q < cb[s + 1] * planes; q += 16)
            ug.push_back(UGroup{(int)s, (int)q, (int)std::min<int64_t>(16, cb[s + 1] * planes - q), 0});
    }
    LHL_TRY(h->gbuf.ensure(sizeof(UGroup) * std::max<size_t>(ug.size(), 1)));
    UGroup *dg = h->gbuf.as<UGroup>();
    LHL_HIP_TRY(hipMemcpy(dg, ug.data(), sizeof(UGroup) * ug.size(), hipMemcpyHostToDevice));

    const dim3 gq((unsigned)((n + 255) / 256), (unsigned)Q);
    LHL_LAUNCH(h, "gather", (k_gather<T, TX>), gq, dim3(256), 0, n, (const int *)h->perm.as<int>(),
               (const double *)h->iscale64.as<double>(), (const TX *)a.xin, a.ldx, a.xplane, planes, Yw, Yo, ldy);
    // y' <- L'^-1 y'  (final rows in Yo)
    const int N1 = n - 1;
    const int nbt = (N1 + kT - 1) / kT;
    const unsigned ct = (unsigned)((Q + kT - 1) / kT);
    const T *Linv = h->Linv.as<T>();
    if (nbt > 0)
        LHL_LAUNCH(h, "trsm", k_ltrsm_cols<T>, dim3((unsigned)((Q + kCN - 1) / kCN)), dim3(64 * kCW), 0, n, A, h->lda,
                   Linv, (const T *)Yw, Yo, ldy, Q);
    // forward interchange chain in place on Yo (columns resident in LDS)
    LHL_TRY(hess_device(h, a, Yo, ldy));
    // x <- U^-1 y per shift (solution in Yw)
    const int nbu = (n + kT - 1) / kT;
    const int64_t uis = a.uistride ? a.uistride : (int64_t)nbu * kT * kT * (a.shift_complex ? 2 : 1);
    const dim3 gu((unsigned)ug.size());
    if (a.shift_complex)
        LHL_LAUNCH(h, "utrsm", (k_utrsm_cols<T, 1>), gu, dim3(64 * kCW), 0, n, (const T *)a.U, a.ustride,
                   (const T *)a.Uinv, uis, (const UGroup *)dg, (const T *)Yo, Yw, ldy);
    else
        LHL_LAUNCH(h, "utrsm", (k_utrsm_cols<T, 0>), gu, dim3(64 * kCW), 0, n, (const T *)a.U, a.ustride,
                   (const T *)a.Uinv, uis, (const UGroup *)dg, (const T *)Yo, Yw, ldy);
    LHL_HIP_TRY(hipGetLastError());
    LHL_LAUNCH(h, "trmm", (k_trmm_scatter<T, TX>), dim3((unsigned)std::max(nbt, 1), ct), dim3(256), 0, n, A, h->lda,
               (const T *)Yw, ldy, Q, planes, (const int *)h->perm.as<int>(), (const double *)h->scale64.as<double>(),
               (TX *)a.xout, a.ldx, a.xplane, a.accumulate);
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

template <typename T>
Status uinv_typed(lhl_handle *h, const void *U, int64_t ustride, const void *rdiag, int64_t lstride, int64_t nshift,
                  int cplx, DevBuf &Uinv, int64_t *uistride) {
    const int n = (int)h->n;
    const int nbu = (n + kT - 1) / kT;
    const int64_t uis = (int64_t)nbu * kT * kT * (cplx ? 2 : 1);
    LHL_TRY(Uinv.ensure(sizeof(T) * (size_t)uis * (size_t)nshift));
    *uistride = uis;
    if (cplx)
        LHL_LAUNCH(h, "shift", (k_uinv<T, 1>), dim3((unsigned)nbu, (unsigned)nshift), dim3(64), 0, n, (const T *)U,
                   ustride, (const T *)rdiag, lstride, Uinv.as<T>(), uis);
    else
        LHL_LAUNCH(h, "shift", (k_uinv<T, 0>), dim3((unsigned)nbu, (unsigned)nshift), dim3(64), 0, n, (const T *)U,
                   ustride, (const T *)rdiag, lstride, Uinv.as<T>(), uis);
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

template <typename TW, typename TX>
Status residual_typed(lhl_handle *h, const ResidArgs &a) {
    const int n = (int)h->n;
    const int planes = a.x_complex ? 2 : 1;
    const int Q = (int)(a.ncol * planes);
    const int64_t ldq = ldy_for(n);
    LHL_TRY(h->ybuf.ensure(sizeof(TX) * (size_t)ldq * Q));
    TX *JX = h->ybuf.as<TX>();
    TX *JXi = nullptr;
    const dim3 gg((unsigned)((n + kT - 1) / kT), (unsigned)((Q + kT - 1) / kT));
    if (a.W) {
        LHL_LAUNCH(h, "residual", (k_gemm_jx<TW, TX>), gg, dim3(256), 0, n, (const TW *)a.W, a.ldw, (const TX *)a.x,
                   a.ld, a.plane, planes, Q, JX, ldq);
        if (a.w_complex) {
            LHL_TRY(h->rbuf.ensure(sizeof(TX) * (size_t)ldq * Q));
            JXi = h->rbuf.as<TX>();
            const TW *Wi = (const TW *)a.W + a.ldw * (int64_t)n;
            LHL_LAUNCH(h, "residual", (k_gemm_jx<TW, TX>), gg, dim3(256), 0, n, Wi, a.ldw, (const TX *)a.x, a.ld,
                       a.plane, planes, Q, JXi, ldq);
        }
    } else {
        LHL_LAUNCH(h, "residual", (k_gemm_jx<TW, TX>), gg, dim3(256), 0, n, (const TW *)h->J.p, h->ldj,
                   (const TX *)a.x, a.ld, a.plane, planes, Q, JX, ldq);
    }
    const dim3 gc((unsigned)((n + 255) / 256), (unsigned)a.ncol);
    LHL_LAUNCH(h, "residual", k_resid_combine<TX>, gc, dim3(256), 0, n, (const TX *)JX, ldq, (const TX *)JXi,
               (const TX *)a.x, (const TX *)a.b, (TX *)a.r, a.ld, a.plane, planes, a.W ? nullptr : a.colshift,
               a.W ? a.w_complex : 0);
    LHL_LAUNCH(h, "residual", k_colnorms<TX>, dim3((unsigned)a.ncol), dim3(256), 0, n, (const TX *)a.r,
               (const TX *)a.b, a.ld, a.plane, planes, (int)a.ncol, a.norms);
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

template <typename T>
Status linv_typed(lhl_handle *h) {
    const int n = (int)h->n;
    const int nbt = (n - 1 + kT - 1) / kT;
    LHL_TRY(h->Linv.ensure(sizeof(T) * (size_t)std::max(nbt, 1) * kT * kT));
    if (nbt > 0)
        LHL_LAUNCH(h, "reduce_finish", k_linv<T>, dim3((unsigned)nbt), dim3(64), 0, n, (const T *)h->A.as<T>(), h->lda,
                   h->Linv.as<T>());
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

}  // namespace

Status uinv_device(lhl_handle *h, const void *U, int64_t ustride, const void *rdiag, int64_t lstride,
                   int64_t nshift, int is_complex, DevBuf &Uinv, int64_t *uistride) {
    if (h->n == 0 || nshift == 0) return {};
    return h->dtype == LHL_F64 ? uinv_typed<double>(h, U, ustride, rdiag, lstride, nshift, is_complex, Uinv, uistride)
                               : uinv_typed<float>(h, U, ustride, rdiag, lstride, nshift, is_complex, Uinv, uistride);
}

Status linv_device(lhl_handle *h) {
    if (h->n < 2) return {};
    return h->dtype == LHL_F64 ? linv_typed<double>(h) : linv_typed<float>(h);
}

Status solve_device(lhl_handle *h, const SolveArgs &a) {
    if (h->n == 0 || a.ncol == 0) return {};
    if (h->dtype == LHL_F64) {
        if (a.xdt != LHL_F64) return {LHL_ERR_ARG, "lhl_ldiv: a Float32 vector cannot take a Float64 solve"};
        return solve_batch_typed<double, double>(h, a);
    }
    if (a.xdt == LHL_F64) return solve_batch_typed<float, double>(h, a);
    return solve_batch_typed<float, float>(h, a);
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

}  // namespace lhl
