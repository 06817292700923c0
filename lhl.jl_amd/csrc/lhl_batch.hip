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
#include <rocblas/rocblas.h>

#include <type_traits>
#include "lhl_common.h"

#include <cstring>
#include "lhl_mfma.h"

namespace lhl {

constexpr int kT = 64;   // tile edge

// Y[i + q*ldy] = x[perm[i], col] * iscale[perm[i]] for internal column q = col*planes + pl;
// row 0 is also copied to Y2 (the second buffer of the triangular solve).
template <typename T, typename TX>
__global__ void k_gather(int n, const int *__restrict__ perm, const double *__restrict__ iscale,
                         const TX *__restrict__ x, int64_t ldx, int64_t xplane, int planes, T *__restrict__ Y,
                         T *__restrict__ Y2, int64_t ldy, const int *__restrict__ cact) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (i >= n) return;
    const int col = q / planes, pl = q % planes;
    if (cact && !cact[col]) return;   // a converged column (refinement rounds)
    const int p = perm[i];
    const T v = (T)(x[p + (int64_t)col * ldx + pl * xplane] * (TX)iscale[p]);
    Y[i + (int64_t)q * ldy] = v;
    if (i == 0) Y2[(int64_t)q * ldy] = v;
}

// The same with one workgroup per internal column: the column of x read into
// LDS in one coalesced pass, then permuted from there (the row-per-thread
// form reads x at permuted rows straight from memory).
template <typename T, typename TX>
__global__ void __launch_bounds__(256) k_gather_col(int n, const int *__restrict__ perm,
                                                    const double *__restrict__ iscale, const TX *__restrict__ x,
                                                    int64_t ldx, int64_t xplane, int planes, T *__restrict__ Y,
                                                    T *__restrict__ Y2, int64_t ldy, const int *__restrict__ cact) {
    const int q = blockIdx.x;
    const int col = q / planes, pl = q % planes;
    if (cact && !cact[col]) return;   // a converged column (whole workgroup)
    extern __shared__ __align__(16) unsigned char lds_dyn[];
    TX *xs = reinterpret_cast<TX *>(lds_dyn);
    const TX *xc = x + (int64_t)col * ldx + pl * xplane;
#pragma unroll 8
    for (int i = threadIdx.x; i < n; i += 256) xs[i] = xc[i];
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 256) {
        const int p = perm[i];
        const T v = (T)(xs[p] * (TX)iscale[p]);
        Y[i + (int64_t)q * ldy] = v;
        if (i == 0) Y2[(int64_t)q * ldy] = v;
    }
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
        for (int i = m + 1; i < kT; i++) x[i] = fma_t(-Ls[i][m], xm, x[i]);
    }
    T *dst = Linv + (int64_t)b * kT * kT + (int64_t)c * kT;
#pragma unroll
    for (int i = 0; i < kT; i++) dst[i] = x[i];
}

// x[perm[p], col] (+)= y_new[p] * d[perm[p]]: the scatter that ends a solve
// (src:3092-3107), from the swept rows y_new' (rows 1.. of Yn) and row 0 of Y
// (the sweep leaves y[0] as it is).  Internal column q = col*planes + plane.
// Run as a gather: thread r takes p = iperm[r], so the read-modify-writes of x
// are contiguous and only the reads of y are permuted.
template <typename T, typename TX>
__global__ void k_scatter(int n, const T *__restrict__ Yn, const T *__restrict__ Y, int64_t ldy, int planes,
                          const int *__restrict__ iperm, const double *__restrict__ scale, TX *__restrict__ x,
                          int64_t ldx, int64_t xplane, int accumulate, const int *__restrict__ cact) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (r >= n) return;
    const int col = q / planes, pl = q % planes;
    if (cact && !cact[col]) return;   // a converged column keeps its x
    const int p = iperm[r];
    const T yn = p == 0 ? Y[(int64_t)q * ldy] : Yn[p + (int64_t)q * ldy];
    TX *dst = x + r + (int64_t)col * ldx + pl * xplane;
    const TX v = (TX)yn * (TX)scale[r];
    if (accumulate) *dst += v;
    else *dst = v;
}

// The explicit inverse of L' (n-1 x n-1, unit lower): column j solves
// L' z = e_j.  One workgroup per 16 columns, left-looking from the block
// holding the first column (the rows above are zero and never stored); each
// block is finished with its 64x64 diagonal inverse.  Lf is column-major with
// leading dimension ldf.
template <typename T>
__global__ void __launch_bounds__(64 * 8) k_linv_full(int n, const T *__restrict__ A, int64_t lda,
                                                      const T *__restrict__ Linv, T *__restrict__ Lf, int64_t ldf) {
    constexpr int NW = 8, NC = 16;
    const int N1 = n - 1, nbt = (N1 + kT - 1) / kT;
    const int c0 = blockIdx.x * NC;
    const int b0 = c0 / kT;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const T *Xc = Lf + (int64_t)min(c0 + li, N1 - 1) * ldf;   // solved rows m at Xc[m]
    __shared__ T red[NW][kT][NC + 1];
    __shared__ T Ys[kT][NC + 1];
    for (int b = b0; b < nbt; b++) {
        const int r0 = b * kT, nb = min(kT, N1 - r0);
        typename Mfma<T>::acc_t acc[4];
#pragma unroll
        for (int rr = 0; rr < 4; rr++) acc[rr] = Mfma<T>::zero();
        const T *Ar = A + r0 + 4 * li + 1;   // MFMA row i of accumulator rr is block row 4i + rr
        const bool full = nb == kT;
        for (int j = 4 * b0 + w; j < 4 * b; j += NW) {
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
        for (int t = threadIdx.x; t < kT * NC; t += 64 * NW) {
            const int i = t % kT, c = t / kT;
            T v = 0;
            if (i < nb && c0 + c < N1) {
                T sum = 0;
#pragma unroll
                for (int ww = 0; ww < NW; ww++) sum += red[ww][i][c];
                v = (r0 + i == c0 + c ? (T)1 : (T)0) - sum;
            }
            Ys[i][c] = v;
        }
        __syncthreads();
        if (w < 4) {
            const T *Lb = Linv + (int64_t)b * kT * kT + 16 * w + li;
            typename Mfma<T>::acc_t xo = Mfma<T>::zero();
#pragma unroll
            for (int sl = 0; sl < 4; sl++)
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const int k = 16 * sl + 4 * lk + t;
                    xo = Mfma<T>::mac(Lb[(int64_t)k * kT], Ys[k][li], xo);
                }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = 16 * w + Mfma<T>::row(lane, r);
                if (i < nb && c0 + li < N1) Lf[(r0 + i) + (int64_t)(c0 + li) * ldf] = xo[r];
            }
        }
        __syncthreads();   // the block's rows are visible to every wave
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
                                             int64_t uistride, int ut) {
#pragma clang fp contract(off)
    constexpr int UP = UC ? 2 : 1;
    const int b = blockIdx.x, s = blockIdx.y;
    const int r0 = b * kT, nb = min(kT, n - r0);
    const T *U = Ubase + (int64_t)s * ustride;
    const T *rd = rdbase + (int64_t)s * lstride;
    const int64_t up = ustride / 2, lp = lstride / 2;
    // the block's strict upper triangle, packed by rows (half the LDS of a
    // square block: twice the workgroups per CU in FP64); every read is a
    // broadcast of one element to the whole wave
    constexpr int NUP = kT * (kT - 1) / 2;
    __shared__ T Us[UP][NUP];
    auto upk = [](int i, int j) -> int { return i * kT - (i * (i + 1)) / 2 + (j - i - 1); };   // j > i
    __shared__ T ds[UP][kT];
    // the block's strict upper triangle: lane c loads column c of every row,
    // all loads issued before the first is used (a select on each value right
    // after its load would wait for it: 64 round trips)
    {
        const int c = threadIdx.x;
        constexpr int UV = utile_gw<T>();
        if constexpr (UC) {   // complex: one row at a time (registers)
            for (int r = 0; r < kT; r++) {
                const bool ok = r < nb && c < nb && c > r;
                const int rr = ok ? r0 + r : 0, cc = ok ? r0 + c : 0;
                const int64_t o = ut ? utile_row<UV>((n + 63) / 64, rr) + utile_col<UV>(cc) : upack_off(n, rr) - rr + cc;
                if (c > r) {
                    Us[0][upk(r, c)] = ok ? U[o] : (T)0;
                    Us[UP - 1][upk(r, c)] = ok ? U[o + up] : (T)0;
                }
            }
        } else {
            T v[kT];
#pragma unroll
            for (int r = 0; r < kT; r++) {
                const bool ok = r < nb && c < nb && c > r;
                const int rr = ok ? r0 + r : 0, cc = ok ? r0 + c : 0;
                const int64_t o = ut ? utile_row<UV>((n + 63) / 64, rr) + utile_col<UV>(cc) : upack_off(n, rr) - rr + cc;
                v[r] = U[o];
            }
#pragma unroll
            for (int r = 0; r < kT; r++)
                if (c > r) Us[0][upk(r, c)] = (r < nb && c < nb) ? v[r] : (T)0;
        }
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
        // real shifts: four partial sums over j mod 4 (a quarter of the
        // dependent chain); complex: one (registers)
        constexpr int NPS = UC ? 1 : 4;
        T pr[4] = {(i == c) ? (T)1 : (T)0, 0, 0, 0}, pim[4] = {0, 0, 0, 0};
#pragma unroll
        for (int j = i + 1; j < kT; j++) {
            const int q = (j - i - 1) % NPS;
            pr[q] = fma_t(-Us[0][upk(i, j)], xr[j], pr[q]);
            if constexpr (UC) {
                pr[q] = fma_t(Us[UP - 1][upk(i, j)], xi[j], pr[q]);
                pim[q] = fma_t(-Us[0][upk(i, j)], xi[j], pim[q]);
                pim[q] = fma_t(-Us[UP - 1][upk(i, j)], xr[j], pim[q]);
            }
        }
        const T ar = (pr[0] + pr[1]) + (pr[2] + pr[3]), ai = (pim[0] + pim[1]) + (pim[2] + pim[3]);
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
    int shift, q0, nq, pad;   // internal columns q0 .. q0+nq-1 (nq <= kNC), all of one shift; nq = 0: idle slot
};

// ---- back substitution with U: kNC internal columns per workgroup ---------
// With few columns per workgroup the batch spreads over every CU (1024
// internal columns fill 256 CUs) and each shift's U streams from HBM once,
// shared through the L2 of the XCD its workgroups are placed on.  VALU FMAs:
// four columns per row leave no MFMA shape worth using.
constexpr int kNC = 4;

// Sum of v[0..V-1] over the 64 lanes, scattered: lane l returns the total of
// v[l % V] (V a power of two <= 64): fold the lane halves down to V lanes,
// then a butterfly reduce-scatter (63 exchanges in all).
template <typename T, int V>
__device__ __forceinline__ T wave_reduce_scatter(T (&v)[V]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int span = 32; span >= V; span >>= 1)
#pragma unroll
        for (int j = 0; j < V; j++) v[j] += __shfl_xor(v[j], span, 64);
#pragma unroll
    for (int half = V / 2; half >= 1; half >>= 1) {
        const bool hi = (lane & half) != 0;
#pragma unroll
        for (int j = 0; j < half; j++) {
            const T keep = hi ? v[j + half] : v[j];
            const T send = hi ? v[j] : v[j + half];
            v[j] = keep + __shfl_xor(send, half, 64);
        }
    }
    return v[0];
}

// x <- U_s^-1 y for kNC internal columns of one shift (two complex columns,
// re/im interleaved, when UC).  U's rows are packed, so lanes run along a row
// (each lane a 16-byte run, coalesced) and each wave owns RW rows of the
// 64-row block; per-lane partial sums are combined by a reduce-scatter.  The
// chunk loop keeps two register sets: the loads of chunk c+1 are in flight
// while chunk c's FMAs run.  The block's right-hand side and diagonal inverse
// are fetched when the block starts and used after its K loop.  Workgroups of
// one shift are placed on one XCD by the host so U comes from HBM once per
// shift.  RW = 8 rows per wave, 8 waves; complex shifts (twice the loads per
// chunk) keep one register set.
template <typename T> constexpr int utrsm4_rows() { return 8; }

template <typename T> struct V16T;
template <> struct V16T<float> { typedef float type __attribute__((ext_vector_type(4), aligned(4))); };
template <> struct V16T<double> { typedef double type __attribute__((ext_vector_type(2), aligned(8))); };
template <typename T> using V16 = typename V16T<T>::type;   // 16 bytes, element-aligned loads
template <typename T> __device__ __forceinline__ V16<T> ld16u(const T *p) { return *(const V16<T> *)p; }

template <typename T, int UC>
__global__ void __launch_bounds__(64 * (64 / utrsm4_rows<T>()))
k_utrsm4(int n, const T *__restrict__ Ubase, int64_t ustride, const T *__restrict__ Uinv, int64_t uistride,
         const UGroup *__restrict__ groups, const T *__restrict__ Yin, T *__restrict__ Yf, int64_t ldy,
         const int *__restrict__ sact) {
    constexpr int UP = UC ? 2 : 1;
    constexpr int RW = utrsm4_rows<T>();
    constexpr int NV = RW * kNC;           // partial sums per lane
    constexpr int NT = 64 * (64 / RW);     // threads
    constexpr int RL = 16 / sizeof(T);     // elements per lane per load
    constexpr int CH = 64 * RL;            // chunk width
    constexpr int NPRE = kT * kT * UP / NT;   // diagonal-inverse elements per thread
    const UGroup g = groups[blockIdx.x];
    if (g.nq <= 0) return;   // padding slot of the XCD placement (whole workgroup)
    if (sact && !sact[g.shift]) return;   // a converged shift (whole workgroup)
    const int nbu = (n + kT - 1) / kT;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const T *Us = Ubase + (int64_t)g.shift * ustride;
    const int64_t up = ustride / 2;
    const T *Xc[kNC];
#pragma unroll
    for (int c = 0; c < kNC; c++) Xc[c] = Yf + (int64_t)(g.q0 + min(c, g.nq - 1)) * ldy;
    __shared__ T Ysum[kT][kNC];
    __shared__ T Ys[kT][kNC];
    __shared__ T UiS[UP][kT][kT + 1];   // UiS[p][k][i] = Uinv_b(i, k), plane p
    const int di = tid >> 2, dc = tid & 3;   // the (row, column) of the block's solve this thread owns (tid < 256)
    for (int b = nbu - 1; b >= 0; b--) {
        const int rb = b * kT, nbk = min(kT, n - rb);
        const int kbeg = rb + kT;
        // the block's own operands, consumed after the K loop
        T pre[NPRE];
        {
            const T *Ui = Uinv + (int64_t)g.shift * uistride + (int64_t)b * kT * kT * UP;
#pragma unroll
            for (int u = 0; u < NPRE; u++) pre[u] = Ui[tid + NT * u];
        }
        T yin = 0;
        if (tid < kT * kNC) {
            const T yv = Yin[min(rb + di, n - 1) + (int64_t)(g.q0 + min(dc, g.nq - 1)) * ldy];
            yin = (di < nbk && dc < g.nq) ? yv : (T)0;
        }
        T v[NV];
#pragma unroll
        for (int j = 0; j < NV; j++) v[j] = 0;
        if (kbeg < n) {
            int64_t ro[RW];   // rows rb + RW*w + p (all < n here)
#pragma unroll
            for (int p = 0; p < RW; p++) {
                const int row = rb + RW * w + p;
                ro[p] = upack_off(n, row) - row;
            }
            const int nfull = (n - kbeg) / CH;
            V16<T> xA[kNC], uA[RW], iA[UC ? RW : 1], xB[kNC], uB[RW], iB[UC ? RW : 1];
            // chunk c into a register set (c clamped to the last full chunk: such loads are never used)
            auto load_set = [&](int c, V16<T>(&xs)[kNC], V16<T>(&us)[RW], V16<T>(&is)[UC ? RW : 1]) __attribute__((always_inline)) {
                const int m = kbeg + min(c, nfull - 1) * CH + RL * lane;
#pragma unroll
                for (int cc = 0; cc < kNC; cc++) xs[cc] = ld16u(Xc[cc] + m);
#pragma unroll
                for (int p = 0; p < RW; p++) {
                    us[p] = ld16u(Us + ro[p] + m);
                    if constexpr (UC) is[p] = ld16u(Us + up + ro[p] + m);
                }
            };
            auto fma_set = [&](const V16<T>(&xs)[kNC], const V16<T>(&us)[RW], const V16<T>(&is)[UC ? RW : 1]) __attribute__((always_inline)) {
#pragma unroll
                for (int p = 0; p < RW; p++) {
#pragma unroll
                    for (int t = 0; t < RL; t++) {
                        const T ur = us[p][t];
                        if constexpr (UC) {
                            const T ui = is[p][t];
                            v[4 * p + 0] = fma_t(ur, xs[0][t], fma_t(-ui, xs[1][t], v[4 * p + 0]));
                            v[4 * p + 1] = fma_t(ur, xs[1][t], fma_t(ui, xs[0][t], v[4 * p + 1]));
                            v[4 * p + 2] = fma_t(ur, xs[2][t], fma_t(-ui, xs[3][t], v[4 * p + 2]));
                            v[4 * p + 3] = fma_t(ur, xs[3][t], fma_t(ui, xs[2][t], v[4 * p + 3]));
                        } else {
#pragma unroll
                            for (int cc = 0; cc < kNC; cc++) v[4 * p + cc] = fma_t(ur, xs[cc][t], v[4 * p + cc]);
                        }
                    }
                }
            };
            if (nfull > 0 && !UC) {
                load_set(0, xA, uA, iA);
                load_set(1, xB, uB, iB);
                int c = 0;
                // scheduling barriers keep each set's next loads after its FMAs,
                // so only two sets are ever live
                for (; c + 2 <= nfull; c += 2) {
                    fma_set(xA, uA, iA);
                    __builtin_amdgcn_sched_barrier(0);
                    load_set(c + 2, xA, uA, iA);
                    __builtin_amdgcn_sched_barrier(0);
                    fma_set(xB, uB, iB);
                    __builtin_amdgcn_sched_barrier(0);
                    load_set(c + 3, xB, uB, iB);
                    __builtin_amdgcn_sched_barrier(0);
                }
                if (c < nfull) fma_set(xA, uA, iA);
            } else {
                for (int c = 0; c < nfull; c++) {
                    load_set(c, xA, uA, iA);
                    fma_set(xA, uA, iA);
                }
            }
            // the last partial chunk: clamped addresses, out-of-range terms selected to zero
            const int m0 = kbeg + nfull * CH;
#pragma unroll
            for (int t = 0; t < RL; t++) {
                const int m = m0 + RL * lane + t;
                const bool ok = m < n;
                const int mc = min(m, n - 1);
                T xv[kNC];
#pragma unroll
                for (int cc = 0; cc < kNC; cc++) {
                    const T x = Xc[cc][mc];
                    xv[cc] = ok ? x : (T)0;
                }
#pragma unroll
                for (int p = 0; p < RW; p++) {
                    const T u0 = Us[ro[p] + mc];
                    const T ur = ok ? u0 : (T)0;
                    if constexpr (UC) {
                        const T u1 = Us[up + ro[p] + mc];
                        const T ui = ok ? u1 : (T)0;
                        v[4 * p + 0] = fma_t(ur, xv[0], fma_t(-ui, xv[1], v[4 * p + 0]));
                        v[4 * p + 1] = fma_t(ur, xv[1], fma_t(ui, xv[0], v[4 * p + 1]));
                        v[4 * p + 2] = fma_t(ur, xv[2], fma_t(-ui, xv[3], v[4 * p + 2]));
                        v[4 * p + 3] = fma_t(ur, xv[3], fma_t(ui, xv[2], v[4 * p + 3]));
                    } else {
#pragma unroll
                        for (int cc = 0; cc < kNC; cc++) v[4 * p + cc] = fma_t(ur, xv[cc], v[4 * p + cc]);
                    }
                }
            }
        }
        const T tot = wave_reduce_scatter<T, NV>(v);   // lane l < NV: row RW*w + l/4, column l%4
        if (lane < NV) Ysum[RW * w + (lane >> 2)][lane & 3] = tot;
#pragma unroll
        for (int u = 0; u < NPRE; u++) {
            const int e = tid + NT * u;   // plane p, column-major (i, k)
            const int pl = e / (kT * kT), ik = e % (kT * kT);
            UiS[pl][ik / kT][ik % kT] = pre[u];
        }
        __syncthreads();
        if (tid < kT * kNC) Ys[di][dc] = yin - Ysum[di][dc];
        __syncthreads();
        if (tid < kT * kNC) {
            T x = 0;
            if constexpr (UC) {
                const int cr = dc & ~1;   // the complex column's real part
                const bool im = (dc & 1) != 0;
#pragma unroll 8
                for (int k = 0; k < kT; k++) {
                    const T ar = UiS[0][k][di], ai = UiS[UP - 1][k][di];
                    const T yr = Ys[k][cr], yi = Ys[k][cr + 1];
                    x = im ? fma_t(ar, yi, fma_t(ai, yr, x))
                           : fma_t(ar, yr, fma_t(-ai, yi, x));
                }
            } else {
#pragma unroll 16
                for (int k = 0; k < kT; k++) x = fma_t(UiS[0][k][di], Ys[k][dc], x);
            }
            if (di < nbk && dc < g.nq) Yf[(rb + di) + (int64_t)(g.q0 + dc) * ldy] = x;
        }
        __syncthreads();
    }
}

// ---- U in packed columns -----------------------------------------------------
// UT[ucol_off(j) + i] = U[i, j] for i <= j: each 64x64 tile of the upper
// triangle is read along U's packed rows and written along columns through
// LDS (both sides contiguous).  One workgroup per tile (blockIdx.x,
// upper-triangle index) and shift-plane (blockIdx.y = shift * planes + plane).
// Each lane writes V consecutive rows of a column (V = 4 FP32, 2 FP64) as one
// aligned 16-byte store, 64 / V lanes per 64-row column segment; the rows past
// the diagonal in a diagonal tile land in the column's padding (its slots up
// to round_up(j+1, V)), which the back substitution never reads.  Offsets are
// 32-bit: UT is built for n <= 8192, where every packed offset is below 2^26.
template <typename T>
__global__ void __launch_bounds__(256) k_ut_pack(int n, const T *__restrict__ U, T *__restrict__ UT, int64_t ustride,
                                                 int planes) {
    constexpr int V = ut_vec<T>();
    // tile t of the upper triangle, column-block major: J = the largest with
    // J(J+1)/2 <= t, I = t - J(J+1)/2 <= J
    const int t = blockIdx.x;
    int J = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
    while ((J + 1) * (J + 2) / 2 <= t) J++;
    while (J * (J + 1) / 2 > t) J--;
    const int I = t - J * (J + 1) / 2;
    const int s = blockIdx.y / planes, pl = blockIdx.y % planes;
    const T *Us = U + (int64_t)s * ustride + (int64_t)pl * (ustride / 2);
    T *Ut = UT + (int64_t)s * ustride + (int64_t)pl * (ustride / 2);
    __shared__ T tile[kT][kT + 1];
    const int i0 = I * kT, j0 = J * kT, tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int i = i0 + ty + 4 * r, j = j0 + tx;
        // U[i, j] at i n - i (i + 1) / 2 + j (row i of the packed rows starts at U[i, i])
        if (i < n && j < n && j >= i) tile[ty + 4 * r][tx] = Us[i * n - (i * (i + 1)) / 2 + j];
    }
    __syncthreads();
    if constexpr (V > 1) {
        // kT / V lanes per 64-row column segment, each storing V rows as one
        // aligned 16-byte store; 256 V / kT columns per pass
        constexpr int LC = kT / V, CP = 256 / LC;
        const int l = threadIdx.x % LC, cb = threadIdx.x / LC;
        using FV = T __attribute__((ext_vector_type(V)));
#pragma unroll
        for (int r = 0; r < kT / CP; r++) {
            const int c = cb + CP * r, j = j0 + c, ir = V * l, i = i0 + ir;
            if (j < n && i <= j) {
                FV v;
#pragma unroll
                for (int e = 0; e < V; e++) v[e] = tile[ir + e][c];
                *reinterpret_cast<FV *>(Ut + (int)ucol_off<V>(j) + i) = v;
            }
        }
    } else {
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int j = j0 + ty + 4 * r, i = i0 + tx;
            if (j < n && i <= j) Ut[(int)ucol_off<V>(j) + i] = tile[tx][ty + 4 * r];
        }
    }
}

// x <- U_s^-1 y for kNC internal columns of one shift from U in packed
// columns, right-looking over 64-row blocks from the bottom.  Rows come in
// groups of V consecutive rows (V = ut_vec<T>()); thread t keeps the
// right-hand sides of groups t + NT j (j < NG) in registers.  Per block: its
// owner threads publish the block's rows, the 64x64 diagonal inverse gives
// x_b, and every thread subtracts U[i, block] x_b from its groups above the
// block: one aligned V-wide load per group and column (lanes run down a
// column of U, contiguous), MB columns per batch, the next batch in flight.
// Complex shifts: columns alternate (re, im), U is planar.  n <= V NT NG.
template <typename T, int UC, int NG, int NT>
__global__ void __launch_bounds__(NT) k_utrsm_t(int n, const T *__restrict__ UTbase, int64_t ustride,
                                                 const T *__restrict__ Uinv, int64_t uistride,
                                                 const UGroup *__restrict__ groups, const T *__restrict__ Yin,
                                                 T *__restrict__ Yf, int64_t ldy, const int *__restrict__ sact) {
    constexpr int V = ut_vec<T>();
    using VT = T __attribute__((ext_vector_type(V)));
    constexpr int UP = UC ? 2 : 1;
    constexpr int NPRE = (kT * kT * UP + NT - 1) / NT;
    // columns per batch: 16 one-element loads (V = 1), or 8 (4 with two planes)
    // V-wide loads
    constexpr int MB = V == 1 ? 16 : (UC ? 4 : 8);
    constexpr int NB = kT / MB;        // batches per 64-column block
    constexpr int GB = kT / V;         // groups per 64-row block
    const UGroup g = groups[blockIdx.x];
    if (g.nq <= 0) return;   // padding slot of the XCD placement (whole workgroup)
    if (sact && !sact[g.shift]) return;   // a converged shift (whole workgroup)
    const int nbu = (n + kT - 1) / kT;
    const int tid = threadIdx.x;
    const T *Ut = UTbase + (int64_t)g.shift * ustride;
    const int64_t up = ustride / 2;
    __shared__ T Ys[kT][kNC];
    __shared__ T Xb[kT][kNC];
    __shared__ T UiS[UP][kT][kT + 1];   // UiS[p][k][i] = Uinv_b(i, k), plane p
    T yv[NG][V][kNC];
#pragma unroll
    for (int j = 0; j < NG; j++)
#pragma unroll
        for (int r = 0; r < V; r++) {
            const int i = V * (tid + NT * j) + r;
#pragma unroll
            for (int c = 0; c < kNC; c++) {
                const T v = Yin[min(i, n - 1) + (int64_t)(g.q0 + min(c, g.nq - 1)) * ldy];
                yv[j][r][c] = (i < n && c < g.nq) ? v : (T)0;
            }
        }
    const int di = tid >> 2, dc = tid & 3;
    // batch t of block bb's update: MB columns of U for the thread's group
    // slot t / NB (threads whose group is not above the block read a valid
    // one, unused; group 0 for block 0, which has none above it)
    auto load_at = [&](int bb, int t, VT(&ur)[MB], VT(&ui)[UC ? MB : 1]) __attribute__((always_inline)) {
        const int rbb = bb * kT, gbb = rbb / V;
        const int gc = max(min(tid + NT * (t / NB), gbb - 1), 0);
        const int col0 = rbb + (t % NB) * MB;
        // one running offset: column col+1 starts round_up(col+1, V)
        // elements after column col; past the last column it stays there
        // (those x_b rows are zero)
        const T *pr = Ut + ucol_off<V>(min(col0, n - 1)) + V * gc;
#pragma unroll
        for (int m = 0; m < MB; m++) {
            ur[m] = *reinterpret_cast<const VT *>(pr);
            if constexpr (UC) ui[m] = *reinterpret_cast<const VT *>(pr + up);
            const int col = col0 + m;
            // branch-free (a branch here makes the compiler wait for every load)
            pr += ((col + V) / V) * V & -(int)(col + 1 < n);
        }
    };
    VT uA[MB], iA[UC ? MB : 1], uB[MB], iB[UC ? MB : 1];
    // the diagonal block inverse of block bb, one block ahead: the next
    // block's is loaded while this block's update runs (its latency was
    // exposed at every block start when loaded there)
    T pre[NPRE];
    auto load_pre = [&](int bb, T(&dst)[NPRE]) __attribute__((always_inline)) {
        const T *Ui = Uinv + (int64_t)g.shift * uistride + (int64_t)bb * kT * kT * UP;
#pragma unroll
        for (int u = 0; u < NPRE; u++) dst[u] = Ui[min(tid + NT * u, kT * kT * UP - 1)];
    };
    load_pre(nbu - 1, pre);
    load_at(nbu - 1, 0, uA, iA);
    for (int b = nbu - 1; b >= 0; b--) {
        const int rb = b * kT, nbk = min(kT, n - rb);
        const int gb = rb / V;                 // the block's first group; groups < gb are above it
        // Step 3 below (groups above the block: y_i -= U[i, rb:rb+nbk] x_b) is a
        // sequence of batches t (slot t/NB, columns MB (t%NB) .. +MB-1); the
        // loads of batch t+1 are in flight while batch t's FMAs run, and the
        // last batch's successor is the next block's batch 0, so U loads stay
        // in flight across steps 1-2 and their LDS-only barriers.
        const int nbat = NB * ((gb + NT - 1) / NT);   // slots holding groups < gb (nbat is even)
        // 1. the block's rows: groups gb .. gb+GB-1, slot gb / NT of threads
        //    gb % NT .. (one wave: GB divides 64 and gb % NT is a multiple of GB)
        {
            const int jb = gb / NT, tb = gb % NT;
            if (tid >= tb && tid < tb + GB) {
                const int m = tid - tb;
#pragma unroll
                for (int r = 0; r < V; r++)
#pragma unroll
                    for (int c = 0; c < kNC; c++) {
                        T yb = 0;
#pragma unroll
                        for (int j = 0; j < NG; j++) yb = j == jb ? yv[j][r][c] : yb;
                        Ys[V * m + r][c] = V * m + r < nbk ? yb : (T)0;
                    }
            }
        }
#pragma unroll
        for (int u = 0; u < NPRE; u++) {
            const int e = tid + NT * u;   // plane p, column-major (i, k)
            const int pl = e / (kT * kT), ik = e % (kT * kT);
            if (e < kT * kT * UP) UiS[pl][ik / kT][ik % kT] = pre[u];
        }
        load_pre(max(b - 1, 0), pre);   // block b-1's (block 0: a repeat, unused)
        lds_barrier();   // LDS only: the early U loads stay in flight
        // 2. x_b = Uinv_b y_b
        if (tid < kT * kNC) {
            T x = 0;
            if constexpr (UC) {
                const int cr = dc & ~1;
                const bool im = (dc & 1) != 0;
#pragma unroll 8
                for (int k = 0; k < kT; k++) {
                    const T ar = UiS[0][k][di], ai = UiS[UP - 1][k][di];
                    const T yr = Ys[k][cr], yi = Ys[k][cr + 1];
                    x = im ? fma_t(ar, yi, fma_t(ai, yr, x)) : fma_t(ar, yr, fma_t(-ai, yi, x));
                }
            } else {
#pragma unroll 16
                for (int k = 0; k < kT; k++) x = fma_t(UiS[0][k][di], Ys[k][dc], x);
            }
            Xb[di][dc] = di < nbk ? x : (T)0;
            if (di < nbk && dc < g.nq) Yf[(rb + di) + (int64_t)(g.q0 + dc) * ldy] = x;
        }
        lds_barrier();   // LDS only: the early U loads stay in flight
        auto fma_b = [&](int t, const VT(&ur)[MB], const VT(&ui)[UC ? MB : 1], T(&acc)[V][kNC])
                         __attribute__((always_inline)) {
            const int m0 = (t % NB) * MB;
#pragma unroll
            for (int m = 0; m < MB; m++) {
                // x_b is read from LDS in groups of 8 rows (hoisting all of it would spill)
                if (m % 8 == 0) __builtin_amdgcn_sched_barrier(0);
                const T *xb = Xb[m0 + m];   // zero beyond the block
#pragma unroll
                for (int r = 0; r < V; r++) {
                    const T u = ur[m][r];
                    if constexpr (UC) {
                        const T v = ui[m][r];
                        acc[r][0] = fma_t(u, xb[0], fma_t(-v, xb[1], acc[r][0]));
                        acc[r][1] = fma_t(u, xb[1], fma_t(v, xb[0], acc[r][1]));
                        acc[r][2] = fma_t(u, xb[2], fma_t(-v, xb[3], acc[r][2]));
                        acc[r][3] = fma_t(u, xb[3], fma_t(v, xb[2], acc[r][3]));
                    } else {
#pragma unroll
                        for (int c = 0; c < kNC; c++) acc[r][c] = fma_t(u, xb[c], acc[r][c]);
                    }
                }
            }
        };
        // after the last batch of slot js: y -= acc for its groups above the block
        auto apply = [&](int t, T(&acc)[V][kNC]) __attribute__((always_inline)) {
            if (t % NB != NB - 1) return;
            const int js = t / NB;
            // a select on every slot: an `if (j == js)` chain would be folded
            // into a dynamic index of yv, which would move yv to scratch
#pragma unroll
            for (int j = 0; j < NG; j++) {
                const bool hit = j == js && tid + NT * j < gb;
#pragma unroll
                for (int r = 0; r < V; r++)
#pragma unroll
                    for (int c = 0; c < kNC; c++) yv[j][r][c] = hit ? yv[j][r][c] - acc[r][c] : yv[j][r][c];
            }
#pragma unroll
            for (int r = 0; r < V; r++)
#pragma unroll
                for (int c = 0; c < kNC; c++) acc[r][c] = 0;
        };
        if (nbat > 0) {
            T acc[V][kNC];
#pragma unroll
            for (int r = 0; r < V; r++)
#pragma unroll
                for (int c = 0; c < kNC; c++) acc[r][c] = 0;
            for (int t = 0; t < nbat; t += 2) {
                __builtin_amdgcn_sched_barrier(0);
                load_at(b, t + 1, uB, iB);
                __builtin_amdgcn_sched_barrier(0);
                fma_b(t, uA, iA, acc);
                apply(t, acc);
                __builtin_amdgcn_sched_barrier(0);
                // after the last batch: the next block's batch 0 (block 0 has no
                // next: a repeat of its batch 0, unused)
                const bool last = t + 2 >= nbat;
                load_at(last ? max(b - 1, 0) : b, last ? 0 : t + 2, uA, iA);
                __builtin_amdgcn_sched_barrier(0);
                fma_b(t + 1, uB, iB, acc);
                apply(t + 1, acc);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
}

constexpr int kRsCols = 16;            // internal columns per group of the row-split back substitution
constexpr int kSplitSpinRs = 1 << 20;  // bound on the polls of one hand-off wait

// ---- back substitution split by rows (real shifts, U in 64x64 tiles) --------
// x <- U_s^-1 y for up to 16 internal columns of one shift on G workgroups.
// The 64-row blocks are dealt round-robin: block B belongs to workgroup
// B mod G, which keeps its rows' right-hand sides in registers as MFMA
// accumulators (wave v, slot j holds block G (v + 8 j) + w).  Going up from
// the last block, x_b is computed by the owner of block b from its rows once
// x_{b+1} is known, and handed to the other workgroups of the group through
// a slot in uncached device memory and a counter (device-scope atomics, a
// full memory wait before the count; the wait for it is bounded).  Every
// workgroup subtracts U[rows, block b] x_b from its own rows above block b,
// reading only its tiles of U (each a contiguous 64x64 piece), so U streams
// from HBM once per group.  The owner of block b-1 takes that block's rows
// out of the bulk: from LDS, 64 threads per 16-row tile run
// y_{b-1} -= U[b-1, b] x_b and x_{b-1} = Uinv_{b-1} y_{b-1} on the matrix
// cores right after x_b arrives (the two tiles were loaded into registers G
// blocks earlier), publish x_{b-1}, and then run their bulk.  LDS holds -x
// (the MFMA's B operand, so the accumulators subtract).
// The bulk's unit of work is a lane's row and four consecutive columns of a
// tile (16 or 32 bytes) feeding four MFMAs, one per column (their k index runs
// over the lanes' column quadruples); a ring of RG units keeps RG - 1 in
// flight, running on across blocks (U does not depend on x), so a block starts
// with its first loads done.
// (FP64 complex: 2, one slot only; the two planes' tiles, preloads and units
// fill the registers)
template <typename T, int NG, int UC> constexpr int rs_ring() {
    return UC ? (sizeof(T) == 4 ? 8 : 2) : (sizeof(T) == 4 ? 16 : (NG >= 2 ? 4 : 8));
}
// the most slots per thread of a shape: FP64 complex shifts one, else two
template <typename T> constexpr int rs_max_slots(int uc) { return (uc && sizeof(T) == 8) ? 1 : 2; }

template <typename T, int G, int NG, int UC>
__global__ void __launch_bounds__(512) k_utrsm_rs(int n, const T *__restrict__ Ubase, int64_t ustride,
                                                  const T *__restrict__ Uinv, int64_t uistride,
                                                  const UGroup *__restrict__ groups, int g0, const T *__restrict__ Yin,
                                                  T *__restrict__ Yf, int64_t ldy, const int *__restrict__ sact,
                                                  T *__restrict__ hand, int *__restrict__ cnt, int64_t hslot) {
    constexpr int NW = 8, C = kRsCols;
    constexpr int RG = rs_ring<T, NG, UC>();
    constexpr int PT = kT * kT / 512;  // tile elements per thread
    constexpr int UP = UC ? 2 : 1;     // planes of U and Uinv (complex shifts: real, imaginary)
    using M = Mfma<T>;
    using acc_t = typename M::acc_t;
    const int gl = blockIdx.x / G, w = blockIdx.x % G;
    const UGroup gr = groups[g0 + gl];
    if (gr.nq <= 0) return;
    if (sact && !sact[gr.shift]) return;   // a converged shift (every workgroup of the group)
    const int nbu = (n + kT - 1) / kT;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const T *Ut = Ubase + (int64_t)gr.shift * ustride;
    const T *Ui = Uinv + (int64_t)gr.shift * uistride;
    const int64_t up = ustride / 2;    // the imaginary plane of U (complex shifts)
    auto tile = [&](int I, int J) -> const T * {
        return Ut + ((int64_t)I * nbu - ((int64_t)I * (I - 1)) / 2 + (J - I)) * (kT * kT);
    };
    // element e of a tile: granule column e / (64 UV), row (e / UV) % 64
    constexpr int UV = utile_gw<T>();
    auto trow = [](int e) { return (e / UV) & 63; };
    auto tcol = [](int e) { return (e / (64 * UV)) * UV + (e % UV); };
    T *hx = hand + (int64_t)gl * hslot;
    int *hc = cnt + gl;
    __shared__ T Xs[2][kT][C];        // -x_b, -x_{b-1}
    __shared__ T Ys[kT][C + 1];       // the rows of the block whose x is computed here next
    __shared__ T Us[UP][kT][kT + 1];  // Us[p][m][r] = U(64 cb + r, 64 (cb + 1) + m), plane p
    __shared__ T Is[UP][kT][kT + 1];  // Is[p][k][i] = Uinv_cb(i, k), plane p
    // Complex shifts: internal columns 2c, 2c+1 are one complex column's real
    // and imaginary parts, so U x = Ur X + Ui X' with X' = (-x_im, x_re) per
    // column: the imaginary plane's products take their B operand from the
    // partner column (li ^ 1) with the sign of -X' (rot below).
    auto rot = [&](T v) -> T { return (li & 1) ? v : -v; };

    // the owned rows' right-hand sides, in the accumulator layout: tile rt of
    // slot j, accumulator v holds row 16 rt + M::row(lane, v) of the block, column li
    acc_t yv[NG][4];
#pragma unroll
    for (int j = 0; j < NG; j++) {
        const int B = G * (wv + NW * j) + w;
#pragma unroll
        for (int rt = 0; rt < 4; rt++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = kT * B + 16 * rt + M::row(lane, v);
                const T y = Yin[min(row, n - 1) + (int64_t)(gr.q0 + min(li, gr.nq - 1)) * ldy];
                yv[j][rt][v] = (B < nbu && row < n && li < gr.nq) ? y : (T)0;
            }
    }
    // the two tiles of the next block this workgroup solves, one block ahead
    // (columns past n, in the last block's tile, read as zero)
    T pu[UP][PT], pi[UP][PT];
    auto preload = [&](int cb) __attribute__((always_inline)) {
        if (cb + 1 < nbu) {
            const T *tp = tile(cb, cb + 1);
#pragma unroll
            for (int p = 0; p < UP; p++)
#pragma unroll
                for (int u = 0; u < PT; u++) pu[p][u] = tp[p * up + tid + 512 * u];   // masked when written to LDS
        } else {
#pragma unroll
            for (int p = 0; p < UP; p++)
#pragma unroll
                for (int u = 0; u < PT; u++) pu[p][u] = 0;   // the last block: no update tile
        }
#pragma unroll
        for (int p = 0; p < UP; p++)
#pragma unroll
            for (int u = 0; u < PT; u++) pi[p][u] = Ui[((int64_t)cb * UP + p) * kT * kT + tid + 512 * u];
    };
    {
        const int cb0 = nbu - 1 - ((nbu - 1 - w) % G + G) % G;   // the last block of this workgroup
        if (cb0 >= 0) preload(cb0);
    }
    // Unit u of block bb with NA slots active: column quadruple g16 = u / (4 NA)
    // (columns 16 g16 + 4 lk .. + 3), slot j = (u / 4) % NA, row tile rt = u % 4.
    T A[RG][4 * UP];
    auto load_unit = [&](auto nac, int bb, int u, T(&a)[4 * UP]) __attribute__((always_inline)) {
        constexpr int NA = decltype(nac)::value;
        const int g16 = u / (4 * NA), j = (u / 4) % NA, rt = u % 4;
        const int B = min(G * (wv + NW * j) + w, bb);   // below bb for an active slot
        const int row = 16 * rt + li, m0 = 16 * g16 + 4 * lk;   // this lane: row, columns m0 .. m0 + 3
        const T *tp = tile(B, bb);
        static_assert(UV % 4 == 0, "a lane's four columns lie in one granule");
#pragma unroll
        for (int p = 0; p < UP; p++) {
            const V4<T> v = ld4u(tp + p * up + ((m0 / UV) * kT + row) * UV + (m0 % UV));
#pragma unroll
            for (int q = 0; q < 4; q++) a[4 * p + q] = v[q];   // columns past n are masked where used
        }
    };
    // active slots of this wave in the bulk of block b: owned blocks below
    // b (below b - 1 where block b-1's x is computed here)
    auto nact = [&](int b) -> int {
        if (b >= nbu) return 0;
        const int bl = ((b - 1 - w) % G) == 0 ? b - 1 : b;
        const int first = bl - w - G * wv;
        return first > 0 ? min(NG, (first + G * NW - 1) / (G * NW)) : 0;
    };
    if (nbu >= 2) {   // the first bulk block's first units
        const int na0 = nact(nbu - 1);
        auto pro = [&](auto nac) __attribute__((always_inline)) {
#pragma unroll
            for (int u = 0; u < RG - 1; u++) load_unit(nac, nbu - 1, u, A[u]);
        };
        if (na0 >= 2 && NG >= 2) pro(std::integral_constant<int, NG >= 2 ? 2 : 1>{});
        else if (na0 >= 1) pro(std::integral_constant<int, 1>{});
    }
    bool failed = false;
    for (int b = nbu; b >= 1; b--) {
        const bool crit = ((b - 1 - w) % G) == 0;   // x_{b-1} is computed here
        const bool have = b < nbu;                  // x_b exists
        const bool mine = have && ((b - w) % G) == 0;
        const int bl = crit ? b - 1 : b;            // bulk rows: owned blocks below index bl
        const int na = nact(b);
        // na drops by at most one from block to block (owned blocks of a wave are 8 G apart)
        const int nax = na > 0 && b > 1 ? nact(b - 1) : 0;   // the next block's
        auto iter = [&](auto nac, auto nnc) __attribute__((always_inline)) {
            constexpr int NA = decltype(nac)::value, NAX = decltype(nnc)::value;
            constexpr int NU = 16 * NA;   // units of block b
            // 1. x_b into LDS (from its owner, unless computed here)
            if (have && !mine && (crit || w < bl)) {   // (a workgroup with no rows left skips it)
                const int need = 4 * (nbu - b);   // four publishing waves per block
                for (int it = 0; !failed && __hip_atomic_load(hc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need;
                     it++) {
                    if (it >= kSplitSpinRs) failed = true;
                    __builtin_amdgcn_s_sleep(1);
                }
#pragma unroll
                for (int u = 0; u < kT * C / 512; u++) {
                    const int e = tid + 512 * u;
                    const T v = __hip_atomic_load(hx + (int64_t)b * kT * C + e, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
                    Xs[b & 1][e / C][e % C] = failed ? __builtin_nan("") : -v;
                }
            }
            // 2. block b-1's rows and tiles into LDS
            if (crit) {
                const int lb = (b - 1 - w) / G;
                if (wv == lb % NW) {
                    const int os = lb / NW;
#pragma unroll
                    for (int rt = 0; rt < 4; rt++)
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            T y = 0;
#pragma unroll
                            for (int j = 0; j < NG; j++) y = j == os ? yv[j][rt][v] : y;
                            Ys[16 * rt + M::row(lane, v)][li] = y;
                        }
                }
#pragma unroll
                for (int p = 0; p < UP; p++)
#pragma unroll
                    for (int u = 0; u < PT; u++) {
                        const int e = tid + 512 * u;   // columns past n (the last block's tile) as zero
                        Us[p][tcol(e)][trow(e)] = kT * b + tcol(e) < n ? pu[p][u] : (T)0;
                        Is[p][e >> 6][e & 63] = pi[p][u];
                    }
            }
            lds_barrier();
            // 3. x_{b-1} (waves 0-3, one 16-row tile each), published
            if (crit) {
                const int rt = wv;
                if (have) {
                    if (wv < 4) {
                        acc_t acc;
#pragma unroll
                        for (int v = 0; v < 4; v++) acc[v] = Ys[16 * rt + M::row(lane, v)][li];
#pragma unroll
                        for (int ks = 0; ks < 16; ks++) {
                            acc = M::mac(Us[0][4 * ks + lk][16 * rt + li], Xs[b & 1][4 * ks + lk][li], acc);
                            if constexpr (UC)
                                acc = M::mac(Us[1][4 * ks + lk][16 * rt + li], rot(Xs[b & 1][4 * ks + lk][li ^ 1]),
                                             acc);
                        }
#pragma unroll
                        for (int v = 0; v < 4; v++) Ys[16 * rt + M::row(lane, v)][li] = acc[v];
                    }
                    lds_barrier();
                }
                if (wv < 4) {
                    acc_t acc = M::zero();
#pragma unroll
                    for (int ks = 0; ks < 16; ks++) {
                        acc = M::mac(Is[0][4 * ks + lk][16 * rt + li], Ys[4 * ks + lk][li], acc);
                        // x = Ir Y + Ii Y' with Y' = (-y_im, y_re) = rot of the partner column
                        if constexpr (UC)
                            acc = M::mac(Is[1][4 * ks + lk][16 * rt + li], rot(Ys[4 * ks + lk][li ^ 1]), acc);
                    }
                    T *dst = hx + (int64_t)(b - 1) * kT * C;
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int r = 16 * rt + M::row(lane, v), row = kT * (b - 1) + r;
                        const T x = row < n ? acc[v] : (T)0;
                        Xs[(b - 1) & 1][r][li] = -x;
                        __hip_atomic_store(dst + r * C + li, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (row < n && li < gr.nq) Yf[row + (int64_t)(gr.q0 + li) * ldy] = x;
                    }
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    __builtin_amdgcn_s_waitcnt(0);   // the values have landed
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    if (lane == 0) __hip_atomic_fetch_add(hc, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                if (b - 1 - G >= 0) preload(b - 1 - G);
            }
            // 4. the bulk: y -= U[rows, block b] x_b for the active slots, the
            //    loads running RG - 1 units ahead into block b-1 (its active
            //    slots are among this block's)
            if constexpr (NA > 0) {
                static_assert(NU % RG == 0 && (NAX == 0 || (16 * NAX) % RG == 0), "the ring turns whole blocks");
                T bx[4], bq[4];
#pragma unroll
                for (int u = 0; u < NU; u++) {
                    constexpr int L = RG - 1;
                    __builtin_amdgcn_sched_barrier(0);
                    if (u + L < NU) load_unit(nac, b, u + L, A[(u + L) % RG]);
                    else if constexpr (NAX > 0) load_unit(nnc, b - 1, u + L - NU, A[(u + L) % RG]);
                    __builtin_amdgcn_sched_barrier(0);
                    const int g16 = u / (4 * NA), j = (u / 4) % NA, rt = u % 4;
                    const int lim = n - (kT * b + 16 * g16 + 4 * lk);   // columns past n (the last block): zero
                    if (u % (4 * NA) == 0) {
#pragma unroll
                        for (int q = 0; q < 4; q++) {
                            bx[q] = Xs[b & 1][16 * g16 + 4 * lk + q][li];
                            if constexpr (UC) bq[q] = rot(Xs[b & 1][16 * g16 + 4 * lk + q][li ^ 1]);
                        }
                    }
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        yv[j][rt] = M::mac(q < lim ? A[u % RG][q] : (T)0, bx[q], yv[j][rt]);
                        if constexpr (UC) yv[j][rt] = M::mac(q < lim ? A[u % RG][4 + q] : (T)0, bq[q], yv[j][rt]);
                    }
                }
            }
        };
        using I0 = std::integral_constant<int, 0>;
        using I1 = std::integral_constant<int, 1>;
        using I2 = std::integral_constant<int, NG >= 2 ? 2 : 1>;
        if constexpr (NG >= 2) {
            if (na == 2) {
                if (nax == 2) iter(I2{}, I2{});
                else iter(I2{}, I1{});
            } else if (na == 1) {
                if (nax == 1) iter(I1{}, I1{});
                else iter(I1{}, I0{});
            } else {
                iter(I0{}, I0{});
            }
        } else {
            if (na == 1 && nax == 1) iter(I1{}, I1{});
            else if (na == 1) iter(I1{}, I0{});
            else iter(I0{}, I0{});
        }
    }
}

// Forward interchange chain of the Hessenberg solve (src:2806-2816) as a scan.
// Per column, with c the carried value (c = y[0] at the start), step k reads
// b = y[k+1] and writes y[k] = (s_k ? b : c), then c <- (s_k ? c : b) - l_k y[k];
// y[n-1] = c at the end.  Step k is the affine map c -> alpha c + beta with
// (alpha, beta) = (1, -l_k b) when s_k, else (-l_k, b).  One wave per chain:
// lane L takes steps [L bs, (L+1) bs), composes its block's map, the 64 maps
// are scanned across the wave (Hillis-Steele with lane shuffles), and each
// lane replays its block from its start value with the step formula above.
// Partial pivoting keeps |l_k| <= 1, so the composed maps stay bounded.  The
// column is updated in place: the only element one lane reads that another
// writes is the first of the next block, read once before any write.  Real
// shifts: one chain per internal column.  Complex shifts: columns (re, im)
// pair up into one complex chain.
template <typename T, int UC, int SY>
__global__ void __launch_bounds__(256) k_fwd_scan(int n, const T *__restrict__ lbase, int64_t lstride,
                                                  const uint8_t *__restrict__ swbase, int64_t sstride,
                                                  const UGroup *__restrict__ groups, T *__restrict__ Y, int64_t ldy,
                                                  const int *__restrict__ sact) {
#pragma clang fp contract(off)
    constexpr int PL = UC ? 2 : 1;               // internal columns per chain
    const UGroup g = groups[blockIdx.x];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (PL * wv >= g.nq) return;                 // whole wave: no chain here
    if (sact && !sact[g.shift]) return;          // a converged shift
    const int steps = n - 1;
    const int bs = (steps + 63) / 64;
    const T *lr = lbase + (int64_t)g.shift * lstride;
    const T *li = lr + lstride / 2;              // imaginary parts (UC)
    const uint8_t *sw = swbase + (int64_t)g.shift * sstride;
    T *yr = Y + (int64_t)(g.q0 + PL * wv) * ldy;
    T *yi = yr + ldy;                            // imaginary column (UC)
    const int k0 = lane * bs, k1 = min(k0 + bs, steps);   // this lane's steps [k0, k1)
    // SY (real chains): the wave's column staged in LDS (read coalesced, then
    // each lane walks its own block there; element k at k + k/64, so the lanes'
    // blocks fall in different banks) and written back coalesced at the end
    extern __shared__ __align__(16) unsigned char lds_dyn[];
    T *ys = reinterpret_cast<T *>(lds_dyn) + (int64_t)wv * (n + (n >> 6) + 1);
    auto yget = [&](int k) -> T {
        if constexpr (SY) return ys[k + (k >> 6)];
        else return yr[k];
    };
    auto yset = [&](int k, T v) {
        if constexpr (SY) ys[k + (k >> 6)] = v;
        else yr[k] = v;
    };
    if constexpr (SY) {
#pragma unroll 16
        for (int i = lane; i < n; i += 64) ys[i + (i >> 6)] = yr[i];
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
    }
    // 1. the block's composed map (A, B), and the next block's first element
    T Ar = 1, Ai = 0, Br = 0, Bi = 0;
#pragma unroll 4
    for (int k = k0; k < k1; k++) {
        const bool s = sw[k] != 0;
        const T l = lr[k], b = yget(k + 1);
        if constexpr (UC) {
            const T lm = li[k], bim = yi[k + 1];
            if (s) {   // A unchanged, B -= l b
                Br = Br - (l * b - lm * bim);
                Bi = Bi - (l * bim + lm * b);
            } else {   // (A, B) <- (-l A, b - l B)
                const T ar = -(l * Ar - lm * Ai), ai = -(l * Ai + lm * Ar);
                const T nr = b - (l * Br - lm * Bi), ni = bim - (l * Bi + lm * Br);
                Ar = ar; Ai = ai; Br = nr; Bi = ni;
            }
        } else {
            if (s) {
                Br = Br - l * b;
            } else {
                Ar = -l * Ar;
                Br = b - l * Br;
            }
        }
    }
    const int kb = min(k0 + bs, steps);          // y[kb]: read by this lane's last step
    const T bndr = yget(kb), bndi = UC ? yi[kb] : (T)0;
    // 2. inclusive scan of the maps in lane order: (A, B) <- (A A', A B' + B)
    //    with (A', B') the composition of the lanes before
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const T pAr = __shfl_up(Ar, o, 64), pBr = __shfl_up(Br, o, 64);
        const T pAi = UC ? __shfl_up(Ai, o, 64) : (T)0, pBi = UC ? __shfl_up(Bi, o, 64) : (T)0;
        if (lane >= o) {
            if constexpr (UC) {
                const T nAr = Ar * pAr - Ai * pAi, nAi = Ar * pAi + Ai * pAr;
                const T nBr = (Ar * pBr - Ai * pBi) + Br, nBi = (Ar * pBi + Ai * pBr) + Bi;
                Ar = nAr; Ai = nAi; Br = nBr; Bi = nBi;
            } else {
                Br = Ar * pBr + Br;
                Ar = Ar * pAr;
            }
        }
    }
    // the start value of this lane's block: the lanes before applied to y[0]
    const T c0r = yget(0), c0i = UC ? yi[0] : (T)0;
    T eAr = __shfl_up(Ar, 1, 64), eBr = __shfl_up(Br, 1, 64);
    T eAi = UC ? __shfl_up(Ai, 1, 64) : (T)0, eBi = UC ? __shfl_up(Bi, 1, 64) : (T)0;
    if (lane == 0) { eAr = 1; eAi = 0; eBr = 0; eBi = 0; }
    T cr, ci = 0;
    if constexpr (UC) {
        cr = (eAr * c0r - eAi * c0i) + eBr;
        ci = (eAr * c0i + eAi * c0r) + eBi;
    } else {
        cr = eAr * c0r + eBr;
    }
    __builtin_amdgcn_wave_barrier();   // every lane's reads above precede the writes below
    // 3. replay the block with the step formula, writing y[k]
    for (int k = k0; k < k1; k++) {
        const bool s = sw[k] != 0;
        const T l = lr[k];
        const bool last = k + 1 == kb;
        const T b = last ? bndr : yget(k + 1);
        if constexpr (UC) {
            const T lm = li[k];
            const T bim = last ? bndi : yi[k + 1];
            const T xr = s ? b : cr, xi = s ? bim : ci;
            yset(k, xr);
            yi[k] = xi;
            const T orr = s ? cr : b, oi = s ? ci : bim;
            cr = orr - (l * xr - lm * xi);
            ci = oi - (l * xi + lm * xr);
        } else {
            const T xk = s ? b : cr;
            yset(k, xk);
            cr = (s ? cr : b) - l * xk;
        }
    }
    // y[n-1]: the carried value after the last step (the lane whose block holds step n-2)
    if (k0 < steps && k1 == steps) {
        yset(n - 1, cr);
        if constexpr (UC) yi[n - 1] = ci;
    }
    if constexpr (SY) {
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
#pragma unroll 16
        for (int i = lane; i < n; i += 64) yr[i] = ys[i + (i >> 6)];
    }
}

// ---- the refinement residual: JX = J X (rocBLAS GEMM in one precision, the
// tiled MFMA GEMM below for FP32 J with FP64 x), then
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
                                int64_t plane, int planes, const double *__restrict__ colshift, int w_complex,
                                const int *__restrict__ cact) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int col = blockIdx.y;
    if (i >= n) return;
    if (cact && !cact[col]) return;
    const int64_t o = i + (int64_t)col * ld;
    if (colshift) {
        const TX sr = (TX)colshift[4 * col], si = (TX)colshift[4 * col + 1];
        const TX tr = (TX)colshift[4 * col + 2], ti = (TX)colshift[4 * col + 3];
        const TX xr = X[o], xi = planes > 1 ? X[o + plane] : (TX)0;
        TX jr = JX[i + (int64_t)(col * planes) * ldq];
        TX ji = planes > 1 ? JX[i + (int64_t)(col * planes + 1) * ldq] : (TX)0;
        if (w_complex) {   // a complex J: J x = (Jr xr - Ji xi) + i (Jr xi + Ji xr)
            const TX c = JXi[i + (int64_t)(col * 2) * ldq], d = JXi[i + (int64_t)(col * 2 + 1) * ldq];
            jr = jr - d;
            ji = ji + c;
        }
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
                           int planes, int ncol, double *__restrict__ norms, const int *__restrict__ cact) {
    const int c = blockIdx.x;
    if (cact && !cact[c]) return;   // whole workgroup: the column's norms stay as they were
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

// The solves' dense products on the library's MFMA kernels or on rocBLAS
// (tuning knob gemm_impl; rocBLAS by default, faster at full width).  Column
// masks need the library's kernels, unless the caller asks for the products
// over every column (full = 1: a round where most columns are still active).
inline bool own_products(const lhl_handle *h, const int *cact, int full) {
    return h->tune_gemm == 1 || (cact != nullptr && !full);
}


Status blas_handle(lhl_handle *h, rocblas_handle *out) {
    if (!h->blas) {
        rocblas_handle b = nullptr;
        if (rocblas_create_handle(&b) != rocblas_status_success) return {LHL_ERR_HIP, "rocblas_create_handle failed"};
        if (rocblas_set_stream(b, h->stream) != rocblas_status_success ||
            rocblas_set_pointer_mode(b, rocblas_pointer_mode_host) != rocblas_status_success) {
            rocblas_destroy_handle(b);
            return {LHL_ERR_HIP, "rocblas_set_stream failed"};
        }
        h->blas = b;
        h->blas_stream = h->stream;
    }
    if (h->blas_stream != (void *)h->stream) {   // follows the handle's current stream
        if (rocblas_set_stream((rocblas_handle)h->blas, h->stream) != rocblas_status_success)
            return {LHL_ERR_HIP, "rocblas_set_stream failed"};
        h->blas_stream = h->stream;
    }
    *out = (rocblas_handle)h->blas;
    return {};
}

// JX[:, q] = W X[:, q] for every plane through rocBLAS (same element type for
// W and X): one GEMM per plane, the output columns interleaved by plane.
template <typename T>
Status blas_jx(lhl_handle *h, const T *W, int64_t ldw, const T *X, int64_t ld, int64_t plane, int planes, int ncol,
               T *JX, int64_t ldq, const int *cact = nullptr) {
    const int n = (int)h->n;
    if (own_products(h, cact, 0)) {
        for (int pl = 0; pl < planes; pl++)
            LHL_TRY(mgemm(h, "gemm_jx", 0, n, ncol, n, W, ldw, X + pl * plane, ld, JX + pl * ldq, planes * ldq, cact,
                          1));
        return {};
    }
    rocblas_handle b;
    LHL_TRY(blas_handle(h, &b));
    const T one = 1, zero = 0;
    ProfScope ps(h, "gemm_jx");
    for (int pl = 0; pl < planes; pl++) {
        rocblas_status st;
        if constexpr (std::is_same<T, double>::value)
            st = rocblas_dgemm(b, rocblas_operation_none, rocblas_operation_none, n, ncol, n, &one, W, (int)ldw,
                               X + pl * plane, (int)ld, &zero, JX + pl * ldq, (int)(planes * ldq));
        else
            st = rocblas_sgemm(b, rocblas_operation_none, rocblas_operation_none, n, ncol, n, &one, W, (int)ldw,
                               X + pl * plane, (int)ld, &zero, JX + pl * ldq, (int)(planes * ldq));
        if (st != rocblas_status_success) return {LHL_ERR_HIP, "rocBLAS GEMM of the residual failed"};
    }
    return {};
}

template <typename T>
static rocblas_status gemm_nn(rocblas_handle b, int m, int ncol, int k, const T *A, int64_t lda, const T *B,
                              int64_t ldb, T beta, T *C, int64_t ldc) {
    const T one = 1;
    if constexpr (std::is_same<T, double>::value)
        return rocblas_dgemm(b, rocblas_operation_none, rocblas_operation_none, m, ncol, k, &one, A, (int)lda, B,
                             (int)ldb, &beta, C, (int)ldc);
    else
        return rocblas_sgemm(b, rocblas_operation_none, rocblas_operation_none, m, ncol, k, &one, A, (int)lda, B,
                             (int)ldb, &beta, C, (int)ldc);
}

// C = L B + beta C for the m x m lower-triangular L in dense storage, split
// `depth` times in halves: each diagonal half recursively, the block below
// the first half as one GEMM into the second half of C.
template <typename T>
static rocblas_status lower_mul(rocblas_handle b, int depth, int m, int ncol, const T *L, int64_t ldl, const T *B,
                                int64_t ldb, T beta, T *C, int64_t ldc) {
    if (depth <= 0 || m < 512) return gemm_nn<T>(b, m, ncol, m, L, ldl, B, ldb, beta, C, ldc);
    const int m1 = (m / 2 + 63) / 64 * 64, m2 = m - m1;
    rocblas_status st = lower_mul<T>(b, depth - 1, m1, ncol, L, ldl, B, ldb, beta, C, ldc);
    if (st == rocblas_status_success) st = gemm_nn<T>(b, m2, ncol, m1, L + m1, ldl, B, ldb, beta, C + m1, ldc);
    if (st == rocblas_status_success)
        st = lower_mul<T>(b, depth - 1, m2, ncol, L + m1 + (int64_t)m1 * ldl, ldl, B + m1, ldb, (T)1, C + m1, ldc);
    return st;
}

// C = L B for the unit lower-triangular L stored densely (zeros above the
// diagonal, ones on it), as dense GEMMs: rocBLAS's GEMM runs faster on this
// chip than its triangular product.  Two levels of halving skip 3/8 of the
// dense product's flops (the blocks above the diagonal).  Measured per FP32
// step at n = 4096, 64 x 16 (forward + backward sweeps): one GEMM 2.02 ms,
// one level 1.57 ms, two 1.57 ms, three 2.03 ms (512-row blocks run
// slowly); FP64 2.04, 1.82, 1.72, 2.18 ms.
template <typename T>
Status blas_gemm_lower(lhl_handle *h, const char *name, int m, int ncol, const T *L, int64_t ldl, const T *B,
                       int64_t ldb, T *C, int64_t ldc) {
    if (m <= 0 || ncol <= 0) return {};
    rocblas_handle b;
    LHL_TRY(blas_handle(h, &b));
    ProfScope ps(h, name);
    const int depth = h->tune_lsweep >= 0 ? h->tune_lsweep : 2;
    if (lower_mul<T>(b, depth, m, ncol, L, ldl, B, ldb, (T)0, C, ldc) != rocblas_status_success)
        return {LHL_ERR_HIP, "rocBLAS GEMM of the solve failed"};
    return {};
}

// B <- L^-1 B for the unit lower-triangular L (dense, m x m) by substitution
// (rocBLAS TRSM: forward substitution over blocks, backward stable, unlike
// the product with an explicit inverse, whose error grows with cond(L)).
template <typename T>
static rocblas_status trsm_unit_lower(rocblas_handle b, int m, int ncol, const T *L, int64_t ldl, T *B, int64_t ldb) {
    const T one = 1;
    if constexpr (std::is_same<T, double>::value)
        return rocblas_dtrsm(b, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_unit,
                             m, ncol, &one, L, (int)ldl, B, (int)ldb);
    else
        return rocblas_strsm(b, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_unit,
                             m, ncol, &one, L, (int)ldl, B, (int)ldb);
}

// Dense unit lower-triangular m x m matrix in dst: dst(i, c) = src(i, c) below
// the diagonal, 1 on it, 0 above.  src == dst fixes the diagonal and upper part
// in place.
template <typename T>
__global__ void k_unit_lower(int m, const T *__restrict__ src, int64_t lds, T *__restrict__ dst, int64_t ldd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
    if (i >= m) return;
    if (i > c) {
        if (src != dst) dst[i + (int64_t)c * ldd] = src[i + (int64_t)c * lds];
    } else {
        dst[i + (int64_t)c * ldd] = i == c ? (T)1 : (T)0;
    }
}

// I + L' (L'(i, m) = A(i+1, m)) as a dense unit lower matrix, the rocBLAS
// backward sweep's operand, built once per reduction on first use.
template <typename T>
Status lp1_typed(lhl_handle *h) {
    const int n = (int)h->n;
    if (h->lp1_ready || n < 2) return {};
    LHL_TRY(h->Lp1.ensure(sizeof(T) * (size_t)h->ldlf * (size_t)std::max(n - 1, 1)));
    const dim3 gl((unsigned)((n - 1 + 255) / 256), (unsigned)(n - 1));
    LHL_LAUNCH(h, "reduce_finish", k_unit_lower<T>, gl, dim3(256), 0, n - 1, (const T *)h->A.as<T>() + 1, h->lda,
               h->Lp1.as<T>(), h->ldlf);
    LHL_HIP_TRY(hipGetLastError());
    h->lp1_ready = 1;
    return {};
}

// U solve groups: kNC internal columns of one shift each.  Workgroups are
// dealt to the 8 XCDs round-robin (blockIdx % 8), so all groups of a shift go
// to slots of one residue class and share that XCD's L2 copy of U; empty
// slots (nq = 0) pad the classes to equal length.
std::vector<UGroup> build_groups(int64_t nshift, const int64_t *cb, int planes) {
    constexpr int kXcd = 8;
    std::vector<std::vector<UGroup>> per(kXcd);
    for (int64_t s = 0; s < nshift; s++) {
        for (int64_t q = cb[s] * planes; q < cb[s + 1] * planes; q += kNC)
            per[s % kXcd].push_back(UGroup{(int)s, (int)q, (int)std::min<int64_t>(kNC, cb[s + 1] * planes - q), 0});
    }
    size_t depth = 0;
    for (const auto &v : per) depth = std::max(depth, v.size());
    std::vector<UGroup> ug(depth * kXcd, UGroup{0, 0, 0, 0});
    for (int x = 0; x < kXcd; x++)
        for (size_t j = 0; j < per[x].size(); j++) ug[j * kXcd + x] = per[x][j];
    return ug;
}

// Groups of the row-split back substitution: up to kRsCols internal columns
// of one shift each, in shift order (its workgroups read disjoint rows of U,
// so no XCD placement).
std::vector<UGroup> build_wgroups(int64_t nshift, const int64_t *cb, int planes) {
    std::vector<UGroup> wg;
    for (int64_t s = 0; s < nshift; s++)
        for (int64_t q = cb[s] * planes; q < cb[s + 1] * planes; q += kRsCols)
            wg.push_back(UGroup{(int)s, (int)q, (int)std::min<int64_t>(kRsCols, cb[s + 1] * planes - q), 0});
    return wg;
}

__global__ void k_rs_reset(int *__restrict__ cnt, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cnt[i] = 0;
}

// Workgroups of k_utrsm_rs<T, G, NG> resident on the whole device at once
// (0 if the query fails, which leaves the row split off).
template <typename T, int G, int NG, int UC> int rs_capacity() {
    static const int cap = [] {
        int nb = 0, dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_utrsm_rs<T, G, NG, UC>, 512, 0) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
            (void)hipGetLastError();
            return 0;
        }
        return nb * cus;
    }();
    return cap;
}

// Whether some G <= gmax in {8, 4, 2} gives every workgroup two blocks or more
// and at most maxng slots per thread (before any capacity check).
static bool rs_shape_ok(int nbu, int gmax, int maxng) {
    for (int gc : {8, 4, 2})
        if (gc <= gmax && nbu >= 2 * gc && (nbu + 8 * gc - 1) / (8 * gc) <= maxng) return true;
    return false;
}

// f(G, NG) with the pair as integral constants (G in {2, 4, 8}, NG in {1, 2})
template <typename T, typename F> Status with_rs_shape(int G, int NG, F &&f) {
    using std::integral_constant;
    auto ng = [&](auto g) -> Status {
        if (NG == 1) return f(g, integral_constant<int, 1>{});
        return f(g, integral_constant<int, 2>{});
    };
    if (G == 2) return ng(integral_constant<int, 2>{});
    if (G == 4) return ng(integral_constant<int, 4>{});
    return ng(integral_constant<int, 8>{});
}

// The row-split back substitution of real shifts whose U is in tiles
// (n <= 8192, 256 rows and up): G workgroups per group of 16 columns, G as large as the knob allows
// with at least two blocks per workgroup, at most two slots per thread, and
// the whole launch resident at once (else fewer groups per launch).  Sets
// *used to false, launching nothing, where it does not apply.
template <typename T>
Status utrsm_rs(lhl_handle *h, const SolveArgs &a, int n, const UGroup *wg, int64_t nwide, const T *Yin, T *Yf,
                int64_t ldy, int64_t uis, bool *used) {
    *used = false;
    const int gmax = h->tune_utrsm_split > 0 ? h->tune_utrsm_split : 8;
    const int nbu = (n + kT - 1) / kT;
    if (gmax < 2 || nwide <= 0 || !a.ut || n > 8192) return {};
    const int uc = a.shift_complex ? 1 : 0;
    int G = 0, NG = 0, cap = 0;
    for (int gc : {8, 4, 2}) {
        if (gc > gmax || nbu < 2 * gc) continue;
        const int ng = (nbu + 8 * gc - 1) / (8 * gc);   // slots per thread
        if (ng > rs_max_slots<T>(uc)) continue;
        int cc = 0;
        LHL_TRY(with_rs_shape<T>(gc, ng, [&](auto g, auto s) -> Status {
            constexpr int GC = decltype(g)::value, NGC = decltype(s)::value;
            if constexpr (sizeof(T) == 4 || NGC == 1)
                cc = uc ? rs_capacity<T, GC, NGC, 1>() : rs_capacity<T, GC, NGC, 0>();
            else
                cc = rs_capacity<T, GC, NGC, 0>();
            return {};
        }));
        if (cc < gc) continue;
        const bool fits = nwide * gc <= cc;
        if (G == 0 || fits) {
            G = gc;
            NG = ng;
            cap = cc;
        }
        if (fits) break;
    }
    if (G == 0) return {};
    const int64_t chunk = std::min<int64_t>(nwide, cap / G);
    const int64_t hslot = (int64_t)nbu * kT * kRsCols;
    const size_t hb = ((sizeof(T) * (size_t)(chunk * hslot) + 255) / 256) * 256;
    if (!uc_ensure(h, hb + sizeof(int) * (size_t)chunk).ok()) {
        (void)hipGetLastError();
        return {};
    }
    T *hand = static_cast<T *>(h->uc.back().first);
    int *cnt = reinterpret_cast<int *>(static_cast<char *>(h->uc.back().first) + hb);
    for (int64_t g0 = 0; g0 < nwide; g0 += chunk) {
        const int cg = (int)std::min<int64_t>(chunk, nwide - g0);
        LHL_LAUNCH(h, "utrsm_reset", k_rs_reset, dim3((unsigned)((cg + 255) / 256)), dim3(256), 0, cnt, cg);
        LHL_TRY(with_rs_shape<T>(G, NG, [&](auto g, auto s) -> Status {
            constexpr int GC = decltype(g)::value, NGC = decltype(s)::value;
            if constexpr (sizeof(T) == 4 || NGC == 1) {
                if (uc) {
                    LHL_LAUNCH(h, "utrsm", (k_utrsm_rs<T, GC, NGC, 1>), dim3((unsigned)(cg * GC)), dim3(512), 0, n,
                               (const T *)a.U, a.ustride, (const T *)a.Uinv, uis, wg, (int)g0, Yin, Yf, ldy, a.sact,
                               hand, cnt, hslot);
                    return {};
                }
            }
            LHL_LAUNCH(h, "utrsm", (k_utrsm_rs<T, GC, NGC, 0>), dim3((unsigned)(cg * GC)), dim3(512), 0, n,
                       (const T *)a.U, a.ustride, (const T *)a.Uinv, uis, wg, (int)g0, Yin, Yf, ldy, a.sact, hand, cnt,
                       hslot);
            return {};
        }));
    }
    *used = true;
    return {};
}

template <typename T, typename TX>
Status solve_batch_typed(lhl_handle *h, const SolveArgs &a) {
    const int n = (int)h->n;
    const int planes = a.x_complex ? 2 : 1;
    const int Q = (int)(a.ncol * planes);
    const int64_t ldy = ldy_for(n);
    DevBuf &yb = a.ybuf ? *a.ybuf : h->ybuf, &rb = a.rbuf ? *a.rbuf : h->rbuf;
    LHL_TRY(yb.ensure(sizeof(T) * (size_t)ldy * Q));
    LHL_TRY(rb.ensure(sizeof(T) * (size_t)ldy * Q));
    T *Yw = yb.as<T>(), *Yo = rb.as<T>();
    ProfWork pw(h, a.nshift, Q);
    const T *A = h->A.as<T>();
    if (a.phase != 2) {
        // y = P^-1 D^-1 x, then y' <- L'^-1 y' as a product with the explicit
        // inverse formed by lhl() (final rows in Yo)
        const size_t xlds = sizeof(TX) * (size_t)n;
        if (xlds <= lds_limit() / 2) {   // a column of x in LDS
            LHL_TRY(allow_dynamic_lds(k_gather_col<T, TX>));
            LHL_LAUNCH(h, "gather", (k_gather_col<T, TX>), dim3((unsigned)Q), dim3(256), xlds, n,
                       (const int *)h->perm.as<int>(), (const double *)h->iscale64.as<double>(), (const TX *)a.xin,
                       a.ldx, a.xplane, planes, Yw, Yo, ldy, a.cact);
        } else {
            const dim3 gq((unsigned)((n + 255) / 256), (unsigned)Q);
            LHL_LAUNCH(h, "gather", (k_gather<T, TX>), gq, dim3(256), 0, n, (const int *)h->perm.as<int>(),
                       (const double *)h->iscale64.as<double>(), (const TX *)a.xin, a.ldx, a.xplane, planes, Yw, Yo,
                       ldy, a.cact);
        }
        if (h->cplx)   // a complex L: the blocked sweep (every column; the gather skipped none of Yo's rows)
            LHL_TRY(zsweep_device<T>(h, 1, Yw, Yo, ldy, a.ncol));
        else if (own_products(h, a.cact, a.full_products))
            LHL_TRY(mgemm(h, "trsm", 1, n - 1, Q, n - 1, (const T *)h->Lfull.as<T>(), h->ldlf, (const T *)Yw + 1, ldy,
                          Yo + 1, ldy, a.cact, planes));
        else if (a.nshift == 1 && !a.col_begin && !a.groups && n >= 2) {
            // one shift's solve (the drop-in lhl_ldiv! and lhl_refine!
            // corrections): forward substitution with I + L', as the reference
            // sweeps (src:2148-2195); batches keep the product with the
            // explicit inverse, whose extra error their refinement absorbs
            LHL_TRY(lp1_typed<T>(h));
            LHL_HIP_TRY(hipMemcpy2DAsync(Yo + 1, sizeof(T) * (size_t)ldy, (const T *)Yw + 1, sizeof(T) * (size_t)ldy,
                                         sizeof(T) * (size_t)(n - 1), (size_t)Q, hipMemcpyDeviceToDevice, h->stream));
            rocblas_handle b;
            LHL_TRY(blas_handle(h, &b));
            ProfScope ps(h, "trsm");
            if (trsm_unit_lower<T>(b, n - 1, Q, (const T *)h->Lp1.as<T>(), h->ldlf, Yo + 1, ldy) !=
                rocblas_status_success)
                return {LHL_ERR_HIP, "rocBLAS TRSM of the solve failed"};
        } else
            LHL_TRY(blas_gemm_lower<T>(h, "trsm", n - 1, Q, (const T *)h->Lfull.as<T>(), h->ldlf,
                                       (const T *)Yw + 1, ldy, Yo + 1, ldy));
        if (a.phase == 1) {
            LHL_HIP_TRY(hipGetLastError());
            return {};
        }
    }
    // the back substitution's column groups (within one shift)
    const UGroup *dg;
    size_t ngroups;
    int64_t nwide = 0;
    if (a.groups) {   // uploaded once by the caller (a plan)
        dg = (const UGroup *)a.groups;
        ngroups = (size_t)a.ngroups;
        nwide = a.nwide;
    } else {
        std::vector<int64_t> cb((size_t)a.nshift + 1);
        if (a.col_begin) {
            for (int64_t s = 0; s <= a.nshift; s++) cb[s] = a.col_begin[s];
        } else {
            cb[0] = 0;
            cb[1] = a.ncol;
        }
        std::vector<UGroup> ug = build_groups(a.nshift, cb.data(), planes);
        ngroups = ug.size();
        if (a.ut) {
            const std::vector<UGroup> wgv = build_wgroups(a.nshift, cb.data(), planes);
            nwide = (int64_t)wgv.size();
            ug.insert(ug.end(), wgv.begin(), wgv.end());
        }
        LHL_TRY(h->gbuf.ensure(sizeof(UGroup) * std::max<size_t>(ug.size(), 1)));
        // ordered on the handle's stream: an earlier solve's back substitution
        // may still be reading the table
        const size_t ugb = sizeof(UGroup) * ug.size();
        unsigned char *hug = host_keep(h, ugb);
        std::memcpy(hug, ug.data(), ugb);
        LHL_HIP_TRY(hipMemcpyAsync(h->gbuf.p, hug, ugb, hipMemcpyHostToDevice, h->stream));
        dg = h->gbuf.as<UGroup>();
    }

    const int N1 = n - 1;
    // forward interchange chain in place on Yo: a scan per column
    if (n >= 2) {
        // real chains: each wave's column staged in LDS where four fit
        const size_t ylds = 4 * sizeof(T) * (size_t)(n + (n >> 6) + 1);
        if (a.shift_complex) {
            LHL_LAUNCH(h, "fwd_chain", (k_fwd_scan<T, 1, 0>), dim3((unsigned)ngroups), dim3(256), 0, n,
                       (const T *)a.l, a.lstride, a.swp, a.sstride, (const UGroup *)dg, Yo, ldy, a.sact);
        } else if (ylds <= lds_limit()) {
            LHL_TRY(allow_dynamic_lds(k_fwd_scan<T, 0, 1>));
            LHL_LAUNCH(h, "fwd_chain", (k_fwd_scan<T, 0, 1>), dim3((unsigned)ngroups), dim3(256), ylds, n,
                       (const T *)a.l, a.lstride, a.swp, a.sstride, (const UGroup *)dg, Yo, ldy, a.sact);
        } else {
            LHL_LAUNCH(h, "fwd_chain", (k_fwd_scan<T, 0, 0>), dim3((unsigned)ngroups), dim3(256), 0, n,
                       (const T *)a.l, a.lstride, a.swp, a.sstride, (const UGroup *)dg, Yo, ldy, a.sact);
        }
        LHL_HIP_TRY(hipGetLastError());
    }
    // x <- U^-1 y per shift (solution in Yw)
    const int nbu = (n + kT - 1) / kT;
    const int64_t uis = a.uistride ? a.uistride : (int64_t)nbu * kT * kT * (a.shift_complex ? 2 : 1);
    const dim3 gu((unsigned)ngroups);
    // real shifts: rows split over several workgroups per 16 columns where it applies
    bool rs_used = false;
    LHL_TRY(utrsm_rs<T>(h, a, n, dg + ngroups, nwide, (const T *)Yo, Yw, ldy, uis, &rs_used));
    if (rs_used) {
        // launched
    } else if (a.ut) {
        return {LHL_ERR_HIP, "solve: U is in tiles but the row-split U solve could not run (hand-off memory)"};
    } else if (a.UT && n <= 512 * 16) {
        // U in packed columns, right-hand sides in registers (512 threads:
        // measured faster than 1024 on this shape)
        constexpr int NTR = 512;
        auto go = [&](auto uc, auto np) -> Status {
            constexpr int UC = decltype(uc)::value, NP = decltype(np)::value;
            constexpr int NTK = UC ? 512 : NTR;
            LHL_LAUNCH(h, "utrsm", (k_utrsm_t<T, UC, NP, NTK>), gu, dim3(NTK), 0, n, (const T *)a.UT, a.ustride,
                       (const T *)a.Uinv, uis, (const UGroup *)dg, (const T *)Yo, Yw, ldy, a.sact);
            return {};
        };
        using F = std::integral_constant<int, 0>;
        using C = std::integral_constant<int, 1>;
        Status st;
        // groups of V rows per thread
        constexpr int VR = ut_vec<T>();
        int nper = (n + VR * 512 - 1) / (VR * 512);
        if (a.shift_complex) {
            if (nper <= 1) st = go(C{}, std::integral_constant<int, 1>{});
            else if (nper <= 2) st = go(C{}, std::integral_constant<int, 2>{});
            else if (nper <= 4) st = go(C{}, std::integral_constant<int, 4>{});
            else if (nper <= 8) st = go(C{}, std::integral_constant<int, 8>{});
            else st = go(C{}, std::integral_constant<int, 16>{});
        } else {
            nper = (n + VR * NTR - 1) / (VR * NTR);
            if (nper <= 1) st = go(F{}, std::integral_constant<int, 1>{});
            else if (nper <= 2) st = go(F{}, std::integral_constant<int, 2>{});
            else if (nper <= 4) st = go(F{}, std::integral_constant<int, 4>{});
            else if (nper <= 8) st = go(F{}, std::integral_constant<int, 8>{});
            else st = go(F{}, std::integral_constant<int, 16>{});
        }
        LHL_TRY(st);
    } else if (a.shift_complex) {
        LHL_LAUNCH(h, "utrsm", (k_utrsm4<T, 1>), gu, dim3(64 * (64 / utrsm4_rows<T>())), 0, n, (const T *)a.U,
                   a.ustride, (const T *)a.Uinv, uis, (const UGroup *)dg, (const T *)Yo, Yw, ldy, a.sact);
    } else {
        LHL_LAUNCH(h, "utrsm", (k_utrsm4<T, 0>), gu, dim3(64 * (64 / utrsm4_rows<T>())), 0, n, (const T *)a.U,
                   a.ustride, (const T *)a.Uinv, uis, (const UGroup *)dg, (const T *)Yo, Yw, ldy, a.sact);
    }
    LHL_HIP_TRY(hipGetLastError());
    // y_new' = (I + L') y' (Yo is free again), then x = D P y_new
    if (h->cplx) {
        LHL_TRY(zsweep_device<T>(h, 0, Yw, Yo, ldy, a.ncol));
    } else if (own_products(h, a.cact, a.full_products)) {
        // (I + L') y' straight from the factors' strict lower triangle
        LHL_TRY(mgemm(h, "trmm", 2, N1, Q, N1, (const T *)h->A.as<T>() + 1, h->lda, (const T *)Yw + 1, ldy, Yo + 1,
                      ldy, a.cact, planes));
    } else {
        LHL_TRY(lp1_typed<T>(h));
        LHL_TRY(blas_gemm_lower<T>(h, "trmm", N1, Q, (const T *)h->Lp1.as<T>(), h->ldlf, (const T *)Yw + 1, ldy,
                                   Yo + 1, ldy));
    }
    LHL_LAUNCH(h, "scatter", (k_scatter<T, TX>), dim3((unsigned)((n + 255) / 256), (unsigned)Q), dim3(256), 0, n,
               (const T *)Yo, (const T *)Yw, ldy, planes, (const int *)h->iperm.as<int>(),
               (const double *)h->scale64.as<double>(), (TX *)a.xout, a.ldx, a.xplane, a.accumulate, a.cact);
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

template <typename T>
Status uinv_typed(lhl_handle *h, const void *U, int64_t ustride, const void *rdiag, int64_t lstride, int64_t nshift,
                  int cplx, DevBuf &Uinv, int64_t *uistride, DevBuf &UT, int ut) {
    const int n = (int)h->n;
    const int nbu = (n + kT - 1) / kT;
    const int64_t uis = (int64_t)nbu * kT * kT * (cplx ? 2 : 1);
    LHL_TRY(Uinv.ensure(sizeof(T) * (size_t)uis * (size_t)nshift));
    ProfWork pw(h, nshift, 0);
    *uistride = uis;
    if (cplx)
        LHL_LAUNCH(h, "uinv", (k_uinv<T, 1>), dim3((unsigned)nbu, (unsigned)nshift), dim3(64), 0, n, (const T *)U,
                   ustride, (const T *)rdiag, lstride, Uinv.as<T>(), uis, ut);
    else
        LHL_LAUNCH(h, "uinv", (k_uinv<T, 0>), dim3((unsigned)nbu, (unsigned)nshift), dim3(64), 0, n, (const T *)U,
                   ustride, (const T *)rdiag, lstride, Uinv.as<T>(), uis, ut);
    // U in packed columns for the other back substitutions (n <= 8192; U in
    // tiles is read as it is)
    if (!ut && n <= 512 * 16) {
        LHL_TRY(UT.ensure(sizeof(T) * (size_t)ustride * (size_t)nshift));
        const int planes = cplx ? 2 : 1;
        LHL_LAUNCH(h, "ut_pack", k_ut_pack<T>, dim3((unsigned)(nbu * (nbu + 1) / 2), (unsigned)(nshift * planes)),
                   dim3(256), 0, n, (const T *)U, UT.as<T>(), ustride, planes);
    }
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

template <typename TW, typename TX>
Status residual_typed(lhl_handle *h, const ResidArgs &a) {
    const int n = (int)h->n;
    const int planes = a.x_complex ? 2 : 1;
    const int Q = (int)(a.ncol * planes);
    const int64_t ldq = ldy_for(n);
    ProfWork pw(h, 0, Q);
    DevBuf &jb = a.jxbuf ? *a.jxbuf : h->ybuf, &jib = a.jxibuf ? *a.jxibuf : h->rbuf;
    LHL_TRY(jb.ensure(sizeof(TX) * (size_t)ldq * Q));
    TX *JX = jb.as<TX>();
    TX *JXi = nullptr;
    const dim3 gg((unsigned)((n + kT - 1) / kT), (unsigned)((Q + kT - 1) / kT));
    // a complex resident J: JX = Jr X and JXi = Ji X (both planes), combined below
    const bool jc = !a.W && h->cplx;
    if constexpr (std::is_same<TW, TX>::value) {
        // plain GEMM in one precision: the library's tuned kernels
        const TW *W = a.W ? (const TW *)a.W : (const TW *)h->J.p;
        const int64_t ldw = a.W ? a.ldw : h->ldj;
        LHL_TRY(blas_jx<TX>(h, W, ldw, (const TX *)a.x, a.ld, a.plane, planes, (int)a.ncol, JX, ldq,
                            a.full_products ? nullptr : a.cact));
        if ((a.W && a.w_complex) || jc) {
            LHL_TRY(jib.ensure(sizeof(TX) * (size_t)ldq * Q));
            JXi = jib.as<TX>();
            LHL_TRY(blas_jx<TX>(h, W + ldw * (int64_t)n, ldw, (const TX *)a.x, a.ld, a.plane, planes, (int)a.ncol,
                                JXi, ldq));
        }
    } else if (a.W) {
        LHL_LAUNCH(h, "gemm_jx", (k_gemm_jx<TW, TX>), gg, dim3(256), 0, n, (const TW *)a.W, a.ldw, (const TX *)a.x,
                   a.ld, a.plane, planes, Q, JX, ldq);
        if (a.w_complex) {
            LHL_TRY(jib.ensure(sizeof(TX) * (size_t)ldq * Q));
            JXi = jib.as<TX>();
            const TW *Wi = (const TW *)a.W + a.ldw * (int64_t)n;
            LHL_LAUNCH(h, "gemm_jx", (k_gemm_jx<TW, TX>), gg, dim3(256), 0, n, Wi, a.ldw, (const TX *)a.x, a.ld,
                       a.plane, planes, Q, JXi, ldq);
        }
    } else {
        LHL_LAUNCH(h, "gemm_jx", (k_gemm_jx<TW, TX>), gg, dim3(256), 0, n, (const TW *)h->J.p, h->ldj,
                   (const TX *)a.x, a.ld, a.plane, planes, Q, JX, ldq);
        if (jc) {
            LHL_TRY(jib.ensure(sizeof(TX) * (size_t)ldq * Q));
            JXi = jib.as<TX>();
            LHL_LAUNCH(h, "gemm_jx", (k_gemm_jx<TW, TX>), gg, dim3(256), 0, n,
                       (const TW *)h->J.p + h->ldj * (int64_t)n, h->ldj, (const TX *)a.x, a.ld, a.plane, planes, Q,
                       JXi, ldq);
        }
    }
    const dim3 gc((unsigned)((n + 255) / 256), (unsigned)a.ncol);
    LHL_LAUNCH(h, "resid_combine", k_resid_combine<TX>, gc, dim3(256), 0, n, (const TX *)JX, ldq, (const TX *)JXi,
               (const TX *)a.x, (const TX *)a.b, (TX *)a.r, a.ld, a.plane, planes, a.W ? nullptr : a.colshift,
               a.W ? a.w_complex : (jc ? 1 : 0), a.cact);
    LHL_LAUNCH(h, "colnorms", k_colnorms<TX>, dim3((unsigned)a.ncol), dim3(256), 0, n, (const TX *)a.r,
               (const TX *)a.b, a.ld, a.plane, planes, (int)a.ncol, a.norms, a.cact);
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
    // the explicit inverse of L' used by every solve (lower part; the rows
    // above each column's diagonal block are never read)
    const int64_t ldf = std::max<int64_t>(64, ((int64_t)(n - 1) + 63) / 64 * 64);
    LHL_TRY(h->Lfull.ensure(sizeof(T) * (size_t)ldf * (size_t)std::max(n - 1, 1)));
    h->ldlf = ldf;
    if (nbt > 0)
        LHL_LAUNCH(h, "reduce_finish", k_linv_full<T>, dim3((unsigned)((n - 1 + 15) / 16)), dim3(512), 0, n,
                   (const T *)h->A.as<T>(), h->lda, (const T *)h->Linv.as<T>(), h->Lfull.as<T>(), ldf);
    // the forward sweep as one product: L'^-1 with its diagonal and upper part
    // set (the backward sweep's I + L' is read from the factors, or built as a
    // dense copy for rocBLAS on first use)
    h->lp1_ready = 0;
    if (n >= 2) {
        const dim3 gl((unsigned)((n - 1 + 255) / 256), (unsigned)(n - 1));
        LHL_LAUNCH(h, "reduce_finish", k_unit_lower<T>, gl, dim3(256), 0, n - 1, (const T *)h->Lfull.as<T>(), ldf,
                   h->Lfull.as<T>(), ldf);
    }
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

}  // namespace

namespace {

__global__ void k_fill_i32(int *__restrict__ p, int64_t n, int v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// The refinement's convergence decision on the device, after round `it`'s
// residual: a shift still active whose columns all satisfy
// ||r_c|| <= tol ||b_c|| (norms[c], norms[ncol + c]; NaN fails) is retired
// (sact[s] = 0 and its columns' cact = 0, so later rounds skip it); a shift
// that fails is refined once more when it < max_iter (iters[s] = it + 1) and
// otherwise stays flagged as not converged.  One thread per shift.
__global__ void k_conv(int it, int max_iter, double tol, int64_t nshift, const int64_t *__restrict__ cb,
                       const double *__restrict__ norms, int64_t ncol, int *__restrict__ sact, int *__restrict__ cact,
                       int *__restrict__ iters) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshift || !sact[s]) return;
    bool bad = false;
    for (int64_t c = cb[s]; c < cb[s + 1]; c++)
        if (!(norms[c] <= tol * norms[ncol + c])) bad = true;
    if (!bad) {
        sact[s] = 0;
        for (int64_t c = cb[s]; c < cb[s + 1]; c++) cact[c] = 0;
    } else if (it < max_iter) {
        iters[s] = it + 1;
    }
}

}  // namespace

Status fill_i32_device(lhl_handle *h, int *p, int64_t n, int v) {
    if (n <= 0) return {};
    LHL_LAUNCH(h, "conv", k_fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p, n, v);
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

Status conv_device(lhl_handle *h, int it, int max_iter, double tol, int64_t nshift, const int64_t *cb,
                   const double *norms, int64_t ncol, int *sact, int *cact, int *iters) {
    if (nshift <= 0) return {};
    LHL_LAUNCH(h, "conv", k_conv, dim3((unsigned)((nshift + 255) / 256)), dim3(256), 0, it, max_iter, tol, nshift, cb,
               norms, ncol, sact, cact, iters);
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

bool pool_tiles(const lhl_handle *h, int64_t n, int is_complex) {
    // complex shifts on a real H: tiles in two planes (FP64: where one slot
    // per thread reaches every block, n <= 4096)
    if (h->cplx || n > 8192) return false;
    const int gmax = h->tune_utrsm_split > 0 ? h->tune_utrsm_split : 8;
    const int maxng = h->dtype == LHL_F64 ? rs_max_slots<double>(is_complex) : rs_max_slots<float>(is_complex);
    return rs_shape_ok((int)((n + kT - 1) / kT), gmax, maxng);
}

Status uinv_device(lhl_handle *h, const void *U, int64_t ustride, const void *rdiag, int64_t lstride,
                   int64_t nshift, int is_complex, DevBuf &Uinv, int64_t *uistride, DevBuf &UT, int ut) {
    if (h->n == 0 || nshift == 0) return {};
    return h->dtype == LHL_F64
               ? uinv_typed<double>(h, U, ustride, rdiag, lstride, nshift, is_complex, Uinv, uistride, UT, ut)
               : uinv_typed<float>(h, U, ustride, rdiag, lstride, nshift, is_complex, Uinv, uistride, UT, ut);
}

Status linv_device(lhl_handle *h) {
    if (h->cplx) return {};   // a complex L is swept directly (lhl_cplx.hip)
    if (h->n < 2) return {};
    return h->dtype == LHL_F64 ? linv_typed<double>(h) : linv_typed<float>(h);
}

Status groups_device(lhl_handle *h, int64_t nshift, const int64_t *cb, int planes, DevBuf &out, int64_t *ngroups,
                     int64_t *nwide) {
    std::vector<UGroup> ug = build_groups(nshift, cb, planes);
    const std::vector<UGroup> wg = build_wgroups(nshift, cb, planes);
    *ngroups = (int64_t)ug.size();
    *nwide = (int64_t)wg.size();
    ug.insert(ug.end(), wg.begin(), wg.end());
    LHL_TRY(out.ensure(sizeof(UGroup) * std::max<size_t>(ug.size(), 1)));
    if (!ug.empty()) LHL_HIP_TRY(hipMemcpy(out.p, ug.data(), sizeof(UGroup) * ug.size(), hipMemcpyHostToDevice));
    (void)h;
    return {};
}

// internal columns <-> planar staged columns (no permutation): the ends of
// applyZ / applyZinv on a complex L
template <typename T, typename TX>
__global__ void k_zstage(int n, const TX *__restrict__ x, int64_t ldx, int64_t xplane, T *__restrict__ Y,
                         int64_t ldy, int to_y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
    if (i >= n) return;
    const int col = q >> 1, pl = q & 1;
    TX *xp = const_cast<TX *>(x) + i + (int64_t)col * ldx + pl * xplane;
    T *yp = Y + i + (int64_t)q * ldy;
    if (to_y) *yp = (T)*xp;
    else *xp = (TX)*yp;
}

// applyZinv (inverse = 1): x <- L^-1 P^-1 D^-1 x; applyZ: x <- D P L x, on
// staged planar columns of element type xdt.
template <typename T, typename TX>
Status zapplyz_typed(lhl_handle *h, int inverse, void *x, int64_t ncol, int64_t ldx, int64_t xplane) {
    const int n = (int)h->n;
    const int64_t ldy = std::max<int64_t>(64, ((n + 63) / 64) * 64);
    const int Q = (int)(2 * ncol);
    LHL_TRY(h->ybuf.ensure(sizeof(T) * (size_t)ldy * Q));
    LHL_TRY(h->rbuf.ensure(sizeof(T) * (size_t)ldy * Q));
    T *Yw = h->ybuf.as<T>(), *Yo = h->rbuf.as<T>();
    const dim3 gq((unsigned)((n + 255) / 256), (unsigned)Q);
    if (inverse) {
        LHL_LAUNCH(h, "gather", (k_gather<T, TX>), gq, dim3(256), 0, n, (const int *)h->perm.as<int>(),
                   (const double *)h->iscale64.as<double>(), (const TX *)x, ldx, xplane, 2, Yw, Yo, ldy,
                   (const int *)nullptr);
        LHL_TRY(zsweep_device<T>(h, 1, Yw, Yo, ldy, ncol));
        LHL_LAUNCH(h, "scatter", (k_zstage<T, TX>), gq, dim3(256), 0, n, (TX *)x, ldx, xplane, Yo, ldy, 0);
    } else {
        LHL_LAUNCH(h, "gather", (k_zstage<T, TX>), gq, dim3(256), 0, n, (const TX *)x, ldx, xplane, Yw, ldy, 1);
        LHL_TRY(zsweep_device<T>(h, 0, Yw, Yo, ldy, ncol));
        LHL_LAUNCH(h, "scatter", (k_scatter<T, TX>), gq, dim3(256), 0, n, (const T *)Yo, (const T *)Yo, ldy, 2,
                   (const int *)h->iperm.as<int>(), (const double *)h->scale64.as<double>(), (TX *)x, ldx, xplane, 0,
                   (const int *)nullptr);
    }
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

Status zapplyz_device(lhl_handle *h, int inverse, void *x, int64_t ncol, int64_t ldx, int64_t xplane, int xdt) {
    if (h->n == 0 || ncol == 0) return {};
    if (h->dtype == LHL_F64) {
        if (xdt == LHL_F64) return zapplyz_typed<double, double>(h, inverse, x, ncol, ldx, xplane);
        return zapplyz_typed<double, float>(h, inverse, x, ncol, ldx, xplane);
    }
    if (xdt == LHL_F64) return zapplyz_typed<float, double>(h, inverse, x, ncol, ldx, xplane);
    return zapplyz_typed<float, float>(h, inverse, x, ncol, ldx, xplane);
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

void blas_release(lhl_handle *h) {
    if (h->blas) rocblas_destroy_handle((rocblas_handle)h->blas);
    h->blas = nullptr;
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
