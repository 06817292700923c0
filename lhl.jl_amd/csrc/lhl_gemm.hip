// The solves' dense products on the matrix cores: C = A B for column-major
// operands, A m x K, B K x q, C m x q, in FP32 or FP64 with the same
// precision accumulate.  Modes: 0 a full A; 1 a lower-triangular A stored with
// zeros above the diagonal (the chunks above a row panel's diagonal block are
// skipped); 2 C = B + tril(A, -1) B, reading only A's strict lower triangle
// (whatever A holds on and above the diagonal), for the unit lower (I + L').
// They replace the L sweeps' triangular products (y' <- L'^-1 y' and
// y' <- (I + L') y', the reference's hand-written sweeps,
// src/LHLFactorization.jl:2245-2434) and the refinement residual's J x
// (src:3151, a BLAS mul! in the reference).
//
// Each workgroup (4 waves) owns a 128 x 64 tile of C; wave w holds the 64 x 32
// quadrant (rows 64*(w&1), columns 32*(w>>1)) as 4 x 2 blocks of 16 x 16
// (v_mfma_{f32,f64}_16x16x4).  K runs in chunks of BK (32 FP32, 16 FP64)
// staged through two LDS buffers: the next chunk's global loads are in flight
// while the current chunk's MFMAs run.  The LDS rows are padded so that a
// row's stride is 16 words mod 32: the two 16-lane halves of a 32-lane LDS
// read touch disjoint banks.
//
// Column masks: with `cact` set, a workgroup whose 64 columns all belong to
// batch columns with cact[c] == 0 returns at once (internal column j is batch
// column j / planes).  The refinement rounds pass the converged columns' flags,
// so a round costs what its unconverged columns need.
//
// Workgroup placement: the 8 XCDs receive workgroups round-robin (blockIdx %
// 8), each with its own L2.  Row panels are dealt to the XCDs so that every
// workgroup of one row panel (which all read the same A panel) runs on one
// XCD.  For a triangular A a workgroup takes two row panels, the long
// nrt - 1 - p and then the short p: every workgroup then has the same work.
#include "lhl_common.h"

#include "lhl_mfma.h"

#include <type_traits>

namespace lhl {
namespace {

constexpr int kBM = 128, kBN = 64;

template <typename T> constexpr int gemm_bk() { return sizeof(T) == 8 ? 16 : 32; }
template <typename T> constexpr int gemm_pad() { return sizeof(T) == 8 ? 8 : 16; }

template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_mgemm(int m, int q, int K, const T *__restrict__ A, int64_t lda,
                                               const T *__restrict__ B, int64_t ldb, T *__restrict__ C, int64_t ldc,
                                               const int *__restrict__ cact, int planes, int nrt, int nct) {
    constexpr int BK = gemm_bk<T>(), PAD = gemm_pad<T>();
    constexpr int AE = kBM * BK / 256;          // A elements per thread (a run down one column)
    constexpr int ATC = kBM / AE;               // threads per A column
    constexpr int BE = BK * kBN / 256;          // B elements per thread (a run down one column)
    constexpr int BTC = BK / BE;                // threads per B column
    static_assert(AE % 4 == 0 && BE % 4 == 0, "vector runs of four");
    __shared__ T As[2][BK][kBM + PAD];
    __shared__ T Bs[2][BK][kBN + PAD];
    using M = Mfma<T>;
    constexpr bool LOWER = MODE != 0;

    // tiles of this workgroup: one row panel, or for a triangular A the pair
    // (nrt - 1 - pp, pp), whose lengths add up to the same for every pair
    const int x = blockIdx.x & 7, s = blockIdx.x >> 3;
    const int u = s / nct, ct = s - u * nct;
    int rts[2], npan;
    if (LOWER) {
        const int pp = 8 * u + x;
        if (pp >= (nrt + 1) / 2) return;
        rts[0] = nrt - 1 - pp;
        rts[1] = pp;
        npan = rts[0] != rts[1] ? 2 : 1;
    } else {
        rts[0] = rts[1] = 8 * u + x;
        if (rts[0] >= nrt) return;
        npan = 1;
    }
    const int j0 = ct * kBN;
    const int tid = threadIdx.x;
    if (cact) {
        int any = 0;
        if (tid < kBN) {
            const int j = j0 + tid;
            any = j < q && cact[j / planes] != 0;
        }
        if (!__syncthreads_or(any)) return;
    }
    for (int pi = 0; pi < npan; pi++) {
    const int rt = rts[pi];
    const int i0 = rt * kBM;
    const int kend = LOWER ? min(K, i0 + kBM) : K;
    const int nch = (kend + BK - 1) / BK;

    // this thread's load slots
    const int ak = tid / ATC, ai = (tid % ATC) * AE;
    const int bj = tid / BTC, bk = (tid % BTC) * BE;
    const bool arow_full = i0 + kBM <= m;
    const bool bcol_ok = j0 + bj < q;
    // two register sets: chunk c + 2's loads are issued while chunk c is
    // multiplied and chunk c + 1 (loaded one chunk earlier) goes to LDS
    T ra[2][AE], rb[2][BE];
    auto load = [&](auto par, int k0) {
        constexpr int P = decltype(par)::value;
        const int ka = k0 + ak;
        const T *ap = A + (int64_t)ka * lda + i0 + ai;
        if (ka < kend && arow_full) {
#pragma unroll
            for (int v = 0; v < AE; v += 4) {
                const V4<T> t4 = ld4u(ap + v);
                ra[P][v] = t4[0]; ra[P][v + 1] = t4[1]; ra[P][v + 2] = t4[2]; ra[P][v + 3] = t4[3];
            }
        } else {
#pragma unroll
            for (int v = 0; v < AE; v++) ra[P][v] = (ka < kend && i0 + ai + v < m) ? ap[v] : (T)0;
        }
        if (MODE == 2 && ka >= i0 + ai) {   // keep the strict lower triangle only
#pragma unroll
            for (int v = 0; v < AE; v++)
                if (i0 + ai + v <= ka) ra[P][v] = (T)0;
        }
        const int kb = k0 + bk;
        const T *bp = B + (int64_t)(j0 + bj) * ldb + kb;
        if (bcol_ok && kb + BE <= kend) {
#pragma unroll
            for (int v = 0; v < BE; v += 4) {
                const V4<T> t4 = ld4u(bp + v);
                rb[P][v] = t4[0]; rb[P][v + 1] = t4[1]; rb[P][v + 2] = t4[2]; rb[P][v + 3] = t4[3];
            }
        } else {
#pragma unroll
            for (int v = 0; v < BE; v++) rb[P][v] = (bcol_ok && kb + v < kend) ? bp[v] : (T)0;
        }
    };
    auto store = [&](auto par) {
        constexpr int P = decltype(par)::value;
#pragma unroll
        for (int v = 0; v < AE; v++) As[P][ak][ai + v] = ra[P][v];
#pragma unroll
        for (int v = 0; v < BE; v++) Bs[P][bk + v][bj] = rb[P][v];
    };

    typename M::acc_t acc[4][2];
#pragma unroll
    for (int bi = 0; bi < 4; bi++)
#pragma unroll
        for (int bjj = 0; bjj < 2; bjj++) acc[bi][bjj] = M::zero();
    const int lane = tid & 63, w = tid >> 6;
    const int wr = (w & 1) * 64, wc = (w >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    using P0 = std::integral_constant<int, 0>;
    using P1 = std::integral_constant<int, 1>;

    // chunk c sits in LDS buffer c & 1 and register set c & 1
    auto body = [&](auto par, int c) {
        constexpr int P = decltype(par)::value;
        using PN = std::integral_constant<int, 1 - P>;
        if (c + 2 < nch) load(par, (c + 2) * BK);
#pragma unroll
        for (int k4 = 0; k4 < BK; k4 += 4) {
            const int k = k4 + lk;
            T a[4], b[2];
#pragma unroll
            for (int bi = 0; bi < 4; bi++) a[bi] = As[P][k][wr + 16 * bi + li];
#pragma unroll
            for (int bjj = 0; bjj < 2; bjj++) b[bjj] = Bs[P][k][wc + 16 * bjj + li];
#pragma unroll
            for (int bi = 0; bi < 4; bi++)
#pragma unroll
                for (int bjj = 0; bjj < 2; bjj++) acc[bi][bjj] = M::mac(a[bi], b[bjj], acc[bi][bjj]);
        }
        if (c + 1 < nch) store(PN{});
        lds_barrier();   // LDS only: chunk c + 2's loads stay in flight
    };
    load(P0{}, 0);
    store(P0{});
    if (nch > 1) load(P1{}, BK);
    lds_barrier();
    for (int c = 0; c < nch; c += 2) {
        body(P0{}, c);
        if (c + 1 < nch) body(P1{}, c + 1);
    }
#pragma unroll
    for (int bi = 0; bi < 4; bi++)
#pragma unroll
        for (int bjj = 0; bjj < 2; bjj++) {
            const int col = j0 + wc + 16 * bjj + li;
            if (col >= q) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = i0 + wr + 16 * bi + M::row(lane, r);
                if (row < m) {
                    T v = acc[bi][bjj][r];
                    if (MODE == 2) v += B[row + (int64_t)col * ldb];
                    C[row + (int64_t)col * ldc] = v;
                }
            }
        }
    }   // row panels
}

template <typename T>
Status mgemm_t(lhl_handle *h, const char *name, int mode, int m, int q, int K, const T *A, int64_t lda, const T *B,
               int64_t ldb, T *C, int64_t ldc, const int *cact, int planes) {
    if (m <= 0 || q <= 0) return {};
    if (mode == 2 && K != m) return {LHL_ERR_ARG, "mgemm: the unit lower product needs a square A"};
    const bool lower = mode != 0;
    const int nrt = (m + kBM - 1) / kBM, nct = (q + kBN - 1) / kBN;
    // row panels (pairs of them for a triangular A) per XCD; every XCD gets
    // the same slot count, idle slots return
    const int per = lower ? (((nrt + 1) / 2) + 7) / 8 : (nrt + 7) / 8;
    const unsigned grid = 8u * (unsigned)per * (unsigned)nct;
    if (mode == 2)
        LHL_LAUNCH(h, name, (k_mgemm<T, 2>), dim3(grid), dim3(256), 0, m, q, K, A, lda, B, ldb, C, ldc, cact,
                   planes, nrt, nct);
    else if (mode == 1)
        LHL_LAUNCH(h, name, (k_mgemm<T, 1>), dim3(grid), dim3(256), 0, m, q, K, A, lda, B, ldb, C, ldc, cact,
                   planes, nrt, nct);
    else
        LHL_LAUNCH(h, name, (k_mgemm<T, 0>), dim3(grid), dim3(256), 0, m, q, K, A, lda, B, ldb, C, ldc, cact,
                   planes, nrt, nct);
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

}  // namespace

Status mgemm(lhl_handle *h, const char *name, int mode, int m, int q, int K, const float *A, int64_t lda,
             const float *B, int64_t ldb, float *C, int64_t ldc, const int *cact, int planes) {
    return mgemm_t<float>(h, name, mode, m, q, K, A, lda, B, ldb, C, ldc, cact, planes);
}
Status mgemm(lhl_handle *h, const char *name, int mode, int m, int q, int K, const double *A, int64_t lda,
             const double *B, int64_t ldb, double *C, int64_t ldc, const int *cact, int planes) {
    return mgemm_t<double>(h, name, mode, m, q, K, A, lda, B, ldb, C, ldc, cact, planes);
}

}  // namespace lhl
