// The blocked (delayed-update) ELMHES reduction on one GPU (SciML/LHL.jl
// _lhl_reduce_blocked!, src:765-2001), on the same row-major working copy as
// the row-resident step loop of lhl_reduce.hip: physical row r holds logical
// row ilog[r], so every row interchange is a relabelling.
//
// Steps k0..kb form a panel.  The trailing block T = A[:, kb+2:n] is read,
// never written, while the panel runs; each step
//   * takes the pivot decision from the previous step's per-workgroup
//     candidates (largest |A[i,k]| over logical rows i >= k+1, ties to the
//     smaller row, NaN never wins, src:470-477);
//   * interchanges columns k+1 and p in the panel's own columns, which live in
//     per-row records (the panel columns and their pre-panel copies, ping-pong
//     by step parity so any workgroup may read any row's record of the
//     previous step).  A trailing column p pulled into the panel gets the
//     panel's missed left transforms (src:865-885); the panel column pushed
//     out lands, in its pre-panel state, in a side slot that stands in for
//     column p of T until the panel ends;
//   * scales the multipliers and applies the left transform to the panel
//     columns (src:811-824) and the panel part of the right update (src:825-831);
//   * forms w = T0 l against the pre-panel T (a streaming GEMV, read-only),
//     then the correction of src:832-846: the panel's interchanges are the
//     relabelling, and its left transforms need w at the panel's pivot rows,
//     which every workgroup forms itself (the same dot products, from L2)
//     before it finishes column k+1 on its own rows.
// At the panel's end: the side slots are written into T, the top rows take
// the panel's column interchanges, the pivot rows' trailing part is solved
// with the panel's unit lower triangle (src:1914-2001, U12), the trailing
// rows take the rank-nb update T -= L21 U12 on the matrix cores
// (src:1674-1911), and the rows above the panel take the deferred right
// updates A[1:k0, k0+1:kb+1] += A[1:k0, k0+2:n] L (src:1585-1672) on the
// matrix cores as well.  The deferred row interchanges of earlier multipliers
// (src:1381-1415) are the relabelling.
#include "lhl_common.h"
#include "lhl_mfma.h"

#include <algorithm>
#include <climits>

namespace lhl {
namespace {

template <typename T> __device__ __forceinline__ T babs(T x) { return x < 0 ? -x : x; }

template <typename T>
struct BCand {
    T v, x;           // |A[r, k]| (-1: none) and A[r, k]
    int i, r, pos;    // logical index (INT_MAX: none), physical row, position in the panel's row list
    int pad;
};
template <typename T>
struct BNext {
    T vk1;            // A[k+1, k] of the row at logical position k+1
    int r1, pos1;     // its physical row and position
};

template <typename T>
__device__ __forceinline__ bool bbetter(T v2, int i2, T v, int i) {
    return v2 > v || (v2 == v && i2 < i);
}

constexpr int kBWaves = 8;        // waves of the step kernel
constexpr int kBThreads = kBWaves * 64;
constexpr int kBMaxRows = 64;     // own rows per workgroup of the step kernel
constexpr int kBMaxCand = 256;    // workgroups whose candidates one decision reads

// Per-row record of a panel, indexed by the row's position in the panel's row
// list: the panel columns k0..k0+NB (NB+1 entries), then their pre-panel
// copies for columns k0+1..k0+NB (NB entries).
template <int NB> constexpr int rec_stride() { return 2 * NB + 2; }
// 16-byte vectors of the trailing GEMV
template <typename T> constexpr int bvec() { return 16 / (int)sizeof(T); }

template <typename T>
struct BArgs {
    T *B;
    int64_t ldb;
    int n, k, k0, kb;
    const int *act;   // position -> physical row (logical order at the panel start)
    int *plog;        // position -> logical index
    int m, R;         // rows in the list, rows per workgroup
    const T *rin;     // records read (previous step)
    T *rout;          // records written
    T *PB;            // side slots of pushed columns: PB[s * ldpb + r]
    int64_t ldpb;
    const BCand<T> *cin;
    BCand<T> *cout;
    int ncand;
    const BNext<T> *nin;
    BNext<T> *nout;
    const T *gin;     // column k in logical order
    T *gout;          // column k+1 in logical order
    int *ilog, *ipiv, *perm;
    int *ppos;        // ppos[q]: position of the panel's q-th pivot row
    T *Lc;            // Lc[j * NB + c]: multiplier of step k0+c at logical row j (written at the last step)
};

// One elimination step of a panel.  Workgroup g owns the positions
// g*R .. g*R+R-1 of the row list.  Dynamic LDS holds l_j = A[j, k] / piv for
// j >= k+2.  The trailing GEMV runs over all rows (own + the panel's pivot
// rows) with the columns split over the 512 threads: thread t owns the
// 16-byte vectors t, t+512, ... (NI of them), so every load instruction is
// one coalesced 1 KiB run per wave, and RB rows' loads are in flight at once.
template <typename T, int NB, int NI>
__global__ void __launch_bounds__(kBThreads) k_bstep(BArgs<T> a) {
#pragma clang fp contract(off)
    constexpr int PS = rec_stride<NB>();
    constexpr int CPL = kBMaxCand / 64;
    constexpr int FILL = 8;
    constexpr int VEC = bvec<T>();
    constexpr int RB = NI >= 16 ? 1 : 16 / NI;     // rows per load batch
    constexpr int GT = kBThreads - 64;                // GEMV threads: waves 1..7 (wave 0 steers)
    constexpr int SMAX = kBMaxRows + NB + 1;
    using V = T __attribute__((ext_vector_type(VEC)));
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, tid = threadIdx.x;
    const int n = a.n, k = a.k, k0 = a.k0, kb = a.kb;
    const int q0 = k - k0;            // this step's index in the panel
    const int ck1 = q0 + 1;           // record index of column k+1
    const int clast = kb + 1 - k0;    // record index of the panel's last column
    const int jt = kb + 2;            // first trailing column
    const int jl = k + 2;             // first column of lv
    extern __shared__ __align__(16) unsigned char lds_dyn[];
    T *lv = reinterpret_cast<T *>(lds_dyn);
    __shared__ int s_p, s_rp, s_rpos, s_skip, s_nred, s_oldslot;
    __shared__ T s_piv, s_a1, s_pk;
    __shared__ int s_prr[NB + 1], s_ppos[NB + 1];   // physical rows / positions of the panel's pivots 0..q0
    __shared__ int s_ipv[NB];                      // ipiv[k0 .. k-1]
    __shared__ T s_L[NB + 1][NB + 1];              // records (panel columns) of pivot rows 0..q0
    __shared__ T s_prow[NB + 1];                   // the pivot row's panel columns after the interchange
    __shared__ T s_z[NB + 1], s_x[NB + 1], s_xp[NB + 1];
    __shared__ int s_rcol[NB + 1], s_rslot[NB + 1];
    __shared__ int s_row[kBMaxRows], s_il[kBMaxRows];
    __shared__ T s_rec[kBMaxRows][PS];
    __shared__ T s_zown[kBMaxRows], s_t[kBMaxRows];
    __shared__ T s_cpart[SMAX][NB];
    __shared__ BCand<T> s_cand[kBThreads / 32];
    __shared__ T s_part[SMAX][kBWaves];

    const int rbeg = blockIdx.x * a.R;
    const int rcnt = min(a.R, a.m - rbeg);
    // the decision's inputs first: a wave waits for its loads in issue order
    BCand<T> ce[CPL];
    BNext<T> nx;
    if (wv == 0) {
#pragma unroll
        for (int u = 0; u < CPL; u++) ce[u] = a.cin[min(lane + 64 * u, a.ncand - 1)];
        nx = *a.nin;
    }
    if (tid < rcnt) {
        s_row[tid] = a.act[rbeg + tid];
        s_il[tid] = a.plog[rbeg + tid];
    } else if (tid >= 64 && tid < 64 + q0) {
        s_prr[tid - 64] = a.perm[k0 + 1 + (tid - 64)];
        s_ppos[tid - 64] = a.ppos[tid - 64];
    } else if (tid >= 128 && tid < 128 + q0) {
        s_ipv[tid - 128] = a.ipiv[k0 + (tid - 128)];
    }
    for (int t = tid; t < rcnt * PS; t += kBThreads) {
        const int i = t / PS, c = t - i * PS;
        s_rec[i][c] = a.rin[(int64_t)(rbeg + i) * PS + c];
    }
    // the trailing GEMV's first batch of rows, issued now so it overlaps the
    // pivot decision: the own rows, then the panel's earlier pivot rows
    const int Sk = rcnt + q0;         // GEMV rows known before the decision
    const int jta = jt - jt % VEC;
    const int gt = tid - 64;          // this thread's GEMV index (waves 1..7)
    V v0[RB][NI];
    if (wv != 0) {
#pragma unroll
        for (int b = 0; b < RB; b++) {
            const int r = b < Sk ? (b < rcnt ? a.act[rbeg + b] : a.perm[k0 + 1 + (b - rcnt)]) : 0;
            const T *row = a.B + (int64_t)r * a.ldb;
#pragma unroll
            for (int i = 0; i < NI; i++) {
                const int j0 = jta + VEC * (gt + GT * i);
                V z = {};
                if (b < Sk && j0 < n) z = *reinterpret_cast<const V *>(row + j0);
                v0[b][i] = z;
            }
        }
    }
    // the multipliers' numerators (column k in logical order), first fill pass
    T gpre[FILL];
#pragma unroll
    for (int u = 0; u < FILL; u++) gpre[u] = a.gin[min(k + 1 + kBThreads * u + tid, n - 1)];

    // the pivot decision (src:470-477)
    if (wv == 0) {
        const BCand<T> *e = ce;
        T gv = (T)-1, gx = 0;
        int gi = INT_MAX, gr = 0, gp = 0;
#pragma unroll
        for (int u = 0; u < CPL; u++)
            if (lane + 64 * u < a.ncand && bbetter(e[u].v, e[u].i, gv, gi)) {
                gv = e[u].v; gi = e[u].i; gx = e[u].x; gr = e[u].r; gp = e[u].pos;
            }
        int gl = lane;
        for (int o = 32; o > 0; o >>= 1) {
            const T v2 = __shfl_xor(gv, o, 64);
            const int i2 = __shfl_xor(gi, o, 64), l2 = __shfl_xor(gl, o, 64);
            if (bbetter(v2, i2, gv, gi)) { gv = v2; gi = i2; gl = l2; }
        }
        gx = __shfl(gx, gl, 64);
        gr = __shfl(gr, gl, 64);
        gp = __shfl(gp, gl, 64);
        if (lane == 0) {
            const T a1 = nx.vk1;
            int p = k + 1, rp = nx.r1, rpos = nx.pos1;
            T piv = a1;
            if (a1 == a1 && gi != INT_MAX) {
                p = gi;
                piv = gx;
                if (p != k + 1) { rp = gr; rpos = gp; }
            }
            // the candidates of a panel's first step carry positions in the
            // previous panel's list; this panel's list is in logical order
            if (q0 == 0) rpos = p - k0 - 1;
            s_p = p;
            s_rp = rp;
            s_rpos = rpos;
            s_piv = piv;
            s_a1 = a1;
            s_skip = (piv == (T)0);
            if (blockIdx.x == 0) {
                a.ipiv[k] = p;
                a.perm[k + 1] = rp;
                if (p != k + 1) a.perm[p] = nx.r1;
                a.ppos[q0] = rpos;
            }
        }
    }
    // LDS only: the GEMV prefetch stays in flight
    lds_barrier();
    const int p = s_p, rp = s_rp, skip = s_skip;
    const T piv = s_piv, a1 = s_a1;
    const bool pull = p > kb + 1;
    // this step's pivot row for the GEMV
    V vc[NI];
    if (wv != 0) {
        const T *row = a.B + (int64_t)rp * a.ldb;
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const int j0 = jta + VEC * (gt + GT * i);
            V z = {};
            if (j0 < n) z = *reinterpret_cast<const V *>(row + j0);
            vc[i] = z;
        }
    }
    // the pivot rows' records (panel columns only): wave 0 loads them now and
    // uses them while waves 1..7 run the GEMV
    constexpr int LPT = ((NB + 1) * (NB + 1) + 63) / 64;
    T lval[LPT];
    if (wv == 0) {
#pragma unroll
        for (int u = 0; u < LPT; u++) {
            const int t = lane + 64 * u;
            lval[u] = 0;
            if (t < (q0 + 1) * (NB + 1)) {
                const int q = t / (NB + 1), c = t - q * (NB + 1);
                const int pos = q < q0 ? s_ppos[q] : s_rpos;
                lval[u] = a.rin[(int64_t)pos * PS + c];
            }
        }
    }
    if (tid == 0) {
        s_prr[q0] = rp;
        s_ppos[q0] = s_rpos;
    }
    if (wv == 2) {
        // the side slots standing in for trailing columns (lane q: the push of
        // step q, and this step's push of p to slot q0 on a pull); the latest
        // push of each column wins.  Also the slot holding column p before
        // this step (the latest earlier push of p).
        const int q = lane;
        int pc = -1;
        if (q < q0) pc = s_ipv[q];
        else if (q == q0 && pull) pc = p;
        bool eff = pc > kb + 1;
        int old = (q < q0 && pc == p) ? q : -1;
        for (int u = 0; u <= q0; u++) {
            const int x = __builtin_amdgcn_readlane(pc, u);
            if (u > q && x == pc) eff = false;
        }
        const unsigned long long mk = __ballot(eff);
        const int idx = __popcll(mk & ((1ull << q) - 1ull));
        if (eff) { s_rcol[idx] = pc; s_rslot[idx] = q; }
        for (int o = 32; o > 0; o >>= 1) old = max(old, __shfl_xor(old, o, 64));
        if (lane == 0) {
            s_nred = __popcll(mk);
            s_oldslot = pull ? old : -1;
        }
    }
    // l_j = A[j, k] / piv for j >= k+2; logical p now holds the row that was at k+1
    {
        auto fill = [&](int jb, const T(&g)[FILL]) __attribute__((always_inline)) {
#pragma unroll
            for (int u = 0; u < FILL; u++) {
                const int j = jb + kBThreads * u + tid;
                if (j >= jl && j < n) lv[j - jl] = skip ? (T)0 : ((j == p) ? a1 : g[u]) / piv;
            }
        };
        fill(k + 1, gpre);
        for (int jb = k + 1 + kBThreads * FILL; jb < n; jb += kBThreads * FILL) {
            T g[FILL];
#pragma unroll
            for (int u = 0; u < FILL; u++) g[u] = a.gin[min(jb + kBThreads * u + tid, n - 1)];
            fill(jb, g);
        }
    }
    lds_barrier();
    const int S = rcnt + q0 + 1;      // GEMV rows: own rows, then pivot rows 0..q0
    const int nred = s_nred;
    // a pulled column's values before this step's push: own rows (tvl) and the
    // pivot rows (wave 1, lane q), issued before the GEMV and used after it
    T pv = 0, tvl = 0;
    if (pull) {
        const int so = s_oldslot;
        if (tid < rcnt) {
            const int r = s_row[tid];
            tvl = so >= 0 ? a.PB[(int64_t)so * a.ldpb + r] : a.B[(int64_t)r * a.ldb + p];
        }
        if (wv == 0 && lane <= q0) {
            const int r = s_prr[lane];
            pv = so >= 0 ? a.PB[(int64_t)so * a.ldpb + r] : a.B[(int64_t)r * a.ldb + p];
        }
    }
    // the side-slot terms (slot value - T0 value) l_j of GEMV row s and slot u,
    // one (s, u) pair per thread and pass: loaded now, used after the GEMV
    constexpr int CPP = (SMAX * NB + kBThreads - 1) / kBThreads;
    T sv[CPP], tv[CPP];
    const int npair = skip ? 0 : S * nred;
#pragma unroll
    for (int w = 0; w < CPP; w++) {
        const int t = tid + kBThreads * w;
        sv[w] = 0;
        tv[w] = 0;
        if (t < npair) {
            const int sl = t / nred, u = t - sl * nred;
            const int r = sl < rcnt ? s_row[sl] : s_prr[sl - rcnt];
            const int pos = sl < rcnt ? rbeg + sl : s_ppos[sl - rcnt];
            const int j = s_rcol[u], so = s_rslot[u];
            sv[w] = so < q0 ? a.PB[(int64_t)so * a.ldpb + r] : a.rin[(int64_t)pos * PS + NB + 1 + q0];   // this step's push
            tv[w] = a.B[(int64_t)r * a.ldb + j];
        }
    }
    // the GEMV against the pre-panel trailing block: z_r = sum_{j >= kb+2} T0[r, j] l_j
    if (!skip && wv != 0) {
        T lr[NI][VEC];
#pragma unroll
        for (int i = 0; i < NI; i++)
#pragma unroll
            for (int x = 0; x < VEC; x++) {
                const int j = jta + VEC * (gt + GT * i) + x;
                lr[i][x] = (j >= jt && j < n) ? lv[j - jl] : (T)0;
            }
        auto reduce = [&](int sl, const V(&vv)[NI]) __attribute__((always_inline)) {
            T acc = 0;
#pragma unroll
            for (int i = 0; i < NI; i++)
#pragma unroll
                for (int x = 0; x < VEC; x++) acc = fma_t(lr[i][x], vv[i][x], acc);
            acc = wave_sum_fast(acc);
            if (lane == 0) s_part[sl][wv] = acc;
        };
        // batches of RB rows, the next batch's loads in flight while one is
        // reduced (batch 0 was issued at entry, the step's pivot row after
        // the decision)
        auto load = [&](int s0, V(&vv)[RB][NI]) __attribute__((always_inline)) {
#pragma unroll
            for (int b = 0; b < RB; b++) {
                const int sl = s0 + b;
                const int r = sl < Sk ? (sl < rcnt ? s_row[sl] : s_prr[sl - rcnt]) : 0;
                const T *row = a.B + (int64_t)r * a.ldb;
#pragma unroll
                for (int i = 0; i < NI; i++) {
                    const int j0 = jta + VEC * (gt + GT * i);
                    V z = {};
                    if (sl < Sk && j0 < n) z = *reinterpret_cast<const V *>(row + j0);
                    vv[b][i] = z;
                }
            }
        };
        if constexpr (NI >= 16) {
            // the widest rows: one buffer (registers)
#pragma unroll
            for (int b = 0; b < RB; b++)
                if (b < Sk) reduce(b, v0[b]);
            reduce(Sk, vc);
            for (int s0 = RB; s0 < Sk; s0 += RB) {
                load(s0, v0);
#pragma unroll
                for (int b = 0; b < RB; b++)
                    if (s0 + b < Sk) reduce(s0 + b, v0[b]);
            }
        } else {
            V v1[RB][NI];
            if (RB < Sk) load(RB, v1);
#pragma unroll
            for (int b = 0; b < RB; b++)
                if (b < Sk) reduce(b, v0[b]);   // uniform over the workgroup
            reduce(Sk, vc);
            for (int s0 = RB; s0 < Sk; s0 += 2 * RB) {
                if (s0 + RB < Sk) load(s0 + RB, v0);
#pragma unroll
                for (int b = 0; b < RB; b++)
                    if (s0 + b < Sk) reduce(s0 + b, v1[b]);
                if (s0 + 2 * RB < Sk) load(s0 + 2 * RB, v1);
#pragma unroll
                for (int b = 0; b < RB; b++)
                    if (s0 + RB + b < Sk) reduce(s0 + RB + b, v0[b]);
            }
        }
    }
#pragma unroll
    for (int w = 0; w < CPP; w++) {
        const int t = tid + kBThreads * w;
        if (t < npair) {
            const int sl = t / nred, u = t - sl * nred;
            s_cpart[sl][u] = (sv[w] - tv[w]) * lv[s_rcol[u] - jl];
        }
    }
    if (wv == 0) {
        // wave 0, while waves 1..7 stream: the pivot rows' records into LDS;
        // a pulled column's missed left transforms solved on the pivot rows
        // (x'_q = v_q - sum_{q' < q} L[q][q'] x'_q', q < q0) and the pivot
        // row's new entry pk (q = q0); the pivot row's panel columns after the
        // interchange
#pragma unroll
        for (int u = 0; u < LPT; u++) {
            const int t = lane + 64 * u;
            if (t < (q0 + 1) * (NB + 1)) {
                const int q = t / (NB + 1), c = t - q * (NB + 1);
                s_L[q][c] = lval[u];
            }
        }
        if (pull) {
            if (tid < rcnt) s_t[tid] = tvl;
            const int q = lane;
            T xq = pv;
            for (int u = 0; u < q0; u++) {
                const T xu = wsum_detail::lane_read(xq, u);
                if (q > u && q <= q0) xq = xq - s_L[q][u] * xu;
            }
            if (q < q0) s_xp[q] = xq;
            if (q == q0) s_pk = xq;
        }
        if (lane <= NB) {
            const int c = lane;
            T v = s_L[q0][c];
            if (!pull && p != k + 1) {
                const int cp = p - k0;
                if (c == ck1) v = s_L[q0][cp];
                if (c == cp) v = s_L[q0][ck1];
            }
            if (pull && c == ck1) v = s_pk;
            s_prow[c] = v;
        }
    }
    __syncthreads();
    if (wv == 0 && !skip) {
        // the GEMV sums of every row (with the side-slot terms), then the
        // correction's right-hand side at the pivot rows (src:836-843):
        // x_q = z_q - sum_{q' < q} L[q][q'] x_q'
        for (int sl = lane; sl < S; sl += 64) {
            T z = s_part[sl][1];
#pragma unroll
            for (int w = 2; w < kBWaves; w++) z = z + s_part[sl][w];
            for (int u = 0; u < nred; u++) z = z + s_cpart[sl][u];
            if (sl < rcnt) s_zown[sl] = z;
            else s_z[sl - rcnt] = z;
        }
        const int q = lane;
        T xq = q <= q0 ? s_z[q] : (T)0;
        for (int u = 0; u < q0; u++) {
            const T xu = wsum_detail::lane_read(xq, u);
            if (q > u && q <= q0) xq = fma_t(-s_L[q][u], xu, xq);
        }
        if (q <= q0) s_x[q] = xq;
    }
    __syncthreads();
    // own rows: interchange, scaling, left transform, right update of column
    // k+1.  A row is a group of 32 lanes, lane c holding panel column k0+c
    // (e) and its pre-panel copy (b0, c >= 1); the sums over a row's columns
    // are group reductions.
    {
        const int g = tid >> 5, c = tid & 31, gb = lane & 32;
        const int cp = p - k0;
        BCand<T> best;
        best.v = (T)-1; best.x = 0; best.i = INT_MAX; best.r = 0; best.pos = 0; best.pad = 0;
        for (int t0 = 0; t0 < rcnt; t0 += kBThreads / 32) {
            const int t = t0 + g;
            const bool ok = t < rcnt;
            const int tt = ok ? t : 0;
            const int r = s_row[tt], il = s_il[tt];
            const int nl = (r == rp) ? k + 1 : (il == k + 1 ? p : il);
            T ev = (ok && c <= NB) ? s_rec[tt][c] : (T)0;
            T bv = (ok && c >= 1 && c <= NB) ? s_rec[tt][NB + c] : (T)0;
            if (ok && !pull && p != k + 1) {
                if (c == ck1) { ev = s_rec[tt][cp]; bv = s_rec[tt][NB + cp]; }
                else if (c == cp) { ev = s_rec[tt][ck1]; bv = s_rec[tt][NB + ck1]; }
            }
            if (pull) {
                // the pulled column's missed left transforms (src:878-884)
                T term = (c < q0 && nl >= k0 + c + 2) ? ev * s_xp[c] : (T)0;
                for (int o = 16; o > 0; o >>= 1) term += __shfl_xor(term, o, 64);
                if (ok && c == ck1) {
                    const T tv = s_t[tt];
                    a.PB[(int64_t)q0 * a.ldpb + r] = bv;
                    bv = tv;
                    ev = tv - term;
                }
            }
            if (!skip) {
                const T eq0 = __shfl(ev, gb + q0, 64);
                const bool below = nl >= k + 2;
                const T lq = below ? eq0 / piv : (T)0;
                if (below) {
                    if (c == q0) ev = lq;
                    if (c >= ck1 && c <= clast) ev = fma_t(lq, -s_prow[c], ev);
                }
                // l_j e_j over the panel columns beyond k+1 (src:825-831), and
                // the correction -e_q x_q (src:836-846)
                T term = 0;
                if (c > ck1 && c <= clast) term = lv[k0 + c - jl] * ev;
                if (c <= q0 && nl >= k0 + c + 2) term = term - ev * s_x[c];
                for (int o = 16; o > 0; o >>= 1) term += __shfl_xor(term, o, 64);
                if (c == ck1) ev = ev + (term + s_zown[tt]);
            }
            const T v1 = __shfl(ev, gb + ck1, 64);
            if (ok) {
                if (c <= NB) s_rec[t][c] = ev;
                if (c >= 1 && c <= NB) s_rec[t][NB + c] = bv;
                if (k == kb && c < NB) a.Lc[(int64_t)nl * NB + c] = (c <= q0 && nl >= k0 + c + 2) ? ev : (T)0;
                if (c == 0) {
                    a.gout[nl] = v1;
                    if (nl != il) { a.ilog[r] = nl; a.plog[rbeg + t] = nl; }
                    if (nl == k + 2) { BNext<T> x; x.vk1 = v1; x.r1 = r; x.pos1 = rbeg + t; *a.nout = x; }
                    const T av = babs(v1);
                    if (nl >= k + 2 && av == av && bbetter(av, nl, best.v, best.i)) {
                        best.v = av; best.x = v1; best.i = nl; best.r = r; best.pos = rbeg + t;
                    }
                }
            }
        }
        if (c == 0) s_cand[g] = best;
    }
    __syncthreads();
    // the workgroup's candidate for step k+1
    if (wv == 0) {
        constexpr int NG = kBThreads / 32;   // row groups
        T cv = (T)-1;
        int ci = INT_MAX, cl = lane;
        if (lane < NG) { cv = s_cand[lane].v; ci = s_cand[lane].i; }
        for (int o = NG / 2; o > 0; o >>= 1) {
            const T v2 = __shfl_xor(cv, o, 64);
            const int i2 = __shfl_xor(ci, o, 64), l2 = __shfl_xor(cl, o, 64);
            if (bbetter(v2, i2, cv, ci)) { cv = v2; ci = i2; cl = l2; }
        }
        if (lane == 0) a.cout[blockIdx.x] = s_cand[min(cl, NG - 1)];
    }
    for (int t = tid; t < rcnt * PS; t += kBThreads) {
        const int i = t / PS, c = t - i * PS;
        a.rout[(int64_t)(rbeg + i) * PS + c] = s_rec[i][c];
    }
}

// Zeroes the pitch padding of the row-major copy (columns n .. ldb-1), which
// the vector loads of the trailing GEMV read (and multiply by zero).
template <typename T>
__global__ void k_bpad(T *__restrict__ B, int64_t ldb, int n) {
    const int r = blockIdx.x;
    for (int c = n + (int)threadIdx.x; c < ldb; c += blockDim.x) B[(int64_t)r * ldb + c] = 0;
}

// Step 0's inputs: the first panel's row list (rows 1..n-1) and records,
// labels, column 0 in logical order, candidates over rows >= 1 (one per
// workgroup), the row at logical 1.
template <typename T, int NB>
__global__ void __launch_bounds__(256) k_binit(const T *__restrict__ B, int64_t ldb, int n, int rows_per_wg,
                                               T *__restrict__ rec, int *__restrict__ act, int *__restrict__ plog,
                                               int *__restrict__ ilog, T *__restrict__ gout,
                                               BCand<T> *__restrict__ cout, BNext<T> *__restrict__ nout) {
    constexpr int PS = rec_stride<NB>();
    const int r0 = blockIdx.x * rows_per_wg, r1 = min(n, r0 + rows_per_wg);
    BCand<T> best;
    best.v = (T)-1; best.x = 0; best.i = INT_MAX; best.r = 0; best.pos = 0; best.pad = 0;
    for (int r = r0 + (int)threadIdx.x; r < r1; r += blockDim.x) {
        const T *row = B + (int64_t)r * ldb;
        ilog[r] = r;
        const T x = row[0];
        gout[r] = x;
        if (r >= 1) {
            const int pos = r - 1;
            act[pos] = r;
            plog[pos] = r;
            T *o = rec + (int64_t)pos * PS;
            for (int c = 0; c <= NB; c++) o[c] = c < n ? row[c] : (T)0;
            for (int c = 1; c <= NB; c++) o[NB + c] = c < n ? row[c] : (T)0;
            if (r == 1) { BNext<T> q; q.vk1 = x; q.r1 = 1; q.pos1 = 0; *nout = q; }
            const T av = babs(x);
            if (av == av && bbetter(av, r, best.v, best.i)) { best.v = av; best.x = x; best.i = r; best.r = r; best.pos = pos; }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const T v2 = __shfl_xor(best.v, o, 64), x2 = __shfl_xor(best.x, o, 64);
        const int i2 = __shfl_xor(best.i, o, 64), r2 = __shfl_xor(best.r, o, 64), p2 = __shfl_xor(best.pos, o, 64);
        if (bbetter(v2, i2, best.v, best.i)) { best.v = v2; best.x = x2; best.i = i2; best.r = r2; best.pos = p2; }
    }
    __shared__ BCand<T> sw[4];
    if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        BCand<T> b = sw[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++)
            if (bbetter(sw[w].v, sw[w].i, b.v, b.i)) b = sw[w];
        cout[blockIdx.x] = b;
    }
}

// Panel end, first pass.  blockIdx.y = 0: each row of the panel's list gets
// its panel columns back from its final record and the side slots written
// into its trailing columns (the last push of each column wins).
// blockIdx.y = 1: the rows above the panel (logical 0..k0) take the panel's
// column interchanges (src:1344-1379) as one permutation of the columns
// involved: each such column's final value is read from the column the
// interchange sequence traced backwards leads to, all loads before any store.
// The lists are built by one wave (ballots and prefix counts).
template <typename T, int NB>
__global__ void __launch_bounds__(256) k_bfix(T *__restrict__ B, int64_t ldb, int k0, int kb, const int *__restrict__ act,
                                              int m, const T *__restrict__ rec, const T *__restrict__ PB, int64_t ldpb,
                                              const int *__restrict__ ipiv, const int *__restrict__ perm) {
    constexpr int PS = rec_stride<NB>();
    __shared__ int s_ipv[NB], s_col[NB], s_slot[NB];
    __shared__ int s_pc[2 * NB], s_src[2 * NB];
    __shared__ int s_nr, s_np;
    const int nbp = kb - k0 + 1;
    const int lane = threadIdx.x & 63;
    if (threadIdx.x < nbp) s_ipv[threadIdx.x] = ipiv[k0 + threadIdx.x];
    __syncthreads();
    if (threadIdx.x < 64) {
        // side slots: lane q pushed column pc (a pull), effective unless a later step pushed it again
        const int q = lane;
        const int pc = q < nbp ? s_ipv[q] : -1;
        bool eff = pc > kb + 1;
        for (int u = 0; u < nbp; u++) {
            const int x = __builtin_amdgcn_readlane(pc, u);
            if (u > q && x == pc) eff = false;
        }
        const unsigned long long mk = __ballot(eff);
        const int idx = __popcll(mk & ((1ull << q) - 1ull));
        if (eff) { s_col[idx] = pc; s_slot[idx] = q; }
        // the columns the interchanges touch: lane q < nbp offers k0+q+1, lane
        // NB+q offers ipiv (both only when step q interchanged); first offer wins
        const int qq = q < NB ? q : q - NB;
        int c = -1;
        if (qq < nbp && q < 2 * NB) {
            const int pq = s_ipv[qq];
            if (pq != k0 + qq + 1) c = q < NB ? k0 + qq + 1 : pq;
        }
        bool keep = c >= 0;
        for (int u = 0; u < 2 * NB; u++) {
            const int x = __shfl(c, u, 64);
            if (u < q && x == c) keep = false;
        }
        const unsigned long long mc = __ballot(keep);
        const int ic = __popcll(mc & ((1ull << q) - 1ull));
        if (keep) {
            int pos = c;   // traced backwards through the interchanges
            for (int w = nbp - 1; w >= 0; w--) {
                const int a1 = k0 + w + 1, b1 = s_ipv[w];
                if (b1 != a1) {
                    if (pos == a1) pos = b1;
                    else if (pos == b1) pos = a1;
                }
            }
            s_pc[ic] = c;
            s_src[ic] = pos;
        }
        if (lane == 0) {
            s_nr = __popcll(mk);
            s_np = __popcll(mc);
        }
    }
    __syncthreads();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.y == 0) {
        if (t < m) {
            const int r = act[t];
            T *row = B + (int64_t)r * ldb;
            const T *o = rec + (int64_t)t * PS;
            const int nr = s_nr;
            T sv[NB];
#pragma unroll
            for (int u = 0; u < NB; u++)
                if (u < nr) sv[u] = PB[(int64_t)s_slot[u] * ldpb + r];
            for (int c = 0; c <= nbp; c++) row[k0 + c] = o[c];
#pragma unroll
            for (int u = 0; u < NB; u++)
                if (u < nr) row[s_col[u]] = sv[u];
        }
    } else if (t <= k0) {
        T *row = B + (int64_t)perm[t] * ldb;
        const int np = s_np;
        T v[2 * NB];
#pragma unroll
        for (int u = 0; u < 2 * NB; u++)
            if (u < np) v[u] = row[s_src[u]];
#pragma unroll
        for (int u = 0; u < 2 * NB; u++)
            if (u < np) row[s_pc[u]] = v[u];
    }
}

// U12: the pivot rows' trailing part after the panel's left transforms,
// U12[q][j - kb - 2] = T[piv q, j] - sum_{q' < q} L[q][q'] U12[q'][j]
// (src:1914-2001), one column per thread.
template <typename T, int NB>
__global__ void __launch_bounds__(256) k_bu12(const T *__restrict__ B, int64_t ldb, int n, int k0, int kb,
                                              const int *__restrict__ perm, const T *__restrict__ Lc,
                                              T *__restrict__ U12, int64_t W) {
    __shared__ T sL[NB][NB];
    __shared__ int sr[NB];
    const int nbp = kb - k0 + 1;
    for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
        const int q = t / NB, c = t - q * NB;
        sL[q][c] = q < nbp ? Lc[(int64_t)(k0 + 1 + q) * NB + c] : (T)0;
    }
    if (threadIdx.x < nbp) sr[threadIdx.x] = perm[k0 + 1 + threadIdx.x];
    __syncthreads();
    const int64_t jj = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (jj >= W) return;
    const int j = kb + 2 + (int)jj;
    T u[NB];
#pragma unroll
    for (int q = 0; q < NB; q++) u[q] = q < nbp ? B[(int64_t)sr[q] * ldb + j] : (T)0;
#pragma unroll
    for (int q = 1; q < NB; q++)
#pragma unroll
        for (int c = 0; c < q; c++) u[q] = fma_t(-sL[q][c], u[c], u[q]);
#pragma unroll
    for (int q = 0; q < NB; q++)
        if (q < nbp) U12[(int64_t)q * W + jj] = u[q];
}

// The rank-nb update of the trailing block on the matrix cores (src:1674-1911):
// for a listed row r with logical index nl >= kb+2,
//   T[r, j] -= sum_c L[r][c] U12[c][j],
// and for the pivot rows (k0+1 <= nl <= kb+1) T[r, :] = U12[nl - k0 - 1, :].
// A workgroup (4 waves) owns 64 list positions x 64 columns; wave w the
// positions 16w .. 16w+15, as four 16x16 output blocks (v_mfma_*_16x16x4, K = NB).
template <typename T, int NB>
__global__ void __launch_bounds__(256) k_bupd(T *__restrict__ B, int64_t ldb, int k0, int kb,
                                              const int *__restrict__ act, int m, const int *__restrict__ ilog,
                                              const T *__restrict__ rec, const T *__restrict__ U12, int64_t W) {
    constexpr int PS = rec_stride<NB>();
    using M = Mfma<T>;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int64_t j0 = (int64_t)blockIdx.x * 64;
    const int rbase = blockIdx.y * 64 + wv * 16;
    const int nbp = kb - k0 + 1;
    __shared__ int s_r[4][16], s_nl[4][16];
    if (lane < 16) {
        const int ri = rbase + lane;
        const int r = ri < m ? act[ri] : -1;
        s_r[wv][lane] = r;
        s_nl[wv][lane] = r >= 0 ? ilog[r] : -1;
    }
    __syncthreads();
    // A operand: -L of list position rbase + li, K index 4*kq + lk
    T am[NB / 4];
    {
        const int r = s_r[wv][li], nl = s_nl[wv][li];
        const bool trail = r >= 0 && nl >= kb + 2;
#pragma unroll
        for (int kq = 0; kq < NB / 4; kq++) {
            const int c = 4 * kq + lk;
            am[kq] = (trail && c < nbp) ? -rec[(int64_t)(rbase + li) * PS + c] : (T)0;
        }
    }
    typename M::acc_t acc[4];
#pragma unroll
    for (int bj = 0; bj < 4; bj++) {
        const int64_t jj = j0 + 16 * bj + li;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int rw = M::row(lane, i);
            const int r = s_r[wv][rw], nl = s_nl[wv][rw];
            T v = 0;
            if (r >= 0 && jj < W)
                v = nl >= kb + 2 ? B[(int64_t)r * ldb + kb + 2 + jj] : U12[(int64_t)(nl - k0 - 1) * W + jj];
            acc[bj][i] = v;
        }
    }
#pragma unroll
    for (int kq = 0; kq < NB / 4; kq++) {
        const int c = 4 * kq + lk;
#pragma unroll
        for (int bj = 0; bj < 4; bj++) {
            const int64_t jj = j0 + 16 * bj + li;
            const T bv = (c < nbp && jj < W) ? U12[(int64_t)c * W + jj] : (T)0;
            acc[bj] = M::mac(am[kq], bv, acc[bj]);
        }
    }
#pragma unroll
    for (int bj = 0; bj < 4; bj++) {
        const int64_t jj = j0 + 16 * bj + li;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int rw = M::row(lane, i);
            const int r = s_r[wv][rw];
            if (r >= 0 && jj < W) B[(int64_t)r * ldb + kb + 2 + jj] = acc[bj][i];
        }
    }
}

// The deferred right updates of the rows above the panel (src:1585-1672): for
// the rows at logical 0..k0, column k0+1+c += sum_{j >= k0+2} A[r, j] L[j][c],
// on the rows' values after the panel's column interchanges (each step of the
// panel reads only columns no later step of it has changed, so all columns
// come from the same values).  A workgroup (8 waves) owns 16 rows; the sum
// over j is split in 8 contiguous ranges, one per wave, each a chain of
// 16x16x4 matrix-core products over chunks of 64 (lane quarter g takes
// j = chunk + 16g .. +15, so a lane reads 16 consecutive elements of its row),
// and the 8 partial tiles are added in wave order.
template <typename T, int NB>
__global__ void __launch_bounds__(512) k_btop(T *__restrict__ B, int64_t ldb, int n, int k0, int kb,
                                              const int *__restrict__ perm, const T *__restrict__ Lc) {
    using M = Mfma<T>;
    constexpr int NC = NB >= 16 ? NB / 16 : 1;     // 16-wide output blocks
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int t0 = blockIdx.x * 16;
    const int nbp = kb - k0 + 1;
    const int ta = t0 + li;
    const int ra = ta <= k0 ? perm[ta] : -1;
    const int jb = k0 + 2, K = n - jb;
    const int KW = ((K + 8 * 64 - 1) / (8 * 64)) * 64;   // per wave, a multiple of 64
    const int ka = wv * KW, ke = min(K, ka + KW);
    typename M::acc_t acc[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) acc[c] = M::zero();
    const T *rowa = B + (int64_t)(ra >= 0 ? ra : 0) * ldb + jb;
    for (int kc = ka; kc < ke; kc += 64) {
        const int kl = kc + 16 * lk;
        T av[16], bv[NC][16];
#pragma unroll
        for (int q = 0; q < 16; q++) av[q] = (ra >= 0 && kl + q < ke) ? rowa[kl + q] : (T)0;
#pragma unroll
        for (int q = 0; q < 16; q++)
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const int cc = 16 * c + li;
                bv[c][q] = (kl + q < ke && cc < NB) ? Lc[(int64_t)(jb + kl + q) * NB + cc] : (T)0;
            }
#pragma unroll
        for (int q = 0; q < 16; q++)
#pragma unroll
            for (int c = 0; c < NC; c++) acc[c] = M::mac(av[q], bv[c][q], acc[c]);
    }
    __shared__ T s_red[8][NC][4][64];
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int i = 0; i < 4; i++) s_red[wv][c][i][lane] = acc[c][i];
    __syncthreads();
    if (wv != 0) return;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int cc = 16 * c + li;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            T s = s_red[0][c][i][lane];
#pragma unroll
            for (int w = 1; w < 8; w++) s = s + s_red[w][c][i][lane];
            const int t = t0 + M::row(lane, i);
            if (t <= k0 && cc < nbp) {
                T *o = B + (int64_t)perm[t] * ldb + k0 + 1 + cc;
                *o = *o + s;
            }
        }
    }
}

// Panel end, last pass: the next panel's row list (logical order), positions'
// logical indices and records, from the rows' values after the update.
template <typename T, int NB>
__global__ void __launch_bounds__(256) k_bload(const T *__restrict__ B, int64_t ldb, int n, int k0n,
                                               const int *__restrict__ perm, int *__restrict__ act,
                                               int *__restrict__ plog, T *__restrict__ rec) {
    constexpr int PS = rec_stride<NB>();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n - k0n - 1) return;
    const int r = perm[k0n + 1 + t];
    act[t] = r;
    plog[t] = k0n + 1 + t;
    const T *row = B + (int64_t)r * ldb;
    T *o = rec + (int64_t)t * PS;
    T v[NB + 1];
#pragma unroll
    for (int c = 0; c <= NB; c++) v[c] = (k0n + c < n) ? row[k0n + c] : (T)0;
#pragma unroll
    for (int c = 0; c <= NB; c++) o[c] = v[c];
#pragma unroll
    for (int c = 1; c <= NB; c++) o[NB + c] = v[c];
}

Status ensure_side_stream(lhl_handle *h) {
    if (!h->stream2) LHL_HIP_TRY(hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking));
    if (!h->ev_fork) LHL_HIP_TRY(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
    if (!h->ev_join) LHL_HIP_TRY(hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
    return {};
}

template <typename T, int NB, int NI>
Status launch_bstep(lhl_handle *h, const BArgs<T> &a, int G, size_t lds) {
    auto fn = k_bstep<T, NB, NI>;
    LHL_TRY(allow_dynamic_lds(fn));
    LHL_LAUNCH(h, "reduce_step", fn, dim3(G), dim3(kBThreads), lds, a);
    return {};
}

template <typename T, int NB>
Status blocked_steps_nb(lhl_handle *h, int64_t ldb, int rmin, int kswitch, int *kstart) {
    const int n = (int)h->n;
    constexpr int PS = rec_stride<NB>();
    constexpr int VEC = bvec<T>();
    T *B = h->rowA.as<T>();
    const int64_t ldpb = ((n + 63) / 64) * 64;
    LHL_TRY(h->brec.ensure(sizeof(T) * 2 * (size_t)n * PS + 256));
    LHL_TRY(h->bpb.ensure(sizeof(T) * (size_t)NB * ldpb + 256));
    LHL_TRY(h->bact.ensure(sizeof(int) * 3 * (size_t)n + 256));
    LHL_TRY(h->blc.ensure(sizeof(T) * 2 * (size_t)n * NB + 256));
    LHL_TRY(h->bu12.ensure(sizeof(T) * (size_t)NB * n + 256));
    LHL_TRY(h->cand.ensure(sizeof(BCand<T>) * 2 * kBMaxCand + sizeof(BNext<T>) * 2 + 64));
    LHL_TRY(h->gcol.ensure(sizeof(T) * 2 * (size_t)n));
    LHL_TRY(h->ilog.ensure(sizeof(int) * (size_t)n));
    LHL_TRY(ensure_side_stream(h));
    T *rec = h->brec.as<T>();
    T *PB = h->bpb.as<T>();
    int *act = h->bact.as<int>();
    int *plog = act + n;
    int *ppos = act + 2 * (int64_t)n;
    T *Lc = h->blc.as<T>();
    T *U12 = h->bu12.as<T>();
    BCand<T> *cand = h->cand.as<BCand<T>>();
    BNext<T> *nxt = reinterpret_cast<BNext<T> *>(cand + 2 * kBMaxCand);
    T *g = h->gcol.as<T>();
    int *ilog = h->ilog.as<int>();
    int *ipiv = h->ipiv.as<int>(), *perm = h->perm.as<int>();
    LHL_LAUNCH(h, "reduce_copy", k_bpad<T>, dim3(n), dim3(64), 0, B, ldb, n);
    const int G0 = std::max(1, std::min(kBMaxCand, (n + 255) / 256));
    const int rpw0 = (n + G0 - 1) / G0;
    LHL_LAUNCH(h, "reduce_copy", (k_binit<T, NB>), dim3(G0), dim3(256), 0, (const T *)B, ldb, n, rpw0, rec, act, plog,
               ilog, g, cand, nxt);
    const size_t lds = sizeof(T) * (size_t)n;
    int ncand = G0;
    bool top_pending = false;
    *kstart = n - 2;
    for (int k0 = 0; k0 <= n - 3; k0 += NB) {
        const int kb = std::min(k0 + NB - 1, n - 3);
        const int m = n - k0 - 1;
        int G = std::min(kBMaxCand, std::max(1, (m + rmin - 1) / rmin));
        int R = std::min(kBMaxRows, (m + G - 1) / G);
        G = (m + R - 1) / R;
        // vectors of the trailing GEMV per thread
        const int jt = kb + 2, jta = jt - jt % VEC;
        const int nvec = (n - jta + VEC - 1) / VEC;
        const int per = (nvec + kBThreads - 64 - 1) / (kBThreads - 64);   // GEMV on waves 1..7
        T *Lcp = Lc + (int64_t)((k0 / NB) & 1) * n * NB;
        for (int k = k0; k <= kb; k++) {
            const int pa = k & 1, pb = pa ^ 1;
            BArgs<T> a;
            a.B = B; a.ldb = ldb; a.n = n; a.k = k; a.k0 = k0; a.kb = kb;
            a.act = act; a.plog = plog; a.m = m; a.R = R;
            a.rin = rec + (int64_t)pa * n * PS; a.rout = rec + (int64_t)pb * n * PS;
            a.PB = PB; a.ldpb = ldpb;
            a.cin = cand + pa * kBMaxCand; a.cout = cand + pb * kBMaxCand; a.ncand = ncand;
            a.nin = nxt + pa; a.nout = nxt + pb;
            a.gin = g + (int64_t)pa * n; a.gout = g + (int64_t)pb * n;
            a.ilog = ilog; a.ipiv = ipiv; a.perm = perm; a.ppos = ppos; a.Lc = Lcp;
            if (per <= 1) LHL_TRY((launch_bstep<T, NB, 1>(h, a, G, lds)));
            else if (per <= 2) LHL_TRY((launch_bstep<T, NB, 2>(h, a, G, lds)));
            else if (per <= 4) LHL_TRY((launch_bstep<T, NB, 4>(h, a, G, lds)));
            else if (per <= 8) LHL_TRY((launch_bstep<T, NB, 8>(h, a, G, lds)));
            else LHL_TRY((launch_bstep<T, NB, 16>(h, a, G, lds)));
            ncand = G;
        }
        const int64_t Wt = n - kb - 2;
        const T *recf = rec + (int64_t)((kb + 1) & 1) * n * PS;
        // the previous panel's top update (side stream) writes rows this pass permutes
        if (top_pending) LHL_HIP_TRY(hipStreamWaitEvent(h->stream, h->ev_join, 0));
        LHL_LAUNCH(h, "reduce_panel", (k_bfix<T, NB>),
                   dim3((unsigned)std::max((m + 255) / 256, (k0 + 1 + 255) / 256), 2), dim3(256), 0, B, ldb, k0, kb,
                   (const int *)act, m, recf, (const T *)PB, ldpb, (const int *)ipiv, (const int *)perm);
        // the rows above the panel: on the side stream, beside the next panel's steps
        LHL_HIP_TRY(hipEventRecord(h->ev_fork, h->stream));
        LHL_HIP_TRY(hipStreamWaitEvent(h->stream2, h->ev_fork, 0));
        {
            hipStream_t keep = h->stream;
            h->stream = h->stream2;
            LHL_LAUNCH(h, "reduce_top", (k_btop<T, NB>), dim3((unsigned)((k0 + 1 + 15) / 16)), dim3(512), 0, B, ldb,
                       n, k0, kb, (const int *)perm, (const T *)Lcp);
            h->stream = keep;
        }
        LHL_HIP_TRY(hipEventRecord(h->ev_join, h->stream2));
        top_pending = true;
        LHL_LAUNCH(h, "reduce_panel", (k_bu12<T, NB>), dim3((unsigned)((Wt + 255) / 256)), dim3(256), 0, (const T *)B,
                   ldb, n, k0, kb, (const int *)perm, (const T *)Lcp, U12, Wt);
        LHL_LAUNCH(h, "reduce_gemm", (k_bupd<T, NB>), dim3((unsigned)((Wt + 63) / 64), (unsigned)((m + 63) / 64)),
                   dim3(256), 0, B, ldb, k0, kb, (const int *)act, m, (const int *)ilog, recf, (const T *)U12, Wt);
        if (kb + 1 <= n - 3 && kb + 1 >= kswitch) {
            // the row-resident loop takes the remaining steps
            *kstart = kb + 1;
            break;
        }
        if (kb + 1 <= n - 3)
            LHL_LAUNCH(h, "reduce_panel", (k_bload<T, NB>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (const T *)B, ldb, n, kb + 1, (const int *)perm, act, plog,
                       rec + (int64_t)((kb + 1) & 1) * n * PS);
    }
    if (top_pending) LHL_HIP_TRY(hipStreamWaitEvent(h->stream, h->ev_join, 0));
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

template <typename T>
Status blocked_steps(lhl_handle *h, int64_t ldb, int nb, int rmin, int kswitch, int *kstart) {
    if (nb <= 8) return blocked_steps_nb<T, 8>(h, ldb, rmin, kswitch, kstart);
    return blocked_steps_nb<T, 16>(h, ldb, rmin, kswitch, kstart);
}

}  // namespace

// The steps of the blocked reduction on h->rowA (row-major, pitch ldb); the
// caller has loaded the rows and finishes with the labels and the factors.
Status blocked_device(lhl_handle *h, int64_t ldb, int nb, int rmin, int kswitch, int *kstart) {
    if (h->dtype == LHL_F64) return blocked_steps<double>(h, ldb, nb, rmin, kswitch, kstart);
    return blocked_steps<float>(h, ldb, nb, rmin, kswitch, kstart);
}

}  // namespace lhl
