// The per-shift Hessenberg LU of G = sigma*I + tau*H (SciML/LHL.jl lhl_shift!,
// src:2456-2594; bit-for-bit spec test/core_tests.jl:119-171).
//
// One workgroup per shift.  The pending row r (the row of G not yet chosen as
// a pivot row) lives in registers, a strided share of its columns per thread.
// Step k compares r[k] with tau*H[k+1,k], writes U row k (coalesced, packed
// rows of U) and updates the owned r[j], j > k, against row k+1 of H (a
// contiguous column of Ht).  Two forms:
//  - k_shift_panel (real shifts up to n = 4480 in FP64, 8512 in FP32): one
//    wave decides 63 or 64 steps at a time on a window of as many columns with
//    no barrier
//    while the other threads apply the previous block to the rest of the row;
//  - k_shift_real / k_shift_cplx (larger n, complex shifts): one step per
//    barrier; the owner of r[k+1] publishes it through a double-buffered LDS
//    slot and the next rows of H are prefetched into registers.
//
// Every element sees the reference's operations in the reference's order with
// no contraction (mul, mul, then sub), and IEEE division, so the real state
// (U, multipliers, swap, rdiag, info) equals the reference's bit for bit.
#include "lhl_common.h"
#include "lhl_cdiv.h"

#include <climits>
#include <cstring>
#include <mutex>
#include <type_traits>

namespace lhl {

// Uncached blocks on each device.  Memory of this kind is never freed once a
// kernel has read it: a handle keeps every block it was given until it is
// destroyed (a captured graph may hold an older one), and then the blocks go
// to this list for later handles on the same device.
namespace {
std::mutex uc_mu;
std::vector<std::pair<int, std::pair<void *, size_t>>> uc_free;   // (device, block)
}  // namespace

Status uc_ensure(lhl_handle *h, size_t bytes) {
    if (!h->uc.empty() && h->uc.back().second >= bytes) return {};
    std::lock_guard<std::mutex> lk(uc_mu);
    size_t best = uc_free.size();
    for (size_t i = 0; i < uc_free.size(); i++)
        if (uc_free[i].first == h->device && uc_free[i].second.second >= bytes &&
            (best == uc_free.size() || uc_free[i].second.second < uc_free[best].second.second))
            best = i;
    if (best < uc_free.size()) {
        h->uc.push_back(uc_free[best].second);
        uc_free.erase(uc_free.begin() + (long)best);
        return {};
    }
    const size_t want = std::max<size_t>(bytes, size_t(4) << 20);
    void *p = nullptr;
    LHL_HIP_TRY(hipExtMallocWithFlags(&p, want, hipDeviceMallocUncached));
    h->uc.push_back({p, want});
    return {};
}

void uc_release(lhl_handle *h) {
    std::lock_guard<std::mutex> lk(uc_mu);
    for (auto &b : h->uc) uc_free.push_back({h->device, b});
    h->uc.clear();
}

namespace {

constexpr int kShiftThreads = 1024;  // the most threads per shift (the spare slots per U plane)

template <typename T> __device__ __forceinline__ T dabs(T x) { return x < 0 ? -x : x; }

// Shift parameters per workgroup: sigma, tau (re, im) in the working type.
template <typename T>
struct ShiftParam {
    T sr, si, tr, ti;
};

// ---------------------------------------------------------------- real shifts
// Per step, each owned element costs two products, a difference and one
// store: rows of H come in through buffer loads at per-lane offsets fixed for
// the whole kernel, and U row k is written through a buffer resource that
// starts at U[k,k] and ends with the row, so lanes past the end (j >= n) fall
// outside it and store nothing, with no select on the address.  Only the
// chunk holding column k has lanes of finished columns (j < k); they get the
// out-of-range offset.  The lane of column k writes the diagonal itself: its
// element is a (no swap) or tau*H[k+1,k] = b (swap).
// Only the chunk holding column k+1 (a uniform branch) adds sigma and
// publishes r[k+1].  The reciprocal pivots are formed after the loop from the
// diagonals, kept in the subdiagonal's LDS slots once those have been read.
// There is no branch between a prefetched row's load and its use, so the rows
// of H loaded three steps ahead stream in while the current steps run.
template <typename T, int NPER, int NT>
__global__ void __launch_bounds__(NT) k_shift_real(const T *__restrict__ Ht, int64_t ldh, int n,
                                                              const ShiftParam<T> *__restrict__ prm,
                                                              T *__restrict__ Ubase, int64_t ustride,
                                                              T *__restrict__ lbase, T *__restrict__ rdbase,
                                                              int64_t lstride, uint8_t *__restrict__ swbase,
                                                              int64_t sstride, int *__restrict__ infobase) {
#pragma clang fp contract(off)
    constexpr int ES = (int)sizeof(T);
    const int s = blockIdx.x;
    const T sigma = prm[s].sr, tau = prm[s].tr;
    T *U = Ubase + (int64_t)s * ustride;
    T *lv = lbase + (int64_t)s * lstride;
    T *rd = rdbase + (int64_t)s * lstride;
    uint8_t *sw = swbase + (int64_t)s * sstride;
    const int tid = threadIdx.x;
    __shared__ T slot[2];
    extern __shared__ __align__(16) unsigned char lds_dyn[];
    // subd[k] = H[k+1, k], read by every lane at step k; afterwards U[k, k]
    T *subd = reinterpret_cast<T *>(lds_dyn);
    for (int i = tid; i + 1 < n; i += NT) subd[i] = Ht[i + (int64_t)(i + 1) * ldh];

    T r[NPER];
    int jb[NPER];   // byte offset of owned column j = tid + m NT within a row
#pragma unroll
    for (int m = 0; m < NPER; m++) {
        const int j = tid + m * NT;
        jb[m] = j * ES;
        r[m] = (j < n) ? tau * Ht[j] : (T)0;          // r = tau * H[0, :]
    }
    if (tid == 0) r[0] = r[0] + sigma;
    if (tid == 0) slot[0] = r[0];
    // Rows of H are prefetched three steps ahead into three register sets
    // used in rotation (the k loop is unrolled by three so no set is copied);
    // columns j >= n read 0 and are never used.
    T hA[NPER], hB[NPER], hC[NPER];
    auto load_row = [&](int row, T(&hd)[NPER]) {
        const BufRsrc rs = buf_rsrc(Ht + (int64_t)min(row, n - 1) * ldh, n * ES);
#pragma unroll
        for (int m = 0; m < NPER; m++) hd[m] = buf_ld<T>(rs, jb[m]);
    };
    load_row(1, hA);
    load_row(2, hB);
    load_row(3, hC);
    __syncthreads();
    int info = 0;
    T ukp = 0;   // U[k-1, k-1], stored into subd[k-1] at step k (lane 0)
    auto step = [&](int k, const T(&h)[NPER]) {
        const T a = slot[k & 1];
        const T b = tau * subd[k];
        // the decision (src:2545-2565): swap iff |b| > |a|; l = a/b, b/a, or 0 for a zero pivot
        const bool sw_k = dabs(b) > dabs(a);
        const bool zero = !sw_k && a == (T)0;
        const T num = sw_k ? a : b, den = sw_k ? b : a;
        const T q = num / den;
        const T l = zero ? (T)0 : q;
        info = (zero && info == 0) ? k + 1 : info;
        if (tid == 0) {
            lv[k] = l;
            sw[k] = (uint8_t)sw_k;
            if (k > 0) subd[k - 1] = ukp;   // every lane read subd[k-1] before the last barrier
        }
        ukp = sw_k ? b : a;
        // row k of U from its diagonal: column j at byte (j - k) ES
        const BufRsrc ur = buf_rsrc(U + upack_off(n, k), (n - k) * ES);
        const int kb = k * ES;
        const int dcm = k / NT;         // the chunk holding column k
        const int fcm = (k + 1) / NT;   // the chunk holding column k + 1
        // U[k, j] at byte (j - k) ES; j < k (chunk dcm only) stores nothing
        auto uoff = [&](int m) {
            const int x = jb[m] - kb;
            return m == dcm ? (x >= 0 ? x : kBufNone) : x;
        };
        // the elimination with the reference's operation order (src:2550-2572).
        // a and b are the same in every lane (one LDS slot, one address), so the
        // decision is read from the first lane and taken as a scalar branch.
        const T ls = l * sigma;
        if (__builtin_amdgcn_readfirstlane((int)sw_k)) {
#pragma unroll
            for (int m = 0; m < NPER; m++) {
                if ((m + 1) * NT <= k) continue;   // every column of the chunk is finished (uniform)
                const T rj = r[m];
                const T g = tau * h[m];
                const T rs = rj - l * g;           // swap: pending row minus l * (new row)
                if (m == fcm) {
                    const bool first = tid + m * NT == k + 1;
                    const T rnew = first ? rs - ls : rs;
                    r[m] = rnew;
                    buf_st(first ? g + sigma : g, ur, uoff(m));
                    if (first) slot[(k + 1) & 1] = rnew;
                } else {
                    r[m] = rs;
                    buf_st(g, ur, uoff(m));
                }
            }
        } else {
#pragma unroll
            for (int m = 0; m < NPER; m++) {
                if ((m + 1) * NT <= k) continue;
                const T rj = r[m];
                const T g = tau * h[m];
                const T rn = g - l * rj;           // no swap: new row minus l * pending row
                if (m == fcm) {
                    const bool first = tid + m * NT == k + 1;
                    const T rnew = first ? rn + sigma : rn;
                    r[m] = rnew;
                    buf_st(rj, ur, uoff(m));
                    if (first) slot[(k + 1) & 1] = rnew;
                } else {
                    r[m] = rn;
                    buf_st(rj, ur, uoff(m));
                }
            }
        }
        lds_barrier();   // the slot is the only data shared between steps
    };
    int k = 0;
    for (; k + 3 <= n - 1; k += 3) {
        step(k, hA);
        load_row(k + 4, hA);
        step(k + 1, hB);
        load_row(k + 5, hB);
        step(k + 2, hC);
        load_row(k + 6, hC);
    }
    if (k < n - 1) step(k, hA);
    if (k + 1 < n - 1) step(k + 1, hB);
    // the last diagonal: U[n-1, n-1] = r[n-1]
    const int own = (n - 1) % NT, mm = (n - 1) / NT;
    if (tid == 0 && n >= 2) subd[n - 2] = ukp;
    if (tid == own) {
        T rn = 0;
#pragma unroll
        for (int m = 0; m < NPER; m++) if (m == mm) rn = r[m];
        U[upack_off(n, n - 1)] = rn;
        subd[n - 1] = rn;
        sw[n - 1] = 0;
        if (rn == (T)0 && info == 0) info = n;
        infobase[s] = info;
    }
    __syncthreads();
    for (int i = tid; i < n; i += NT) rd[i] = (T)1 / subd[i];
}

// Pipelined panel + bulk form of the same elimination (real shifts).  The
// decisions of a block of KP steps k0 .. k0+KP-1 (block b; KP = 63 or 64)
// depend only on the pending row's columns k0 .. k0+KP-1 (window b), so wave 0
// (the panel wave) runs those steps alone with no barrier: lane L holds column
// k0+L, the step's pivot candidates are read from lane i (a) and from the row
// of H (b), the window is updated and its part of the block's U rows stored.
// It then applies the block, with the multipliers still in its registers, to
// window b+1 (the next KP columns, handed over through LDS) and goes straight
// on to block b+1's decisions.  Meanwhile the other NTB threads apply block b
// to every column beyond window b+1 (each column's recurrence needs only the
// block's multipliers and swap flags, published in LDS, and its rows of H,
// prefetched R steps ahead across blocks) and hand window b+2 over.  One
// barrier per block separates the two; the LDS buffers alternate by block
// parity.
//
// A bulk thread owns groups of W adjacent columns (W = 2 for FP64), so each of
// its loads and stores moves W elements: the bulk is bound by its memory
// operations (loads and stores share one in-order counter, so the stores'
// completion limits the prefetch).  With KP even, the first bulk column, a
// multiple of KP, never splits a pair.  U is written with ordinary stores
// (measured as fast as non-temporal ones with packed rows and tiles alike).
// Each element sees the reference's operations in the
// reference's order, as in the one-step kernel above, so the results are
// bit-identical to it.

// A shift split over G workgroups (G = 1: one workgroup per shift): the
// window values each owner hands to the others' panel waves.  hand and cnt are
// uncached device memory, every access to them a device-scope atomic (coherent
// across the XCDs wherever the workgroups run): per shift, nwin windows of KP
// values and nwin counts of the values that have landed.
template <typename T> struct SplitArgs {
    int G = 1, nsh = 0, nwin = 0;
    T *hand = nullptr;
    int *cnt = nullptr;
    int ut = 0;   // U in 64x64 tiles (utile_row / utile_col) instead of packed rows
};
// polls of a hand-off count before it is given up (about a second; never
// reached while every workgroup of the shift is resident, which the launch
// checks), and the default number of workgroups per shift
constexpr int kSplitSpin = 1 << 20;
constexpr int kSplitDefault = 4;

// W-column groups: the scalar (W = 1) or a two-element vector (W = 2)
template <typename T, int W> struct ColGroup;
template <typename T> struct ColGroup<T, 1> {
    using V = T;
    static __device__ __forceinline__ V ld(BufRsrc r, int off) { return buf_ld<T>(r, off); }
    static __device__ __forceinline__ void st(V v, BufRsrc r, int off) { buf_st(v, r, off); }
    static __device__ __forceinline__ T get(V v, int) { return v; }
    static __device__ __forceinline__ void set(V &v, int, T x) { v = x; }
};
template <typename T> struct ColGroup<T, 2> {
    using V = Vec2<T>;
    static __device__ __forceinline__ V ld(BufRsrc r, int off) { return buf_ld2<T>(r, off); }
    static __device__ __forceinline__ void st(V v, BufRsrc r, int off) { buf_st2(v, r, off); }
    static __device__ __forceinline__ T get(V v, int e) { return v[e]; }
    static __device__ __forceinline__ void set(V &v, int e, T x) { v[e] = x; }
};

// f(integral_constant<int, v>) for the runtime value v in [M, N)
template <int M, int N, typename F> __device__ __forceinline__ void dispatch_const(int v, F &f) {
    if constexpr (M < N) {
        if (v == M) f(std::integral_constant<int, M>{});
        else dispatch_const<M + 1, N>(v, f);
    }
}

template <typename T> __device__ __forceinline__ T rdlane(T v, int l);
template <> __device__ __forceinline__ float rdlane<float>(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
template <> __device__ __forceinline__ double rdlane<double>(double v, int l) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

template <typename T, int NP, int NTB, int PD, int R, int W, int KP>
__global__ void __launch_bounds__(64 + NTB) k_shift_panel(const T *__restrict__ Ht, int64_t ldh, int n,
                                                          const ShiftParam<T> *__restrict__ prm,
                                                          T *__restrict__ Ubase, int64_t ustride,
                                                          T *__restrict__ lbase, T *__restrict__ rdbase,
                                                          int64_t lstride, uint8_t *__restrict__ swbase,
                                                          int64_t sstride, int *__restrict__ infobase,
                                                          SplitArgs<T> sp) {
#pragma clang fp contract(off)
    constexpr int kPanel = KP;   // steps per block = window columns
    static_assert(KP <= 64 && kPanel % R == 0 && kPanel % PD == 0, "a block of steps in whole rotations");
    static_assert(W == 1 || KP % W == 0, "the bulk's first column starts a group");
    using G = ColGroup<T, W>;
    using T2 = typename G::V;
    constexpr int ES = (int)sizeof(T);
    // the shift and this workgroup's part of it: with sp.G workgroups per
    // shift, the shift's G workgroups sit on one XCD (workgroups are dealt to
    // the 8 XCDs in turn) and workgroup g owns the windows w = g (mod G)
    const int GS = sp.G;
    int s = blockIdx.x, g = 0;
    if (GS > 1) {
        const int q = (int)(blockIdx.x >> 3);
        s = (int)(blockIdx.x & 7) + 8 * (q / GS);
        g = q % GS;
        if (s >= sp.nsh) return;   // padding (whole workgroup)
    }
    // owned index o -> column: window g + GS (o / KP), offset o % KP
    auto col_of = [&](int o) -> int {
        if (GS == 1) return o;
        const int wl = o / KP;
        return (g + GS * wl) * KP + (o - wl * KP);
    };
    auto owns = [&](int w) -> bool { return GS == 1 || w % GS == g; };
    // U[k, j] (j >= k) at urow(k) + ucol(j): packed rows, or 64x64 tiles
    const int nbu = (n + 63) / 64;
    constexpr int UV = utile_gw<T>();   // columns per granule of the tiles
    auto urow = [&](int k) -> int64_t { return sp.ut ? utile_row<UV>(nbu, k) : upack_off(n, k) - k; };
    auto ucol = [&](int j) -> int { return sp.ut ? (int)utile_col<UV>(j) : j; };
    bool failed = false;
    const T sigma = prm[s].sr, tau = prm[s].tr;
    T *U = Ubase + (int64_t)s * ustride;
    // The panel wave's stores address U[k, c] as one buffer over this shift's
    // U at byte urow(k) ES + ucol(c) ES: per block, lane L holds row k0 + L's
    // offset (read from lane i at step i) and its column's; lanes that store
    // nothing carry kBufNone, which stays out of range after the row offset is
    // added (U of one shift is below 2 GiB: n <= 13440).
    const BufRsrc uall = buf_rsrc(U, (int)min(ustride * (int64_t)ES, (int64_t)0x7fffffff));
    auto urow_b = [&](int k) -> int { return (int)(urow(min(k, n - 1)) * ES); };
    T *lv = lbase + (int64_t)s * lstride;
    T *rd = rdbase + (int64_t)s * lstride;
    uint8_t *sw = swbase + (int64_t)s * sstride;
    const int tid = threadIdx.x, lane = tid & 63;
    const bool is_panel = tid < 64;
    const int tb = tid - 64;   // bulk thread index (tid >= 64)
    __shared__ T win[2][64];   // pending values of a window, by block parity
    __shared__ T lsh[2][64];   // a block's multipliers
    __shared__ int ssh[2][64]; // a block's swap flags
    __shared__ T plast;        // the pending value of column n-1 after the last step
    // The panel wave reads the rows of H for its steps from LDS tiles that the
    // bulk threads stage one block ahead, straight from global memory into LDS
    // (no registers), so the decision chain never waits on global memory (with
    // rows prefetched into registers it ran at the load latency over the
    // prefetch depth).  Per iteration parity: A (window_apply: H[k0+1+i,
    // c1+lane]), B (the next block's decisions: H[c1+1+i, c1+lane]) and, where
    // windows are handed over two blocks ahead (FAR), C (block b-1 on window
    // b+1: H[k0-KP+1+i, c1+lane]).  Windows are handed over one block ahead:
    // FAR (two ahead, the panel wave applying block b-1 to the window itself)
    // measured 2-5 % slower in FP32, and FP64's three tile pairs would not fit
    // in LDS; it stays a compile-time switch.
    constexpr bool FAR = false;
    constexpr int NTL = FAR ? 3 : 2;
    __shared__ __align__(16) T htile[2][NTL][KP][64];
    const int nsteps = n - 1;

    // bulk threads: the column groups j .. j+W-1 with j = W (tb + m NTB) of
    // the pending row r = tau H[0, :] + sigma e_0
    T2 r[NP];
    int jb[NP];   // byte offset of column j within a row
#pragma unroll
    for (int m = 0; m < NP; m++) {
        const int j0 = col_of(W * (tb + m * NTB));
        jb[m] = j0 * ES;
#pragma unroll
        for (int e = 0; e < W; e++) {
            const int j = j0 + e;
            T v = (!is_panel && j < n) ? tau * Ht[j] : (T)0;
            if (j == 0) v = v + sigma;
            G::set(r[m], e, v);
        }
    }
    // windows 0 and 1 (every workgroup of the shift forms them, as the bulk
    // forms its columns)
    if (is_panel) {
#pragma unroll
        for (int w = 0; w < 2; w++) {
            const int j = w * kPanel + lane;
            T v = (lane < kPanel && j < n) ? tau * Ht[j] : (T)0;
            if (j == 0) v = v + sigma;
            win[w][lane] = v;
            if (j == 0 && n == 1) plast = v;
        }
    }
    // rows of H for the bulk steps, R steps ahead in rotation (a group's
    // elements at columns >= n read 0)
    T2 hs[R][NP];
    auto load_rows = [&](int row, T2(&hd)[NP], int ms) __attribute__((always_inline)) {
        const BufRsrc rs = buf_rsrc(Ht + (int64_t)min(row, n - 1) * ldh, n * ES);
#pragma unroll
        for (int m = 0; m < NP; m++)
            if (m >= ms) hd[m] = G::ld(rs, jb[m]);
    };
    if (!is_panel) {
#pragma unroll
        for (int q = 0; q < R; q++) load_rows(1 + q, hs[q], 0);
    }
    // Barriers in this kernel order LDS only (windows, multipliers, swap flags,
    // plast): no wave reads global memory another wave writes, so loads in
    // flight (the rows of H prefetched across a block boundary) and U stores
    // stay outstanding through them.
    lds_barrier();

    // ---- panel wave state: the current window (lane L = column c0 + L)
    int info = 0;
    T p = 0;                       // pending value of the lane's window column
    T lvec = 0;                    // the last panel block's multipliers (lane i: step i)
    int svec = 0;                  // and swap flags
    T hw[PD];
    // row i of a staged tile (H[r0+i, c+lane]), PD ahead
    auto load_t = [&](const T *tl, int i, T &hd) __attribute__((always_inline)) {
        hd = tl[min(i, kPanel - 1) * 64 + lane];
    };
    // decisions of block k0 on window k0 (p holds it): steps k0 .. k0+nb-1.  A
    // full block (nb = kPanel) is straight-line code, so the waits for the
    // prefetched rows count the stores between them exactly.
    auto panel_nb = [&](int k0, int nb, bool wr, int set, auto fullc) __attribute__((always_inline)) {
        constexpr bool FULL = decltype(fullc)::value;
        const T *tl = &htile[set][1][0][0];
        T uvec = 1;
        // lane L: row k0 + L and column k0 + L; lane i leaves the stores after
        // step i (the diagonal), so every step stores columns c >= k
        const int rowb = urow_b(k0 + lane);
        int colb = (wr && lane < kPanel && k0 + lane <= n - 1) ? ucol(k0 + lane) * ES : kBufNone;
#pragma unroll
        for (int i = 0; i < PD; i++) load_t(tl, i, hw[i]);
#pragma nounroll
        for (int ib = 0; ib < kPanel; ib += PD)
#pragma unroll
        for (int ii = 0; ii < PD; ii++) {
            const int i = ib + ii;
            if (FULL || i < nb) {
                const T h = hw[ii];
                const T a = rdlane(p, i);
                const T b = tau * rdlane(h, i);
                // the decision (src:2545-2565): swap iff |b| > |a|; l = a/b, b/a, or 0 for a zero pivot
                const bool swk = dabs(b) > dabs(a);
                const bool zero = !swk && a == (T)0;
                const T num = swk ? a : b, den = swk ? b : a;
                const T q = num / den;
                const T l = zero ? (T)0 : q;
                // the elimination with the reference's operation order (src:2550-2572)
                const bool first = lane == i + 1;
                const T g = tau * h;
                const T ls = l * sigma;
                const T rs = p - l * g;       // swap: pending row minus l * (new row)
                const T rn = g - l * p;       // no swap: new row minus l * pending row
                const T u = swk ? (first ? g + sigma : g) : p;
                const T pn = swk ? (first ? rs - ls : rs) : (first ? rn + sigma : rn);
                // U[k, c] for window columns c = k0 + lane >= k (lane i: the
                // diagonal, a or b); columns past n - 1 store nothing
                buf_st(u, uall, colb + __builtin_amdgcn_readlane(rowb, i));
                colb = lane == i ? kBufNone : colb;
                lvec = lane == i ? l : lvec;
                svec = lane == i ? (int)swk : svec;
                uvec = lane == i ? (swk ? b : a) : uvec;
                p = pn;   // lanes <= i are finished: their value is no longer read
            }
            load_t(tl, i + PD, hw[ii]);   // past the block: a clamped row, unused
        }
        // a zero pivot at step k (no swap and a == 0) is exactly a zero in
        // uvec (a swap's pivot b has |b| > |a| >= 0): the first one is info
        const unsigned long long zm = __ballot(lane < nb && uvec == (T)0);
        if (info == 0 && zm != 0ull) info = k0 + __builtin_ctzll(zm) + 1;
        const int bp = (k0 / kPanel) & 1;
        if (lane < nb) {
            lsh[bp][lane] = lvec;
            ssh[bp][lane] = svec;
            if (wr) {
                lv[k0 + lane] = lvec;
                sw[k0 + lane] = (uint8_t)svec;
                rd[k0 + lane] = (T)1 / uvec;
            }
        }
    };
    // the owner of the window writes its part of U and the block's records
    auto panel = [&](int k0, int nb, int set) __attribute__((always_inline)) {
        const bool wr = owns(k0 / kPanel);
        if (nb == kPanel) panel_nb(k0, nb, wr, set, std::true_type{});
        else panel_nb(k0, nb, wr, set, std::false_type{});
    };
    // block k0 (full, multipliers in lvec/svec) applied to window c1 = k0+kPanel,
    // whose pending values are in p; the last step's new row has its
    // diagonal (sigma) in column c1
    auto window_apply = [&](int k0, int set) __attribute__((always_inline)) {
        const int c1 = k0 + kPanel;
        const int rowb = urow_b(k0 + lane);
        const int colb =
            (owns(c1 / kPanel) && lane < kPanel && c1 + lane <= n - 1) ? ucol(c1 + lane) * ES : kBufNone;
        const T *tl = &htile[set][0][0][0];
#pragma unroll
        for (int i = 0; i < PD; i++) load_t(tl, i, hw[i]);
#pragma nounroll
        for (int ib = 0; ib < kPanel; ib += PD)
#pragma unroll
        for (int ii = 0; ii < PD; ii++) {
            const int i = ib + ii;
            const T l = rdlane(lvec, i);
            const bool swk = __builtin_amdgcn_readlane(svec, i) != 0;
            const bool first = lane == 0 && i == kPanel - 1;
            const T g = tau * hw[ii];
            const T rs = p - l * g;
            const T rn = g - l * p;
            const T u = swk ? (first ? g + sigma : g) : p;
            p = swk ? (first ? rs - l * sigma : rs) : (first ? rn + sigma : rn);
            buf_st(u, uall, colb + __builtin_amdgcn_readlane(rowb, i));
            load_t(tl, i + PD, hw[ii]);
        }
    };

    // block k0 applied to the window at column c >= k0 + 2 kPanel (a window two
    // ahead, handed over one block early): no diagonal and no stores; the
    // block's multipliers from LDS.  Each element takes the bulk's operations.
    auto window_apply_far = [&](int k0, int set) __attribute__((always_inline)) {
        const int bq = (k0 / kPanel) & 1;
        const T lvv = lsh[bq][lane];
        const int svv = ssh[bq][lane];
        const T *tl = &htile[set][NTL - 1][0][0];
#pragma unroll
        for (int i = 0; i < PD; i++) load_t(tl, i, hw[i]);
#pragma nounroll
        for (int ib = 0; ib < kPanel; ib += PD)
#pragma unroll
        for (int ii = 0; ii < PD; ii++) {
            const int i = ib + ii;
            const T l = rdlane(lvv, i);
            const bool swk = __builtin_amdgcn_readlane(svv, i) != 0;
            const T g = tau * hw[ii];
            const T rs = p - l * g;
            const T rn = g - l * p;
            p = swk ? rs : rn;
            load_t(tl, i + PD, hw[ii]);
        }
    };

    // The panel wave's decisions are the kernel's critical path (a chain of
    // dependent steps); the bulk waves beside it on its SIMD would otherwise
    // take most of the issue slots.  It runs at the highest wave priority.
    if (is_panel) __builtin_amdgcn_s_setprio(3);

    // bulk threads: stage the tiles of iteration b (into set b & 1), one row
    // per load instruction per wave (FP64: two dword halves of a row)
    typedef __attribute__((address_space(1))) void gvoid_t;
    typedef __attribute__((address_space(3))) void lvoid_t;
    const int wv = tb >> 6, nwv = NTB / 64;
    auto stage_row = [&](int set, int t, int r0, int c, int i) __attribute__((always_inline)) {
        const T *src = Ht + (int64_t)min(r0 + i, n - 1) * ldh;
        char *dst = reinterpret_cast<char *>(&htile[set][t][i][0]);
        if constexpr (sizeof(T) == 4) {
            __builtin_amdgcn_global_load_lds((gvoid_t *)(src + min(c + lane, n - 1)), (lvoid_t *)dst, 4, 0, 0);
        } else {
#pragma unroll
            for (int part = 0; part < 2; part++) {
                const int d = part * 64 + lane;   // dword d of the row
                const char *a = reinterpret_cast<const char *>(src + min(c + (d >> 1), n - 1)) + 4 * (d & 1);
                __builtin_amdgcn_global_load_lds((gvoid_t *)a, (lvoid_t *)(dst + 256 * part), 4, 0, 0);
            }
        }
    };
    auto stage_iter = [&](int b) __attribute__((always_inline)) {
        const int k0 = b * kPanel, c1 = k0 + kPanel;
        if (k0 >= nsteps || c1 >= n) return;   // no window work in iteration b
        const int set = b & 1;
        const int ntl = (FAR && GS > 1 && b >= 1) ? 3 : 2;
        for (int rr = wv; rr < ntl * kPanel; rr += nwv) {
            const int t = rr / kPanel, i = rr - t * kPanel;
            const int r0 = t == 0 ? k0 + 1 : (t == 1 ? c1 + 1 : k0 - kPanel + 1);
            stage_row(set, t, r0, c1, i);
        }
    };

    // block 0's rows (set 1, tile B) and iteration 0's tiles
    if (!is_panel) {
        if (nsteps > 0)
            for (int rr = wv; rr < kPanel; rr += nwv) stage_row(1, 1, 1, 0, rr);
        stage_iter(0);
        __builtin_amdgcn_s_waitcnt(0);   // the staged rows have landed in LDS
    }
    lds_barrier();

    if (is_panel && nsteps > 0) {
        p = win[0][lane];
        panel(0, min(kPanel, nsteps), 1);
        if (nsteps < kPanel && lane == nsteps) plast = p;   // one partial block: column n-1 is in window 0
    }
    lds_barrier();
    for (int k0 = 0; k0 < nsteps; k0 += kPanel) {
        const int bp = (k0 / kPanel) & 1;
        const int c1 = k0 + kPanel;     // window b+1
        const int c2 = c1 + kPanel;     // the bulk's first column
        if (!is_panel) stage_iter(k0 / kPanel + 1);   // the next iteration's tiles
        if (is_panel) {
            if (c1 < n) {   // block b is full
                const int w1 = c1 / kPanel;
                if (w1 < 2 || owns(w1)) {
                    p = win[bp ^ 1][lane];
                } else if constexpr (!FAR) {
                    // window b+1 of another workgroup after blocks <= b-1, from
                    // its owner (published at the end of its bulk pass of block b-1)
                    const int nval = min(kPanel, n - c1);
                    int *cnt = sp.cnt + (int64_t)s * sp.nwin + w1;
                    for (int it = 0; !failed && __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < nval;
                         it++) {
                        if (it >= kSplitSpin) failed = true;
                        __builtin_amdgcn_s_sleep(1);
                    }
                    T *src = sp.hand + ((int64_t)s * sp.nwin + w1) * kPanel;
                    p = (!failed && lane < nval)
                            ? __hip_atomic_load(src + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                            : (T)0;
                } else {
                    // window b+1 of another workgroup: its state after blocks
                    // <= b-2 (window 2: its initial values; later ones from the
                    // owner, published at the end of its bulk pass of block b-2,
                    // a block before it is needed), then block b-1 applied here
                    if (w1 == 2) {
                        const int j = c1 + lane;
                        p = (lane < kPanel && j < n) ? tau * Ht[j] : (T)0;
                    } else {
                        const int nval = min(kPanel, n - c1);
                        int *cnt = sp.cnt + (int64_t)s * sp.nwin + w1;
                        for (int it = 0;
                             !failed && __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < nval; it++) {
                            if (it >= kSplitSpin) failed = true;
                            __builtin_amdgcn_s_sleep(1);
                        }
                        T *src = sp.hand + ((int64_t)s * sp.nwin + w1) * kPanel;
                        p = (!failed && lane < nval)
                                ? __hip_atomic_load(src + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                : (T)0;
                    }
                    window_apply_far(k0 - kPanel, bp);
                }
                window_apply(k0, bp);
                if (c1 < nsteps) {
                    const int nb1 = min(kPanel, nsteps - c1);
                    panel(c1, nb1, bp);
                    // a partial last block: column n-1 is in window b+1 (a full one
                    // leaves it to the next window's pass)
                    if (c1 + nb1 == nsteps && nb1 < kPanel && lane == nb1) plast = p;
                } else if (lane == 0) {
                    plast = p;   // c1 = n-1: no step after block b
                }
            }
        } else if (c2 < n) {
            // block b on columns c2 .. n-1, instantiated per first live group
            // chunk MS = c2 / (W NTB) so that every step is straight-line code
            // (exact waits for the prefetched rows)
            auto bulk = [&](auto msc) __attribute__((always_inline)) {
                constexpr int MS = decltype(msc)::value;
                // U[k, j] at byte (ucol(j) - ucol(c2)) ES from U[k, c2]; groups
                // below c2 store nothing, and a group's elements at columns >= n
                // fall past the row
                const int uc2 = ucol(c2);
                int boff[NP];
#pragma unroll
                for (int m = MS; m < NP; m++) {
                    const int j0 = col_of(W * (tb + m * NTB));
                    boff[m] = j0 >= c2 ? (ucol(j0) - uc2) * ES : kBufNone;
                }
                const T lvv = lsh[bp][lane];
                const int svv = ssh[bp][lane];
                const int ubytes = (ucol(n - 1) - uc2 + 1) * ES;
                // row k0 + L's byte offset in lane L, read per step from lane i
                const int rowb = urow_b(k0 + lane);
                const char *uc2p = reinterpret_cast<const char *>(U + uc2);
                for (int i0 = 0; i0 < kPanel; i0 += R) {
#pragma unroll
                    for (int q = 0; q < R; q++) {
                        const int i = i0 + q;
                        const int k = k0 + i;
                        const T l = rdlane(lvv, i);
                        const int sk = __builtin_amdgcn_readlane(svv, i);
                        const BufRsrc ur = buf_rsrc(uc2p + __builtin_amdgcn_readlane(rowb, i), ubytes);
                        // values only in the two branches, then the stores: every
                        // path issues the same memory operations
                        T2 u[NP];
                        if (sk) {
#pragma unroll
                            for (int m = MS; m < NP; m++) {
                                const T2 g = tau * hs[q][m];
                                u[m] = g;
                                r[m] = r[m] - l * g;
                            }
                        } else {
#pragma unroll
                            for (int m = MS; m < NP; m++) {
                                const T2 g = tau * hs[q][m];
                                u[m] = r[m];
                                r[m] = g - l * r[m];
                            }
                        }
#pragma unroll
                        for (int m = MS; m < NP; m++) G::st(u[m], ur, boff[m]);
                        load_rows(k + R + 1, hs[q], MS);
                    }
                }
                // window b+2 = columns c2 .. c2+63 for this workgroup's panel
                // wave, and with several workgroups per shift the same window
                // for the others' (uncached slots, then a count of the values
                // written once they have landed; with FAR, window b+3, to which
                // they apply block b+1 themselves)
                if (owns(c2 / kPanel)) {
#pragma unroll
                    for (int m = MS; m < NP; m++)
#pragma unroll
                        for (int e = 0; e < W; e++) {
                            const int j = col_of(W * (tb + m * NTB)) + e;
                            if (j >= c2 && j < c2 + kPanel) win[bp][j - c2] = G::get(r[m], e);
                        }
                }
                // the window handed over: b+2 (b+3 two blocks ahead, FAR)
                const int c3 = FAR ? c2 + kPanel : c2;
                if (GS > 1 && c3 < n && owns(c3 / kPanel)) {
                    int put = 0;
                    T *dst = sp.hand + ((int64_t)s * sp.nwin + c3 / kPanel) * kPanel;
#pragma unroll
                    for (int m = MS; m < NP; m++)
#pragma unroll
                        for (int e = 0; e < W; e++) {
                            const int j = col_of(W * (tb + m * NTB)) + e;
                            if (j >= c3 && j < c3 + kPanel && j < n) {
                                __hip_atomic_store(dst + (j - c3), G::get(r[m], e), __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                                put++;
                            }
                        }
                    if (put > 0) {
                        __atomic_signal_fence(__ATOMIC_SEQ_CST);
                        __builtin_amdgcn_s_waitcnt(0);   // the values have landed
                        __atomic_signal_fence(__ATOMIC_SEQ_CST);
                        __hip_atomic_fetch_add(sp.cnt + (int64_t)s * sp.nwin + c3 / kPanel, put, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
            };
            // the first group chunk holding a column >= c2: owned windows of
            // index >= ceil((b + 2 - g) / G)
            const int wl0 = GS == 1 ? c2 / kPanel : max(0, (c2 / kPanel - g + GS - 1) / GS);
            dispatch_const<0, NP>(wl0 * kPanel / (W * NTB), bulk);
        }
        if (!is_panel) __builtin_amdgcn_s_waitcnt(0);   // the staged rows have landed in LDS
        lds_barrier();
    }

    // the last diagonal: U[n-1, n-1] = p[n-1]
    if (tid == 0 && g == 0) {
        const T pl = plast;
        U[urow(n - 1) + ucol(n - 1)] = pl;
        rd[n - 1] = (T)1 / pl;
        sw[n - 1] = 0;
        if (pl == (T)0 && info == 0) info = n;
        if (GS == 1) infobase[s] = info;
        else atomicMin(infobase + s, info);   // the reset left INT_MAX
    }
    // a hand-off that did not arrive (a bound, never reached when every
    // workgroup of the shift is resident): info -1, reported as an error
    if (GS > 1 && lane == 0 && is_panel && failed) atomicMin(infobase + s, -1);
}

// ------------------------------------------------------------- complex shifts
// Complex shift on a real H, planar: plane 0 = real parts, plane 1 = imaginary
// parts, at +ustride/2 (U) and +lstride/2 (l, rdiag).  Products follow Julia's
// term order; division and 1/x use the Baudin-Smith robust algorithm
// ("A Robust Complex Division in Scilab", 2012).
template <typename T, int NPER, int NT>
__global__ void __launch_bounds__(NT) k_shift_cplx(const T *__restrict__ Ht, int64_t ldh, int n,
                                                              const ShiftParam<T> *__restrict__ prm,
                                                              T *__restrict__ Ubase, int64_t ustride,
                                                              T *__restrict__ lbase, T *__restrict__ rdbase,
                                                              int64_t lstride, uint8_t *__restrict__ swbase,
                                                              int64_t sstride, int *__restrict__ infobase, int ut) {
#pragma clang fp contract(off)
    const int s = blockIdx.x;
    const T sr = prm[s].sr, si = prm[s].si, tr = prm[s].tr, ti = prm[s].ti;
    // U[k, j] (j >= k) at urow(k) + ucol(j) of each plane: packed rows, or
    // 64x64 tiles (the batched pool; the spare slots past the tiles hold the
    // idle threads' stores as with packed rows)
    const int nbu = (n + 63) / 64;
    constexpr int UV = utile_gw<T>();
    auto urow = [&](int k) -> int64_t { return ut ? utile_row<UV>(nbu, k) : upack_off(n, k) - k; };
    auto ucol = [&](int j) -> int64_t { return ut ? utile_col<UV>(j) : (int64_t)j; };
    const int64_t up = ustride / 2, lp = lstride / 2;
    T *Ur = Ubase + (int64_t)s * ustride, *Ui = Ur + up;
    T *dump = Ur + ustride - NT;      // spare tail of the imaginary plane
    T *lr_ = lbase + (int64_t)s * lstride, *li_ = lr_ + lp;
    T *rdr = rdbase + (int64_t)s * lstride, *rdi = rdr + lp;
    uint8_t *sw = swbase + (int64_t)s * sstride;
    const int tid = threadIdx.x;
    __shared__ T slot[2][2];
    __shared__ T ldump[2][NT];
    extern __shared__ __align__(16) unsigned char lds_dyn[];
    T *subd = reinterpret_cast<T *>(lds_dyn);     // subd[k] = H[k+1, k]
    for (int i = tid; i + 1 < n; i += NT) subd[i] = Ht[i + (int64_t)(i + 1) * ldh];

    int64_t uc[NPER];   // this thread's columns' offsets within a row
#pragma unroll
    for (int m = 0; m < NPER; m++) uc[m] = ucol(min(tid + m * NT, n - 1));
    T rr[NPER], ri[NPER];
#pragma unroll
    for (int m = 0; m < NPER; m++) {
        const int j = tid + m * NT;
        const T hv = (j < n) ? Ht[j] : (T)0;
        rr[m] = tr * hv;
        ri[m] = ti * hv;
    }
    if (tid == 0) { rr[0] = rr[0] + sr; ri[0] = ri[0] + si; }
    if (tid == 0) { slot[0][0] = rr[0]; slot[0][1] = ri[0]; }
    // rows of H three steps ahead in three register sets, the subdiagonal in
    // LDS, as in k_shift_real
    T hA[NPER], hB[NPER], hC[NPER];
    auto load_row = [&](int row, T(&hd)[NPER]) {
        const int64_t col = (int64_t)min(row, n - 1) * ldh;
#pragma unroll
        for (int m = 0; m < NPER; m++) hd[m] = Ht[min(tid + m * NT, n - 1) + col];
    };
    load_row(1, hA);
    load_row(2, hB);
    load_row(3, hC);
    __syncthreads();
    int info = 0;
    auto step = [&](int k, const T(&h)[NPER]) {
        const T hk = subd[k];
        const T ar = slot[k & 1][0], ai = slot[k & 1][1];
        const T br = tr * hk, bi = ti * hk;
        const bool sw_k = (dabs(br) + dabs(bi)) > (dabs(ar) + dabs(ai));
        const bool zero = !sw_k && ar == (T)0 && ai == (T)0;
        T qr, qi;
        cdiv_sel(sw_k ? ar : br, sw_k ? ai : bi, sw_k ? br : ar, sw_k ? bi : ai, qr, qi);
        const T lre = zero ? (T)0 : qr, lim = zero ? (T)0 : qi;
        info = (zero && info == 0) ? k + 1 : info;
        const int64_t ro = urow(k);
        const T ukr = sw_k ? br : ar, uki = sw_k ? bi : ai;
        if (tid < kWave) {   // wave 0 alone: row k's diagonal, reciprocal pivot, multiplier, swap flag
            T dr, di;
            cdiv_sel((T)1, (T)0, ukr, uki, dr, di);
            const bool t0 = tid == 0;
            *(t0 ? Ur + ro + ucol(k) : dump + tid) = ukr;
            *(t0 ? Ui + ro + ucol(k) : dump + tid) = uki;
            *(t0 ? rdr + k : dump + tid) = dr;
            *(t0 ? rdi + k : dump + tid) = di;
            *(t0 ? lr_ + k : dump + tid) = lre;
            *(t0 ? li_ + k : dump + tid) = lim;
            *(t0 ? sw + k : reinterpret_cast<uint8_t *>(dump + tid)) = (uint8_t)sw_k;
        }
        // l * sigma for the swap branch's first element
        const T qsr = lre * sr - lim * si, qsi = lre * si + lim * sr;
#pragma unroll
        for (int m = 0; m < NPER; m++) {
            if ((m + 1) * NT <= k + 1) continue;   // chunk finished for every thread (uniform)
            const int j = tid + m * NT;
            const bool active = j > k && j < n;
            const bool first = j == k + 1;
            const T xr0 = rr[m], xi0 = ri[m];
            const T gr = tr * h[m], gi = ti * h[m];
            // swap: r - l*g, then - l*sigma on the first element; U row = g (+ sigma)
            const T pr = lre * gr - lim * gi, pi = lre * gi + lim * gr;
            const T s1r = xr0 - pr, s1i = xi0 - pi;
            const T s2r = s1r - qsr, s2i = s1i - qsi;
            const T usr = first ? gr + sr : gr, usi = first ? gi + si : gi;
            // no swap: U row = r; r = g - l*r (+ sigma on the first element)
            const T pr2 = lre * xr0 - lim * xi0, pi2 = lre * xi0 + lim * xr0;
            const T n1r = gr - pr2, n1i = gi - pi2;
            const T n2r = n1r + sr, n2i = n1i + si;
            const T ur = sw_k ? usr : xr0, ui = sw_k ? usi : xi0;
            const T xr = sw_k ? (first ? s2r : s1r) : (first ? n2r : n1r);
            const T xi = sw_k ? (first ? s2i : s1i) : (first ? n2i : n1i);
            rr[m] = active ? xr : xr0;
            ri[m] = active ? xi : xi0;
            *(active ? Ur + ro + uc[m] : dump + tid) = ur;
            *(active ? Ui + ro + uc[m] : dump + tid) = ui;
            *(first ? &slot[(k + 1) & 1][0] : &ldump[0][tid]) = xr;
            *(first ? &slot[(k + 1) & 1][1] : &ldump[1][tid]) = xi;
        }
        lds_barrier();   // the slot is the only data shared between steps
    };
    int k = 0;
    for (; k + 3 <= n - 1; k += 3) {
        step(k, hA);
        load_row(k + 4, hA);
        step(k + 1, hB);
        load_row(k + 5, hB);
        step(k + 2, hC);
        load_row(k + 6, hC);
    }
    if (k < n - 1) step(k, hA);
    if (k + 1 < n - 1) step(k + 1, hB);
    if (n >= 1) {
        const int own = (n - 1) % NT, mm = (n - 1) / NT;
        if (tid == own) {
            T xr = 0, xi = 0;
#pragma unroll
            for (int m = 0; m < NPER; m++) if (m == mm) { xr = rr[m]; xi = ri[m]; }
            const int64_t o = urow(n - 1) + ucol(n - 1);
            Ur[o] = xr; Ui[o] = xi;
            T qr, qi;
            cdiv((T)1, (T)0, xr, xi, qr, qi);
            rdr[n - 1] = qr; rdi[n - 1] = qi;
            sw[n - 1] = 0;
            if (xr == (T)0 && xi == (T)0 && info == 0) info = n;
            infobase[s] = info;
        }
    }
}

template <typename T, int NPER, int NT>
Status launch_shift(lhl_handle *h, int cplx, int64_t nshift, const T *Ht, int64_t ldh, int n,
                    const ShiftParam<T> *prm, T *U, int64_t us, T *l, T *rd, int64_t ls, uint8_t *sw, int64_t ss,
                    int *info, int ut = 0) {
    const size_t lds = sizeof(T) * (size_t)std::max(n, 1);   // the subdiagonal of H
    if (cplx) {
        auto fn = k_shift_cplx<T, NPER, NT>;
        LHL_TRY(allow_dynamic_lds(fn));
        LHL_LAUNCH(h, "shift", fn, dim3((unsigned)nshift), dim3(NT), lds, Ht, ldh, n, prm, U, us, l, rd, ls, sw, ss,
                   info, ut);
    } else {
        if (ut) return {LHL_ERR_ARG, "lhl_shift: U in tiles is written by the panel kernel for real shifts"};
        auto fn = k_shift_real<T, NPER, NT>;
        LHL_TRY(allow_dynamic_lds(fn));
        LHL_LAUNCH(h, "shift", fn, dim3((unsigned)nshift), dim3(NT), lds, Ht, ldh, n, prm, U, us, l, rd, ls, sw, ss,
                   info);
    }
    return {};
}

// The pipelined panel kernel: one panel wave + kPanelBulk bulk threads, NP
// column groups per bulk thread.  FP32: single columns, 63-step blocks, the
// panel's rows 21 steps ahead; FP64: column pairs, 64-step blocks (even, so
// no pair straddles the bulk's first column), the panel's rows 16 steps ahead
// (measured at n = 4096: pairs take the FP64 shift LUs 3.74 -> 3.20 ms and
// would cost FP32 10 %).  The bulk's rows of H are prefetched 9 (FP32) or 8
// (FP64) steps ahead where the registers hold them (448 threads, up to 6 / 3
// column groups each), else 3 / 2: at n = 4096 with 64 shifts split 4 ways,
// FP32 1.71 -> 1.49 ms and FP64 2.50 -> 2.06 ms per launch against 3 / 2
// (21 / 16 steps measured no faster and spill).  448 bulk threads are faster
// than 960 up to n = 4096 (FP64 3.19 vs 3.74 ms, FP32 2.02 vs 2.25 ms, one
// workgroup per shift; split 4 ways, 1.49 vs 1.68 ms F32); beyond 448 threads'
// reach (FP64 4480, FP32 8512) the panel kernel takes 960, up to n = 13440
// (64 shifts at n = 8192 FP64: 10.6 ms against 18.2 ms in the one-step
// kernel; 32 at n = 13440: FP32 18.6 vs 55.0 ms, FP64 32.8 vs 163.4 ms).
// FP32 with 19 columns per thread at 960 spills 573 registers and loses to
// the one-step kernel (n = 16384: 78.1 vs 70.9 ms).
constexpr int kPanelBulk = 448;
constexpr int kPanelBulkWide = 960;
template <typename T> constexpr int panel_w() { return sizeof(T) == 8 ? 2 : 1; }
template <typename T> constexpr int panel_kp() { return sizeof(T) == 8 ? 64 : 63; }
template <typename T, int NP, int NTB> constexpr int panel_r() {
    if constexpr (sizeof(T) == 8) return (NTB == kPanelBulk && NP <= 3) ? 8 : 2;
    else return (NTB == kPanelBulk && NP <= 6) ? 9 : 3;
}
template <typename T, int NP, int NTB> constexpr auto panel_fn() {
    constexpr bool F64 = sizeof(T) == 8;
    return k_shift_panel<T, NP, NTB, F64 ? 16 : 21, panel_r<T, NP, NTB>(), panel_w<T>(), F64 ? 64 : 63>;
}

// Workgroups of one panel-kernel shape resident on the whole device at once
// (queried once per shape; 0 if the query fails, which turns the split off).
template <typename T, int NP, int NTB> int panel_capacity() {
    static const int cap = [] {
        int nb = 0, dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, panel_fn<T, NP, NTB>(), 64 + NTB, 0) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
            (void)hipGetLastError();
            return 0;
        }
        return nb * cus;
    }();
    return cap;
}

template <typename T, int NP, int NTB>
Status launch_panel(lhl_handle *h, int64_t nshift, const T *Ht, int64_t ldh, int n, const ShiftParam<T> *prm,
                    T *U, int64_t us, T *l, T *rd, int64_t ls, uint8_t *sw, int64_t ss, int *info,
                    const SplitArgs<T> &sp) {
    const unsigned grid = sp.G > 1 ? 8u * (unsigned)((nshift + 7) / 8) * (unsigned)sp.G : (unsigned)nshift;
    LHL_LAUNCH(h, "shift", (panel_fn<T, NP, NTB>()), dim3(grid), dim3(64 + NTB), 0, Ht, ldh, n, prm, U, us, l, rd,
               ls, sw, ss, info, sp);
    return {};
}

// f(NP, NTB) for the panel kernel's shape that holds `ncols` columns per
// workgroup: 448 bulk threads while they reach, 960 beyond (or from the
// tuning knob's n on, `wide`)
template <typename T, typename F> Status with_panel_shape(int ncols, bool wide, F &&f) {
    using std::integral_constant;
    constexpr int B = panel_w<T>() * kPanelBulk;   // columns per group chunk
    constexpr int BW = panel_w<T>() * kPanelBulkWide;
    if constexpr (sizeof(T) == 4) {
        if (wide) {
            if (ncols <= 10 * BW) return f(integral_constant<int, 10>{}, integral_constant<int, kPanelBulkWide>{});
            return f(integral_constant<int, 14>{}, integral_constant<int, kPanelBulkWide>{});
        }
        if (ncols <= B) return f(integral_constant<int, 1>{}, integral_constant<int, kPanelBulk>{});
        if (ncols <= 2 * B) return f(integral_constant<int, 2>{}, integral_constant<int, kPanelBulk>{});
        if (ncols <= 4 * B) return f(integral_constant<int, 4>{}, integral_constant<int, kPanelBulk>{});
        if (ncols <= 6 * B) return f(integral_constant<int, 6>{}, integral_constant<int, kPanelBulk>{});
        if (ncols <= 10 * B) return f(integral_constant<int, 10>{}, integral_constant<int, kPanelBulk>{});
        if (ncols <= 14 * B) return f(integral_constant<int, 14>{}, integral_constant<int, kPanelBulk>{});
        if (ncols <= 19 * B) return f(integral_constant<int, 19>{}, integral_constant<int, kPanelBulk>{});
        if (ncols <= 10 * BW) return f(integral_constant<int, 10>{}, integral_constant<int, kPanelBulkWide>{});
        return f(integral_constant<int, 14>{}, integral_constant<int, kPanelBulkWide>{});
    } else {
        if (wide) {
            if (ncols <= 3 * BW) return f(integral_constant<int, 3>{}, integral_constant<int, kPanelBulkWide>{});
            if (ncols <= 5 * BW) return f(integral_constant<int, 5>{}, integral_constant<int, kPanelBulkWide>{});
            return f(integral_constant<int, 7>{}, integral_constant<int, kPanelBulkWide>{});
        }
        if (ncols <= B) return f(integral_constant<int, 1>{}, integral_constant<int, kPanelBulk>{});
        if (ncols <= 2 * B) return f(integral_constant<int, 2>{}, integral_constant<int, kPanelBulk>{});
        if (ncols <= 3 * B) return f(integral_constant<int, 3>{}, integral_constant<int, kPanelBulk>{});
        if (ncols <= 5 * B) return f(integral_constant<int, 5>{}, integral_constant<int, kPanelBulk>{});
        if (ncols <= 3 * BW) return f(integral_constant<int, 3>{}, integral_constant<int, kPanelBulkWide>{});
        if (ncols <= 5 * BW) return f(integral_constant<int, 5>{}, integral_constant<int, kPanelBulkWide>{});
        return f(integral_constant<int, 7>{}, integral_constant<int, kPanelBulkWide>{});
    }
}

// counts of landed window values to 0, info to INT_MAX (the split kernel
// lowers it with atomicMin)
__global__ void k_split_reset(int *__restrict__ cnt, int64_t ncnt, int *__restrict__ info, int64_t nshift) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ncnt) cnt[i] = 0;
    if (i < nshift) info[i] = INT_MAX;
}

// Real shifts: the panel kernel up to n = 13440, split over G workgroups per
// shift when they can all be resident (G from 4 down; the knob shift_split
// caps it, 1 turns it off), its column groups per bulk thread and its bulk
// width chosen from the columns each workgroup holds.
// Otherwise (and complex shifts) the one-step kernels: 512 threads up to
// n = 8192, 1024 beyond.
template <typename T>
Status launch_shift_n(lhl_handle *h, int cplx, int64_t nshift, const T *Ht, int64_t ldh, int n,
                      const ShiftParam<T> *prm, T *U, int64_t us, T *l, T *rd, int64_t ls, uint8_t *sw,
                      int64_t ss, int *info, int ut = 0) {
    static_assert(sizeof(ShiftParam<T>) == 4 * sizeof(T), "ShiftParam is four T: (sr, si, tr, ti)");
    if (ut && (h->cplx || n > 8192))
        return {LHL_ERR_ARG, "lhl_shift: U in tiles is written for a real H up to n = 8192 only"};
    if (h->cplx)   // a complex H (a complex J's reduction)
        return zshift_launch<T>(h, nshift, reinterpret_cast<const T *>(prm), U, us, l, rd, ls, sw, ss, info);
    constexpr bool F32 = sizeof(T) == 4;
    constexpr int BW = panel_w<T>() * kPanelBulkWide;
    constexpr int KP = panel_kp<T>();
    // the tuning knob: the 960-thread bulk from n = tune_shift_wide on
    const bool wide = h->tune_shift_wide > 0 && n >= h->tune_shift_wide;
    if (!cplx && n <= (F32 ? 14 : 7) * BW) {
        SplitArgs<T> sp;
        sp.ut = ut;
        sp.nsh = (int)nshift;
        sp.nwin = (n + KP - 1) / KP;
        const int gmax = h->tune_shift_split > 0 ? h->tune_shift_split : kSplitDefault;
        for (int gc = std::min(gmax, 8); gc > 1; gc /= 2) {
            if (sp.nwin < 2 * gc + 2) continue;   // too few windows to share
            const int ncols = KP * ((sp.nwin + gc - 1) / gc);
            int cap = 0;
            (void)with_panel_shape<T>(ncols, wide, [&](auto np, auto ntb) -> Status {
                cap = panel_capacity<T, decltype(np)::value, decltype(ntb)::value>();
                return {};
            });
            if (8 * ((nshift + 7) / 8) * gc > cap) continue;
            const size_t hb = ((sizeof(T) * (size_t)nshift * sp.nwin * KP + 255) / 256) * 256;
            if (!uc_ensure(h, hb + sizeof(int) * (size_t)nshift * sp.nwin).ok()) {
                (void)hipGetLastError();
                break;
            }
            sp.G = gc;
            sp.hand = static_cast<T *>(h->uc.back().first);
            sp.cnt = reinterpret_cast<int *>(static_cast<char *>(h->uc.back().first) + hb);
            break;
        }
        if (sp.G > 1) {
            const int64_t nc = nshift * sp.nwin;
            LHL_LAUNCH(h, "shift_reset", k_split_reset, dim3((unsigned)((std::max<int64_t>(nc, nshift) + 255) / 256)),
                       dim3(256), 0, sp.cnt, nc, info, nshift);
        }
        const int ncols = sp.G > 1 ? KP * ((sp.nwin + sp.G - 1) / sp.G) : n;
        LHL_TRY(with_panel_shape<T>(ncols, wide, [&](auto np, auto ntb) -> Status {
            return launch_panel<T, decltype(np)::value, decltype(ntb)::value>(h, nshift, Ht, ldh, n, prm, U, us, l,
                                                                               rd, ls, sw, ss, info, sp);
        }));
        LHL_HIP_TRY(hipGetLastError());
        return {};
    }
    if ((size_t)n * sizeof(T) + 16 * 1024 > lds_limit())
        return {LHL_ERR_ARG, "lhl_shift: n too large for the LDS-resident subdiagonal of the shift kernel"};
    if (n <= 512) LHL_TRY((launch_shift<T, 1, 512>(h, cplx, nshift, Ht, ldh, n, prm, U, us, l, rd, ls, sw, ss, info, ut)));
    else if (n <= 1024) LHL_TRY((launch_shift<T, 2, 512>(h, cplx, nshift, Ht, ldh, n, prm, U, us, l, rd, ls, sw, ss, info, ut)));
    else if (n <= 2048) LHL_TRY((launch_shift<T, 4, 512>(h, cplx, nshift, Ht, ldh, n, prm, U, us, l, rd, ls, sw, ss, info, ut)));
    else if (n <= 4096) LHL_TRY((launch_shift<T, 8, 512>(h, cplx, nshift, Ht, ldh, n, prm, U, us, l, rd, ls, sw, ss, info, ut)));
    else if (n <= 8192) LHL_TRY((launch_shift<T, 16, 512>(h, cplx, nshift, Ht, ldh, n, prm, U, us, l, rd, ls, sw, ss, info, ut)));
    else if (n <= 16384) LHL_TRY((launch_shift<T, 16, 1024>(h, cplx, nshift, Ht, ldh, n, prm, U, us, l, rd, ls, sw, ss, info, ut)));
    else return {LHL_ERR_ARG, "lhl_shift: n above 16384 is not supported by the register-resident shift kernel"};
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

// ShiftParam<T> from device doubles (sr, si, tr, ti per shift): the parameter
// upload of a captured batched step, whose shifts change between replays.
template <typename T>
__global__ void k_shift_params(int64_t nshift, const double *__restrict__ src, ShiftParam<T> *__restrict__ dst) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshift) return;
    dst[s].sr = (T)src[4 * s];
    dst[s].si = (T)src[4 * s + 1];
    dst[s].tr = (T)src[4 * s + 2];
    dst[s].ti = (T)src[4 * s + 3];
}

template <typename T>
Status shift_into(lhl_handle *h, int cplx, int64_t nshift, const double *sig, const double *tau, DevBuf &params,
                  T *U, int64_t us, T *l, T *rd, int64_t ls, uint8_t *sw, int64_t ss, int *info, int ut = 0) {
    if (h->dev_shiftpar) {
        // the shifts are on the device already (a plan's step)
        LHL_TRY(params.ensure(sizeof(ShiftParam<T>) * (size_t)nshift));
        LHL_LAUNCH(h, "shift_params", k_shift_params<T>, dim3((unsigned)((nshift + 255) / 256)), dim3(256), 0, nshift,
                   h->dev_shiftpar, params.as<ShiftParam<T>>());
        ProfWork pw(h, nshift, 0);
        return launch_shift_n<T>(h, cplx, nshift, (const T *)h->Ht.as<T>(), h->lda, (int)h->n,
                                 params.as<ShiftParam<T>>(), U, us, l, rd, ls, sw, ss, info, ut);
    }
    std::vector<ShiftParam<T>> hp((size_t)nshift);
    for (int64_t s = 0; s < nshift; s++) {
        hp[s].sr = (T)sig[2 * s];
        hp[s].si = (T)sig[2 * s + 1];
        hp[s].tr = (T)tau[2 * s];
        hp[s].ti = (T)tau[2 * s + 1];
    }
    LHL_TRY(params.ensure(sizeof(ShiftParam<T>) * (size_t)nshift));
    // ordered on the stream; the host copy is kept by the handle until the
    // next call (every call ends by synchronizing its streams)
    unsigned char *keep = host_keep(h, sizeof(ShiftParam<T>) * (size_t)nshift);
    std::memcpy(keep, hp.data(), sizeof(ShiftParam<T>) * (size_t)nshift);
    LHL_HIP_TRY(hipMemcpyAsync(params.p, keep, sizeof(ShiftParam<T>) * (size_t)nshift, hipMemcpyHostToDevice,
                               h->stream));
    ProfWork pw(h, nshift, 0);
    return launch_shift_n<T>(h, cplx, nshift, (const T *)h->Ht.as<T>(), h->lda, (int)h->n,
                             params.as<ShiftParam<T>>(), U, us, l, rd, ls, sw, ss, info, ut);
}

}  // namespace

Status shift_pool_device(lhl_handle *h, ShiftPool &pool, int64_t nshift, const double *sig, const double *tau,
                         int is_complex) {
    const int64_t n = h->n;
    const size_t es = elt_size(h->dtype);
    pool.n = n;
    pool.count = nshift;
    pool.planes = is_complex ? 2 : 1;
    pool.ut = pool_tiles(h, n, is_complex) ? 1 : 0;
    const int64_t uplane = upack_plane(n);
    // tiles: complex shifts keep kShiftThreads spare slots past each plane's
    // tiles (the idle threads' stores of the one-step kernel)
    const int64_t tplane = ((utile_plane(n) + (is_complex ? kShiftThreads : 0) + 63) / 64) * 64;
    pool.ustride = (pool.ut ? tplane : uplane) * pool.planes;
    const int64_t lplane = ((n + 63) / 64) * 64;
    LHL_TRY(pool.U.ensure(es * (size_t)pool.ustride * (size_t)nshift));
    LHL_TRY(pool.l.ensure(es * (size_t)lplane * pool.planes * (size_t)nshift));
    LHL_TRY(pool.rdiag.ensure(es * (size_t)lplane * pool.planes * (size_t)nshift));
    LHL_TRY(pool.swp.ensure((size_t)lplane * (size_t)nshift));
    LHL_TRY(pool.dinfo.ensure(sizeof(int) * (size_t)nshift));
    if (h->dtype == LHL_F64)
        LHL_TRY(shift_into<double>(h, is_complex, nshift, sig, tau, pool.params, pool.U.as<double>(), pool.ustride,
                                   pool.l.as<double>(), pool.rdiag.as<double>(), lplane * pool.planes,
                                   pool.swp.as<uint8_t>(), lplane, pool.dinfo.as<int>(), pool.ut));
    else
        LHL_TRY(shift_into<float>(h, is_complex, nshift, sig, tau, pool.params, pool.U.as<float>(), pool.ustride,
                                  pool.l.as<float>(), pool.rdiag.as<float>(), lplane * pool.planes,
                                  pool.swp.as<uint8_t>(), lplane, pool.dinfo.as<int>(), pool.ut));
    int64_t uis = 0;
    return uinv_device(h, pool.U.p, pool.ustride, pool.rdiag.p, lplane * pool.planes, nshift, is_complex, pool.Uinv,
                       &uis, pool.UT, pool.ut);
}

// One shift state (ws.shift or an LHLShift): storage is the pool layout with
// a single entry, so the solves read both the same way.
Status shift_device(lhl_handle *h, lhl_shift_state *sh, int64_t /*nshift*/, const double *sig, const double *tau,
                    int64_t *info_out) {
    const int64_t n = h->n;
    const size_t es = elt_size(h->dtype);
    const int planes = sh->is_complex ? 2 : 1;
    const int64_t uplane = upack_plane(n);
    const int64_t lplane = ((n + 63) / 64) * 64;
    LHL_TRY(sh->U.ensure(es * (size_t)uplane * planes));
    LHL_TRY(sh->l.ensure(es * (size_t)lplane * planes));
    LHL_TRY(sh->rdiag.ensure(es * (size_t)lplane * planes));
    LHL_TRY(sh->swp.ensure((size_t)lplane));
    LHL_TRY(sh->dinfo.ensure(64));
    sh->n = n;
    sh->planes = planes;
    sh->dtype = h->dtype;
    DevBuf params;
    Status st;
    if (n > 0) {
        if (h->dtype == LHL_F64)
            st = shift_into<double>(h, sh->is_complex, 1, sig, tau, params, sh->U.as<double>(), uplane * planes,
                                    sh->l.as<double>(), sh->rdiag.as<double>(), lplane * planes,
                                    sh->swp.as<uint8_t>(), lplane, sh->dinfo.as<int>());
        else
            st = shift_into<float>(h, sh->is_complex, 1, sig, tau, params, sh->U.as<float>(), uplane * planes,
                                   sh->l.as<float>(), sh->rdiag.as<float>(), lplane * planes,
                                   sh->swp.as<uint8_t>(), lplane, sh->dinfo.as<int>());
    }
    if (!st.ok()) { params.release(); return st; }
    if (n > 0) {
        int64_t uis = 0;
        st = uinv_device(h, sh->U.p, uplane * planes, sh->rdiag.p, lplane * planes, 1, sh->is_complex, sh->Uinv, &uis,
                         sh->UT);
        if (!st.ok()) { params.release(); return st; }
    }
    int info = 0;
    if (n > 0) {
        hipError_t e = hipMemcpyAsync(&info, sh->dinfo.p, sizeof(int), hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        params.release();
        if (e != hipSuccess) return {LHL_ERR_HIP, std::string("lhl_shift: ") + hipGetErrorString(e)};
        if (info < 0) return {LHL_ERR_HIP, "lhl_shift: a split shift LU's window hand-off did not arrive"};
    }
    sh->info = info;
    sh->sigma[0] = sig[0]; sh->sigma[1] = sig[1];
    sh->tau[0] = tau[0]; sh->tau[1] = tau[1];
    sh->loaded = 1;
    if (info_out) *info_out = info;
    return {};
}

}  // namespace lhl
