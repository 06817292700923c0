// Internal declarations shared by the HIP translation units of liblhl_hip.so.
// Everything here is gfx950 (CDNA4) code: wave64, 160 KB LDS per CU.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <algorithm>
#include <string>
#include <vector>

#include "../../include/lhl_hip.h"

namespace lhl {

// A workgroup barrier that orders LDS only: s_waitcnt lgkmcnt(0) + s_barrier.
// Outstanding global loads (prefetches) and stores stay in flight across it;
// use it only where no thread reads global memory another thread of the same
// workgroup wrote in the kernel.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Fused multiply-add in the operands' own precision: a single-rounding FP32
// FMA for floats (__builtin_fma alone would promote floats to double).
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Wave-wide sum by the xor butterfly v += v[lane ^ o], o = 32, 16, ..., 1 (the
// pairing of a __shfl_xor loop, so bit-identical to it), with one LDS-crossbar
// permute for o = 32, swizzles (bitmask mode, xor within 32 lanes) for 16, 8,
// 4 and DPP quad permutes for 2, 1.  Call with every lane of the wave active.
namespace wsum_detail {
template <int CTRL> __device__ __forceinline__ float dpp(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int CTRL> __device__ __forceinline__ double dpp(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
template <int PAT> __device__ __forceinline__ float swz(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(__builtin_bit_cast(int, v), PAT));
}
template <int PAT> __device__ __forceinline__ double swz(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_ds_swizzle((int)b, PAT);
    const int hi = __builtin_amdgcn_ds_swizzle((int)(b >> 32), PAT);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float lane_read(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ double lane_read(double v, int l) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
}  // namespace wsum_detail

template <typename T> __device__ __forceinline__ T wave_sum_fast(T v) {
    using namespace wsum_detail;
    v = v + __shfl_xor(v, 32, 64);
    v = v + swz<(16 << 10) | 0x1F>(v);
    v = v + swz<(8 << 10) | 0x1F>(v);
    v = v + swz<(4 << 10) | 0x1F>(v);
    v = v + dpp<0x4E>(v);     // quad_perm [2,3,0,1]: lane ^ 2
    v = v + dpp<0xB1>(v);     // quad_perm [1,0,3,2]: lane ^ 1
    return v;
}

// Raw buffer access.  A resource covers `bytes` bytes from `base`; the
// hardware's range check compares the per-lane offset with that size, so a
// load at or beyond it returns 0 and a store there is dropped.  Callers pass
// offsets in [0, bytes) or at/above it, never negative: kBufNone (2^31) marks
// a lane with nothing to load or store.  The offset goes through an empty asm
// so the compiler cannot move a constant part of it into the instruction's
// immediate offset: the range check sees the per-lane register as given, and a
// split that leaves it negative turns an in-range store into a dropped one.
// Word 3 = 0x00020000: 32-bit data format, no swizzle, no thread-id
// addressing.  The scalar offset is 0.
using BufRsrc = __amdgpu_buffer_rsrc_t;
constexpr int kBufNone = (int)0x80000000u;
__device__ __forceinline__ BufRsrc buf_rsrc(const void *base, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, bytes, 0x00020000);
}
__device__ __forceinline__ int buf_opaque(int off) {
    asm volatile("" : "+v"(off));
    return off;
}
template <typename T> __device__ __forceinline__ T buf_ld(BufRsrc r, int off);
template <> __device__ __forceinline__ float buf_ld<float>(BufRsrc r, int off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, buf_opaque(off), 0, 0));
}
template <> __device__ __forceinline__ double buf_ld<double>(BufRsrc r, int off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, buf_opaque(off), 0, 0));
}
__device__ __forceinline__ void buf_st(float v, BufRsrc r, int off) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, buf_opaque(off), 0, 0);
}
__device__ __forceinline__ void buf_st(double v, BufRsrc r, int off) {
    using U2 = unsigned __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, buf_opaque(off), 0, 0);
}
// The same with the non-temporal cache policy (aux bit 1, "nt" on gfx950), for
// data written once that no cache holds until it is read back.
__device__ __forceinline__ void buf_st_nt(float v, BufRsrc r, int off) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, buf_opaque(off), 0, 2);
}
__device__ __forceinline__ void buf_st_nt(double v, BufRsrc r, int off) {
    using U2 = unsigned __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, buf_opaque(off), 0, 2);
}

// Two consecutive elements per lane (one 8- or 16-byte access; the range check
// applies per dword, and the callers' offsets are never negative).  The store
// is non-temporal, as buf_st_nt.
template <typename T> using Vec2 = T __attribute__((ext_vector_type(2)));
template <typename T> __device__ __forceinline__ Vec2<T> buf_ld2(BufRsrc r, int off);
template <> __device__ __forceinline__ Vec2<float> buf_ld2<float>(BufRsrc r, int off) {
    return __builtin_bit_cast(Vec2<float>, __builtin_amdgcn_raw_buffer_load_b64(r, buf_opaque(off), 0, 0));
}
template <> __device__ __forceinline__ Vec2<double> buf_ld2<double>(BufRsrc r, int off) {
    return __builtin_bit_cast(Vec2<double>, __builtin_amdgcn_raw_buffer_load_b128(r, buf_opaque(off), 0, 0));
}
__device__ __forceinline__ void buf_st2(Vec2<float> v, BufRsrc r, int off) {
    using U2 = unsigned __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, buf_opaque(off), 0, 0);
}
__device__ __forceinline__ void buf_st2(Vec2<double> v, BufRsrc r, int off) {
    using U4 = unsigned __attribute__((ext_vector_type(4)));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(U4, v), r, buf_opaque(off), 0, 0);
}
__device__ __forceinline__ void buf_st2_nt(Vec2<float> v, BufRsrc r, int off) {
    using U2 = unsigned __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, buf_opaque(off), 0, 2);
}
__device__ __forceinline__ void buf_st2_nt(Vec2<double> v, BufRsrc r, int off) {
    using U4 = unsigned __attribute__((ext_vector_type(4)));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(U4, v), r, buf_opaque(off), 0, 2);
}

constexpr int kWave = 64;
constexpr int kMaxLds = 160 * 1024;   // bytes of LDS one workgroup may hold on gfx950

// LDS one workgroup may use on the current device (160 KB on gfx950),
// queried once (thread-safe static initialisation).
inline size_t lds_limit() {
    static const int cached = [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = 64 * 1024;
        return v < kMaxLds ? v : kMaxLds;
    }();
    return (size_t)cached;
}

// Status carried out of internal helpers; the ABI layer stores the message.
struct Status {
    int code = LHL_OK;
    std::string msg;
    bool ok() const { return code == LHL_OK; }
};

#define LHL_HIP_TRY(expr)                                                              \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ::lhl::Status{LHL_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
    } while (0)

// Lets kernel `fn` take all the LDS a workgroup may hold beyond its static
// LDS as dynamic LDS.  Every caller sets the same value, so handles driven
// from other host threads cannot lower it between this call and a launch.
template <typename F>
inline Status allow_dynamic_lds(F fn) {
    hipFuncAttributes a{};
    LHL_HIP_TRY(hipFuncGetAttributes(&a, (const void *)fn));
    const size_t lim = lds_limit();
    const int dyn = (int)(lim > a.sharedSizeBytes ? lim - a.sharedSizeBytes : 0);
    LHL_HIP_TRY(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
    return {};
}

#define LHL_TRY(expr)                     \
    do {                                  \
        ::lhl::Status s_ = (expr);        \
        if (!s_.ok()) return s_;          \
    } while (0)

inline size_t elt_size(int dtype) { return dtype == LHL_F64 ? 8 : 4; }

// Leading dimension of a device matrix of n rows: a multiple of 64 elements so
// every column starts on a 256/512-byte boundary, bumped by 64 when the column
// stride is a multiple of 16 KiB (keeps neighbouring columns off the same HBM
// channel group).
inline int64_t pad_ld(int64_t n, int dtype) {
    int64_t ld = ((n + 63) / 64) * 64;
    if (ld == 0) ld = 64;
    if ((ld * (int64_t)elt_size(dtype)) % 16384 == 0) ld += 64;
    return ld;
}

// Offset of row k of the packed upper-triangular U (row k holds U[k, k..n-1]).
__host__ __device__ inline int64_t upack_off(int64_t n, int64_t k) { return k * n - (k * (k - 1)) / 2; }
inline int64_t upack_len(int64_t n) { return n * (n + 1) / 2; }
// Elements per plane of one shift's U: the packed triangle plus 1024 spare
// elements at the end (the complex one-step shift kernel's dump slots, one per
// thread) and 4n for the column padding of the same U in packed columns
// (below), rounded to 64.
inline int64_t upack_plane(int64_t n) { return ((upack_len(n) + 1024 + 4 * n + 63) / 64) * 64; }

// U in packed columns for the back substitution: column j holds U[0:j+1, j]
// from ucol_off<V>(j), a multiple of V, so V consecutive rows of a column are
// one aligned 16-byte load (V = 4 for FP32, V = 2 for FP64).  Column j takes
// round_up(j+1, V) slots; the total over n columns is at most
// n(n+1)/2 + (V-1) n.
template <typename T> constexpr int ut_vec() { return sizeof(T) == 4 ? 4 : 2; }
template <int V> __host__ __device__ inline int64_t ucol_off(int64_t j) {
    const int64_t a = j / V, b = j % V;
    return V * ((V * a * (a + 1)) / 2 + b * (a + 1));
}

// U of a real shift in 64x64 tiles (the batched solves' pool, n <= 8192):
// the upper triangle's tiles (I, J), J >= I, one after another, tile row I
// holding J = I .. nbu-1, each tile row-major.  The shift LU's row stores stay
// whole 256-byte runs per tile, and the U solve reads a tile (one 64-column
// block of 64 rows) as one contiguous piece.  Element (k, j), j >= k, sits at
// utile_row<V>(nbu, k) + utile_col<V>(j) with V = utile_gw<T>() = 64; the
// formulas also describe tiles of V-column granules stored granule column by
// granule column ((64 g + r) V), which read better in the U solve and wrote
// worse in the shift LU (FP32 with 16-byte granules: 118.5 k against 121.4 k
// solves/s on the 4096 x 64 shifts x 16 batch).
template <typename T> constexpr int utile_gw() { return 64; }
template <int V> __host__ __device__ inline int64_t utile_row(int64_t nbu, int64_t k) {
    const int64_t I = k >> 6;
    return (I * nbu - (I * (I - 1)) / 2 - I) * 4096 + (k & 63) * V;
}
template <int V> __host__ __device__ inline int64_t utile_col(int64_t j) {
    return ((j >> 6) << 12) + ((j & 63) / V) * (64 * V) + (j & (V - 1));
}
inline int64_t utile_plane(int64_t n) {
    const int64_t nbu = (n + 63) / 64;
    return nbu * (nbu + 1) / 2 * 4096;
}

// Device blocks kept by a handle between calls (per-call buffers of the
// batched solves: the shift pool is gigabytes, and hipMalloc/hipFree of it on
// every call costs milliseconds).  A block is handed back only after the
// handle's stream has drained, so a later call on the same stream may reuse it.
struct BlockCache {
    std::vector<std::pair<void *, size_t>> blocks;
    size_t bytes = 0;
    static constexpr size_t kCap = size_t(48) << 30;   // beyond this, give memory back to the device
    // the smallest cached block of at least `want` bytes, or nullptr
    void *take(size_t want, size_t *got) {
        size_t best = blocks.size();
        for (size_t i = 0; i < blocks.size(); i++)
            if (blocks[i].second >= want && (best == blocks.size() || blocks[i].second < blocks[best].second)) best = i;
        if (best == blocks.size()) return nullptr;
        void *p = blocks[best].first;
        *got = blocks[best].second;
        bytes -= *got;
        blocks.erase(blocks.begin() + (long)best);
        return p;
    }
    void give(void *p, size_t n) {
        blocks.push_back({p, n});
        bytes += n;
        while (bytes > kCap && !blocks.empty()) {   // oldest first
            (void)hipFree(blocks.front().first);
            bytes -= blocks.front().second;
            blocks.erase(blocks.begin());
        }
    }
    void clear() {
        for (auto &b : blocks) (void)hipFree(b.first);
        blocks.clear();
        bytes = 0;
    }
};

// Growable device allocation.  With `cache` set, blocks come from and return
// to that cache instead of hipMalloc/hipFree.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    BlockCache *cache = nullptr;
    Status ensure(size_t want) {
        if (want <= bytes && p) return {};
        release();
        if (want == 0) want = 256;
        if (cache) {
            size_t got = 0;
            if ((p = cache->take(want, &got)) != nullptr) { bytes = got; return {}; }
        }
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess && cache && !cache->blocks.empty()) {   // out of memory: return the cache and retry
            (void)hipGetLastError();
            cache->clear();
            e = hipMalloc(&p, want);
        }
        if (e != hipSuccess) { p = nullptr; return {LHL_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e)}; }
        bytes = want;
        return {};
    }
    void release() {
        if (p) {
            if (cache) cache->give(p, bytes);
            else (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

// Control words of the reduction's step loop (device memory, ping-pong by step
// parity so a step's consumers never race its producer).
template <typename T>
struct RedCtl {
    int p;        // pivot row of the step (0-based)
    int skip;     // 1 when the pivot is zero: the step's updates are skipped
    int nchunk;   // column chunks of the trailing GEMV partials
    int pad;
    T piv;        // the pivot value A[p, k] (post-interchange A[k+1, k])
    T pk;         // post-interchange A[k+1, k+1]
};

// Opt-in kernel timing: while on, every launch through LHL_LAUNCH is
// bracketed by two HIP events on the handle's stream; durations are summed per
// kernel family when read.
// Each launch also carries the work it was given (`ws` shifts, `wc` internal
// right-hand-side columns; both 0 where a launch does not split that way), so
// per-launch algorithmic bytes can be priced from what the launch processed.
struct ProfRec {
    const char *name;
    hipEvent_t a, b;
    int64_t ws, wc;
};
struct ProfAcc {
    std::string name;
    int64_t launches = 0;
    double ms = 0.0;
    int64_t ws = 0, wc = 0;   // summed work tags
};
struct Profiler {
    bool on = false;
    std::vector<ProfRec> pending;
    std::vector<ProfAcc> acc;
    int64_t ws = 0, wc = 0;   // work tag of the launches being issued
};

}  // namespace lhl

// The shift state (LHLShift{TG}, src:84-96): the packed U of sigma*I + tau*H,
// the multipliers, interchange flags, reciprocal pivots, one plane per real
// component.
struct lhl_shift_state {
    lhl_handle *h = nullptr;
    int64_t n = 0;
    int is_complex = 0;
    int dtype = LHL_F64;
    int planes = 1;
    lhl::DevBuf U, l, rdiag, swp, dinfo, Uinv, UT;   // UT: U in packed columns (the back substitution)
    double sigma[2] = {0, 0}, tau[2] = {0, 0};
    int64_t info = 0;
    int loaded = 0;   // a shift has been factored for the current n
};

// The workspace (LHLWorkspace{T,Tr,TG}, src:144-162).
struct lhl_handle {
    int64_t n = 0;
    int dtype = LHL_F64;        // element type T of the reduction and factors
    int shift_complex = 0;      // element type of ws.shift
    int device = 0;
    hipStream_t stream = nullptr;
    // a second stream (default priority) and two events: the first solve pass's
    // shift-independent part runs there while the shift LUs run on `stream`
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_h2d[2] = {nullptr, nullptr};   // the host-J upload's two staging halves
    void *blas_stream = nullptr;    // the stream the rocBLAS handle is set to
    // pinned host memory for the batched solve's small transfers (zero-pivot
    // flags, per-column shifts, residual norms), grown on demand
    void *hpin = nullptr;
    size_t hpin_bytes = 0;
    int64_t lda = 0;            // leading dimension of A and Ht
    int64_t ldj = 0;            // leading dimension of the resident J
    int j_dtype = LHL_F64;      // element type of the resident J
    int reduced = 0;
    int cplx = 0;               // a complex J (LHLWorkspace{Complex{T}}): A, Ht and J hold two planes each
    lhl::DevBuf A, Ht, J, scale64, iscale64, scaleT, iscaleT, ipiv, perm, iperm;
    lhl::DevBuf part, cand, ctl, counter, flag;   // reduction scratch
    lhl::DevBuf rowA, gcol, ilog;   // row-resident step loop: A row-major, final column k, row -> logical index
    // blocked reduction: per-row panel records (ping-pong), side slots of pushed
    // columns, active rows (two halves), panel multipliers by logical row, U12
    lhl::DevBuf brec, bpb, bact, blc, bu12;
    lhl::DevBuf grp;            // a grouped reduction's per-matrix table (this handle leading the group)
    // tuning knobs (0 = the built-in default): the n from which the blocked
    // reduction runs, its panel width and the fewest rows per step workgroup
    int64_t tune_blocked_min = 0;
    int tune_nb = 0, tune_rmin = 0;
    int tune_switch = 0;   // step (thousandths of n) where the row-resident loop takes over
    int tune_lsweep = -1;  // halvings of the L sweeps' triangular GEMMs (-1: default, 2)
    int64_t tune_shift_wide = 0;   // n from which the shift LU's bulk is 960 threads wide (0: default)
    int tune_shift_split = 0;      // most workgroups per real shift LU (0: default 4, 1: one per shift)
    int tune_utrsm_split = 0;      // most workgroups per 16 columns of a real shift's U solve (0: default 8,
                                   // 1: the kernel with one workgroup per 4 columns)
    // uncached device blocks (the split shift LU's hand-off slots); the last is
    // in use, the others are kept while the handle lives
    std::vector<std::pair<void *, size_t>> uc;
    int64_t tune_pool_bytes = 0;   // bound on one chunk's packed U's in a batched solve (0: default)
    int64_t tune_h2d_chunk = -1;   // bytes per staged chunk of a host J's upload (-1: default, 0: runtime copy)
    int tune_gemm = -1;            // the solves' dense products: 0 rocBLAS, 1 the library's MFMA kernels (-1: default)
    int64_t tune_group_bytes = 0;  // a grouped reduction's sub-group budget of row-major bytes (0: default)
    int lp1_ready = 0;             // Lp1 holds I + L' of the current reduction (the rocBLAS sweep's operand)
    const double *dev_shiftpar = nullptr;   // while set: the batched shift LUs read (sr, si, tr, ti) from here
    lhl::BlockCache cache;          // device blocks of the batched solves, reused between calls
    // host sources of asynchronous uploads (the solves' work tables, the
    // shifts' parameters): each upload reads its own block, kept until the
    // next public call begins (every call ends by synchronizing its streams)
    std::vector<std::vector<unsigned char>> host_kept;
    // pinned host arena for host_keep (asynchronous uploads stay asynchronous):
    // bump-allocated within a call, sized between calls to the last call's use
    unsigned char *pin_arena = nullptr;
    size_t pin_cap = 0, pin_used = 0, pin_want = 0;
    lhl::DevBuf Linv;           // inverses of the 64x64 diagonal blocks of L'
    lhl::DevBuf Lfull;          // the explicit inverse of L' (solve), dense unit lower, leading dimension ldlf
    lhl::DevBuf Lp1;            // I + L' as a dense unit lower matrix (solve), leading dimension ldlf
    int64_t ldlf = 0;
    lhl::DevBuf ybuf, rbuf, gbuf, xstage, bstage;  // solve / refine scratch
    lhl::Profiler prof;
    void *blas = nullptr;       // rocBLAS handle on `stream` (the residual's plain GEMM), created on first use
    lhl_shift_state *own = nullptr;   // ws.shift
    std::vector<lhl_shift_state *> held; // LHLShift states created on this handle
    std::string err;
};

namespace lhl {

struct ProfScope {
    lhl_handle *h;
    const char *name;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(lhl_handle *h_, const char *n_) : h(h_), name(n_) {
        if (h->prof.on && hipEventCreate(&a) == hipSuccess) (void)hipEventRecord(a, h->stream);
    }
    ~ProfScope() {
        if (a && hipEventCreate(&b) == hipSuccess) {
            (void)hipEventRecord(b, h->stream);
            h->prof.pending.push_back(ProfRec{name, a, b, h->prof.ws, h->prof.wc});
        }
    }
};

// Sets the work tag of the launches issued while it lives (restored after).
struct ProfWork {
    lhl_handle *h;
    int64_t ows, owc;
    ProfWork(lhl_handle *h_, int64_t shifts, int64_t cols) : h(h_), ows(h_->prof.ws), owc(h_->prof.wc) {
        h->prof.ws = shifts;
        h->prof.wc = cols;
    }
    ~ProfWork() {
        h->prof.ws = ows;
        h->prof.wc = owc;
    }
};

// A host block of `bytes` bytes that stays valid until the handle's next
// public call: the source of an asynchronous host-to-device copy.
// From the handle's pinned arena when it has room (a copy from pinned memory
// is truly asynchronous; one from pageable memory is staged by the runtime),
// else pageable; the arena grows between calls to what the last call used.
inline unsigned char *host_keep(lhl_handle *h, size_t bytes) {
    const size_t b = ((std::max<size_t>(bytes, 1) + 255) / 256) * 256;
    if (h->pin_arena && h->pin_used + b <= h->pin_cap) {
        unsigned char *p = h->pin_arena + h->pin_used;
        h->pin_used += b;
        return p;
    }
    h->pin_want = std::max(h->pin_want, h->pin_used + b);
    h->host_kept.emplace_back(std::max<size_t>(bytes, 1));
    return h->host_kept.back().data();
}
// Drops the blocks of the previous call (which ended synchronized).
// (The streams are drained first in case an earlier call left by an error.)
inline void host_keep_reset(lhl_handle *h) {
    if (h->host_kept.empty() && h->pin_used == 0) return;
    (void)hipStreamSynchronize(h->stream);
    if (h->stream2) (void)hipStreamSynchronize(h->stream2);
    h->host_kept.clear();
    h->pin_used = 0;
    if (h->pin_want > h->pin_cap) {   // the arena for the next calls
        if (h->pin_arena) (void)hipHostFree(h->pin_arena);
        h->pin_arena = nullptr;
        h->pin_cap = 0;
        void *p = nullptr;
        const size_t want = std::max<size_t>(2 * h->pin_want, size_t(64) << 10);
        if (hipHostMalloc(&p, want, hipHostMallocDefault) == hipSuccess) {
            h->pin_arena = static_cast<unsigned char *>(p);
            h->pin_cap = want;
        } else {
            (void)hipGetLastError();
        }
        h->pin_want = 0;
    }
}

#define LHL_LAUNCH(h, name, kern, grid, block, lds, ...)                            \
    do {                                                                           \
        ::lhl::ProfScope ps_((h), (name));                                         \
        hipLaunchKernelGGL(kern, grid, block, lds, (h)->stream, __VA_ARGS__);      \
    } while (0)

// ---- launchers implemented in the translation units ----
Status reduce_device(lhl_handle *h, int balance);          // lhl_reduce.hip
// many handles' reductions, grouped by size class (lhl_reduce.hip)
Status reduce_group_device(lhl_handle *const *hs, int64_t count, int balance);
// the steps of the blocked reduction on rowA (pitch ldb); nb 8 or 16
// from step kswitch (rounded up to a panel boundary) the caller's row-resident
// loop takes over: *kstart receives that step (n - 2 when the panels finished)
Status blocked_device(lhl_handle *h, int64_t ldb, int nb, int rmin, int kswitch, int *kstart);   // lhl_blocked.hip
Status balance_device(lhl_handle *h);                      // lhl_reduce.hip
// uncached device memory of at least `bytes` in h->uc.back() (lhl_shift.hip),
// and the handle's blocks back to the process-wide list (at destroy)
Status uc_ensure(lhl_handle *h, size_t bytes);
void uc_release(lhl_handle *h);
// a complex J (lhl_cplx.hip): the reduction of the handle's planar J, the
// shift LUs of its complex H (prm: sigma_re, sigma_im, tau_re, tau_im per
// shift, in T), and the sweeps y_out <- L^-1 y_in (fwd) or L y_in of its
// complex L on internal columns (q = 2c real plane, 2c + 1 imaginary)
Status zreduce_device(lhl_handle *h, int balance);
template <typename T>
Status zshift_launch(lhl_handle *h, int64_t nshift, const T *prm, T *U, int64_t us, T *l, T *rd, int64_t ls,
                     uint8_t *sw, int64_t ss, int *info);
template <typename T>
Status zsweep_device(lhl_handle *h, int fwd, const T *Yin, T *Yout, int64_t ldy, int64_t ncol);
// applyZ (inverse 0) / applyZinv (1) of a complex J on staged planar columns (lhl_batch.hip)
Status zapplyz_device(lhl_handle *h, int inverse, void *x, int64_t ncol, int64_t ldx, int64_t xplane, int xdt);
Status shift_device(lhl_handle *h, lhl_shift_state *sh, int64_t nshift, const double *sig, const double *tau,
                    int64_t *info_out);                    // lhl_shift.hip (one state)

// Batched shift LU into a pool of states laid out back to back.
struct ShiftPool {
    int64_t n = 0, count = 0;
    int planes = 1;
    int ut = 0;            // U in 64x64 tiles (utile_row / utile_col) instead of packed rows
    int64_t ustride = 0;   // elements between consecutive U's (per plane stride = ustride/planes)
    DevBuf U, l, rdiag, swp, dinfo, params, Uinv, UT;
};
Status shift_pool_device(lhl_handle *h, ShiftPool &pool, int64_t nshift, const double *sig,
                         const double *tau, int is_complex);

// The solve: x (device, element type xdt, nplanes planes of n x ncol, ld ldx
// per plane) <- W^-1 x, with U from `uptr` etc.  accumulate = 1 adds the
// solution to xout instead of overwriting it.
struct SolveArgs {
    const void *U, *l, *rdiag;
    const void *Uinv = nullptr;   // inverses of U's 64x64 diagonal blocks (per shift, uistride apart)
    const void *UT = nullptr;     // U in packed columns, ustride apart (nullptr: solve from the packed rows)
    int ut = 0;                   // U is in 64x64 tiles (read by the row-split U solve; UT unused)
    int64_t uistride = 0;
    const uint8_t *swp;
    int64_t ustride = 0;          // per shift, elements (both planes)
    int64_t lstride = 0;          // per shift, elements of l / rdiag (both planes)
    int64_t sstride = 0;          // per shift, swap bytes
    int shift_complex = 0;
    int64_t nshift = 1;
    const int64_t *col_begin = nullptr;   // host: nshift+1 column offsets (nullptr: one shift, all cols)
    int64_t ncol = 0;             // total right-hand-side columns
    int x_complex = 0;            // columns hold complex values (planar in device memory)
    int xdt = LHL_F64;            // element type of x in memory
    const void *xin = nullptr;    // input columns (planar: re plane then im plane at +xplane)
    void *xout = nullptr;
    int64_t ldx = 0, xplane = 0;
    int accumulate = 0;
    // 0: the whole solve; 1: only the gather and the L solve (no shift state
    // read); 2: the rest, after a phase-1 call on the same columns
    int phase = 0;
    // device flags per right-hand-side column and per shift (nullptr: all):
    // columns with cact[c] == 0 and shifts with sact[s] == 0 are skipped (the
    // refinement's converged ones); full_products = 1 runs the dense products
    // over every column anyway (a round where most columns are still active)
    const int *cact = nullptr;
    const int *sact = nullptr;
    int full_products = 0;
    // a caller's own buffers (a plan's, whose graph holds their addresses):
    // the back substitution's group table on the device (ngroups entries, as
    // build_groups lays them out) and the two work buffers
    const void *groups = nullptr;
    int64_t ngroups = 0;
    // and the row-split back substitution's table (nwide entries of up to 16
    // internal columns of one shift, after the ngroups entries)
    int64_t nwide = 0;
    DevBuf *ybuf = nullptr, *rbuf = nullptr;
};
Status solve_device(lhl_handle *h, const SolveArgs &a);    // lhl_batch.hip
// the back substitution's group tables for nshift shifts with column offsets
// cb (host, nshift + 1) and planes per column, uploaded into `out`
// (synchronously): *ngroups entries of up to 4 columns, then *nwide of up to 16
Status groups_device(lhl_handle *h, int64_t nshift, const int64_t *cb, int planes, DevBuf &out, int64_t *ngroups,
                     int64_t *nwide);
// after a shift LU: the inverses of the 64x64 diagonal blocks of nshift packed
// U's (per-shift stride returned in *uistride) and U in packed columns (UT,
// ustride apart)
Status uinv_device(lhl_handle *h, const void *U, int64_t ustride, const void *rdiag, int64_t lstride,
                   int64_t nshift, int is_complex, DevBuf &Uinv, int64_t *uistride, DevBuf &UT,
                   int ut = 0);   // lhl_batch.hip (ut: U in 64x64 tiles, no packed columns made)
// whether a batched solve's shift pool keeps U in 64x64 tiles: real shifts
// with 256 <= n <= 8192, where the shift LU writes them and the row-split U
// solve reads them (the knob utrsm_split = 1 keeps packed rows)
bool pool_tiles(const lhl_handle *h, int64_t n, int is_complex);   // lhl_batch.hip
// the inverses of L's 64x64 diagonal blocks, formed once per reduction
Status linv_device(lhl_handle *h);                          // lhl_batch.hip
Status applyz_device(lhl_handle *h, int inverse, void *x, int64_t ncol, int64_t ldx, int64_t xplane,
                     int xdt, int x_complex);              // lhl_solve.hip
// the same on device columns with its staging in the library's static scratch
// array (no allocation); sdt is the staging element type
Status applyz_noalloc_device(lhl_handle *h, int inverse, void *x, int64_t nrhs, int64_t ldx, int xdt,
                             int x_complex, int sdt);      // lhl_solve.hip
size_t applyz_noalloc_limit();                             // bytes of that scratch array

// r = b - W x for ncol columns, W dense (w != nullptr, element type xdt) or
// sigma_c*I + tau_c*J (per-column shift scalars, J resident).  All vectors are
// planar device arrays of element type xdt; norms[c] = ||r_c||_inf,
// norms[ncol+c] = ||b_c||_inf (double).
struct ResidArgs {
    const void *W = nullptr;
    int64_t ldw = 0;
    int w_complex = 0;
    const double *colshift = nullptr;   // device: 4 doubles per column (sr, si, tr, ti)
    const void *b = nullptr;
    const void *x = nullptr;
    void *r = nullptr;
    int64_t ld = 0, plane = 0, ncol = 0;
    int xdt = LHL_F64;
    int x_complex = 0;
    double *norms = nullptr;
    const int *cact = nullptr;    // per column flags (nullptr: all); as in SolveArgs
    int full_products = 0;
    DevBuf *jxbuf = nullptr, *jxibuf = nullptr;   // a caller's own work buffers (as in SolveArgs)
};
void blas_release(lhl_handle *h);                          // lhl_batch.hip
// device-side refinement control (lhl_batch.hip): p[0:n] = v; and the
// convergence step after round `it` (cb: nshift + 1 column offsets on the
// device; norms as residual_device writes them for ncol columns)
Status fill_i32_device(lhl_handle *h, int *p, int64_t n, int v);
Status conv_device(lhl_handle *h, int it, int max_iter, double tol, int64_t nshift, const int64_t *cb,
                   const double *norms, int64_t ncol, int *sact, int *cact, int *iters);
// C = A B on the matrix cores (lhl_gemm.hip), column-major: mode 0 a full A
// (m x K), 1 a lower-triangular A with zeros stored above the diagonal, 2
// C = B + tril(A, -1) B (K == m).  cact (device, nullable): per batch column
// flags, internal column j being batch column j / planes; 64-column tiles with
// no flagged column are skipped (their C is left as it was).
Status mgemm(lhl_handle *h, const char *name, int mode, int m, int q, int K, const float *A, int64_t lda,
             const float *B, int64_t ldb, float *C, int64_t ldc, const int *cact, int planes);
Status mgemm(lhl_handle *h, const char *name, int mode, int m, int q, int K, const double *A, int64_t lda,
             const double *B, int64_t ldb, double *C, int64_t ldc, const int *cact, int planes);
Status residual_device(lhl_handle *h, const ResidArgs &a);  // lhl_batch.hip

// Planar <-> interleaved conversions and precision casts (lhl_solve.hip).
Status to_planar(lhl_handle *h, const void *src, int64_t ld_src, int src_dt, int src_complex,
                 void *dst, int64_t ld_dst, int64_t plane_dst, int dst_dt, int64_t n, int64_t ncol);
Status from_planar(lhl_handle *h, const void *src, int64_t ld_src, int64_t plane_src, int src_dt,
                   void *dst, int64_t ld_dst, int dst_dt, int dst_complex, int64_t n, int64_t ncol);

}  // namespace lhl
