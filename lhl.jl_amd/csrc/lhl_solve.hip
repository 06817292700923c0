// applyZ!/applyZinv! on a few vectors (SciML/LHL.jl src:2069-2126) and the
// planar layout conversions of the staged right-hand sides.  (The Hessenberg
// solves of the batch are in lhl_batch.hip.)
#include "lhl_common.h"

#include <cstring>
#include <mutex>

namespace lhl {
namespace {

constexpr int kSolveThreads = 256;
constexpr int kBlk = 64;

template <typename T> __device__ __forceinline__ T shfl(T v, int src) { return __shfl(v, src, 64); }
template <typename T> __device__ __forceinline__ T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}


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
        LHL_TRY(allow_dynamic_lds(fn));
        const int groups = (int)(ncol * planes);
        LHL_LAUNCH(h, "applyz", fn, dim3(groups), dim3(kSolveThreads), lds1, n, inverse,
                           (const T *)h->A.as<T>(), h->lda, (const int *)h->perm.as<int>(),
                           (const int *)h->iperm.as<int>(), (const double *)h->scale64.as<double>(),
                           (const double *)h->iscale64.as<double>(), (TX *)x, ldx, xplane, planes, (int)ncol);
    } else {
        auto fn = k_applyz<T, TX, NV>;
        LHL_TRY(allow_dynamic_lds(fn));
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

// ---------------------------------------------------------------- allocation-free applyZ

// Device memory of the allocation-free entry point: the planar staging copy of
// x.  One array per GPU, held until the process ends.  In host code the name
// `scratch` is a host-side stand-in, not a device address: every host use goes
// through the per-device address from hipGetSymbolAddress below.
constexpr size_t kScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void *scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (with that device current).
Status scratch_address(int dev, void **out) {
    if (dev < 0 || dev >= kMaxDevices) return {LHL_ERR_ARG, "applyZ (no allocation): device index out of range"};
    std::lock_guard<std::mutex> g(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return {LHL_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
        scratch_ptr[dev] = p;
    }
    *out = scratch_ptr[dev];
    return {};
}
}  // namespace

size_t applyz_noalloc_limit() { return kScratchBytes; }

// applyZ!/applyZinv! on device columns with the planar staging copy in
// `scratch`: the same kernels and layout as lhl_applyZ's device path, no
// device allocation, no library call.  Profiling events are not recorded here.
Status applyz_noalloc_device(lhl_handle *h, int inverse, void *x, int64_t nrhs, int64_t ldx, int xdt, int x_complex,
                             int sdt) {
    const int64_t n = h->n;
    const int planes = x_complex ? 2 : 1;
    const size_t need = elt_size(sdt) * (size_t)n * (size_t)nrhs * planes;
    if (need > kScratchBytes)
        return {LHL_ERR_ARG, "applyZ (no allocation): " + std::to_string(need) + " bytes of staging exceed the " +
                                 std::to_string(kScratchBytes) + "-byte scratch array (n * nrhs * planes * " +
                                 std::to_string(elt_size(sdt)) + ")"};
    void *buf = nullptr;
    LHL_TRY(scratch_address(h->device, &buf));
    const bool prof = h->prof.on;
    h->prof.on = false;
    const int64_t ld = std::max<int64_t>(n, 1), plane = ld * nrhs;
    Status st = to_planar(h, x, ldx, xdt, x_complex, buf, ld, plane, sdt, n, nrhs);
    if (st.ok()) st = applyz_device(h, inverse, buf, nrhs, ld, plane, sdt, x_complex);
    if (st.ok()) st = from_planar(h, buf, ld, plane, sdt, x, ldx, xdt, x_complex, n, nrhs);
    h->prof.on = prof;
    return st;
}

}  // namespace lhl
