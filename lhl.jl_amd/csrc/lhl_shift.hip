// The per-shift Hessenberg LU of G = sigma*I + tau*H (SciML/LHL.jl lhl_shift!,
// src:2456-2594; bit-for-bit spec test/core_tests.jl:119-171).
//
// One workgroup per shift.  The pending row r (the row of G not yet chosen as
// a pivot row) lives in registers: thread t owns elements j = t + m*256.  Step
// k compares r[k] with tau*H[k+1,k], writes U row k (coalesced, packed rows
// of U) and updates the owned r[j], j > k, against row k+1 of H (a
// contiguous column of Ht).  The owner of r[k+1] publishes it through a
// double-buffered LDS slot, so one barrier separates consecutive steps; the
// next step's row of H is prefetched into registers before the barrier.
//
// Every element sees the reference's operations in the reference's order with
// no contraction (mul, mul, then sub), and IEEE division, so the real state
// (U, multipliers, swap, rdiag, info) equals the reference's bit for bit.
#include "lhl_common.h"

namespace lhl {
namespace {

constexpr int kShiftThreads = 256;

template <typename T> __device__ __forceinline__ T dabs(T x) { return x < 0 ? -x : x; }

// Shift parameters per workgroup: sigma, tau (re, im) in the working type.
template <typename T>
struct ShiftParam {
    T sr, si, tr, ti;
};

// ---------------------------------------------------------------- real shifts
template <typename T, int NPER>
__global__ void __launch_bounds__(kShiftThreads) k_shift_real(const T *__restrict__ Ht, int64_t ldh, int n,
                                                              const ShiftParam<T> *__restrict__ prm,
                                                              T *__restrict__ Ubase, int64_t ustride,
                                                              T *__restrict__ lbase, T *__restrict__ rdbase,
                                                              int64_t lstride, uint8_t *__restrict__ swbase,
                                                              int64_t sstride, int *__restrict__ infobase) {
#pragma clang fp contract(off)
    const int s = blockIdx.x;
    const T sigma = prm[s].sr, tau = prm[s].tr;
    T *U = Ubase + (int64_t)s * ustride;
    T *lv = lbase + (int64_t)s * lstride;
    T *rd = rdbase + (int64_t)s * lstride;
    uint8_t *sw = swbase + (int64_t)s * sstride;
    const int tid = threadIdx.x;
    __shared__ T slot[2];

    T r[NPER], h[NPER];
#pragma unroll
    for (int m = 0; m < NPER; m++) {
        const int j = tid + m * kShiftThreads;
        r[m] = (j < n) ? tau * Ht[j] : (T)0;          // r = tau * H[0, :]
    }
    if (tid == 0) r[0] = r[0] + sigma;
    // prefetch row 1 of H (column 1 of Ht)
#pragma unroll
    for (int m = 0; m < NPER; m++) {
        const int j = tid + m * kShiftThreads;
        h[m] = (n > 1 && j < n) ? Ht[j + ldh] : (T)0;
    }
    if (tid == 0) slot[0] = r[0];
    __syncthreads();
    int info = 0;
    for (int k = 0; k < n - 1; k++) {
        const T a = slot[k & 1];
        const T b = tau * Ht[k + (int64_t)(k + 1) * ldh];
        const bool sw_k = dabs(b) > dabs(a);
        T l;
        if (sw_k) l = a / b;
        else if (a == (T)0) { if (info == 0) info = k + 1; l = (T)0; }
        else l = b / a;
        T *Urow = U + upack_off(n, k) - k;     // Urow[j] = U[k, j]
        if (tid == 0) {
            const T ukk = sw_k ? b : a;
            Urow[k] = ukk;
            rd[k] = (T)1 / ukk;
            lv[k] = l;
            sw[k] = (uint8_t)sw_k;
        }
        // prefetch row k+2 of H for the next step
        T hn[NPER];
        const bool more = k + 2 < n;
#pragma unroll
        for (int m = 0; m < NPER; m++) {
            const int j = tid + m * kShiftThreads;
            hn[m] = (more && j > k + 1 && j < n) ? Ht[j + (int64_t)(k + 2) * ldh] : (T)0;
        }
#pragma unroll
        for (int m = 0; m < NPER; m++) {
            const int j = tid + m * kShiftThreads;
            if (j > k && j < n) {
                T u;
                if (sw_k) {
                    const T g = tau * h[m];
                    T rj = r[m] - l * g;
                    u = g;
                    if (j == k + 1) { u = g + sigma; rj = rj - l * sigma; }
                    r[m] = rj;
                } else {
                    u = r[m];
                    T rj = tau * h[m] - l * r[m];
                    if (j == k + 1) rj = rj + sigma;
                    r[m] = rj;
                }
                Urow[j] = u;
                if (j == k + 1) slot[(k + 1) & 1] = r[m];
            }
            h[m] = hn[m];
        }
        __syncthreads();
    }
    if (n >= 1) {
        // the last diagonal: U[n-1, n-1] = r[n-1]
        const int own = (n - 1) % kShiftThreads, mm = (n - 1) / kShiftThreads;
        if (tid == own) {
            T rn = 0;
#pragma unroll
            for (int m = 0; m < NPER; m++) if (m == mm) rn = r[m];
            U[upack_off(n, n - 1)] = rn;
            rd[n - 1] = (T)1 / rn;
            sw[n - 1] = 0;
            if (rn == (T)0 && info == 0) info = n;
            infobase[s] = info;
        }
    }
}

// ------------------------------------------------------------- complex shifts
// Complex shift on a real H, planar: plane 0 = real parts, plane 1 = imaginary
// parts, at +ustride/2 (U) and +lstride/2 (l, rdiag).  Products follow Julia's
// term order; division and 1/x use the Baudin-Smith robust algorithm
// ("A Robust Complex Division in Scilab", 2012).
template <typename T>
__device__ __forceinline__ void cdiv1(T a, T b, T c, T d, T &p, T &q) {
#pragma clang fp contract(off)
    const T r = d / c;
    const T t = (T)1 / (c + d * r);
    if (r != (T)0) {
        const T br = b * r;
        p = (br != (T)0) ? (a + br) * t : a * t + (b * t) * r;
        const T ar = -a * r;
        q = (ar != (T)0) ? (b + ar) * t : b * t + (-a * t) * r;
    } else {
        p = (a + d * (b / c)) * t;
        q = (b + d * (-a / c)) * t;
    }
}
template <typename T>
__device__ __forceinline__ void cdiv(T a, T b, T c, T d, T &p, T &q) {
    if (dabs(d) <= dabs(c)) cdiv1(a, b, c, d, p, q);
    else { cdiv1(b, a, d, c, p, q); q = -q; }
}

template <typename T, int NPER>
__global__ void __launch_bounds__(kShiftThreads) k_shift_cplx(const T *__restrict__ Ht, int64_t ldh, int n,
                                                              const ShiftParam<T> *__restrict__ prm,
                                                              T *__restrict__ Ubase, int64_t ustride,
                                                              T *__restrict__ lbase, T *__restrict__ rdbase,
                                                              int64_t lstride, uint8_t *__restrict__ swbase,
                                                              int64_t sstride, int *__restrict__ infobase) {
#pragma clang fp contract(off)
    const int s = blockIdx.x;
    const T sr = prm[s].sr, si = prm[s].si, tr = prm[s].tr, ti = prm[s].ti;
    const int64_t up = ustride / 2, lp = lstride / 2;
    T *Ur = Ubase + (int64_t)s * ustride, *Ui = Ur + up;
    T *lr_ = lbase + (int64_t)s * lstride, *li_ = lr_ + lp;
    T *rdr = rdbase + (int64_t)s * lstride, *rdi = rdr + lp;
    uint8_t *sw = swbase + (int64_t)s * sstride;
    const int tid = threadIdx.x;
    __shared__ T slot[2][2];

    T rr[NPER], ri[NPER], h[NPER];
#pragma unroll
    for (int m = 0; m < NPER; m++) {
        const int j = tid + m * kShiftThreads;
        const T hv = (j < n) ? Ht[j] : (T)0;
        rr[m] = tr * hv;
        ri[m] = ti * hv;
    }
    if (tid == 0) { rr[0] = rr[0] + sr; ri[0] = ri[0] + si; }
#pragma unroll
    for (int m = 0; m < NPER; m++) {
        const int j = tid + m * kShiftThreads;
        h[m] = (n > 1 && j < n) ? Ht[j + ldh] : (T)0;
    }
    if (tid == 0) { slot[0][0] = rr[0]; slot[0][1] = ri[0]; }
    __syncthreads();
    int info = 0;
    for (int k = 0; k < n - 1; k++) {
        const T ar = slot[k & 1][0], ai = slot[k & 1][1];
        const T hk = Ht[k + (int64_t)(k + 1) * ldh];
        const T br = tr * hk, bi = ti * hk;
        const bool sw_k = (dabs(br) + dabs(bi)) > (dabs(ar) + dabs(ai));
        T lre, lim;
        if (sw_k) cdiv(ar, ai, br, bi, lre, lim);
        else if (ar == (T)0 && ai == (T)0) { if (info == 0) info = k + 1; lre = 0; lim = 0; }
        else cdiv(br, bi, ar, ai, lre, lim);
        const int64_t ro = upack_off(n, k) - k;
        if (tid == 0) {
            const T ur = sw_k ? br : ar, ui = sw_k ? bi : ai;
            Ur[ro + k] = ur; Ui[ro + k] = ui;
            T qr, qi;
            cdiv((T)1, (T)0, ur, ui, qr, qi);
            rdr[k] = qr; rdi[k] = qi;
            lr_[k] = lre; li_[k] = lim;
            sw[k] = (uint8_t)sw_k;
        }
        T hn[NPER];
        const bool more = k + 2 < n;
#pragma unroll
        for (int m = 0; m < NPER; m++) {
            const int j = tid + m * kShiftThreads;
            hn[m] = (more && j > k + 1 && j < n) ? Ht[j + (int64_t)(k + 2) * ldh] : (T)0;
        }
#pragma unroll
        for (int m = 0; m < NPER; m++) {
            const int j = tid + m * kShiftThreads;
            if (j > k && j < n) {
                T ur, ui;
                const T gr = tr * h[m], gi = ti * h[m];
                if (sw_k) {
                    const T pr = lre * gr - lim * gi, pi = lre * gi + lim * gr;
                    T xr = rr[m] - pr, xi = ri[m] - pi;
                    ur = gr; ui = gi;
                    if (j == k + 1) {
                        ur = gr + sr; ui = gi + si;
                        const T qr = lre * sr - lim * si, qi = lre * si + lim * sr;
                        xr = xr - qr; xi = xi - qi;
                    }
                    rr[m] = xr; ri[m] = xi;
                } else {
                    ur = rr[m]; ui = ri[m];
                    const T pr = lre * ur - lim * ui, pi = lre * ui + lim * ur;
                    T xr = gr - pr, xi = gi - pi;
                    if (j == k + 1) { xr = xr + sr; xi = xi + si; }
                    rr[m] = xr; ri[m] = xi;
                }
                Ur[ro + j] = ur; Ui[ro + j] = ui;
                if (j == k + 1) { slot[(k + 1) & 1][0] = rr[m]; slot[(k + 1) & 1][1] = ri[m]; }
            }
            h[m] = hn[m];
        }
        __syncthreads();
    }
    if (n >= 1) {
        const int own = (n - 1) % kShiftThreads, mm = (n - 1) / kShiftThreads;
        if (tid == own) {
            T xr = 0, xi = 0;
#pragma unroll
            for (int m = 0; m < NPER; m++) if (m == mm) { xr = rr[m]; xi = ri[m]; }
            const int64_t o = upack_off(n, n - 1);
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

template <typename T, int NPER>
void launch_shift(lhl_handle *h, int cplx, int64_t nshift, const T *Ht, int64_t ldh, int n,
                  const ShiftParam<T> *prm, T *U, int64_t us, T *l, T *rd, int64_t ls, uint8_t *sw,
                  int64_t ss, int *info) {
    if (cplx)
        LHL_LAUNCH(h, "shift", (k_shift_cplx<T, NPER>), dim3((unsigned)nshift), dim3(kShiftThreads), 0, Ht, ldh, n,
                   prm, U, us, l, rd, ls, sw, ss, info);
    else
        LHL_LAUNCH(h, "shift", (k_shift_real<T, NPER>), dim3((unsigned)nshift), dim3(kShiftThreads), 0, Ht, ldh, n,
                   prm, U, us, l, rd, ls, sw, ss, info);
}

template <typename T>
Status launch_shift_n(lhl_handle *h, int cplx, int64_t nshift, const T *Ht, int64_t ldh, int n,
                      const ShiftParam<T> *prm, T *U, int64_t us, T *l, T *rd, int64_t ls, uint8_t *sw,
                      int64_t ss, int *info) {
    const int nper = (n + kShiftThreads - 1) / kShiftThreads;
    if (nper <= 1) launch_shift<T, 1>(h, cplx, nshift, Ht, ldh, n, prm, U, us, l, rd, ls, sw, ss, info);
    else if (nper <= 2) launch_shift<T, 2>(h, cplx, nshift, Ht, ldh, n, prm, U, us, l, rd, ls, sw, ss, info);
    else if (nper <= 4) launch_shift<T, 4>(h, cplx, nshift, Ht, ldh, n, prm, U, us, l, rd, ls, sw, ss, info);
    else if (nper <= 8) launch_shift<T, 8>(h, cplx, nshift, Ht, ldh, n, prm, U, us, l, rd, ls, sw, ss, info);
    else if (nper <= 16) launch_shift<T, 16>(h, cplx, nshift, Ht, ldh, n, prm, U, us, l, rd, ls, sw, ss, info);
    else if (nper <= 32) launch_shift<T, 32>(h, cplx, nshift, Ht, ldh, n, prm, U, us, l, rd, ls, sw, ss, info);
    else return {LHL_ERR_ARG, "lhl_shift: n above 8192 is not supported by the register-resident shift kernel"};
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

template <typename T>
Status shift_into(lhl_handle *h, int cplx, int64_t nshift, const double *sig, const double *tau, DevBuf &params,
                  T *U, int64_t us, T *l, T *rd, int64_t ls, uint8_t *sw, int64_t ss, int *info) {
    std::vector<ShiftParam<T>> hp((size_t)nshift);
    for (int64_t s = 0; s < nshift; s++) {
        hp[s].sr = (T)sig[2 * s];
        hp[s].si = (T)sig[2 * s + 1];
        hp[s].tr = (T)tau[2 * s];
        hp[s].ti = (T)tau[2 * s + 1];
    }
    LHL_TRY(params.ensure(sizeof(ShiftParam<T>) * (size_t)nshift));
    LHL_HIP_TRY(hipMemcpyAsync(params.p, hp.data(), sizeof(ShiftParam<T>) * (size_t)nshift,
                               hipMemcpyHostToDevice, h->stream));
    // the host vector must outlive the async copy
    LHL_HIP_TRY(hipStreamSynchronize(h->stream));
    return launch_shift_n<T>(h, cplx, nshift, (const T *)h->Ht.as<T>(), h->lda, (int)h->n,
                             params.as<ShiftParam<T>>(), U, us, l, rd, ls, sw, ss, info);
}

}  // namespace

Status shift_pool_device(lhl_handle *h, ShiftPool &pool, int64_t nshift, const double *sig, const double *tau,
                         int is_complex) {
    const int64_t n = h->n;
    const size_t es = elt_size(h->dtype);
    pool.n = n;
    pool.count = nshift;
    pool.planes = is_complex ? 2 : 1;
    const int64_t uplane = ((upack_len(n) + 63) / 64) * 64;
    pool.ustride = uplane * pool.planes;
    const int64_t lplane = ((n + 63) / 64) * 64;
    LHL_TRY(pool.U.ensure(es * (size_t)pool.ustride * (size_t)nshift));
    LHL_TRY(pool.l.ensure(es * (size_t)lplane * pool.planes * (size_t)nshift));
    LHL_TRY(pool.rdiag.ensure(es * (size_t)lplane * pool.planes * (size_t)nshift));
    LHL_TRY(pool.swp.ensure((size_t)lplane * (size_t)nshift));
    LHL_TRY(pool.dinfo.ensure(sizeof(int) * (size_t)nshift));
    if (h->dtype == LHL_F64)
        return shift_into<double>(h, is_complex, nshift, sig, tau, pool.params, pool.U.as<double>(), pool.ustride,
                                  pool.l.as<double>(), pool.rdiag.as<double>(), lplane * pool.planes,
                                  pool.swp.as<uint8_t>(), lplane, pool.dinfo.as<int>());
    return shift_into<float>(h, is_complex, nshift, sig, tau, pool.params, pool.U.as<float>(), pool.ustride,
                             pool.l.as<float>(), pool.rdiag.as<float>(), lplane * pool.planes,
                             pool.swp.as<uint8_t>(), lplane, pool.dinfo.as<int>());
}

// One shift state (ws.shift or an LHLShift): storage is the pool layout with
// a single entry, so the solves read both the same way.
Status shift_device(lhl_handle *h, lhl_shift_state *sh, int64_t /*nshift*/, const double *sig, const double *tau,
                    int64_t *info_out) {
    const int64_t n = h->n;
    const size_t es = elt_size(h->dtype);
    const int planes = sh->is_complex ? 2 : 1;
    const int64_t uplane = ((upack_len(n) + 63) / 64) * 64;
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
    int info = 0;
    if (n > 0) {
        hipError_t e = hipMemcpyAsync(&info, sh->dinfo.p, sizeof(int), hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        params.release();
        if (e != hipSuccess) return {LHL_ERR_HIP, std::string("lhl_shift: ") + hipGetErrorString(e)};
    }
    sh->info = info;
    sh->sigma[0] = sig[0]; sh->sigma[1] = sig[1];
    sh->tau[0] = tau[0]; sh->tau[1] = tau[1];
    sh->loaded = 1;
    if (info_out) *info_out = info;
    return {};
}

}  // namespace lhl
