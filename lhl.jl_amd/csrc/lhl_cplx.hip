// A complex J (LHLWorkspace{Complex{T}}, SciML/LHL.jl's generic paths):
// balancing with the complex modulus (src:297-340), the ELMHES step loop on a
// complex matrix (src:466-538), the Ht fill (src:2024-2039), the shift LU of
// a complex H (src:2496-2594 with TG = Complex{T}) and the sweeps of a complex
// L (src:2128-2243).
//
// Storage is planar: the factors A, Ht and the resident J each hold a real
// plane and, lda * n (ldj * n) elements later, an imaginary plane.  The shift
// states and the staged right-hand sides use the planar layouts of complex
// shifts on a real reduction, so the interchange chain, the U solves, the
// gather and the scatter of lhl_batch.hip serve both.
//
// Every product is written out on the planes in Julia's term order
// (re*re - im*im, re*im + im*re), with no contraction; divisions use the
// Baudin-Smith form of lhl_cdiv.h; the modulus is zabs below.  With these the
// pivots, the scales and the shift states are reproducible bit for bit by a
// one-step CPU form with the same formulas, and so are the factors of
// matrices whose trailing update fits one column chunk.
#include "lhl_common.h"
#include "lhl_cdiv.h"

namespace lhl {
namespace {

constexpr int kZThreads = 1024;   // the single-workgroup kernels (balancing, pivot + interchange + scaling)
constexpr int kZRows = 256;       // rows per workgroup of the trailing update
constexpr int kZCols = 256;       // columns per chunk of the trailing update (even: column pairs stay whole)
constexpr int kZBalChunk = 1024;  // moduli staged per pass of the balancing sums

// |re + i im| = a sqrt(1 + (b/a)^2), a = max(|re|, |im|), b = min(|re|, |im|):
// correctly rounded division and square root, so a CPU form of the same
// formula gives the same bits.
template <typename T> __device__ __forceinline__ T zabs(T re, T im) {
#pragma clang fp contract(off)
    T a = re < 0 ? -re : re, b = im < 0 ? -im : im;
    if (__builtin_isinf(a) || __builtin_isinf(b)) return (T)INFINITY;
    if (a < b) { const T t = a; a = b; b = t; }
    if (a == (T)0) return b;
    const T r = b / a;
    return a * __builtin_sqrt((T)1 + r * r);
}
template <> __device__ __forceinline__ float zabs<float>(float re, float im) {
#pragma clang fp contract(off)
    float a = re < 0 ? -re : re, b = im < 0 ? -im : im;
    if (__builtin_isinf(a) || __builtin_isinf(b)) return INFINITY;
    if (a < b) { const float t = a; a = b; b = t; }
    if (a == 0.f) return b;
    const float r = b / a;
    return a * __builtin_sqrtf(1.f + r * r);
}

// ---------------------------------------------------------------- balancing
// Parlett-Reinsch on the resident planar J with the scales applied on the fly:
// A[j, i] = J[j, i] d[i] / d[j], exact for powers of two, so every sum sees the
// values the reference's in-place scaling produces.  One workgroup walks the
// indices in the reference's order; the moduli of a column and a row are
// formed in parallel and summed left to right by one lane each (the
// reference's order).  d lives in LDS (n doubles) and is written out at the end.
__global__ void __launch_bounds__(kZThreads) k_zbal(const double *__restrict__ Jr, const double *__restrict__ Ji,
                                                    int64_t ldj, int n, double *__restrict__ dout,
                                                    double *__restrict__ iout) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char lds_dyn[];
    double *d = reinterpret_cast<double *>(lds_dyn);
    __shared__ double cv[kZBalChunk], rv[kZBalChunk];
    __shared__ double sums[2];
    __shared__ int conv;
    const int tid = threadIdx.x;
    for (int i = tid; i < n; i += kZThreads) d[i] = 1.0;
    __syncthreads();
    for (int it = 0; it < 20; it++) {
        if (tid == 0) conv = 1;
        for (int i = 0; i < n; i++) {
            const double di = d[i];
            double c = 0.0, r = 0.0;   // live in lanes 0 and 64
            for (int j0 = 0; j0 < n; j0 += kZBalChunk) {
                const int m = min(kZBalChunk, n - j0);
                __syncthreads();
                for (int t = tid; t < m; t += kZThreads) {
                    const int j = j0 + t;
                    double a = 0.0, b = 0.0;
                    if (j != i) {
                        const double dj = d[j];
                        const int64_t oc = j + (int64_t)i * ldj, orow = i + (int64_t)j * ldj;
                        a = zabs((Jr[oc] * di) / dj, (Ji[oc] * di) / dj);
                        b = zabs((Jr[orow] * dj) / di, (Ji[orow] * dj) / di);
                    }
                    cv[t] = a;
                    rv[t] = b;
                }
                __syncthreads();
                if (tid == 0)
                    for (int t = 0; t < m; t++) c += cv[t];
                else if (tid == 64)
                    for (int t = 0; t < m; t++) r += rv[t];
            }
            if (tid == 64) sums[1] = r;
            __syncthreads();
            if (tid == 0) {
                r = sums[1];
                if (!(c == 0.0 || r == 0.0)) {
                    double f = 1.0;
                    const double s = c + r;
                    while (c < r / 2) { c *= 2; r /= 2; f *= 2; }
                    while (c >= 2 * r) { c /= 2; r *= 2; f /= 2; }
                    if (c + r < 0.95 * s) {
                        conv = 0;
                        d[i] = di * f;
                    }
                }
            }
            __syncthreads();
        }
        if (conv) break;
        __syncthreads();
    }
    for (int i = tid; i < n; i += kZThreads) {
        dout[i] = d[i];
        iout[i] = 1.0 / d[i];
    }
}

// A = D^-1 J D in the factor type, both planes.
template <typename T>
__global__ void k_zload(const double *__restrict__ J, int64_t ldj, int64_t jplane, T *__restrict__ A, int64_t lda,
                        int64_t aplane, int n, const double *__restrict__ d, const double *__restrict__ isc) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    const double f = isc[i], g = d[j];
    const int64_t o = i + (int64_t)j * ldj, oa = i + (int64_t)j * lda;
    A[oa] = (T)((J[o] * f) * g);
    A[oa + aplane] = (T)((J[o + jplane] * f) * g);
}

// ---------------------------------------------------------------- the step loop
// Step k, part 1 (one workgroup): the pivot (largest modulus in column k at
// rows >= k+1, first of equals, src:469-477), the interchange of rows and
// columns k+1 and p, the multipliers l_i = A[i,k] / A[k+1,k] and column
// k+1's left update (src:487-497).  ctl[0] = 1 when the pivot is zero.
template <typename T>
__global__ void __launch_bounds__(kZThreads) k_zpivot(T *__restrict__ Ar, T *__restrict__ Ai, int64_t lda, int n,
                                                      int k, int *__restrict__ ipiv, int *__restrict__ ctl) {
#pragma clang fp contract(off)
    __shared__ T bv[kZThreads / 64];
    __shared__ int bi[kZThreads / 64];
    __shared__ int sp;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    auto AR = [&](int i, int j) -> T & { return Ar[i + (int64_t)j * lda]; };
    auto AI = [&](int i, int j) -> T & { return Ai[i + (int64_t)j * lda]; };
    // candidates beyond row k+1 never win on NaN (they must exceed), and a NaN
    // at row k+1 keeps p = k+1
    T best = (T)-1;
    int bidx = n;
    for (int i = k + 2 + tid; i < n; i += kZThreads) {
        const T a = zabs(AR(i, k), AI(i, k));
        if (a > best) { best = a; bidx = i; }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const T ov = __shfl_xor(best, o, 64);
        const int oi = __shfl_xor(bidx, o, 64);
        if (ov > best || (ov == best && oi < bidx)) { best = ov; bidx = oi; }
    }
    if (lane == 0) { bv[wv] = best; bi[wv] = bidx; }
    __syncthreads();
    if (tid == 0) {
        T b = bv[0];
        int ix = bi[0];
        for (int w = 1; w < kZThreads / 64; w++)
            if (bv[w] > b || (bv[w] == b && bi[w] < ix)) { b = bv[w]; ix = bi[w]; }
        const T a0 = zabs(AR(k + 1, k), AI(k + 1, k));
        int p = k + 1;
        if (ix < n && b > a0) p = ix;   // a0 NaN: never exceeded
        sp = p;
        ipiv[k] = p;
    }
    __syncthreads();
    const int p = sp;
    if (p != k + 1) {
        for (int t = tid; t < n; t += kZThreads) {
            if (t == k + 1 || t == p) continue;
            T x = AR(k + 1, t); AR(k + 1, t) = AR(p, t); AR(p, t) = x;
            x = AI(k + 1, t); AI(k + 1, t) = AI(p, t); AI(p, t) = x;
            x = AR(t, k + 1); AR(t, k + 1) = AR(t, p); AR(t, p) = x;
            x = AI(t, k + 1); AI(t, k + 1) = AI(t, p); AI(t, p) = x;
        }
        if (tid == 0) {
            T x = AR(k + 1, k + 1); AR(k + 1, k + 1) = AR(p, p); AR(p, p) = x;
            x = AI(k + 1, k + 1); AI(k + 1, k + 1) = AI(p, p); AI(p, p) = x;
            x = AR(k + 1, p); AR(k + 1, p) = AR(p, k + 1); AR(p, k + 1) = x;
            x = AI(k + 1, p); AI(k + 1, p) = AI(p, k + 1); AI(p, k + 1) = x;
        }
    }
    __syncthreads();
    const T pr = AR(k + 1, k), pi = AI(k + 1, k);
    const bool skip = pr == (T)0 && pi == (T)0;
    if (tid == 0) ctl[0] = skip ? 1 : 0;
    if (skip) return;
    const T kr = AR(k + 1, k + 1), ki = AI(k + 1, k + 1);
    const bool upd = !(kr == (T)0 && ki == (T)0);
    for (int i = k + 2 + tid; i < n; i += kZThreads) {
        T lr, li;
        cdiv(AR(i, k), AI(i, k), pr, pi, lr, li);
        AR(i, k) = lr;
        AI(i, k) = li;
        if (upd) {
            const T qr = lr * kr - li * ki, qi = lr * ki + li * kr;
            AR(i, k + 1) = AR(i, k + 1) - qr;
            AI(i, k + 1) = AI(i, k + 1) - qi;
        }
    }
}

// Step k, part 2: the trailing update in the reference's column-pair order
// (src:503-538).  Workgroup (row block, column chunk): its rows i >= k+2 take
// A[i,j] -= l_i A[k+1,j], and every row accumulates v_j A[i,j] + v_j1 A[i,j+1]
// over the chunk's pairs in order.  Chunk 0 starts from A[i,k+1]; with one
// chunk it writes column k+1 itself, otherwise each chunk's sum goes to
// part[c] and k_zcolfix adds the chunks in order.
template <typename T>
__global__ void __launch_bounds__(kZRows) k_ztrail(T *__restrict__ Ar, T *__restrict__ Ai, int64_t lda, int n, int k,
                                                   const int *__restrict__ ctl, T *__restrict__ part, int64_t pplane,
                                                   int nchunk) {
#pragma clang fp contract(off)
    if (ctl[0]) return;
    __shared__ T pr_[kZCols], pi_[kZCols], vr_[kZCols], vi_[kZCols];
    const int c = blockIdx.y;
    const int j0 = k + 2 + c * kZCols, j1 = min(n, j0 + kZCols);
    const int tid = threadIdx.x;
    for (int t = tid; t < j1 - j0; t += kZRows) {
        const int j = j0 + t;
        pr_[t] = Ar[(k + 1) + (int64_t)j * lda];
        pi_[t] = Ai[(k + 1) + (int64_t)j * lda];
        vr_[t] = Ar[j + (int64_t)k * lda];
        vi_[t] = Ai[j + (int64_t)k * lda];
    }
    __syncthreads();
    const int i = blockIdx.x * kZRows + tid;
    if (i >= n) return;
    const bool low = i >= k + 2;
    T lr = 0, li = 0;
    if (low) { lr = Ar[i + (int64_t)k * lda]; li = Ai[i + (int64_t)k * lda]; }
    T accr = 0, acci = 0;
    if (c == 0) { accr = Ar[i + (int64_t)(k + 1) * lda]; acci = Ai[i + (int64_t)(k + 1) * lda]; }
    int t = 0;
    const int m = j1 - j0;
    for (; t + 1 < m; t += 2) {
        const int64_t o0 = i + (int64_t)(j0 + t) * lda, o1 = o0 + lda;
        T xr = Ar[o0], xi = Ai[o0], yr = Ar[o1], yi = Ai[o1];
        if (low) {
            const T q1r = lr * pr_[t] - li * pi_[t], q1i = lr * pi_[t] + li * pr_[t];
            const T q2r = lr * pr_[t + 1] - li * pi_[t + 1], q2i = lr * pi_[t + 1] + li * pr_[t + 1];
            xr = xr - q1r; xi = xi - q1i;
            yr = yr - q2r; yi = yi - q2i;
            Ar[o0] = xr; Ai[o0] = xi;
            Ar[o1] = yr; Ai[o1] = yi;
        }
        const T t1r = vr_[t] * xr - vi_[t] * xi, t1i = vr_[t] * xi + vi_[t] * xr;
        const T t2r = vr_[t + 1] * yr - vi_[t + 1] * yi, t2i = vr_[t + 1] * yi + vi_[t + 1] * yr;
        accr = accr + (t1r + t2r);
        acci = acci + (t1i + t2i);
    }
    if (t < m) {   // the last column alone (only in the last chunk: kZCols is even)
        const int64_t o0 = i + (int64_t)(j0 + t) * lda;
        T xr = Ar[o0], xi = Ai[o0];
        if (low) {
            const T q1r = lr * pr_[t] - li * pi_[t], q1i = lr * pi_[t] + li * pr_[t];
            xr = xr - q1r; xi = xi - q1i;
            Ar[o0] = xr; Ai[o0] = xi;
        }
        const T t1r = vr_[t] * xr - vi_[t] * xi, t1i = vr_[t] * xi + vi_[t] * xr;
        accr = accr + t1r;
        acci = acci + t1i;
    }
    if (nchunk == 1) {
        Ar[i + (int64_t)(k + 1) * lda] = accr;
        Ai[i + (int64_t)(k + 1) * lda] = acci;
    } else {
        part[i + (int64_t)c * n] = accr;
        part[i + (int64_t)c * n + pplane] = acci;
    }
}

template <typename T>
__global__ void k_zcolfix(T *__restrict__ Ar, T *__restrict__ Ai, int64_t lda, int n, int k,
                          const int *__restrict__ ctl, const T *__restrict__ part, int64_t pplane, int nchunk) {
#pragma clang fp contract(off)
    if (ctl[0]) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T sr = part[i], si = part[i + pplane];
    for (int c = 1; c < nchunk; c++) {
        sr = sr + part[i + (int64_t)c * n];
        si = si + part[i + (int64_t)c * n + pplane];
    }
    Ar[i + (int64_t)(k + 1) * lda] = sr;
    Ai[i + (int64_t)(k + 1) * lda] = si;
}

// perm from ipiv (src:380-387) in LDS by one lane, then iperm; n <= 16384.
__global__ void __launch_bounds__(256) k_zperm(const int *__restrict__ ipiv, int *__restrict__ perm,
                                               int *__restrict__ iperm, int n) {
    extern __shared__ __align__(16) unsigned char lds_dyn[];
    int *p = reinterpret_cast<int *>(lds_dyn);
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = i;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k + 2 < n; k++) {
            const int q = ipiv[k];
            const int t = p[k + 1];
            p[k + 1] = p[q];
            p[q] = t;
        }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        perm[i] = p[i];
        iperm[p[i]] = i;
    }
}

// Ht[j, i] = A[i, j] for i <= j+1, zero elsewhere, both planes (src:2024-2039).
template <typename T>
__global__ void k_zht(const T *__restrict__ A, int64_t lda, int64_t aplane, T *__restrict__ Ht, int64_t ldh,
                      int64_t hplane, int n) {
    __shared__ T tile[2][32][33];
    const int bi = blockIdx.y * 32, bj = blockIdx.x * 32;
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        const int i = bi + threadIdx.x, j = bj + r;
        T v = 0, w = 0;
        if (i < n && j < n && i <= j + 1) {
            v = A[i + (int64_t)j * lda];
            w = A[i + (int64_t)j * lda + aplane];
        }
        tile[0][r][threadIdx.x] = v;
        tile[1][r][threadIdx.x] = w;
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        const int j = bj + threadIdx.x, i = bi + r;
        if (i < n && j < n) {
            Ht[j + (int64_t)i * ldh] = tile[0][threadIdx.x][r];
            Ht[j + (int64_t)i * ldh + hplane] = tile[1][threadIdx.x][r];
        }
    }
}

template <typename T>
__global__ void k_zcast_scale(const double *__restrict__ d, const double *__restrict__ isc, T *__restrict__ dt,
                              T *__restrict__ it, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { dt[i] = (T)d[i]; it[i] = (T)isc[i]; }
}

__global__ void k_zones(double *__restrict__ a, double *__restrict__ b, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { a[i] = 1.0; b[i] = 1.0; }
}

// ---------------------------------------------------------------- shift LU of a complex H
// One workgroup per shift, NPER columns per thread; the pending row r in
// registers, the next row of H prefetched one step ahead, one LDS barrier per
// step (the owner of r[k+1] publishes it through a double-buffered slot).
// Each element takes the reference's operations in its order
// (src:2530-2594 with every product complex by complex).  prm holds
// (sigma_re, sigma_im, tau_re, tau_im) per shift; U, l, rdiag, swap, info in
// the layouts of the real-H kernels (packed rows, planar).
template <typename T, int NPER, int NT>
__global__ void __launch_bounds__(NT) k_zshift(const T *__restrict__ Htr, const T *__restrict__ Hti, int64_t ldh, int n,
                                               const T *__restrict__ prm, T *__restrict__ Ubase, int64_t ustride,
                                               T *__restrict__ lbase, T *__restrict__ rdbase, int64_t lstride,
                                               uint8_t *__restrict__ swbase, int64_t sstride, int *__restrict__ infobase) {
#pragma clang fp contract(off)
    const int s = blockIdx.x;
    const T sr = prm[4 * s], si = prm[4 * s + 1], tr = prm[4 * s + 2], ti = prm[4 * s + 3];
    const int64_t up = ustride / 2, lp = lstride / 2;
    T *Ur = Ubase + (int64_t)s * ustride, *Ui = Ur + up;
    T *lr_ = lbase + (int64_t)s * lstride, *li_ = lr_ + lp;
    T *rdr = rdbase + (int64_t)s * lstride, *rdi = rdr + lp;
    uint8_t *sw = swbase + (int64_t)s * sstride;
    const int tid = threadIdx.x;
    __shared__ T slot[2][2];
    extern __shared__ __align__(16) unsigned char lds_dyn[];
    T *subr = reinterpret_cast<T *>(lds_dyn), *subi = subr + n;   // H[k+1, k] = Ht[k, k+1]
    for (int i = tid; i + 1 < n; i += NT) {
        subr[i] = Htr[i + (int64_t)(i + 1) * ldh];
        subi[i] = Hti[i + (int64_t)(i + 1) * ldh];
    }
    T rr[NPER], ri[NPER], hr[NPER], hi[NPER];
#pragma unroll
    for (int m = 0; m < NPER; m++) {
        const int j = tid + m * NT;
        const T ar = (j < n) ? Htr[j] : (T)0, ai = (j < n) ? Hti[j] : (T)0;
        rr[m] = tr * ar - ti * ai;
        ri[m] = tr * ai + ti * ar;
    }
    if (tid == 0) {
        rr[0] = rr[0] + sr;
        ri[0] = ri[0] + si;
        slot[0][0] = rr[0];
        slot[0][1] = ri[0];
    }
    auto load_row = [&](int row) {
        const int64_t col = (int64_t)min(row, n - 1) * ldh;
#pragma unroll
        for (int m = 0; m < NPER; m++) {
            const int64_t o = min(tid + m * NT, n - 1) + col;
            hr[m] = Htr[o];
            hi[m] = Hti[o];
        }
    };
    load_row(1);
    __syncthreads();
    int info = 0;
    for (int k = 0; k + 1 < n; k++) {
        T cr[NPER], ci[NPER];
#pragma unroll
        for (int m = 0; m < NPER; m++) { cr[m] = hr[m]; ci[m] = hi[m]; }
        load_row(k + 2);
        const T hkr = subr[k], hki = subi[k];
        const T ar = slot[k & 1][0], ai = slot[k & 1][1];
        const T br = tr * hkr - ti * hki, bi = tr * hki + ti * hkr;
        const bool sw_k = (cabs1(br) + cabs1(bi)) > (cabs1(ar) + cabs1(ai));
        const bool zero = !sw_k && ar == (T)0 && ai == (T)0;
        T qr, qi;
        cdiv_sel(sw_k ? ar : br, sw_k ? ai : bi, sw_k ? br : ar, sw_k ? bi : ai, qr, qi);
        const T lre = zero ? (T)0 : qr, lim = zero ? (T)0 : qi;
        if (zero && info == 0) info = k + 1;
        const int64_t ro = upack_off(n, k) - k;
        const T ukr = sw_k ? br : ar, uki = sw_k ? bi : ai;
        if (tid == 0) {
            T dr, di;
            cdiv_sel((T)1, (T)0, ukr, uki, dr, di);
            Ur[ro + k] = ukr;
            Ui[ro + k] = uki;
            rdr[k] = dr;
            rdi[k] = di;
            lr_[k] = lre;
            li_[k] = lim;
            sw[k] = (uint8_t)sw_k;
        }
        const T qsr = lre * sr - lim * si, qsi = lre * si + lim * sr;
#pragma unroll
        for (int m = 0; m < NPER; m++) {
            const int j = tid + m * NT;
            if (!(j > k && j < n)) continue;
            const bool first = j == k + 1;
            const T xr0 = rr[m], xi0 = ri[m];
            const T gr = tr * cr[m] - ti * ci[m], gi = tr * ci[m] + ti * cr[m];
            T ur, ui, xr, xi;
            if (sw_k) {
                const T pr = lre * gr - lim * gi, pi = lre * gi + lim * gr;
                xr = xr0 - pr;
                xi = xi0 - pi;
                ur = gr;
                ui = gi;
                if (first) {
                    xr = xr - qsr;
                    xi = xi - qsi;
                    ur = gr + sr;
                    ui = gi + si;
                }
            } else {
                const T pr = lre * xr0 - lim * xi0, pi = lre * xi0 + lim * xr0;
                xr = gr - pr;
                xi = gi - pi;
                ur = xr0;
                ui = xi0;
                if (first) {
                    xr = xr + sr;
                    xi = xi + si;
                }
            }
            rr[m] = xr;
            ri[m] = xi;
            Ur[ro + j] = ur;
            Ui[ro + j] = ui;
            if (first) {
                slot[(k + 1) & 1][0] = xr;
                slot[(k + 1) & 1][1] = xi;
            }
        }
        lds_barrier();
    }
    const int own = (n - 1) % NT, mm = (n - 1) / NT;
    if (tid == own) {
        T xr = 0, xi = 0;
#pragma unroll
        for (int m = 0; m < NPER; m++)
            if (m == mm) { xr = rr[m]; xi = ri[m]; }
        const int64_t o = upack_off(n, n - 1);
        Ur[o] = xr;
        Ui[o] = xi;
        T qr, qi;
        cdiv((T)1, (T)0, xr, xi, qr, qi);
        rdr[n - 1] = qr;
        rdi[n - 1] = qi;
        sw[n - 1] = 0;
        if (xr == (T)0 && xi == (T)0 && info == 0) info = n;
        infobase[s] = info;
    }
}

// ---------------------------------------------------------------- the sweeps of a complex L
// y <- L^-1 y on the internal columns (q = 2c: real plane, 2c + 1: imaginary),
// one workgroup per complex column, in blocks of 64 steps: the first wave
// finishes the block's own rows (a 64-step triangle, the multipliers' values
// broadcast across lanes) and puts the block's solved values in LDS, then every
// thread applies the block's steps to its rows beyond it.  Each row takes its
// updates in increasing step order, as the one-step sweep (src:2151-2195).
template <typename T>
__global__ void __launch_bounds__(1024) k_zsweep_fwd(int n, const T *__restrict__ Ar, const T *__restrict__ Ai,
                                                     int64_t lda, const T *__restrict__ Yin, T *__restrict__ Yout,
                                                     int64_t ldy) {
#pragma clang fp contract(off)
    const int c = blockIdx.x;
    const T *inr = Yin + (int64_t)(2 * c) * ldy, *ini = inr + ldy;
    T *yr = Yout + (int64_t)(2 * c) * ldy, *yi = yr + ldy;
    const int tid = threadIdx.x, lane = tid & 63;
    __shared__ T ur[64], ui[64];
    if (inr != yr)
        for (int i = tid; i < n; i += 1024) { yr[i] = inr[i]; yi[i] = ini[i]; }
    __syncthreads();
    for (int k0 = 0; k0 + 2 < n; k0 += 64) {
        const int kb = min(64, n - 2 - k0);
        if (tid < 64) {
            const int row = k0 + 1 + lane;
            T vr = 0, vi = 0;
            if (lane < kb) { vr = yr[row]; vi = yi[row]; }
            for (int kk = 0; kk < kb; kk++) {
                const T sr = __shfl(vr, kk, 64), si = __shfl(vi, kk, 64);
                if (lane > kk && lane < kb) {
                    const T lr = Ar[row + (int64_t)(k0 + kk) * lda], li = Ai[row + (int64_t)(k0 + kk) * lda];
                    const T pr = lr * sr - li * si, pi = lr * si + li * sr;
                    vr = vr - pr;
                    vi = vi - pi;
                }
            }
            if (lane < kb) {
                yr[row] = vr;
                yi[row] = vi;
                ur[lane] = vr;
                ui[lane] = vi;
            }
        }
        __syncthreads();
        for (int i = k0 + kb + 1 + tid; i < n; i += 1024) {
            T vr = yr[i], vi = yi[i];
            for (int kk = 0; kk < kb; kk++) {
                const T lr = Ar[i + (int64_t)(k0 + kk) * lda], li = Ai[i + (int64_t)(k0 + kk) * lda];
                const T pr = lr * ur[kk] - li * ui[kk], pi = lr * ui[kk] + li * ur[kk];
                vr = vr - pr;
                vi = vi - pi;
            }
            yr[i] = vr;
            yi[i] = vi;
        }
        __syncthreads();
    }
}

// y_out <- L y_in: every step reads an unchanged y_in[k+1] (the reverse order
// touches y[k+1] only later), so each row is a dot product taken in
// decreasing step order, as the one-step sweep (src:2199-2243).  Workgroup =
// (block of 256 rows, one complex column); y_in staged in LDS in chunks.
template <typename T>
__global__ void __launch_bounds__(256) k_zsweep_bwd(int n, const T *__restrict__ Ar, const T *__restrict__ Ai,
                                                    int64_t lda, const T *__restrict__ Yin, T *__restrict__ Yout,
                                                    int64_t ldy) {
#pragma clang fp contract(off)
    constexpr int CH = 2048;
    __shared__ T sr_[CH], si_[CH];
    const int c = blockIdx.y;
    const T *inr = Yin + (int64_t)(2 * c) * ldy, *ini = inr + ldy;
    T *yr = Yout + (int64_t)(2 * c) * ldy, *yi = yr + ldy;
    const int tid = threadIdx.x;
    const int i = blockIdx.x * 256 + tid;
    T vr = 0, vi = 0;
    if (i < n) { vr = inr[i]; vi = ini[i]; }
    const int rlast = min(n - 1, blockIdx.x * 256 + 255);   // the block's last row
    const int kmax = min(rlast - 2, n - 3);                  // largest step any row of the block takes
    // chunks of steps from the top down: [kc0, kc0 + CH)
    for (int kc1 = kmax + 1; kc1 > 0; kc1 -= CH) {
        const int kc0 = max(0, kc1 - CH);
        __syncthreads();
        for (int t = tid; t < kc1 - kc0; t += 256) {
            sr_[t] = inr[kc0 + t + 1];
            si_[t] = ini[kc0 + t + 1];
        }
        __syncthreads();
        if (i < n) {
            const int ktop = min(kc1 - 1, i - 2);
            for (int k = ktop; k >= kc0; k--) {
                const T lr = Ar[i + (int64_t)k * lda], li = Ai[i + (int64_t)k * lda];
                const T ur = sr_[k - kc0], ui = si_[k - kc0];
                const T pr = lr * ur - li * ui, pi = lr * ui + li * ur;
                vr = vr + pr;
                vi = vi + pi;
            }
        }
    }
    if (i < n) { yr[i] = vr; yi[i] = vi; }
}

template <typename T>
Status zreduce_typed(lhl_handle *h, int balance) {
    const int n = (int)h->n;
    const int64_t aplane = h->lda * (int64_t)n, jplane = h->ldj * (int64_t)n;
    T *Ar = h->A.as<T>(), *Ai = Ar + aplane;
    const double *Jr = h->J.as<double>();
    LHL_LAUNCH(h, "reduce_copy", k_zones, dim3((n + 255) / 256), dim3(256), 0, h->scale64.as<double>(),
               h->iscale64.as<double>(), n);
    if (balance) {
        const size_t lds = sizeof(double) * (size_t)n;
        if (lds + 20 * 1024 > lds_limit()) return {LHL_ERR_ARG, "lhl_reduce: n too large for the complex balancing"};
        LHL_TRY(allow_dynamic_lds(k_zbal));
        LHL_LAUNCH(h, "balance", k_zbal, dim3(1), dim3(kZThreads), lds, Jr, Jr + jplane, h->ldj, n,
                   h->scale64.as<double>(), h->iscale64.as<double>());
    }
    LHL_LAUNCH(h, "reduce_copy", k_zload<T>, dim3((n + 255) / 256, n), dim3(256), 0, Jr, h->ldj, jplane, Ar, h->lda,
               aplane, n, (const double *)h->scale64.as<double>(), (const double *)h->iscale64.as<double>());
    const int maxc = std::max(1, (n + kZCols - 1) / kZCols);
    LHL_TRY(h->part.ensure(sizeof(T) * 2 * (size_t)maxc * n + 64));
    LHL_TRY(h->ctl.ensure(64));
    T *part = h->part.as<T>();
    const int64_t pplane = (int64_t)maxc * n;
    int *ctl = h->ctl.as<int>();
    for (int k = 0; k + 2 < n; k++) {
        LHL_LAUNCH(h, "reduce_panel", k_zpivot<T>, dim3(1), dim3(kZThreads), 0, Ar, Ai, h->lda, n, k,
                   h->ipiv.as<int>(), ctl);
        const int nc = (n - (k + 2) + kZCols - 1) / kZCols;
        LHL_LAUNCH(h, "reduce_step", k_ztrail<T>, dim3((n + kZRows - 1) / kZRows, nc), dim3(kZRows), 0, Ar, Ai,
                   h->lda, n, k, (const int *)ctl, part, pplane, nc);
        if (nc > 1)
            LHL_LAUNCH(h, "reduce_step", k_zcolfix<T>, dim3((n + 255) / 256), dim3(256), 0, Ar, Ai, h->lda, n, k,
                       (const int *)ctl, (const T *)part, pplane, nc);
    }
    LHL_HIP_TRY(hipGetLastError());
    LHL_LAUNCH(h, "reduce_finish", k_zperm, dim3(1), dim3(256), sizeof(int) * (size_t)n,
               (const int *)h->ipiv.as<int>(), h->perm.as<int>(), h->iperm.as<int>(), n);
    const int64_t hplane = h->lda * (int64_t)n;
    LHL_LAUNCH(h, "reduce_finish", k_zht<T>, dim3((n + 31) / 32, (n + 31) / 32), dim3(32, 8), 0, (const T *)Ar,
               h->lda, aplane, h->Ht.as<T>(), h->lda, hplane, n);
    LHL_LAUNCH(h, "reduce_finish", k_zcast_scale<T>, dim3((n + 255) / 256), dim3(256), 0,
               (const double *)h->scale64.as<double>(), (const double *)h->iscale64.as<double>(), h->scaleT.as<T>(),
               h->iscaleT.as<T>(), n);
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

template <typename T, int NPER, int NT>
Status zshift_go(lhl_handle *h, int64_t nshift, int n, const T *prm, T *U, int64_t us, T *l, T *rd, int64_t ls,
                 uint8_t *sw, int64_t ss, int *info) {
    auto fn = k_zshift<T, NPER, NT>;
    LHL_TRY(allow_dynamic_lds(fn));
    const T *Htr = h->Ht.as<T>();
    const T *Hti = Htr + h->lda * (int64_t)n;
    LHL_LAUNCH(h, "shift", fn, dim3((unsigned)nshift), dim3(NT), 2 * sizeof(T) * (size_t)n, Htr, Hti, h->lda, n, prm,
               U, us, l, rd, ls, sw, ss, info);
    LHL_HIP_TRY(hipGetLastError());
    return {};
}

}  // namespace

Status zreduce_device(lhl_handle *h, int balance) {
    if (h->n == 0) return {};
    return h->dtype == LHL_F64 ? zreduce_typed<double>(h, balance) : zreduce_typed<float>(h, balance);
}

template <typename T>
Status zshift_launch(lhl_handle *h, int64_t nshift, const T *prm, T *U, int64_t us, T *l, T *rd, int64_t ls,
                     uint8_t *sw, int64_t ss, int *info) {
    const int n = (int)h->n;
    if (2 * sizeof(T) * (size_t)n + 4096 > lds_limit())
        return {LHL_ERR_ARG, "lhl_shift: n too large for the complex-H shift kernel"};
    if (n <= 1024) return zshift_go<T, 1, 1024>(h, nshift, n, prm, U, us, l, rd, ls, sw, ss, info);
    if (n <= 2048) return zshift_go<T, 2, 1024>(h, nshift, n, prm, U, us, l, rd, ls, sw, ss, info);
    if (n <= 4096) return zshift_go<T, 4, 1024>(h, nshift, n, prm, U, us, l, rd, ls, sw, ss, info);
    if (n <= 8192) return zshift_go<T, 8, 1024>(h, nshift, n, prm, U, us, l, rd, ls, sw, ss, info);
    return {LHL_ERR_ARG, "lhl_shift: a complex J above n = 8192 is not supported"};
}
template Status zshift_launch<double>(lhl_handle *, int64_t, const double *, double *, int64_t, double *, double *,
                                      int64_t, uint8_t *, int64_t, int *);
template Status zshift_launch<float>(lhl_handle *, int64_t, const float *, float *, int64_t, float *, float *,
                                     int64_t, uint8_t *, int64_t, int *);

template <typename T>
Status zsweep_device(lhl_handle *h, int fwd, const T *Yin, T *Yout, int64_t ldy, int64_t ncol) {
    const int n = (int)h->n;
    if (ncol <= 0 || n == 0) return {};
    const T *Ar = h->A.as<T>(), *Ai = Ar + h->lda * (int64_t)n;
    if (fwd)
        LHL_LAUNCH(h, "trsm", k_zsweep_fwd<T>, dim3((unsigned)ncol), dim3(1024), 0, n, Ar, Ai, h->lda, Yin, Yout, ldy);
    else
        LHL_LAUNCH(h, "trmm", k_zsweep_bwd<T>, dim3((unsigned)((n + 255) / 256), (unsigned)ncol), dim3(256), 0, n, Ar,
                   Ai, h->lda, Yin, Yout, ldy);
    LHL_HIP_TRY(hipGetLastError());
    return {};
}
template Status zsweep_device<double>(lhl_handle *, int, const double *, double *, int64_t, int64_t);
template Status zsweep_device<float>(lhl_handle *, int, const float *, float *, int64_t, int64_t);

}  // namespace lhl
