/*
 * CPU oracle, complex J (LHLWorkspace{Complex{T}}) -- TEST INFRASTRUCTURE ONLY.
 * Included by lhl_oracle.c once per real type (REAL, SFX, ABS, SQRT); every
 * complex matrix and vector is planar: a real-part array and an imaginary-part
 * array with the same indexing.
 *
 * The reference reduces a complex J with its generic paths: balancing
 * (src:297-340, abs = the complex modulus), the unblocked step loop
 * (src:466-501) with the column-pair trailing update (src:503-538) below
 * _lhl_block_min(ComplexF64) = 768 (src:705) and the generic blocked path above
 * it (same pivots: ipiv is invariant across blocking, core_tests.jl:76, 90),
 * the Ht fill (src:2024-2039), the shift LU of a complex H (src:2496-2594 with
 * TG = Complex{T}) and the complex solves (src:2128-2243, 2818-2884).
 *
 * Arithmetic: complex products in Julia's term order (re*re - im*im,
 * re*im + im*re) with no contraction; division and reciprocal by the
 * Baudin-Smith robust division (cdiv in lhl_oracle_body.h); the modulus is
 * zabs below, the same formula the GPU kernels use, so pivots and balancing
 * decisions can be compared exactly with the GPU.  Julia's own hypot (a
 * Borges-corrected form) can differ from zabs in the last bit, which moves a
 * pivot only where two candidates agree to about one ulp.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* |re + i im| = a sqrt(1 + (b/a)^2), a = max(|re|, |im|), b = min. */
REAL FN(lhlo_zabs)(REAL re, REAL im)
{
    REAL a = ABS(re), b = ABS(im);
    if (isinf(a) || isinf(b)) return (REAL)INFINITY;
    if (a < b) { REAL t = a; a = b; b = t; }
    if (a == 0) return b;
    REAL r = b / a;
    return a * SQRT(1 + r * r);
}

#define ZR(i, j) Ar[(i) + (size_t)(j) * lda]
#define ZI(i, j) Ai[(i) + (size_t)(j) * lda]

/* Parlett-Reinsch balancing of a complex A in place, src:297-340: column and
 * row sums of moduli (off-diagonal, ascending j), power-of-2 factors. */
void FN(lhlo_zbalance)(int n, REAL *Ar, REAL *Ai, int lda, REAL *d, REAL *isc)
{
    for (int i = 0; i < n; i++) { d[i] = 1; isc[i] = 1; }
    for (int it = 0; it < 20; it++) {
        int converged = 1;
        for (int i = 0; i < n; i++) {
            REAL c = 0, r = 0;
            for (int j = 0; j < n; j++) {
                if (j == i) continue;
                c += FN(lhlo_zabs)(ZR(j, i), ZI(j, i));
                r += FN(lhlo_zabs)(ZR(i, j), ZI(i, j));
            }
            if (c == 0 || r == 0) continue;
            REAL f = 1, s = c + r;
            while (c < r / 2) { c *= 2; r /= 2; f *= 2; }
            while (c >= 2 * r) { c /= 2; r *= 2; f /= 2; }
            if (c + r < (REAL)0.95 * s) {
                converged = 0;
                d[i] *= f;
                isc[i] /= f;
                for (int j = 0; j < n; j++) {
                    ZR(i, j) /= f; ZI(i, j) /= f;
                    ZR(j, i) *= f; ZI(j, i) *= f;
                }
            }
        }
        if (converged) break;
    }
}

/* The unblocked ELMHES step loop on a complex A, src:466-501, with the
 * column-pair trailing update of src:503-538:
 *   A[i,j] -= l_i * A[k+1,j] (rows i >= k+2), then
 *   A[i,k+1] += v_j * A[i,j] + v_j1 * A[i,j+1] (every row), pairs in order. */
void FN(lhlo_zreduce_unblocked)(int n, REAL *Ar, REAL *Ai, int lda, int *ipiv)
{
    for (int k = 0; k + 2 < n; k++) {
        int p = k + 1;
        REAL amax = FN(lhlo_zabs)(ZR(k + 1, k), ZI(k + 1, k));
        for (int i = k + 2; i < n; i++) {
            REAL a = FN(lhlo_zabs)(ZR(i, k), ZI(i, k));
            if (a > amax) { amax = a; p = i; }
        }
        ipiv[k] = p;
        if (p != k + 1) {
            for (int j = 0; j < n; j++) {
                REAL t = ZR(k + 1, j); ZR(k + 1, j) = ZR(p, j); ZR(p, j) = t;
                t = ZI(k + 1, j); ZI(k + 1, j) = ZI(p, j); ZI(p, j) = t;
            }
            for (int i = 0; i < n; i++) {
                REAL t = ZR(i, k + 1); ZR(i, k + 1) = ZR(i, p); ZR(i, p) = t;
                t = ZI(i, k + 1); ZI(i, k + 1) = ZI(i, p); ZI(i, p) = t;
            }
        }
        REAL pr = ZR(k + 1, k), pi = ZI(k + 1, k);
        if (pr == 0 && pi == 0) continue;
        for (int i = k + 2; i < n; i++) FN(cdiv)(ZR(i, k), ZI(i, k), pr, pi, &ZR(i, k), &ZI(i, k));
        REAL kr = ZR(k + 1, k + 1), ki = ZI(k + 1, k + 1);
        if (!(kr == 0 && ki == 0)) {
            for (int i = k + 2; i < n; i++) {
                REAL lr = ZR(i, k), li = ZI(i, k);
                REAL qr = lr * kr - li * ki, qi = lr * ki + li * kr;
                ZR(i, k + 1) = ZR(i, k + 1) - qr;
                ZI(i, k + 1) = ZI(i, k + 1) - qi;
            }
        }
        int j = k + 2;
        for (; j + 1 < n; j += 2) {
            REAL ar = ZR(k + 1, j), ai = ZI(k + 1, j), br = ZR(k + 1, j + 1), bi = ZI(k + 1, j + 1);
            for (int i = k + 2; i < n; i++) {
                REAL lr = ZR(i, k), li = ZI(i, k);
                REAL q1r = lr * ar - li * ai, q1i = lr * ai + li * ar;
                REAL q2r = lr * br - li * bi, q2i = lr * bi + li * br;
                ZR(i, j) = ZR(i, j) - q1r; ZI(i, j) = ZI(i, j) - q1i;
                ZR(i, j + 1) = ZR(i, j + 1) - q2r; ZI(i, j + 1) = ZI(i, j + 1) - q2i;
            }
            REAL vr = ZR(j, k), vi = ZI(j, k), wr = ZR(j + 1, k), wi = ZI(j + 1, k);
            for (int i = 0; i < n; i++) {
                REAL xr = ZR(i, j), xi = ZI(i, j), yr = ZR(i, j + 1), yi = ZI(i, j + 1);
                REAL t1r = vr * xr - vi * xi, t1i = vr * xi + vi * xr;
                REAL t2r = wr * yr - wi * yi, t2i = wr * yi + wi * yr;
                ZR(i, k + 1) = ZR(i, k + 1) + (t1r + t2r);
                ZI(i, k + 1) = ZI(i, k + 1) + (t1i + t2i);
            }
        }
        if (j < n) {
            REAL ar = ZR(k + 1, j), ai = ZI(k + 1, j);
            for (int i = k + 2; i < n; i++) {
                REAL lr = ZR(i, k), li = ZI(i, k);
                REAL q1r = lr * ar - li * ai, q1i = lr * ai + li * ar;
                ZR(i, j) = ZR(i, j) - q1r; ZI(i, j) = ZI(i, j) - q1i;
            }
            REAL vr = ZR(j, k), vi = ZI(j, k);
            for (int i = 0; i < n; i++) {
                REAL xr = ZR(i, j), xi = ZI(i, j);
                REAL t1r = vr * xr - vi * xi, t1i = vr * xi + vi * xr;
                ZR(i, k + 1) = ZR(i, k + 1) + t1r;
                ZI(i, k + 1) = ZI(i, k + 1) + t1i;
            }
        }
    }
}

/* lhl_reduce! for a complex J (src:355-390, 2024-2039): A = balanced copy of
 * J reduced in place (lda = n), perm/iperm from ipiv, Ht = H^T (n x n per
 * plane).  ipiv, perm, iperm are 0-based. */
void FN(lhlo_zreduce)(int n, const REAL *Jr, const REAL *Ji, int ldj, int balance,
                      REAL *Ar, REAL *Ai, int lda, int *ipiv, int *perm, int *iperm,
                      REAL *scale, REAL *iscale, REAL *Htr, REAL *Hti)
{
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) { ZR(i, j) = Jr[i + (size_t)j * ldj]; ZI(i, j) = Ji[i + (size_t)j * ldj]; }
    if (balance) FN(lhlo_zbalance)(n, Ar, Ai, lda, scale, iscale);
    else for (int i = 0; i < n; i++) { scale[i] = 1; iscale[i] = 1; }
    FN(lhlo_zreduce_unblocked)(n, Ar, Ai, lda, ipiv);
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int k = 0; k + 2 < n; k++) { int p = ipiv[k]; int t = perm[k + 1]; perm[k + 1] = perm[p]; perm[p] = t; }
    for (int i = 0; i < n; i++) iperm[perm[i]] = i;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            int keep = i <= j + 1;
            Htr[j + (size_t)i * n] = keep ? ZR(i, j) : 0;
            Hti[j + (size_t)i * n] = keep ? ZI(i, j) : 0;
        }
}

/* lhl_shift! of a complex H, src:2496-2594 with TG = Complex{T}: every
 * tau*Ht product is complex by complex; pivot magnitude |re| + |im|
 * (src:2517-2520); rdiag = 1/diag(U) by cdiv.  Gr/Gi hold Gt (n x n). */
int FN(lhlo_zshift)(int n, const REAL *Htr, const REAL *Hti, REAL sr, REAL si, REAL tr, REAL ti,
                    REAL *Gr, REAL *Gi, unsigned char *swp, REAL *rdr, REAL *rdi)
{
    if (n == 0) return 0;
    REAL *rr = (REAL *)malloc(sizeof(REAL) * n), *ri = (REAL *)malloc(sizeof(REAL) * n);
#define HR(j, i) Htr[(j) + (size_t)(i) * n]
#define HI(j, i) Hti[(j) + (size_t)(i) * n]
#define GR(j, i) Gr[(j) + (size_t)(i) * n]
#define GI(j, i) Gi[(j) + (size_t)(i) * n]
    for (int j = 0; j < n; j++) {
        REAL hr = HR(j, 0), hi = HI(j, 0);
        rr[j] = tr * hr - ti * hi;
        ri[j] = tr * hi + ti * hr;
    }
    rr[0] += sr; ri[0] += si;
    int info = 0;
    for (int k = 0; k < n - 1; k++) {
        REAL ar = rr[k], ai = ri[k];
        REAL hr = HR(k, k + 1), hi = HI(k, k + 1);
        REAL br = tr * hr - ti * hi, bi = tr * hi + ti * hr;
        if (ABS(br) + ABS(bi) > ABS(ar) + ABS(ai)) {
            swp[k] = 1;
            GR(k, k) = br; GI(k, k) = bi;
            REAL lr, li;
            FN(cdiv)(ar, ai, br, bi, &lr, &li);
            GR(k, k + 1) = lr; GI(k, k + 1) = li;
            for (int j = k + 1; j < n; j++) {
                REAL xr = HR(j, k + 1), xi = HI(j, k + 1);
                REAL gr = tr * xr - ti * xi, gi = tr * xi + ti * xr;
                GR(j, k) = gr; GI(j, k) = gi;
                REAL pr = lr * gr - li * gi, pi = lr * gi + li * gr;
                rr[j] = rr[j] - pr; ri[j] = ri[j] - pi;
            }
            GR(k + 1, k) = GR(k + 1, k) + sr; GI(k + 1, k) = GI(k + 1, k) + si;
            REAL pr = lr * sr - li * si, pi = lr * si + li * sr;
            rr[k + 1] = rr[k + 1] - pr; ri[k + 1] = ri[k + 1] - pi;
        } else {
            swp[k] = 0;
            GR(k, k) = ar; GI(k, k) = ai;
            REAL lr, li;
            if (ar == 0 && ai == 0) { if (info == 0) info = k + 1; lr = 0; li = 0; }
            else FN(cdiv)(br, bi, ar, ai, &lr, &li);
            GR(k, k + 1) = lr; GI(k, k + 1) = li;
            for (int j = k + 1; j < n; j++) {
                REAL xr = rr[j], xi = ri[j];
                GR(j, k) = xr; GI(j, k) = xi;
                REAL yr = HR(j, k + 1), yi = HI(j, k + 1);
                REAL gr = tr * yr - ti * yi, gi = tr * yi + ti * yr;
                REAL pr = lr * xr - li * xi, pi = lr * xi + li * xr;
                rr[j] = gr - pr; ri[j] = gi - pi;
            }
            rr[k + 1] += sr; ri[k + 1] += si;
        }
    }
    GR(n - 1, n - 1) = rr[n - 1]; GI(n - 1, n - 1) = ri[n - 1];
    swp[n - 1] = 0;
    if (rr[n - 1] == 0 && ri[n - 1] == 0 && info == 0) info = n;
    for (int j = 0; j < n; j++) FN(cdiv)(1, 0, GR(j, j), GI(j, j), &rdr[j], &rdi[j]);
    free(rr); free(ri);
#undef HR
#undef HI
#undef GR
#undef GI
    return info;
}

/* y <- L^-1 y and y <- L y with a complex L (the multipliers A[i,k], rows
 * i >= k+2), one step per pass (src:2151-2243 restated step by step). */
static void FN(zc_inv_sweep)(int n, const REAL *Ar, const REAL *Ai, int lda, REAL *yr, REAL *yi)
{
    for (int k = 0; k + 2 < n; k++) {
        REAL ur = yr[k + 1], ui = yi[k + 1];
        for (int i = k + 2; i < n; i++) {
            REAL lr = ZR(i, k), li = ZI(i, k);
            REAL pr = lr * ur - li * ui, pi = lr * ui + li * ur;
            yr[i] = yr[i] - pr; yi[i] = yi[i] - pi;
        }
    }
}
static void FN(zc_sweep)(int n, const REAL *Ar, const REAL *Ai, int lda, REAL *yr, REAL *yi)
{
    for (int k = n - 3; k >= 0; k--) {
        REAL ur = yr[k + 1], ui = yi[k + 1];
        for (int i = k + 2; i < n; i++) {
            REAL lr = ZR(i, k), li = ZI(i, k);
            REAL pr = lr * ur - li * ui, pi = lr * ui + li * ur;
            yr[i] = yr[i] + pr; yi[i] = yi[i] + pi;
        }
    }
}

/* lhl_ldiv! for a complex J (src:3054-3075 with complex factors). */
void FN(lhlo_zldiv)(int n, const REAL *Ar, const REAL *Ai, int lda, const int *perm, const int *iperm,
                    const REAL *scale, const REAL *iscale, const REAL *Gr, const REAL *Gi,
                    const unsigned char *swp, const REAL *rdr, const REAL *rdi, REAL *xr, REAL *xi)
{
    if (n == 0) return;
    REAL *yr = (REAL *)malloc(sizeof(REAL) * n), *yi = (REAL *)malloc(sizeof(REAL) * n);
    for (int i = 0; i < n; i++) { int p = perm[i]; yr[i] = xr[p] * iscale[p]; yi[i] = xi[p] * iscale[p]; }
    FN(zc_inv_sweep)(n, Ar, Ai, lda, yr, yi);
    FN(lhlo_hess_solve_c)(n, Gr, Gi, swp, rdr, rdi, yr, yi);
    FN(zc_sweep)(n, Ar, Ai, lda, yr, yi);
    for (int i = 0; i < n; i++) { xr[i] = yr[iperm[i]] * scale[i]; xi[i] = yi[iperm[i]] * scale[i]; }
    free(yr); free(yi);
}

/* applyZ! / applyZinv! (src:2069-2126) with a complex L. */
void FN(lhlo_zapplyZ)(int n, const REAL *Ar, const REAL *Ai, int lda, const int *iperm, const REAL *scale,
                      REAL *xr, REAL *xi)
{
    REAL *yr = (REAL *)malloc(sizeof(REAL) * (n ? n : 1)), *yi = (REAL *)malloc(sizeof(REAL) * (n ? n : 1));
    for (int i = 0; i < n; i++) { yr[i] = xr[i]; yi[i] = xi[i]; }
    FN(zc_sweep)(n, Ar, Ai, lda, yr, yi);
    for (int i = 0; i < n; i++) { xr[i] = yr[iperm[i]] * scale[i]; xi[i] = yi[iperm[i]] * scale[i]; }
    free(yr); free(yi);
}
void FN(lhlo_zapplyZinv)(int n, const REAL *Ar, const REAL *Ai, int lda, const int *perm, const REAL *iscale,
                         REAL *xr, REAL *xi)
{
    REAL *yr = (REAL *)malloc(sizeof(REAL) * (n ? n : 1)), *yi = (REAL *)malloc(sizeof(REAL) * (n ? n : 1));
    for (int i = 0; i < n; i++) { int p = perm[i]; yr[i] = xr[p] * iscale[p]; yi[i] = xi[p] * iscale[p]; }
    FN(zc_inv_sweep)(n, Ar, Ai, lda, yr, yi);
    for (int i = 0; i < n; i++) { xr[i] = yr[i]; xi[i] = yi[i]; }
    free(yr); free(yi);
}

#undef ZR
#undef ZI
#undef CAT2
#undef CAT
#undef FN
