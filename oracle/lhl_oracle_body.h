/*
 * CPU oracle for the LHL hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is a plain-C restatement of the reference algorithm of
 * SciML/LHL.jl (src/LHLFactorization.jl, cited below as "src:LINE").  It is
 * included twice by lhl_oracle.c, once with REAL = double (suffix _d) and once
 * with REAL = float (suffix _s).  Only the tests, the smoke check and the
 * bench's cpu_baseline leg load it; the HIP library never links it.
 *
 * Conventions: column-major storage, 0-based indices.  Reference step k
 * (1-based) is step c = k-1 here; it works on column c, pivots among rows
 * c+1..n-1 and stores its multipliers at rows c+2..n-1 of column c, so every
 * index formula keeps its 1-based shape.  ipiv[c] holds the 0-based pivot row.
 *
 * Rounding: the file is compiled with -ffp-contract=off.  Where the reference
 * writes muladd / llvm.fmuladd the restatement calls FMA() explicitly; a plain
 * "a - b*c" in the reference stays two roundings here.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* Parlett-Reinsch balancing, src:297-340.  Scales A in place to D^-1 A D. */
void FN(lhlo_balance)(int n, REAL *A, int lda, REAL *d, REAL *isc)
{
    for (int i = 0; i < n; i++) { d[i] = 1; isc[i] = 1; }
    for (int it = 0; it < 20; it++) {
        int converged = 1;
        for (int i = 0; i < n; i++) {
            REAL c = 0, r = 0;
            for (int j = 0; j < n; j++) {
                if (j == i) continue;
                c += ABS(A[j + (size_t)i * lda]);
                r += ABS(A[i + (size_t)j * lda]);
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
                    A[i + (size_t)j * lda] /= f;
                    A[j + (size_t)i * lda] *= f;
                }
            }
        }
        if (converged) break;
    }
}

#define AA(i, j) A[(i) + (size_t)(j) * lda]

/* Unblocked ELMHES step loop, src:466-501, with the trailing update of the
 * explicit-vector row-block kernel (src:547-678) for n >= 8: per row i,
 * s = sum_j fma(v_j, a_ij, s) over j = c+2..n-1 in order, where rows i >= c+2
 * first take a_ij = fma(l_i, -p_j, a_ij); then A[i,c+1] += s.  For n < 8 the
 * generic paired sweep src:508-538 applies. */
void FN(lhlo_reduce_unblocked)(int n, REAL *A, int lda, int *ipiv)
{
    for (int k = 0; k + 2 < n; k++) {
        int p = k + 1;
        REAL amax = ABS(AA(k + 1, k));
        for (int i = k + 2; i < n; i++) {
            REAL a = ABS(AA(i, k));
            if (a > amax) { amax = a; p = i; }
        }
        ipiv[k] = p;
        if (p != k + 1) {
            for (int j = 0; j < n; j++) { REAL t = AA(k + 1, j); AA(k + 1, j) = AA(p, j); AA(p, j) = t; }
            for (int i = 0; i < n; i++) { REAL t = AA(i, k + 1); AA(i, k + 1) = AA(i, p); AA(i, p) = t; }
        }
        REAL piv = AA(k + 1, k);
        if (piv == 0) continue;
        for (int i = k + 2; i < n; i++) AA(i, k) /= piv;
        REAL pk = AA(k + 1, k + 1);
        if (pk != 0)
            for (int i = k + 2; i < n; i++) AA(i, k + 1) -= AA(i, k) * pk;
        if (n >= 8) {
            for (int i = 0; i < n; i++) {
                REAL li = (i >= k + 2) ? AA(i, k) : 0;
                REAL s = 0;
                for (int j = k + 2; j < n; j++) {
                    REAL a = AA(i, j);
                    if (i >= k + 2) { a = FMA(li, -AA(k + 1, j), a); AA(i, j) = a; }
                    s = FMA(AA(j, k), a, s);
                }
                AA(i, k + 1) = AA(i, k + 1) + s;
            }
        } else {
            int j = k + 2;
            while (j + 1 < n) {
                REAL pj = AA(k + 1, j), pj1 = AA(k + 1, j + 1);
                for (int i = k + 2; i < n; i++) {
                    REAL l = AA(i, k);
                    AA(i, j) -= l * pj;
                    AA(i, j + 1) -= l * pj1;
                }
                REAL vj = AA(j, k), vj1 = AA(j + 1, k);
                for (int i = 0; i < n; i++) AA(i, k + 1) += vj * AA(i, j) + vj1 * AA(i, j + 1);
                j += 2;
            }
            if (j < n) {
                REAL pj = AA(k + 1, j);
                for (int i = k + 2; i < n; i++) AA(i, j) -= AA(i, k) * pj;
                REAL vj = AA(j, k);
                for (int i = 0; i < n; i++) AA(i, k + 1) += vj * AA(i, j);
            }
        }
    }
}

/* ---- blocked (delayed-update) ELMHES, generic structure src:765-890 ---- */

/* w[r0..r1] = sum_{j=c0..c1} A[j,k] * A[r0..r1, j], four columns per pass,
 * src:1418-1440.  Rows are independent, so the row split is bit-neutral. */
static void FN(gemv_cols)(REAL *w, const REAL *A, int lda, int k, int r0, int r1, int c0, int c1)
{
#pragma omp parallel for schedule(static) if (r1 - r0 > 512)
    for (int i = r0; i <= r1; i++) {
        REAL acc = 0;
        int j = c0;
        for (; j + 3 <= c1; j += 4) {
            REAL l1 = AA(j, k), l2 = AA(j + 1, k), l3 = AA(j + 2, k), l4 = AA(j + 3, k);
            acc += (l1 * AA(i, j) + l2 * AA(i, j + 1)) + (l3 * AA(i, j + 2) + l4 * AA(i, j + 3));
        }
        for (; j <= c1; j++) {
            REAL lj = AA(j, k);
            if (lj != 0) acc += lj * AA(i, j);
        }
        w[i] = acc;
    }
}

/* The interchange of step k inside a panel, src:857-890.  B0 (ldb = n) holds
 * the pre-panel copies of panel columns k0+1..kb+1 at B0[:, j-k0-1]. */
static void FN(step_swaps)(int n, REAL *A, int lda, REAL *B0, const int *ipiv,
                           int k0, int kb, int k, int p)
{
    for (int j = k0; j <= kb + 1; j++) {
        if (j == k) continue;
        REAL t = AA(k + 1, j); AA(k + 1, j) = AA(p, j); AA(p, j) = t;
    }
    REAL *bk = B0 + (size_t)(k - k0) * n;   /* copy of column k+1 */
    if (p <= kb + 1) {
        REAL *bp = B0 + (size_t)(p - k0 - 1) * n;
        for (int i = 0; i < n; i++) {
            REAL t = AA(i, k + 1); AA(i, k + 1) = AA(i, p); AA(i, p) = t;
            t = bk[i]; bk[i] = bp[i]; bp[i] = t;
        }
    } else {
        for (int i = 0; i < n; i++) {
            REAL t = AA(i, p);
            AA(i, p) = bk[i];
            bk[i] = t;
            AA(i, k + 1) = t;
        }
        for (int kk = k0; kk <= k; kk++) {
            int q = ipiv[kk];
            if (q != kk + 1) { REAL t = AA(kk + 1, k + 1); AA(kk + 1, k + 1) = AA(q, k + 1); AA(q, k + 1) = t; }
        }
        for (int kk = k0; kk <= k - 1; kk++) {
            REAL x = AA(kk + 1, k + 1);
            if (x == 0) continue;
            for (int i = kk + 2; i < n; i++) AA(i, k + 1) -= AA(i, kk) * x;
        }
    }
}

static void FN(panel_steps)(int n, REAL *A, int lda, int *ipiv, REAL *B0, REAL *w, int k0, int kb)
{
    for (int j = k0 + 1; j <= kb + 1; j++)
        memcpy(B0 + (size_t)(j - k0 - 1) * n, &AA(0, j), sizeof(REAL) * n);
    for (int k = k0; k <= kb; k++) {
        int p = k + 1;
        REAL amax = ABS(AA(k + 1, k));
        for (int i = k + 2; i < n; i++) {
            REAL a = ABS(AA(i, k));
            if (a > amax) { amax = a; p = i; }
        }
        ipiv[k] = p;
        if (p != k + 1) {
            REAL t = AA(k + 1, k); AA(k + 1, k) = AA(p, k); AA(p, k) = t;
            FN(step_swaps)(n, A, lda, B0, ipiv, k0, kb, k, p);
        }
        REAL piv = AA(k + 1, k);
        if (piv == 0) continue;
        for (int i = k + 2; i < n; i++) AA(i, k) /= piv;
        for (int j = k + 1; j <= kb + 1; j++) {
            REAL pj = AA(k + 1, j);
            if (pj == 0) continue;
            for (int i = k + 2; i < n; i++) AA(i, j) -= AA(i, k) * pj;
        }
        for (int j = k + 2; j <= kb + 1; j++) {
            REAL vj = AA(j, k);
            if (vj == 0) continue;
            for (int i = k0 + 1; i < n; i++) AA(i, k + 1) += vj * AA(i, j);
        }
        if (!(kb + 2 <= n - 1)) continue;
        FN(gemv_cols)(w, A, lda, k, k0 + 1, n - 1, kb + 2, n - 1);
        for (int kk = k0; kk <= k; kk++) {
            int q = ipiv[kk];
            if (q != kk + 1) { REAL t = w[kk + 1]; w[kk + 1] = w[q]; w[q] = t; }
        }
        for (int kk = k0; kk <= k; kk++) {
            REAL x = w[kk + 1];
            if (x == 0) continue;
            for (int i = kk + 2; i < n; i++) w[i] -= AA(i, kk) * x;
        }
        for (int i = k0 + 1; i < n; i++) AA(i, k + 1) += w[i];
    }
}

void FN(lhlo_reduce_blocked)(int n, REAL *A, int lda, int *ipiv, int nb)
{
    REAL *B0 = (REAL *)malloc(sizeof(REAL) * (size_t)n * (nb + 1));
    REAL *w = (REAL *)malloc(sizeof(REAL) * (size_t)(n > 0 ? n : 1));
    for (int k0 = 0; k0 + 2 < n; k0 += nb) {
        int kb = k0 + nb - 1;
        if (kb > n - 3) kb = n - 3;
        FN(panel_steps)(n, A, lda, ipiv, B0, w, k0, kb);
        if (kb + 2 <= n - 1) {
            /* row interchanges on the trailing block, src:1336-1342 */
#pragma omp parallel for schedule(static)
            for (int c = kb + 2; c < n; c++)
                for (int k = k0; k <= kb; k++) {
                    int p = ipiv[k];
                    if (p != k + 1) { REAL t = AA(k + 1, c); AA(k + 1, c) = AA(p, c); AA(p, c) = t; }
                }
            /* unit-lower triangular solve on rows k0+1..kb+1, src:1917-1927 */
#pragma omp parallel for schedule(static)
            for (int c = kb + 2; c < n; c++)
                for (int k = k0; k <= kb; k++) {
                    REAL x = AA(k + 1, c);
                    if (x == 0) continue;
                    for (int i = k + 2; i <= kb + 1; i++) AA(i, c) -= AA(i, k) * x;
                }
            /* rank-nb trailing update, src:1676-1689 with sgn = -1 */
#pragma omp parallel for schedule(static)
            for (int c = kb + 2; c < n; c++)
                for (int j = k0; j <= kb; j++) {
                    REAL b = -AA(j + 1, c);
                    if (b == 0) continue;
                    for (int i = kb + 2; i < n; i++) AA(i, c) += AA(i, j) * b;
                }
        }
        /* deferred right updates of rows 0..k0 (the reference's 1:k0),
         * src:1588-1607 and 1676-1689 */
        {
            int ntop = k0 + 1;
#pragma omp parallel for schedule(static)
            for (int ib = 0; ib < ntop; ib += 64) {
                int ie = ib + 64 < ntop ? ib + 64 : ntop;
                for (int k = k0; k <= kb; k++)
                    for (int j = k + 2; j <= kb + 1; j++) {
                        REAL vj = AA(j, k);
                        if (vj == 0) continue;
                        for (int i = ib; i < ie; i++) AA(i, k + 1) += vj * AA(i, j);
                    }
                for (int c = k0 + 1; c <= kb + 1; c++)
                    for (int j = kb + 2; j < n; j++) {
                        REAL b = AA(j, c - 1);
                        if (b == 0) continue;
                        for (int i = ib; i < ie; i++) AA(i, c) += AA(i, j) * b;
                    }
            }
        }
    }
    /* interchanges deferred on earlier multiplier columns, src:1383-1393 */
#pragma omp parallel for schedule(dynamic, 4)
    for (int c = 0; c < (n > 2 ? n - 2 : 0); c++) {
        int kb = (c / nb) * nb + nb - 1;
        if (kb > n - 3) kb = n - 3;
        for (int k = kb + 1; k <= n - 3; k++) {
            int p = ipiv[k];
            if (p != k + 1) { REAL t = AA(k + 1, c); AA(k + 1, c) = AA(p, c); AA(p, c) = t; }
        }
    }
    free(B0);
    free(w);
}

/* Ht[j,i] = A[i,j] for i <= j+1 (src:2024-2039); the rest of Ht is zeroed. */
void FN(lhlo_ht_fill)(int n, const REAL *A, int lda, REAL *Ht)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            Ht[j + (size_t)i * n] = (i <= j + 1) ? AA(i, j) : 0;
}

/* Driver, src:355-390: copy, optional balance, unblocked below the block
 * minimum (500 double / 1024 float, src:703-704) or when blocked == 0,
 * perm/iperm from ipiv, and Ht.  J is not modified. */
void FN(lhlo_reduce)(int n, const REAL *J, int ldj, int balance, int blocked, int nb,
                     REAL *A, int lda, int *ipiv, int *perm, int *iperm,
                     REAL *scale, REAL *iscale, REAL *Ht)
{
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) AA(i, j) = J[i + (size_t)j * ldj];
    if (balance) FN(lhlo_balance)(n, A, lda, scale, iscale);
    else for (int i = 0; i < n; i++) { scale[i] = 1; iscale[i] = 1; }
    if (blocked < 0) blocked = n >= BLOCK_MIN;
    if (blocked) FN(lhlo_reduce_blocked)(n, A, lda, ipiv, nb);
    else FN(lhlo_reduce_unblocked)(n, A, lda, ipiv);
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int k = 0; k + 2 < n; k++) {
        int p = ipiv[k];
        int t = perm[k + 1]; perm[k + 1] = perm[p]; perm[p] = t;
    }
    for (int i = 0; i < n; i++) iperm[perm[i]] = i;
    if (Ht) FN(lhlo_ht_fill)(n, A, lda, Ht);
}
#undef AA

/* ---- the shift LU, src:2529-2594 == test/core_tests.jl:119-171 ---- */

#define HT(j, i) Ht[(j) + (size_t)(i) * n]
#define GT(j, i) Gt[(j) + (size_t)(i) * n]

/* LU of G = sigma*I + tau*H with one pending row; U row k in Gt[k.., k], the
 * multiplier in Gt[k, k+1], rdiag = 1/diag(U).  Returns info (1-based index of
 * the first zero pivot, 0 if none).  Separate multiply and subtract. */
int FN(lhlo_shift)(int n, const REAL *Ht, REAL sigma, REAL tau,
                   REAL *Gt, unsigned char *swp, REAL *rdiag)
{
    if (n == 0) return 0;
    REAL *r = (REAL *)malloc(sizeof(REAL) * n);
    for (int j = 0; j < n; j++) r[j] = tau * HT(j, 0);
    r[0] += sigma;
    int info = 0;
    for (int k = 0; k < n - 1; k++) {
        REAL a = r[k];
        REAL b = tau * HT(k, k + 1);
        if (ABS(b) > ABS(a)) {
            swp[k] = 1;
            GT(k, k) = b;
            REAL l = a / b;
            GT(k, k + 1) = l;
            for (int j = k + 1; j < n; j++) {
                REAL g = tau * HT(j, k + 1);
                GT(j, k) = g;
                r[j] -= l * g;
            }
            GT(k + 1, k) += sigma;
            r[k + 1] -= l * sigma;
        } else {
            swp[k] = 0;
            GT(k, k) = a;
            REAL l;
            if (a == 0) { if (info == 0) info = k + 1; l = 0; }
            else l = b / a;
            GT(k, k + 1) = l;
            for (int j = k + 1; j < n; j++) {
                REAL rj = r[j];
                GT(j, k) = rj;
                r[j] = tau * HT(j, k + 1) - l * rj;
            }
            r[k + 1] += sigma;
        }
    }
    GT(n - 1, n - 1) = r[n - 1];
    swp[n - 1] = 0;
    if (r[n - 1] == 0 && info == 0) info = n;
    for (int j = 0; j < n; j++) rdiag[j] = 1 / GT(j, j);
    free(r);
    return info;
}

/* The same elimination R steps per pass (src:2601-2744): every element sees the
 * operations of lhlo_shift in the same order, so the state is bit-identical. */
int FN(lhlo_shift_fused)(int n, int R, const REAL *Ht, REAL sigma, REAL tau,
                         REAL *Gt, unsigned char *swp, REAL *rdiag)
{
    if (n == 0) return 0;
    REAL *r = (REAL *)malloc(sizeof(REAL) * n);
    for (int j = 0; j < n; j++) r[j] = tau * HT(j, 0);
    r[0] += sigma;
    int info = 0;
    int S[8];
    REAL L[8];
    int k = 0;
    while (k <= n - 2) {
        int RR = (k + R - 1 <= n - 2) ? R : 1;
        /* decide step k */
        REAL b = tau * HT(k, k + 1), rk = r[k];
        int s = ABS(b) > ABS(rk);
        REAL l;
        if (s) l = rk / b;
        else if (rk == 0) { if (info == 0) info = k + 1; l = 0; }
        else l = b / rk;
        swp[k] = (unsigned char)s;
        GT(k, k) = s ? b : rk;
        GT(k, k + 1) = l;
        for (int m = 0; m < RR; m++) { S[m] = s; L[m] = l; }
        for (int m = 1; m <= RR; m++) {
            int j = k + m;
            REAL rj = r[j];
            for (int i = 1; i <= m; i++) {
                int si = S[i - 1];
                REAL li = L[i - 1], g = tau * HT(j, k + i), u;
                if (si) {
                    u = g;
                    rj -= li * g;
                    if (i == m) { u += sigma; rj -= li * sigma; }
                } else {
                    u = rj;
                    rj = g - li * rj;
                    if (i == m) rj += sigma;
                }
                GT(j, k + i - 1) = u;
            }
            r[j] = rj;
            if (m < RR) {
                REAL bb = tau * HT(j, j + 1);
                int ss = ABS(bb) > ABS(rj);
                REAL ll;
                if (ss) ll = rj / bb;
                else if (rj == 0) { if (info == 0) info = j + 1; ll = 0; }
                else ll = bb / rj;
                swp[j] = (unsigned char)ss;
                GT(j, j) = ss ? bb : rj;
                GT(j, j + 1) = ll;
                S[m] = ss;
                L[m] = ll;
            }
        }
        for (int j = k + RR + 1; j < n; j++) {
            REAL rj = r[j];
            for (int i = 1; i <= RR; i++) {
                REAL g = tau * HT(j, k + i);
                REAL u;
                if (S[i - 1]) { u = g; rj = rj - L[i - 1] * g; }
                else { u = rj; rj = g - L[i - 1] * rj; }
                GT(j, k + i - 1) = u;
            }
            r[j] = rj;
        }
        k += RR;
    }
    GT(n - 1, n - 1) = r[n - 1];
    swp[n - 1] = 0;
    if (r[n - 1] == 0 && info == 0) info = n;
    for (int j = 0; j < n; j++) rdiag[j] = 1 / GT(j, j);
    free(r);
    return info;
}

/* ---- solves ---- */

/* x <- L^-1 x (the forward sweep of the multipliers, src:2151-2195 as one
 * step per pass) and x <- L x (src:2199-2243). */
static void FN(zinv_sweep)(int n, const REAL *A, int lda, REAL *x)
{
    for (int k = 0; k + 2 < n; k++) {
        REAL xk = x[k + 1];
        for (int i = k + 2; i < n; i++) x[i] -= A[i + (size_t)k * lda] * xk;
    }
}
static void FN(z_sweep)(int n, const REAL *A, int lda, REAL *x)
{
    for (int k = n - 3; k >= 0; k--) {
        REAL xk = x[k + 1];
        for (int i = k + 2; i < n; i++) x[i] += A[i + (size_t)k * lda] * xk;
    }
}

/* Hessenberg solve with the shift's LU, src:2746-2816 (one dot product per
 * row, summed in ascending row order). */
void FN(lhlo_hess_solve)(int n, const REAL *Gt, const unsigned char *swp, const REAL *rdiag, REAL *x)
{
    for (int k = 0; k < n - 1; k++) {
        int s = swp[k];
        REAL a = x[k], b = x[k + 1];
        REAL xk = s ? b : a;
        x[k] = xk;
        x[k + 1] = (s ? a : b) - GT(k, k + 1) * xk;
    }
    for (int j = n - 1; j >= 0; j--) {
        REAL s = 0;
        for (int i = j + 1; i < n; i++) s += GT(i, j) * x[i];
        x[j] = (x[j] - s) * rdiag[j];
    }
}

/* lhl_ldiv! for a real shift, src:3054-3075: gather P^-1 D^-1, L^-1 sweep,
 * Hessenberg solve, L sweep, scatter D P. */
void FN(lhlo_ldiv)(int n, const REAL *A, int lda, const int *perm, const int *iperm,
                   const REAL *scale, const REAL *iscale,
                   const REAL *Gt, const unsigned char *swp, const REAL *rdiag, REAL *x)
{
    if (n == 0) return;
    REAL *y = (REAL *)malloc(sizeof(REAL) * n);
    for (int i = 0; i < n; i++) { int p = perm[i]; y[i] = x[p] * iscale[p]; }
    FN(zinv_sweep)(n, A, lda, y);
    FN(lhlo_hess_solve)(n, Gt, swp, rdiag, y);
    FN(z_sweep)(n, A, lda, y);
    for (int i = 0; i < n; i++) x[i] = y[iperm[i]] * scale[i];
    free(y);
}

/* applyZ! / applyZinv! (src:2069-2126) on one vector. */
void FN(lhlo_applyZ)(int n, const REAL *A, int lda, const int *iperm, const REAL *scale, REAL *x)
{
    REAL *y = (REAL *)malloc(sizeof(REAL) * (n ? n : 1));
    for (int i = 0; i < n; i++) y[i] = x[i];
    FN(z_sweep)(n, A, lda, y);
    for (int i = 0; i < n; i++) x[i] = y[iperm[i]] * scale[i];
    free(y);
}
void FN(lhlo_applyZinv)(int n, const REAL *A, int lda, const int *perm, const REAL *iscale, REAL *x)
{
    REAL *y = (REAL *)malloc(sizeof(REAL) * (n ? n : 1));
    for (int i = 0; i < n; i++) { int p = perm[i]; y[i] = x[p] * iscale[p]; }
    FN(zinv_sweep)(n, A, lda, y);
    for (int i = 0; i < n; i++) x[i] = y[i];
    free(y);
}

/* lhl_refine!, src:3147-3157: r = b - W x ; ldiv(r) ; x += r, with W dense
 * (ldw) and the residual summed in ascending column order. */
void FN(lhlo_refine)(int n, const REAL *W, int ldw, const REAL *b,
                     const REAL *A, int lda, const int *perm, const int *iperm,
                     const REAL *scale, const REAL *iscale,
                     const REAL *Gt, const unsigned char *swp, const REAL *rdiag,
                     REAL *x, int steps)
{
    REAL *r = (REAL *)malloc(sizeof(REAL) * (n ? n : 1));
    for (int s = 0; s < steps; s++) {
        for (int i = 0; i < n; i++) r[i] = 0;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) r[i] += W[i + (size_t)j * ldw] * x[j];
        for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
        FN(lhlo_ldiv)(n, A, lda, perm, iperm, scale, iscale, Gt, swp, rdiag, r);
        for (int i = 0; i < n; i++) x[i] += r[i];
    }
    free(r);
}

/* ---- complex shifts on a real reduction (planar storage), src:2496-2594,
 * 2818-2884, 3080-3107.  Complex products follow Julia's term order
 * (re*re - im*im, re*im + im*re); division and inv use the Baudin-Smith
 * robust algorithm (cdiv, below). ---- */

/* (a+ib)/(c+id), Baudin & Smith, "A Robust Complex Division in Scilab" (2012),
 * the unscaled fast path. */
static void FN(cdiv1)(REAL a, REAL b, REAL c, REAL d, REAL *p, REAL *q)
{
    REAL r = d / c;
    REAL t = 1 / (c + d * r);
    if (r != 0) {
        REAL br = b * r;
        *p = (br != 0) ? (a + br) * t : a * t + (b * t) * r;
        REAL ar = -a * r;
        *q = (ar != 0) ? (b + ar) * t : b * t + (-a * t) * r;
    } else {
        *p = (a + d * (b / c)) * t;
        *q = (b + d * (-a / c)) * t;
    }
}
static void FN(cdiv)(REAL a, REAL b, REAL c, REAL d, REAL *p, REAL *q)
{
    if (ABS(d) <= ABS(c)) FN(cdiv1)(a, b, c, d, p, q);
    else { FN(cdiv1)(b, a, d, c, p, q); *q = -*q; }
}

int FN(lhlo_shift_c)(int n, const REAL *Ht, REAL sr, REAL si, REAL tr, REAL ti,
                     REAL *Gr, REAL *Gi, unsigned char *swp, REAL *rdr, REAL *rdi)
{
    if (n == 0) return 0;
    REAL *rr = (REAL *)malloc(sizeof(REAL) * n), *ri = (REAL *)malloc(sizeof(REAL) * n);
#define GR(j, i) Gr[(j) + (size_t)(i) * n]
#define GI(j, i) Gi[(j) + (size_t)(i) * n]
    for (int j = 0; j < n; j++) { rr[j] = tr * HT(j, 0); ri[j] = ti * HT(j, 0); }
    rr[0] += sr; ri[0] += si;
    int info = 0;
    for (int k = 0; k < n - 1; k++) {
        REAL ar = rr[k], ai = ri[k];
        REAL h = HT(k, k + 1);
        REAL br = tr * h, bi = ti * h;
        if (ABS(br) + ABS(bi) > ABS(ar) + ABS(ai)) {
            swp[k] = 1;
            GR(k, k) = br; GI(k, k) = bi;
            REAL lr, li;
            FN(cdiv)(ar, ai, br, bi, &lr, &li);
            GR(k, k + 1) = lr; GI(k, k + 1) = li;
            for (int j = k + 1; j < n; j++) {
                REAL hh = HT(j, k + 1);
                REAL gr = tr * hh, gi = ti * hh;
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
                REAL hh = HT(j, k + 1);
                REAL gr = tr * hh, gi = ti * hh;
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
    return info;
}

void FN(lhlo_hess_solve_c)(int n, const REAL *Gr, const REAL *Gi, const unsigned char *swp,
                           const REAL *rdr, const REAL *rdi, REAL *yr, REAL *yi)
{
    for (int k = 0; k < n - 1; k++) {
        int s = swp[k];
        REAL ar = yr[k], ai = yi[k], br = yr[k + 1], bi = yi[k + 1];
        REAL xr = s ? br : ar, xi = s ? bi : ai;
        yr[k] = xr; yi[k] = xi;
        REAL lr = GR(k, k + 1), li = GI(k, k + 1);
        REAL pr = lr * xr - li * xi, pi = lr * xi + li * xr;
        yr[k + 1] = (s ? ar : br) - pr; yi[k + 1] = (s ? ai : bi) - pi;
    }
    for (int j = n - 1; j >= 0; j--) {
        REAL sr = 0, si = 0;
        for (int i = j + 1; i < n; i++) {
            REAL gr = GR(i, j), gi = GI(i, j), xr = yr[i], xi = yi[i];
            sr += gr * xr - gi * xi;
            si += gr * xi + gi * xr;
        }
        REAL dr = yr[j] - sr, di = yi[j] - si;
        yr[j] = dr * rdr[j] - di * rdi[j];
        yi[j] = dr * rdi[j] + di * rdr[j];
    }
#undef GR
#undef GI
}

void FN(lhlo_ldiv_c)(int n, const REAL *A, int lda, const int *perm, const int *iperm,
                     const REAL *scale, const REAL *iscale,
                     const REAL *Gr, const REAL *Gi, const unsigned char *swp,
                     const REAL *rdr, const REAL *rdi, REAL *xr, REAL *xi)
{
    if (n == 0) return;
    REAL *yr = (REAL *)malloc(sizeof(REAL) * n), *yi = (REAL *)malloc(sizeof(REAL) * n);
    for (int i = 0; i < n; i++) { int p = perm[i]; yr[i] = xr[p] * iscale[p]; yi[i] = xi[p] * iscale[p]; }
    FN(zinv_sweep)(n, A, lda, yr);
    FN(zinv_sweep)(n, A, lda, yi);
    FN(lhlo_hess_solve_c)(n, Gr, Gi, swp, rdr, rdi, yr, yi);
    FN(z_sweep)(n, A, lda, yr);
    FN(z_sweep)(n, A, lda, yi);
    for (int i = 0; i < n; i++) { xr[i] = yr[iperm[i]] * scale[i]; xi[i] = yi[iperm[i]] * scale[i]; }
    free(yr); free(yi);
}
#undef HT
#undef GT
#undef CAT2
#undef CAT
#undef FN
