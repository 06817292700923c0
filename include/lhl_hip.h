/*
 * lhl_hip.h -- C ABI of the MI355X (gfx950) implementation of the LHL hot path:
 * the ELMHES similarity reduction J = Z H Z^-1 (Z = D P L), the per-shift
 * Hessenberg LU of sigma*I + tau*H, and the Z / U solves with refinement.
 *
 * Every entry point replaces one public name of SciML/LHL.jl
 * (src/LHLFactorization.jl, "src:LINE" below) and is what a Julia `ccall`
 * wrapper binds.  Conventions:
 *   - every call returns an int status: LHL_OK (0), or an error code; the
 *     message is in lhl_last_error(h).  LHL_ERR_ARG maps to ArgumentError,
 *     LHL_ERR_DIM to DimensionMismatch, the rest to ErrorException.
 *   - a zero pivot of the shift LU is not an error: it is reported in *info
 *     (1-based index of the first zero pivot, 0 if none), as src:2450.
 *   - matrices are column-major.  Indices returned by the getters are 0-based.
 *   - complex vectors are interleaved (re, im) pairs, Julia's ComplexF64 /
 *     ComplexF32 memory layout.
 *   - host pointers are borrowed for the duration of a call, which is
 *     synchronous with respect to them; device pointers (flag *_on_device) must
 *     live on the handle's device.
 *   - the handle owns its device buffers and its HIP stream.  A handle is not
 *     thread-safe; separate handles may be used concurrently (src:141-142).
 */
#ifndef LHL_HIP_H
#define LHL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lhl_handle lhl_handle;       /* LHLWorkspace{T,Tr,TG}, src:144-162 */
typedef struct lhl_shift_state lhl_shift_state;   /* LHLShift{TG,Tr}, src:84-96 */

enum {
    LHL_OK = 0,
    LHL_ERR_ARG = 1,    /* ArgumentError */
    LHL_ERR_DIM = 2,    /* DimensionMismatch */
    LHL_ERR_HIP = 3,    /* a HIP runtime error */
    LHL_ERR_STATE = 4,  /* call out of order (e.g. shift before reduce) */
    LHL_ERR_NOMEM = 5,
    /* Not an error of the call: with tol > 0, a column's relative residual
     * is still above tol after the allowed refinement rounds.  Every output is
     * written (x, X, info, iters) as on LHL_OK.  The reference's lhl_refine!
     * runs a fixed number of steps and has no such outcome (src:3147-3157). */
    LHL_NOT_CONVERGED = 6
};

/* Element types.  LHL_C64 / LHL_C32 (ComplexF64 / ComplexF32, interleaved in
 * host and device memory) make a complex workspace, LHLWorkspace{Complex{T}}
 * (src:164-178 with a complex T): its reduction, shifts and solves are
 * complex, its scales are real. */
enum { LHL_F64 = 0, LHL_F32 = 1, LHL_C64 = 2, LHL_C32 = 3 };

/* Working precision of a batched solve. */
enum {
    LHL_PREC_WORKING = 0,  /* plain solves in the handle's precision */
    LHL_PREC_REFINE64 = 1  /* solves in the handle's precision, FP64 residual refinement to tol */
};

/* Library version and the number of visible devices (no GPU work). */
int lhl_version(void);

/* LHLWorkspace{T}(n; shift) (src:164-178) with T = dtype (LHL_F64/LHL_F32)
 * and shift = T (shift_complex = 0) or Complex{T} (1).  With dtype LHL_C64 or
 * LHL_C32 the workspace is complex (T = Complex{Tr}) and its shifts are
 * complex whatever shift_complex says (the reference's TG = T).  Allocates on
 * the current HIP device. */
int lhl_create(lhl_handle **h, int64_t n, int dtype, int shift_complex);
int lhl_destroy(lhl_handle *h);
const char *lhl_last_error(const lhl_handle *h);

/* lhl_reduce!(ws, J, balance) / lhl!(ws, J; balance) (src:355-390, 3192-3195).
 * J (n x n, leading dimension ldj) is not modified.  j_dtype is the element
 * type of J (LHL_F64 or LHL_F32); an FP64 J on an FP32 handle is balanced in
 * FP64, reduced in FP32, and kept resident in FP64 for lhl_refine's residual.
 * A complex workspace takes a complex J (LHL_C64 / LHL_C32, interleaved) or a
 * real one (a zero imaginary part), keeps it resident in FP64 and reduces it
 * with the generic step loop (src:466-538; n <= 8192); a complex J on a real
 * workspace is LHL_ERR_ARG.
 * n may differ from the handle's current size: the handle resizes
 * (src:256-275), and held shift states follow on their next lhl_shift. */
int lhl_reduce(lhl_handle *h, const void *J, int64_t n, int64_t ldj, int j_dtype,
               int balance, int j_on_device);

/* LHLShift{TG}(ws) (src:107-116, 180): an extra shift state on the handle's
 * reduction; is_complex selects TG = Complex{T}.  The handle's own state
 * (ws.shift) is used when a NULL shift state is passed below. */
int lhl_shift_create(lhl_handle *h, lhl_shift_state **sh, int is_complex);
int lhl_shift_destroy(lhl_shift_state *sh);

/* lhl_shift!(ws, sigma, tau) / lhl_shift!(sh, ws, sigma, tau) (src:2456-2489):
 * LU of sigma*I + tau*H.  A complex sigma or tau on a real shift state is
 * LHL_ERR_ARG (src:2466-2473).  *info receives the first zero pivot (1-based)
 * or 0. */
int lhl_shift(lhl_handle *h, lhl_shift_state *sh, double sigma_re, double sigma_im,
              double tau_re, double tau_im, int64_t *info);

/* lhl_ldiv!(x, ws) / lhl_ldiv!(x, sh, ws) (src:3052-3123): x <- W^-1 x for
 * nrhs right-hand sides (columns of x, leading dimension ldx).  x_complex says
 * x holds interleaved complex values; a real x with a complex shift state is
 * LHL_ERR_ARG (src:3084-3085).  x_dtype is the element type of x's real
 * parts. */
int lhl_ldiv(lhl_handle *h, lhl_shift_state *sh, void *x, int64_t n, int64_t nrhs,
             int64_t ldx, int x_dtype, int x_complex, int x_on_device);

/* lhl_refine!(x, A, b, ws, steps) / lhl_refine!(x, A, b, sh, ws, steps)
 * (src:3144-3157): up to `steps` rounds of r = b - W x ; ldiv(r) ; x += r.
 * W is the dense operator (n x n, ldw, the element type of x, x_dtype, with
 * interleaved complex values if x_complex) or NULL, which means
 * sigma*I + tau*J with the handle's resident J (FP64 when the handle was
 * reduced from an FP64 J).  b has x's element type, layout and location.
 * With tol > 0 the loop stops once every column has
 * ||b - W x||_inf / ||b||_inf <= tol, and the solution of the last round is
 * checked too: LHL_NOT_CONVERGED if a column is still above tol.  *iters (may
 * be NULL) receives the rounds done.  The residual and the update are
 * computed in the precision of x. */
int lhl_refine(lhl_handle *h, lhl_shift_state *sh, void *x, const void *W, int64_t ldw,
               const void *b, int64_t n, int64_t nrhs, int64_t ldx, int x_dtype, int x_complex,
               int on_device, int steps, double tol, int *iters);

/* Many shifts and right-hand sides in one call (no reference counterpart as a
 * single name: it is the loop `for (sigma,tau): lhl_shift!; lhl_ldiv!;
 * lhl_refine!` of README.md:47-51 batched).  Shift s is (sigma[2s] +
 * i sigma[2s+1], tau[2s] + i tau[2s+1]); shifts_complex selects complex shifts
 * (on a real reduction).  Its right-hand sides are columns
 * rhs_offsets[s] .. rhs_offsets[s+1]-1 of B (FP64, interleaved complex when
 * shifts_complex, ldb) and the solutions go to the same columns of X (FP64,
 * ldx).  precision: LHL_PREC_WORKING or LHL_PREC_REFINE64 (refine with the
 * FP64 J until tol, at most max_iter rounds).  info[s] (may be NULL) receives
 * each shift's zero-pivot index, iters[s] (may be NULL) its refinement rounds.
 * B and X are host or device memory (on_device).  The shifts run in chunks of
 * equal size whose packed U factors take at most 64 GiB, or a third of the
 * device's free memory if less (twice that with their transposed copies); the
 * environment variable LHL_SHIFT_POOL_BYTES, read on each call, replaces that
 * bound.  Results do not depend on the chunking. */
int lhl_shift_solve_batched(lhl_handle *h, int64_t nshift, const double *sigma, const double *tau,
                            int shifts_complex, const int64_t *rhs_offsets, const double *B,
                            int64_t ldb, double *X, int64_t ldx, int on_device, int precision,
                            double tol, int max_iter, int64_t *info, int *iters);

/* (lhl_shift_solve_batched, continued) With precision LHL_PREC_REFINE64 the
 * solutions of the last round are checked as well: the call returns
 * LHL_NOT_CONVERGED, with every output written, when some shift still has a
 * column above tol after max_iter rounds.  rhs_offsets must start at 0 or
 * above and never decrease (LHL_ERR_DIM otherwise). */

/* Many reductions at once: lhl_reduce on each of hs[0..count-1] (distinct
 * handles on one device, each with its own J[m], n[m], ldj[m]; one J element
 * type and location for all).  No counterpart in the reference as one name:
 * it serves batches of independent matrices of uneven n.  The matrices the
 * row-resident step loop takes alone are grouped by element type and size
 * class, and each group runs one launch per step for all of its matrices
 * (instead of one per step per matrix); the results are those of lhl_reduce,
 * bit for bit.  An error names the first handle it concerns. */
int lhl_reduce_group(int64_t count, lhl_handle *const *hs, const void *const *J, const int64_t *n,
                     const int64_t *ldj, int j_dtype, int balance, int j_on_device);

/* Plans: one batched step (lhl_shift_solve_batched's work for fixed shapes)
 * captured in a HIP graph and replayed for new shifts.  No counterpart in the
 * reference: it is the loop of README.md:47-51 as one graph launch, for
 * callers that sweep many shift sets over one reduction (resolvent sweeps,
 * implicit ODE steps).  lhl_plan_create fixes the handle, the shift count and
 * kind, the right-hand-side offsets, B and X (FP64 device memory, interleaved
 * complex when shifts_complex, the layout lhl_shift_solve_batched takes), the
 * precision, tol and max_iter; every shift's factors must fit one chunk of the
 * batched call (LHL_ERR_ARG otherwise).  lhl_plan_run(sigma, tau) computes
 * X for those shifts (sigma/tau as in lhl_shift_solve_batched), with info[s]
 * and iters[s] (may be NULL) and LHL_NOT_CONVERGED as there.  The refinement
 * is decided on the device: converged shifts drop out of later rounds without
 * a host round trip.  The first run learns the round count (one round at a
 * time); the next captures the whole step in a graph; later runs replay it,
 * continuing round by round if a batch needs more (the graph is then
 * captured again).  B is read and X written at each run; the handle must stay
 * reduced at the same n, and a plan must be destroyed before its handle.
 * lhl_plan_stats reports the rounds in the current graph (-1: none), the
 * replays, the captures, and the rounds run beyond a graph. */
typedef struct lhl_plan lhl_plan;
int lhl_plan_create(lhl_handle *h, lhl_plan **plan, int64_t nshift, int shifts_complex,
                    const int64_t *rhs_offsets, const double *B, int64_t ldb, double *X, int64_t ldx,
                    int precision, double tol, int max_iter);
int lhl_plan_run(lhl_plan *plan, const double *sigma, const double *tau, int64_t *info, int *iters);
int lhl_plan_stats(const lhl_plan *plan, int *graph_rounds, int64_t *replays, int64_t *captures,
                   int64_t *extra_rounds);
int lhl_plan_destroy(lhl_plan *plan);

/* applyZ!(x, ws) / applyZinv!(x, ws) (src:2069-2126) on nrhs columns of x
 * (element type x_dtype; complex interleaved if x_complex). */
int lhl_applyZ(lhl_handle *h, void *x, int64_t n, int64_t nrhs, int64_t ldx, int x_dtype,
               int x_complex, int x_on_device);
int lhl_applyZinv(lhl_handle *h, void *x, int64_t n, int64_t nrhs, int64_t ldx, int x_dtype,
                  int x_complex, int x_on_device);

/* applyZ!(x, ws) (inverse = 0) / applyZinv!(x, ws) (inverse = 1) without
 * device allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same kernels and results as
 * lhl_applyZ / lhl_applyZinv, with the planar staging copy of x in a static
 * 16 MiB device array of the library, one per GPU, held until the process
 * ends; the call makes no hipMalloc, hipFree or other allocating call.
 * x must be device memory on the workspace's device (LHL_ERR_ARG otherwise),
 * and n * nrhs * (2 if x_complex) * 8 bytes (4 when both the workspace and x
 * are FP32) must fit the 16 MiB array (LHL_ERR_ARG otherwise).  The call
 * returns with x written.  Every call on one GPU uses that GPU's array: calls
 * that overlap on one GPU (several host threads) race.  Nothing in the array
 * needs to start at zero: each call overwrites what it reads. */
int lhl_applyZ_noalloc(lhl_handle *h, int inverse, void *x, int64_t n, int64_t nrhs, int64_t ldx, int x_dtype,
                       int x_complex);

/* ---- getters (host copies, for parity checks and the wrapper's properties) ---- */

/* n, element type, shift element type and whether a reduction is held
 * (ws.reduced, src:158-160). */
int lhl_get_info(const lhl_handle *h, int64_t *n, int *dtype, int *shift_complex, int *reduced);
/* ws.ipiv (n-2 entries, 0-based pivot rows), ws.perm, ws.iperm (n, 0-based). */
int lhl_get_ipiv(lhl_handle *h, int32_t *ipiv);
int lhl_get_perm(lhl_handle *h, int32_t *perm, int32_t *iperm);
/* ws.scale, ws.iscale as doubles (powers of two). */
int lhl_get_scale(lhl_handle *h, double *scale, double *iscale);
/* ws.factors (n x n, ld = n, handle's element type; interleaved complex for a
 * complex workspace): H in triu(.,-1), step-c multipliers in column c below
 * the subdiagonal (src:127-129). */
int lhl_get_factors(lhl_handle *h, void *factors);
/* ws.Ht: H transposed, n x n with ld = n (interleaved complex for a complex
 * workspace); entries below the first superdiagonal of Ht (i > j+1 of H) are
 * zero. */
int lhl_get_ht(lhl_handle *h, void *ht);
/* The shift state (sh NULL = ws.shift): U (n x n, ld = n, U[k, j] at
 * index k + j*n, zero below the diagonal), the multipliers l (n-1), swap
 * flags (n), rdiag (n), sigma, tau, info.  For a complex state the
 * arrays hold interleaved complex values. */
int lhl_get_shift_state(lhl_handle *h, lhl_shift_state *sh, void *U, void *l, uint8_t *swap,
                        void *rdiag, double *sigma, double *tau, int64_t *info);

/* The scalar fields of a shift state (sh NULL = ws.shift) without a device
 * copy: ws.info (1-based first zero pivot, 0 if none), ws.sigma, ws.tau (each
 * {re, im}) and whether a shift is factored for the current reduction.  The
 * property forwarding of src:182-191 (ws.info, ws.sigma, ws.tau) reads
 * these.  Any pointer may be NULL. */
int lhl_get_shift_info(lhl_handle *h, lhl_shift_state *sh, int64_t *info, double *sigma, double *tau,
                       int *loaded);

/* ---- tuning (no reference counterpart: the knobs the reference fixes as
 * constants, _lhl_block_min / _lhl_panel_width (src:703-706), _LHL_GEMV_GROUP
 * (src:1451), _lhl_shift_fused_min (src:2494), become per-handle settings the
 * autotuner searches) ----
 * knob                 meaning                                          default
 * reduce_blocked_min   n from which lhl() runs delayed-update panels    3073 (FP64), off (FP32)
 * reduce_nb            panel width (8 or 16)                            16
 * reduce_rmin          fewest rows per workgroup of a panel step        16
 * reduce_switch        step (thousandths of n) from which the
 *                      row-resident loop takes over from the panels     400
 *                      (0: the row-resident loop from the first step)
 * lsweep_depth         halvings of the L sweeps' triangular products    2
 * shift_wide_min       n from which the shift LU's bulk is 960 wide     4481 (FP64), 8513 (FP32)
 * shift_pool_bytes     bound on one chunk's packed U's (batched solve)  min(64 GiB, free/3)
 * h2d_chunk_bytes      bytes per pinned staging chunk of a host J's     16 MiB
 *                      upload in lhl_reduce (0: the runtime's copy)
 * gemm_impl            the solves' dense products (L sweeps, J x):      0
 *                      0 rocBLAS, 1 the library's MFMA kernels (a
 *                      plan's masked refinement rounds always use them)
 * group_bytes          lhl_reduce_group: row-major bytes per sub-group   256 MiB
 *                      (each size class reads its first handle's knob)
 * shift_split          most workgroups per real shift LU (1: one)       4
 * utrsm_split          most workgroups per 16 columns of a real shift's 8
 *                      U back substitution (1: the one-workgroup-per-
 *                      4-columns kernel)
 * A value of -1 restores the default; lhl_tune_get reports -1 for a knob at
 * its default.  Results do not depend on the knobs beyond rounding (the
 * pivots of a reduction match for every panel setting). */
int lhl_tune_set(lhl_handle *h, const char *knob, int64_t value);
int lhl_tune_get(lhl_handle *h, const char *knob, int64_t *value);

/* Synchronize the handle's stream. */
int lhl_sync(lhl_handle *h);

/* ---- measurement (no reference counterpart; for benchmarks) ---- */

/* The HIP stream (hipStream_t) every kernel of the handle runs on. */
int lhl_get_stream(lhl_handle *h, void **stream);
/* While enabled, every kernel launch is bracketed by HIP events on the
 * handle's stream.  lhl_profile_read synchronizes, then returns the launches
 * and the summed duration (ms) of one kernel since the last reset:
 * reduction "balance", "reduce_copy", "reduce_step" (one launch per
 * elimination step), "reduce_pivot", "reduce_swap", "reduce_trailing" (the
 * three-launch step of n beyond the row-resident loop), "reduce_finish";
 * shift "shift", "uinv", "ut_pack"; solve "gather",
 * "trsm", "fwd_chain", "utrsm", "trmm", "scatter", "applyz", "convert"; refinement
 * "gemm_jx", "resid_combine", "colnorms". */
int lhl_profile_enable(lhl_handle *h, int on);
int lhl_profile_reset(lhl_handle *h);
int lhl_profile_read(lhl_handle *h, const char *kernel, int64_t *launches, double *total_ms);
/* The same, plus the work the launches were given, summed: shift_units
 * counts the shifts each launch processed, col_units the internal
 * right-hand-side columns (complex columns count once per plane); 0 for
 * launches that do not split that way.  Either pointer may be NULL. */
int lhl_profile_read_work(lhl_handle *h, const char *kernel, int64_t *launches, double *total_ms,
                          int64_t *shift_units, int64_t *col_units);

#ifdef __cplusplus
}
#endif
#endif
