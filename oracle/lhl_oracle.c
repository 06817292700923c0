/*
 * CPU oracle for the LHL hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of SciML/LHL.jl's algorithm (src/LHLFactorization.jl),
 * built into oracle/liblhl_oracle.so by oracle/Makefile.  It is loaded only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, always as
 * the checker or the timed CPU baseline, never as a product path.
 *
 * Parity status of each function (how it is pinned, not by retyping it):
 *   lhlo_balance_*          exact-scale invariant D(D^-1 J D)D^-1 == J, power-of-2
 *                           scales, idempotence on a balanced matrix, closed forms.
 *   lhlo_reduce_*           J == Z H Z^-1, Z Z^-1 == I, |multipliers| <= 1,
 *                           eig(H) == eig(J) (LAPACK), blocked == unblocked ipiv,
 *                           upper-Hessenberg and diagonal closed forms, n in {1,2}.
 *   lhlo_shift_* (real)     rows == fused-R variant bit for bit (the reference's own
 *                           same_state test), [1 0;0 2] known answers, (sigma I + tau H)
 *                           == P L U reconstruction, dense-LU solves.
 *   lhlo_shift_c_*          complex: tolerance only (Julia's complex division bits
 *                           cannot be reproduced without Julia) -- dense complex solve.
 *   lhlo_ldiv_* and refine  numpy/LAPACK dense solves at rtol 1e-9 and the bwd bound,
 *                           mpmath high-precision solves on tiny inputs.
 *   lhlo_z* (complex J)     J == Z H Z^-1 and eig(H) == eig(J) (complex LAPACK),
 *                           |multipliers| <= 1, upper-Hessenberg closed form, the
 *                           real restatement on a complex J with zero imaginary part
 *                           (same pivots, factors and scales), P L U of the complex
 *                           shift, dense complex solves, mpmath on tiny inputs.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define REAL double
#define SFX _d
#define ABS fabs
#define FMA fma
#define BLOCK_MIN 500
#include "lhl_oracle_body.h"
#define SQRT sqrt
#include "lhl_oracle_cplx.h"
#undef SQRT
#undef REAL
#undef SFX
#undef ABS
#undef FMA
#undef BLOCK_MIN

#define REAL float
#define SFX _s
#define ABS fabsf
#define FMA fmaf
#define BLOCK_MIN 1024
#include "lhl_oracle_body.h"
#define SQRT sqrtf
#include "lhl_oracle_cplx.h"
#undef SQRT
#undef REAL
#undef SFX
#undef ABS
#undef FMA
#undef BLOCK_MIN

int lhlo_version(void) { return 1; }
