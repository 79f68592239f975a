/*
 * CPU oracle: visco-elastic rotated-staggered-grid propagator of
 * forward/Q_forward (Q_multishot_fd_subroutine_fsbc_zy_rotated.f90 and the
 * time loop of Q_multishot_fd_main_fsbc_zy_rotated.f90), restated in C.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the test-suite and of the
 * benchmark's CPU-baseline leg).  Product code never links it.
 *
 * The Fortran program cannot be built here (no Fortran compiler with an MPI
 * module), so this restatement is pinned by (tests/test_visco_oracle.py):
 *   - the vs = 0 reduction: sxx and szz bitwise equal, sxz identically zero;
 *   - Taylor conditions of the rotated-grid derivative (polynomials of degree
 *     < 2R differentiated exactly along both axes, through the propagator);
 *   - the sponge tables against their closed forms;
 *   - the Gaussian source footprint: sums to one, centred on the source cell
 *     (odd grids) or one cell before it (even grids, the reference's rounding);
 *   - the tau-method: L = 1 equals the reference's closed form; the fitted
 *     Q(w) of the standard-linear-solid model is flat over the band;
 *   - memory-variable relaxation and Q -> infinity reducing to the elastic
 *     (lossless) scheme; mirror symmetry; linearity; FP64 agreement.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define REAL float
#define ORACLE_F32 1
#define FN(x) x##_f32
#include "visco_body.h"
#undef REAL
#undef ORACLE_F32
#undef FN

#define REAL double
#define ORACLE_F32 0
#define FN(x) x##_f64
#include "visco_body.h"
#undef REAL
#undef ORACLE_F32
#undef FN
