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

/* ---- single routines, exported for the oracle's own pins (FP64 / FP32) ---- */

/* calderxabs (dir 0) / calderzabs (dir 1) with xflag/zflag = shift of a field
 * u [nx][nz] that is zero outside the grid. */
int oracle_rsg_deriv_f64(int dir, int shift, int nx, int nz, const double *u, double *der, double h, int radius,
                         const float *coef)
{
    if (nx < 1 || nz < 1 || radius < 1 || radius > 5) return -1;
    const int H = 5;
    double *uh = calloc((size_t)(nx + 2 * H) * (nz + 2 * H), sizeof(double));
    if (!uh) return -2;
    for (int i = 0; i < nx; ++i)
        memcpy(uh + (size_t)(i + H) * (nz + 2 * H) + H, u + (size_t)i * nz, (size_t)nz * sizeof(double));
    double a[5];
    for (int m = 0; m < radius; ++m) a[m] = coef[m];
    deriv_f64(dir, shift, uh, der, h, a, radius, nx, nz, 1);
    free(uh);
    return 0;
}

/* csponge for one side (lo = 1: spg*n and spg1*n; lo = 0: spg*p and spg2*p). */
void oracle_csponge_f32(int np, float h, float r, int lo, float *node, float *half)
{
    csponge_side_f32(np, h, r, node, half, lo ? -1 : +1);
}

/* Weight addsrc gives each cell of an nx x nz grid for a source at (isx, isz)
 * (0-based): w[x][z] = g2c(idx1, idx2) with idx2 = x - isx + hw2,
 * idx1 = z - isz + hw1, hw = nint((n + 1) / 2.0) (Q sub:1354-1389). */
int oracle_visco_source_weights_f32(int nx, int nz, int isx, int isz, float sigma, float *w)
{
    const int n1 = nz, n2 = nx;
    float *g = malloc((size_t)n1 * n2 * sizeof(float));
    if (!g) return -2;
    gauss_filter_f32(n1, n2, sigma, g);
    const int hw1 = (int)lround((n1 + 1) / 2.0), hw2 = (int)lround((n2 + 1) / 2.0);
    for (int x = 0; x < nx; ++x)
        for (int z = 0; z < nz; ++z) {
            const int idx2 = x - isx + hw2, idx1 = z - isz + hw1;
            w[(size_t)x * nz + z] =
                (idx1 >= 1 && idx1 <= n1 && idx2 >= 1 && idx2 <= n2) ? g[(size_t)(idx2 - 1) * n1 + (idx1 - 1)] : 0.f;
        }
    free(g);
    return 0;
}
