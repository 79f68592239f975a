/*
 * CPU oracle: isotropic elastic standard-staggered-grid propagator with a
 * free surface over topography, forward/yang_version_2016_2_2
 * (Iso_multishot_fd_subroutine_fsbc_zy_m2_pub.f90 / _pri.f90, time loop
 * isoewm2dfsbc, pri:57-230), restated in C.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the test-suite and of the
 * benchmark's CPU-baseline leg).  Product code never links it.
 *
 * The Fortran program cannot be built here (no Fortran compiler, and it needs
 * an MPI module), so this restatement is pinned by
 * tests/test_elastic_oracle.py:
 *   - every stencil of calux / caluz (orders 2..10, both staggers) against
 *     polynomials it must differentiate exactly, and the order reduction next
 *     to the surface against the rows where each order applies;
 *   - the staggered moduli and buoyancies of substagger against their
 *     definitions on layered models, and their replication above the surface;
 *   - the free-surface conditions in the output: tau_zz = 0 at and above the
 *     surface, the tau_xz image, vz above the surface from tau_zz = 0;
 *   - the physics: a P wave's arrival time in a homogeneous full space, a
 *     Rayleigh wave's speed along a flat free surface, x-mirror symmetry,
 *     linearity in the wavelet, and FP32 / FP64 agreement.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define REAL float
#define ORACLE_F32 1
#define FN(x) x##_f32
#include "elastic_body.h"
#undef REAL
#undef ORACLE_F32
#undef FN

#define REAL double
#define ORACLE_F32 0
#define FN(x) x##_f64
#include "elastic_body.h"
#undef REAL
#undef ORACLE_F32
#undef FN

/* One application of calux (axis 0) or caluz (axis 1) with flag = xflag /
 * zflag, to u [nx][nz] (zero outside the grid), with (dt/h) pa = 1 and a
 * zero accumulator: out = the stencil sum, rows k >= itopo[i] only (other
 * cells 0).  FP64 arithmetic on the FP32 coefficient tables. */
int oracle_ssg_deriv_f64(int axis, int flag, int nx, int nz, const double *u, const int *itopo, double *out)
{
    if (nx < 1 || nz < 1 || (axis != 0 && axis != 1) || (flag != 0 && flag != 1)) return -1;
    for (int i = 0; i < nx; ++i)
        if (itopo[i] < 0 || itopo[i] > nz - 6) return -1;
    const int Hh = 5;
    const size_t nzh = (size_t)nz + 2 * Hh, nh = (size_t)(nx + 2 * Hh) * nzh;
    double *uh = calloc(nh, sizeof(double)), *acc = calloc(nh, sizeof(double)), *pa = malloc(nh * sizeof(double));
    if (!uh || !acc || !pa) {
        free(uh);
        free(acc);
        free(pa);
        return -2;
    }
    for (size_t c = 0; c < nh; ++c) pa[c] = 1.0;
    for (int i = 0; i < nx; ++i) memcpy(uh + (size_t)(i + Hh) * nzh + Hh, u + (size_t)i * nz, (size_t)nz * sizeof(double));
    double tab[5][5];
    el_tables_f64(tab);
    if (axis == 0) el_calux_f64(flag, nx, nz, 1.0, 1.0, uh, acc, pa, itopo, tab[4], 1);
    else el_caluz_f64(flag, nx, nz, 1.0, 1.0, uh, acc, pa, itopo, itopo, 0, tab, 1);
    for (int i = 0; i < nx; ++i) memcpy(out + (size_t)i * nz, acc + (size_t)(i + Hh) * nzh + Hh, (size_t)nz * sizeof(double));
    free(uh);
    free(acc);
    free(pa);
    return 0;
}

/* substagger (pub:323-547) alone: the staggered fields the time loop uses,
 * [5][nx*nz] in the order busg2e, busg1e, lambmusg, lambsg, musge. */
int oracle_elastic_stagger_f32(int nx, int nz, const float *vp, const float *vs, const float *rho, const int *itopo,
                               float *out)
{
    if (nx < 2 || nz < 2) return -1;
    for (int i = 0; i < nx; ++i)
        if (itopo[i] < 0 || itopo[i] >= nz) return -1;
    return el_stagger_f32(nx, nz, vp, vs, rho, itopo, out);
}

int oracle_elastic_stagger_f64(int nx, int nz, const float *vp, const float *vs, const float *rho, const int *itopo,
                               double *out)
{
    if (nx < 2 || nz < 2) return -1;
    for (int i = 0; i < nx; ++i)
        if (itopo[i] < 0 || itopo[i] >= nz) return -1;
    return el_stagger_f64(nx, nz, vp, vs, rho, itopo, out);
}

/* gauss_filter_2d (pub:592-634) alone, in place on u [nx][nz]: rows 0 .. itopo(i)
 * of every column, serial order, window 2 rgau - 1, sigma 2 (normalised). */
int oracle_elastic_smooth_f32(int nx, int nz, int rgau, const int *itopo, float *u)
{
    if (nx < 1 || nz < 1 || rgau < 1) return -1;
    for (int i = 0; i < nx; ++i)
        if (itopo[i] < 0 || itopo[i] >= nz) return -1;
    const int ng = 2 * rgau - 1, cen = (ng + 1) / 2;
    float *g = malloc((size_t)ng * ng * sizeof(float));
    if (!g) return -2;
    const float PI = 3.1415926f, sigma2 = 4.0f;
    float sum = 0;
    for (int j = 1; j <= ng; ++j)
        for (int i = 1; i <= ng; ++i) {
            const float dis = (float)(abs(j - cen) * abs(j - cen) + abs(i - cen) * abs(i - cen));
            g[(size_t)(j - 1) * ng + (i - 1)] = 1 / (2 * PI * sigma2) * expf(-dis / (2 * sigma2));
        }
    for (int c = 0; c < ng * ng; ++c) sum = sum + g[c];
    for (int c = 0; c < ng * ng; ++c) g[c] = g[c] / sum;
    el_gauss_smooth_f32(nx, nz, rgau, itopo, g, u);
    free(g);
    return 0;
}
