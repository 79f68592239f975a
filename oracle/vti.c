/*
 * CPU oracle: 2-D VTI qP-acoustic propagator (Zhang's coupled p / q scheme,
 * 2nd order in time, 10th order in space, split PML) of
 * de-multiple/convolution_srm/vti_qa_2d_c3_demo1 (VTI_multishot_qacoustic_fd_subroutine_c3.f90
 * and the time loop of VTI_multishot_qacoustic_fd_main_c3.f90), restated in C.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the test-suite and of the
 * benchmark's CPU-baseline leg).  Product code never links it.
 *
 * The Fortran program cannot be built here (no Fortran compiler with an MPI
 * module; its main program also has a trailing comma in two argument lists,
 * main:462-463, 508-509), so this restatement is pinned by tests/test_vti_oracle.py:
 *   - the coefficient solves against the closed-form central-difference
 *     weights (Taylor conditions);
 *   - eps = delta = 0 without PML against an independent FP64 textbook
 *     second-order acoustic scheme (q stays exactly zero);
 *   - eps = delta (elliptic): q stays zero and p equals the isotropic run on
 *     an x axis stretched by sqrt(1 + 2 delta);
 *   - the one-step impulse response of every coupling (p <- p, p <- q,
 *     q <- p, q <- q) against its closed form;
 *   - the qP phase velocities: along x vz sqrt(1 + 2 eps), along z vz;
 *   - the split PML: x-mirror symmetry, energy leaving the grid, small
 *     reflections against a run on a grid large enough to have none;
 *   - FP32 / FP64 agreement.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define REAL float
#define FN(x) x##_f32
#include "vti_body.h"
#undef REAL
#undef FN

#define REAL double
#define FN(x) x##_f64
#include "vti_body.h"
#undef REAL
#undef FN

/* coefficient_2nd (kind 2: a(i,j) = i^(2j) / (2j)!) and coefficient_1st
 * (kind 1: a(i,j) = i^(2j-1) / (2j-1)!), sub:1264-1472: the k x k system
 * inverted by k rank-one (Sherman-Morrison) updates of the identity, all in
 * single precision with the loop variables as reals, and
 * coefficient(i) = first row of inv(a) times 0.5 e_i. */
int oracle_vti_coefficients_f32(int kind, int k, float *out)
{
    if ((kind != 1 && kind != 2) || k < 1 || k > 8) return -1;
    float a[8][8], inv[8][8], temp[8], f4[8], f2[8][8], f3[8][8];
    for (int i = 1; i <= k; ++i)
        for (int j = 1; j <= k; ++j) {
            const float e = kind == 2 ? (float)(2 * j) : (float)(2 * j - 1);
            float fact = 1.0f;
            for (float m = 1; m <= e; m += 1) fact = fact * m;
            a[i - 1][j - 1] = powf((float)i, e) / fact;
        }
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) inv[i][j] = i == j ? 1.0f : 0.0f;
    for (int i = 0; i < k; ++i) {
        for (int j = 0; j < k; ++j) temp[j] = a[i][j];
        temp[i] = temp[i] - 1;
        float s = 0;
        for (int j = 0; j < k; ++j) s = s + temp[j] * inv[j][i];
        const float f1 = s + 1;
        for (int j = 0; j < k; ++j) f4[j] = inv[j][i];
        for (int j1 = 0; j1 < k; ++j1)
            for (int i1 = 0; i1 < k; ++i1) f2[j1][i1] = f4[j1] * temp[i1];
        for (int r = 0; r < k; ++r)
            for (int cc = 0; cc < k; ++cc) {
                float t = 0;
                for (int m = 0; m < k; ++m) t = t + f2[r][m] * inv[m][cc];
                f3[r][cc] = t / f1;
            }
        for (int r = 0; r < k; ++r)
            for (int cc = 0; cc < k; ++cc) inv[r][cc] = inv[r][cc] - f3[r][cc];
    }
    for (int i = 0; i < k; ++i) {
        /* p1 = matmul(inv, b) with b = 0.5 e_i; coefficient(i) = p1(1) */
        float t = 0;
        for (int m = 0; m < k; ++m) t = t + inv[0][m] * (m == i ? 0.5f : 0.0f);
        out[i] = t;
    }
    return 0;
}

/* The driver's wavelet, main:433-437: for it = 1..nwt,
 * x = pi f0 (it - 0.07/dt) dt, w = exp(-x^2) (1 - 2 x^2), single precision,
 * pi = 3.1415926 (module constant); nwt = nint(2 * 0.07 / dt) + 1 (main:277). */
int oracle_vti_wavelet_f32(float dt, float f0, int nmax, float *w)
{
    const float pi = 3.1415926f;
    const int nwt = (int)lroundf(2.0f * 0.07f / dt) + 1;
    for (int it = 1; it <= nwt && it <= nmax; ++it) {
        const float x = pi * f0 * ((float)it - 0.07f / dt) * dt;
        w[it - 1] = expf(-(x * x)) * (1 - 2 * (x * x));
    }
    return nwt;
}

/* get_currshot_parameters (sub:268-338) and the isotropic top rows of
 * single_shot_modelling (main:422-425), 0-based, [nx][nz] z fastest:
 *   grid nx = range + nl + nr - 2, nz = nzv + nu + nd - 2;
 *   window column i (1-based) takes model column min(max(i + xmin, 1), nxv)
 *   (i + xmin, not i + xmin - 1, as the reference writes it), rows
 *   nu+1 .. nu+nzv; vz is then replicated up, down, left and right, eps and
 *   delta are not (zero outside the window); eps = delta = 0 on rows
 *   1 .. nsz0 + nu + 6. */
int oracle_vti_window(const float *vz, const float *eps, const float *del, int nxv, int nzv, int xmin, int range,
                      int nl, int nr, int nu, int nd, int nsz0, float *ovz, float *oeps, float *odel)
{
    const int nx = range + nl + nr - 2, nz = nzv + nu + nd - 2;
    if (nx < range + nl || nz < nzv + nu) return -1;
    memset(ovz, 0, sizeof(float) * (size_t)nx * nz);
    memset(oeps, 0, sizeof(float) * (size_t)nx * nz);
    memset(odel, 0, sizeof(float) * (size_t)nx * nz);
#define G(f, i, j) f[(size_t)((i) - 1) * nz + ((j) - 1)]
#define M(f, i, j) f[(size_t)((i) - 1) * nzv + ((j) - 1)]
    for (int i = 1; i <= range; ++i) {
        int io = i + xmin;
        if (io < 1) io = 1;
        if (io > nxv) io = nxv;
        for (int j = 1; j <= nzv; ++j) {
            G(ovz, i + nl, nu + j) = M(vz, io, j);
            G(oeps, i + nl, nu + j) = M(eps, io, j);
            G(odel, i + nl, nu + j) = M(del, io, j);
        }
    }
    for (int j = 1; j <= nu; ++j)
        for (int i = 1; i <= nx; ++i) G(ovz, i, j) = G(ovz, i, nu + 1);
    for (int j = nu + nzv + 1; j <= nz; ++j)
        for (int i = 1; i <= nx; ++i) G(ovz, i, j) = G(ovz, i, nu + nzv);
    for (int j = 1; j <= nz; ++j)
        for (int i = 1; i <= nl; ++i) G(ovz, i, j) = G(ovz, nl + 1, j);
    for (int j = 1; j <= nz; ++j)
        for (int i = nl + range + 1; i <= nx; ++i) G(ovz, i, j) = G(ovz, nl + range, j);
    for (int j = 1; j <= nsz0 + nu + 6 && j <= nz; ++j)
        for (int i = 1; i <= nx; ++i) {
            G(oeps, i, j) = 0.0f;
            G(odel, i, j) = 0.0f;
        }
#undef G
#undef M
    return 0;
}
