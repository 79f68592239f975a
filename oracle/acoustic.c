/*
 * CPU oracle: acoustic velocity-pressure staggered-grid propagator of
 * forward/Acoustic_FD_by_ss/src/Acoustic_1order_FD.cpp, restated in C.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the test-suite, the smoke check and
 * the benchmark's CPU-baseline leg).  Product code never links it.
 *
 * Parity pins (tests/test_oracle.py):
 *   oracle_acoustic_f32  - bitwise against the reference program compiled here
 *                          (oracle/_ref) and against the outputs the reference
 *                          ships (forward/Acoustic_FD_by_ss/file/record.dat,
 *                          Snap.dat); x-mirror symmetry; linearity; energy
 *                          conservation without the sponge; 1-D dispersion
 *                          relation; FP64 agreement.
 *   oracle_sponge_table  - closed form, and the reference program's table via
 *                          its outputs.
 *   oracle_ricker        - closed form in double.
 *   oracle_staggered_coefs - Taylor conditions, the reference's C1..C8
 *                          (:38-45) and a80..a104
 *                          (yang_version_2016_2_2/Iso_multishot_fd_module_zy_m2.f90:9-26).
 *   C-PML (boundary == 1) - no reference counterpart: pinned by reflection
 *                          measurements against a large-domain run.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

void oracle_sponge_table(int nx, int nz, int width, float *absorb)
{
    const int L = width;
    for (size_t c = 0; c < (size_t)nx * nz; ++c) absorb[c] = 1.0f; /* :126-132 */
    if (L <= 0) return;
    /* z-edges for interior rows, :164-178. */
    for (int i = L; i < nx - L; ++i) {
        for (int j = 0; j < L; ++j)
            absorb[(size_t)i * nz + j] = (float)(-0.38 * (L - j - 1) * (L - j - 1) / (1.0 * L * L) + 1);
        for (int j = nz - L; j < nz; ++j)
            absorb[(size_t)i * nz + j] = (float)(-0.38 * (j - nz + L) * (j - nz + L) / (1.0 * L * L) + 1);
    }
    /* x-edges over every column, :180-194 (the reference bounds the last loop
     * by Nx, equal to Nz in its square grid; the full z range is meant). */
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < nz; ++j)
            absorb[(size_t)i * nz + j] = (float)(-0.38 * (L - i - 1) * (L - i - 1) / (1.0 * L * L) + 1);
    for (int i = nx - L; i < nx; ++i)
        for (int j = 0; j < nz; ++j)
            absorb[(size_t)i * nz + j] = (float)(-0.38 * (i - nx + L) * (i - nx + L) / (1.0 * L * L) + 1);
}

void oracle_ricker(int nt, float dt, float freq, float amp, float *w)
{
    const double PI_REF = 3.1415926; /* :13 */
    const float t1 = (float)(1.0 / freq); /* float t1 = 1.0/freq, :202,206 */
    for (int i = 0; i < nt; ++i) {
        const float t = dt * (float)i; /* :205 */
        const float d = t - t1;
        w[i] = (float)((double)amp * (1 - 2 * PI_REF * PI_REF * freq * freq * d * d) *
                       exp(-PI_REF * PI_REF * freq * freq * d * d));
    }
}

void oracle_staggered_coefs(int radius, double *c)
{
    /* Weights of sum_m c_m (f(x+(m-1/2)h) - f(x-(m-1/2)h)) = h f'(x) + O(h^(2R+1)):
     * with x_m = (2m-1)^2 the moment conditions sum_m c_m (2m-1) x_m^(k-1) = [k==1]
     * are a Vandermonde system whose solution is the Lagrange basis at 0. */
    for (int m = 1; m <= radius; ++m) {
        const double xm = (double)(2 * m - 1) * (2 * m - 1);
        double u = 1.0;
        for (int l = 1; l <= radius; ++l) {
            if (l == m) continue;
            const double xl = (double)(2 * l - 1) * (2 * l - 1);
            u *= xl / (xl - xm);
        }
        c[m - 1] = u / (2 * m - 1);
    }
}

/* C-PML recursive-convolution coefficients (Roden & Gedney 2000; Komatitsch &
 * Martin 2007, kappa = 1) at coordinate x = i + shift (cells) on an axis of n
 * cells whose first and last `w` cells are absorbing.  d(x) = d0 xi^2,
 * d0 = -3 vmax ln(R) / (2 w h), alpha = pi f0 (1 - xi), xi = depth / w. */
void oracle_cpml_coef(int i, double shift, int n, int w, double h, double dt, double R,
                      double vmax, double f0, double *b, double *a)
{
    const double x = i + shift;
    const double lo = w, hi = n - 1 - w;
    double depth = 0.0;
    if (x < lo) depth = lo - x;
    else if (x > hi) depth = x - hi;
    if (w <= 0 || depth <= 0.0) { *b = 1.0; *a = 0.0; return; }
    double xi = depth / w;
    if (xi > 1.0) xi = 1.0;
    const double d0 = -3.0 * vmax * log(R) / (2.0 * w * h);
    const double d = d0 * xi * xi;
    const double alpha = M_PI * f0 * (1.0 - xi);
    *b = exp(-(d + alpha) * dt);
    *a = d * (*b - 1.0) / (d + alpha);
}

#define ORACLE_F32 1
#define REAL float
#define FN(x) x##_f32
#include "acoustic_body.h"
#undef ORACLE_F32
#undef REAL
#undef FN

#define ORACLE_F32 0
#define REAL double
#define FN(x) x##_f64
#include "acoustic_body.h"
#undef ORACLE_F32
#undef REAL
#undef FN
