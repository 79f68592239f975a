/*
 * CPU oracle for the 2-D finite-difference wave propagators of
 * stone0312/Seismic_code_github `forward/`.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithms, used by the test-suite, by the smoke check and by the benchmark's
 * CPU-baseline leg as the checker.  No product code path links or calls it.
 *
 * Layout (as in the reference): every field is [nx][nz], z fastest
 * (forward/Acoustic_FD_by_ss/src/Acoustic_1order_FD.cpp:51-100 allocates
 * `float*[Nx]` rows of `Nz` floats).  Indices are 0-based.
 *
 * Two arithmetics are provided for every propagator:
 *   *_f32 : the reference's own arithmetic, operation for operation, including
 *           the double-promoted sub-expressions of the C++ source (built with
 *           -ffp-contract=off and no -march, so no FMA is formed, as in the
 *           reference's `g++ -O3` build, Makefile:14).
 *   *_f64 : the same algorithm with every operation in double, on the same
 *           float inputs; used to state the FP32 rounding error of a result.
 */
#ifndef SEISFD_ORACLE_H
#define SEISFD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Acoustic velocity-pressure staggered-grid propagator,
 * Acoustic_1order_FD.cpp:126-393, generalised to any grid, model, source
 * position, receiver list, sponge width and stencil radius. */
typedef struct oracle_acoustic_args {
    int nx, nz;              /* full grid incl. sponge (reference Nx, Nz) */
    int nt;                  /* number of time steps to run */
    int k0;                  /* index of the first step (wavelet sample) */
    float dx, dz, dt;
    int radius;              /* R: number of coefficient pairs, 1..8 */
    const float *coef;       /* C_1..C_R (reference C1..C8, :38-45) */
    int sponge;              /* quadratic sponge width on every side (reference PML) */
    const float *absorb;     /* optional [nx*nz] sponge table; NULL = build it */
    const float *vp;         /* [nx*nz] */
    const float *den;        /* [nx*nz] */
    const float *wavelet;    /* [k0+nt] source samples (reference `signal`) */
    int sx, sz;              /* source cell (reference nx_location, nz_location) */
    int nrec;
    const int *rx, *rz;      /* receiver cells */
    int nsnap;
    const int *snap_steps;   /* steps k after which P, Vx, Vz are copied out */
    int nthreads;            /* OpenMP threads for the stencil loops (1 = serial) */
    int boundary;            /* 0 = reference quadratic sponge, 1 = C-PML */
    float pml_r;             /* C-PML target reflection coefficient */
    float pml_vmax;          /* C-PML reference velocity */
    float pml_f0;            /* C-PML alpha frequency (Hz) */
} oracle_acoustic_args;

/* Outputs (any may be NULL):
 *   record [nrec][nt]  trace-major, like record.dat (:385-393)
 *   snaps  [nsnap][3][nx*nz]  P, Vx, Vz after the absorb of step snap_steps[s]
 *   state  [3][nx*nz]  initial P, Vx, Vz on entry (NULL = zero), final on exit
 *   pml    [4][nx*nz]  initial/final C-PML memory fields (psi_px, psi_pz,
 *                      psi_vx, psi_vz), only used when boundary == 1
 * Returns 0, or a negative value on invalid arguments. */
int oracle_acoustic_f32(const oracle_acoustic_args *a, float *record,
                        float *snaps, float *state, float *pml);
int oracle_acoustic_f64(const oracle_acoustic_args *a, double *record,
                        double *snaps, double *state, double *pml);

/* The reference's quadratic sponge table, Acoustic_1order_FD.cpp:126-132,164-194:
 * 1 - 0.38 d^2 / L^2 (computed in double, stored as float).  The x-edge
 * strips (rows) overwrite the z-edge strips in the corners. */
void oracle_sponge_table(int nx, int nz, int width, float *absorb);

/* The reference's Ricker wavelet, Acoustic_1order_FD.cpp:199-209:
 * amp (1 - 2 pi^2 f^2 (t-1/f)^2) exp(-pi^2 f^2 (t-1/f)^2), t = k dt in float,
 * pi = 3.1415926, double arithmetic, stored as float. */
void oracle_ricker(int nt, float dt, float freq, float amp, float *w);

/* Staggered first-derivative Taylor coefficients of half-width R, in double. */
void oracle_staggered_coefs(int radius, double *c);

/* C-PML profile (b, a) at x = i + shift cells on an axis of n cells whose first
 * and last w cells absorb. */
void oracle_cpml_coef(int i, double shift, int n, int w, double h, double dt, double R,
                      double vmax, double f0, double *b, double *a);

/* Discrete adjoint (FP64) of the acoustic step with the sponge boundary:
 * for J = sum_{r,k} seed[r][k] * record[r][k], returns dJ/dvp [nx*nz] and,
 * if record != NULL, the forward record [nrec][nt] (k0 must be 0). */
int oracle_acoustic_gradient_f64(const oracle_acoustic_args *a, const double *seed, double *grad_vp,
                                 double *record);

/* Aperture window + edge padding of one shot, get_currshot_parameters
 * (forward/Q_forward/Q_multishot_fd_subroutine_fsbc_zy_rotated.f90:288-373). */
void oracle_window(const float *model, int nx, int nz, int ixmine, int nxe, int npxn, int npxp,
                   int npzn, int npzp, float *out);


/* Visco-elastic rotated-staggered-grid propagator, forward/Q_forward
 * (Q_multishot_fd_subroutine_fsbc_zy_rotated.f90:434-1444 and the time loop
 * of Q_multishot_fd_main_fsbc_zy_rotated.f90:436-459), with L standard
 * linear solids (the reference: L = 1).  Grid = the reference's padded shot
 * grid (nxea x nzea); indices 0-based. */
#define ORACLE_MAX_MECH 8
typedef struct oracle_visco_args {
    int nx, nz;              /* padded grid: reference nxea (x), nzea (z) */
    int nsteps;              /* steps run (reference nt + decay) */
    int it0;                 /* Fortran index of the first step (1 = fresh run) */
    float dx, dz, dt, f0;
    int radius;              /* 5 = a100..a104 (reference), 1..5 */
    const float *coef;       /* a_0..a_{R-1} */
    int nmech;               /* L */
    float q_scale;           /* Q multiplier before the fit (reference 18/20) */
    float f_low, f_high;     /* fit band in Hz; <= 0 = reference (2 Hz, 2.5 f0) */
    int npxn, npxp, npzn, npzp;  /* sponge table lengths - 1 (reference np*) */
    float sponge_r;          /* reference R = 1e-3 */
    const float *vp, *vs, *rho, *qp, *qs;  /* [nx*nz]; vs, qs may be NULL (0, qp) */
    int isx, isz;            /* source cell (reference isxe-1, isze-1) */
    float src_sigma;         /* Gaussian source width in cells (reference 1.5) */
    const float *wavelet;    /* [nsteps] or NULL = the reference's formula */
    int nrec;
    const int *rx, *rz;      /* receiver cells */
    int rec_delay;           /* steps before recording (reference decay) */
    int nsnap;
    const int *snap_steps;   /* step indices s (0-based) after which fields are copied */
    int nthreads;
    const float *init;       /* NULL, or [5][nx*nz] initial sxx, szz, sxz (time 0) and vx, vz
                                (time -1/2), memory variables zero: for the oracle's own pins
                                (plane waves); the reference always starts from rest */
} oracle_visco_args;

/* rec_zz, rec_xx [nrec][nsteps - rec_delay] (reference record1 = strzz,
 * record2 = strxx); snaps [nsnap][5][nx*nz]: sxx, szz, sxz, vx, vz.  Any
 * output may be NULL.  Returns 0, or a negative value on invalid arguments. */
int oracle_visco_f32(const oracle_visco_args *a, float *rec_zz, float *rec_xx, float *snaps);
int oracle_visco_f64(const oracle_visco_args *a, double *rec_zz, double *rec_xx, double *snaps);

/* Relaxation times tau_sigma[L] and the tau-method's tau for quality factor q
 * (pcalpas, Q sub:506-542, generalised to L mechanisms). */
void oracle_visco_tau_f32(int L, double f0, double q, double q_scale, double f_low, double f_high,
                          float *tau_sigma, float *tau);
void oracle_visco_tau_f64(int L, double f0, double q, double q_scale, double f_low, double f_high,
                          double *tau_sigma, double *tau);

/* Single routines of the visco restatement, for the oracle's own pins:
 * one rotated-grid derivative (calderxabs dir 0 / calderzabs dir 1, xflag =
 * shift) of u [nx][nz] (zero outside), FP64; one side of csponge; and the
 * per-cell weight of addsrc's Gaussian footprint. */
int oracle_rsg_deriv_f64(int dir, int shift, int nx, int nz, const double *u, double *der, double h, int radius,
                         const float *coef);
void oracle_csponge_f32(int np, float h, float r, int lo, float *node, float *half);
int oracle_visco_source_weights_f32(int nx, int nz, int isx, int isz, float sigma, float *w);

/* Isotropic elastic standard-staggered-grid propagator over topography:
 * isoewm2dfsbc (free surface, yang_version_2016_2_2 pri:57-230) or
 * isoewm2dabc (absorbing top following the topography, pri:232-386, with
 * modipapml's smoothing above the surface) and their set-up (substagger,
 * csponge, get_gauss_filter_coe). */
typedef struct oracle_elastic_args {
    int nx, nz;              /* padded shot grid: reference nxea (x), nzea (z) */
    int nsteps;              /* steps run (reference nt + idecay) */
    int delay;               /* reference idecay: step `delay` is the first recorded */
    float dx, dz, dt, f0;
    const float *vp, *vs, *rho;  /* [nx*nz] before substagger */
    const int *itopo;        /* [nx] 0-based free-surface rows (reference itopoe - 1) */
    int npxn, npxp, npzp;    /* sponge table lengths - 1 (reference np*) */
    float sponge_r;          /* reference R = 1e-3 */
    int isx, isz;            /* source cell (reference isxe - 1, isze - 1) */
    int rgau;                /* Gaussian window half-width + 1 (window 2 rgau - 1) */
    const float *wavelet;    /* [nsteps] or NULL = addsrc's Ricker */
    int nrec;
    const int *rx, *rz;      /* receiver cells */
    int nsnap;
    const int *snap_steps;   /* ascending step indices s after which fields are copied */
    int nthreads;
    int mode;                /* 0 = 'fsbc' (isoewm2dfsbc), 1 = 'abc' (isoewm2dabc, pri:232-386) */
    int npzn;                /* 'abc': top sponge table length - 1 (reference npzn) */
} oracle_elastic_args;
/* rec_vx, rec_vz [nrec][nsteps - delay] (reference record1 = vx, record2 = vz);
 * snaps [nsnap][5][nx*nz]: vx, vz, tau_xx, tau_zz, tau_xz. */
int oracle_elastic_f32(const oracle_elastic_args *a, float *rec_vx, float *rec_vz, float *snaps);
int oracle_elastic_f64(const oracle_elastic_args *a, double *rec_vx, double *rec_vz, double *snaps);
/* Single routines for the oracle's own pins: one calux (axis 0) / caluz
 * (axis 1) stencil sum with flag = xflag / zflag, FP64; substagger's staggered
 * fields [5][nx*nz] (busg2e, busg1e, lambmusg, lambsg, musge). */
int oracle_ssg_deriv_f64(int axis, int flag, int nx, int nz, const double *u, const int *itopo, double *out);
int oracle_elastic_stagger_f32(int nx, int nz, const float *vp, const float *vs, const float *rho, const int *itopo,
                               float *out);
int oracle_elastic_stagger_f64(int nx, int nz, const float *vp, const float *vs, const float *rho, const int *itopo,
                               double *out);
/* modipapml's gauss_filter_2d (pub:592-634), serial order, in place. */
int oracle_elastic_smooth_f32(int nx, int nz, int rgau, const int *itopo, float *u);

/* VTI qP-acoustic propagator of de-multiple/convolution_srm/vti_qa_2d_c3_demo1:
 * p (reference u) and q fields, 2nd order in time, 10th-order second
 * derivatives, split PML on the four sides (extrapolation_one_step_1 / _2,
 * VTI_multishot_qacoustic_fd_subroutine_c3.f90:432-1258). */
typedef struct oracle_vti_args {
    int nx, nz;              /* padded shot grid (reference currshot_range_all, nz) */
    int nsteps;              /* steps run (reference nt + nint((nwt - 1) / 2)) */
    int delay;               /* record sample = step - delay (reference iit = it - nint((nwt-1)/2)) */
    int nt;                  /* recorded samples */
    float dx, dz, dt;
    const float *c2;         /* [5] second-derivative weights (fd_coe = fd_coe_pml_2nd) */
    const float *c1;         /* [5] first-derivative weights (fd_coe_pml_1st) */
    int nl, nr, nu, nd;      /* PML widths: left, right, top, bottom (nx_bound_l, _r, nz_bound_u, _d) */
    float r;                 /* target reflection (module constant R); 0 = 1e-3 */
    const float *vz, *eps, *del;  /* [nx*nz] */
    const float *wavelet;    /* [nwav]: added to p at the source before steps 0 .. nwav-1 */
    int nwav;
    int sx, sz;              /* source cell */
    int nrec;
    const int *rx, *rz;      /* receiver cells: p after each step */
    int nthreads;            /* OpenMP threads over x columns (1 = serial) */
} oracle_vti_args;

/* record [nrec][nt]; state (may be NULL) [2][nx*nz]: final p and q. */
int oracle_vti_f32(const oracle_vti_args *a, float *record, float *state);
int oracle_vti_f64(const oracle_vti_args *a, double *record, double *state);
/* coefficient_2nd (kind 2) / coefficient_1st (kind 1) with k weights. */
int oracle_vti_coefficients_f32(int kind, int k, float *out);
/* The driver's wavelet: writes min(nwt, nmax) samples, returns nwt. */
int oracle_vti_wavelet_f32(float dt, float f0, int nmax, float *w);
/* get_currshot_parameters plus the isotropic top rows; outputs
 * [range + nl + nr - 2][nzv + nu + nd - 2]. */
int oracle_vti_window(const float *vz, const float *eps, const float *del, int nxv, int nzv, int xmin, int range,
                      int nl, int nr, int nu, int nd, int nsz0, float *ovz, float *oeps, float *odel);

#ifdef __cplusplus
}
#endif
#endif
