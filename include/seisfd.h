/*
 * seisfd — MI355X (gfx950) finite-difference wave propagators: C ABI.
 *
 * Drop-in for the time loop of stone0312/Seismic_code_github
 * forward/Acoustic_FD_by_ss/src/Acoustic_1order_FD.cpp (the reference program
 * hard-codes its configuration at :28-45,137-216 and runs the loop at
 * :237-380).  Every input the reference program fixes in its constant block is
 * an argument here: grid (Nx, Nz, Nt, dx, dz, dt), stencil coefficients
 * (C1..C8, :38-45), velocity and density (Vp, DEN, :137-142), sponge width
 * (PML, :30), source wavelet and position (signal, nx_location/nz_location,
 * :199-216) and receiver cells (:376-377).
 *
 * Conventions
 *   - Fields are [nx][nz] float32, z fastest, 0-based: the reference's own
 *     layout (`float*` rows of Nz floats, :51-100).  No transpose is needed.
 *   - All host pointers are owned by the caller and only used during the call.
 *     Every call is synchronous with respect to the host buffers it takes.
 *   - Every entry point returns an int status (SFD_OK = 0, negative on
 *     failure) and never exits or aborts; sfd_last_error() gives the text.
 *   - Every run entry point checks its outputs (gathers, and the final
 *     pressure field of acoustic shots) for Inf / NaN and returns SFD_ENAN
 *     when it finds any (the outputs stay fetchable).
 *   - One context per host thread; contexts are independent.
 */
#ifndef SEISFD_H
#define SEISFD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SEISFD_ABI_VERSION 2

enum sfd_status {
    SFD_OK = 0,
    SFD_EINVAL = -1,   /* invalid argument */
    SFD_ENOMEM = -2,   /* device or host allocation failed */
    SFD_EDEVICE = -3,  /* HIP runtime error */
    SFD_ECFL = -4,     /* the time step violates the stencil's stability bound */
    SFD_ENAN = -5,     /* a non-finite value appeared in the wavefield */
    SFD_ENOTSUP = -6   /* combination of options not supported */
};

enum sfd_boundary_kind {
    /* The reference's quadratic sponge, Acoustic_1order_FD.cpp:164-194,
     * 1 - 0.38 d^2/L^2 on every field after each step (:329-340). */
    SFD_BOUNDARY_SPONGE = 0,
    /* Convolutional PML (no reference counterpart). */
    SFD_BOUNDARY_CPML = 1
};

enum sfd_arith {
    /* FP32 with fused multiply-adds and per-cell coefficients folded once. */
    SFD_ARITH_FAST = 0,
    /* The reference's own arithmetic, operation for operation, including its
     * double-promoted sub-expressions (:259,:276,:293-311): results are
     * bitwise equal to the reference program's.  Sponge boundary only. */
    SFD_ARITH_REFERENCE = 1
};

enum sfd_field {
    SFD_FIELD_P = 1, SFD_FIELD_VX = 2, SFD_FIELD_VZ = 4,
    /* stresses of the elastic propagator */
    SFD_FIELD_SXX = 8, SFD_FIELD_SZZ = 16, SFD_FIELD_SXZ = 32
};

typedef struct sfd_context sfd_context;
typedef struct sfd_acoustic_plan sfd_acoustic_plan;

/* One acoustic shot.  Mirrors the reference's constant block. */
typedef struct sfd_acoustic_params {
    int nx, nz;            /* full grid incl. absorbing layers (reference Nx, Nz) */
    int nt;                /* time steps (reference Nt) */
    float dx, dz, dt;      /* metres, seconds */
    int radius;            /* stencil half-width R: 4, 5 or 8 (reference: 8) */
    const float *coef;     /* C_1..C_R; NULL = the reference's table for R
                              (C1..C8 at :38-45; a80..a83 / a100..a104 of
                              yang_version_2016_2_2/Iso_multishot_fd_module_zy_m2.f90:17-26) */
    int boundary;          /* enum sfd_boundary_kind */
    int width;             /* absorbing layer width in cells (reference PML) */
    float pml_r;           /* C-PML: target reflection coefficient (e.g. 1e-3) */
    float pml_f0;          /* C-PML: frequency of the alpha profile (Hz) */
    int sx, sz;            /* source cell (reference nx_location, nz_location) */
    int nrec;
    const int *rx, *rz;    /* receiver cells; reference: rx = Nx/2, rz = PML..Nz-PML-1 */
    int nsnap;
    const int *snap_steps; /* ascending steps k after which fields are copied out */
    int snap_fields;       /* mask of enum sfd_field */
    int arith;             /* enum sfd_arith */
} sfd_acoustic_params;

/* Context: owns the HIP stream and the error text. device = HIP ordinal.
 * Destroy every plan and batch made in a context before the context. */
int sfd_context_create(int device, sfd_context **ctx);
int sfd_context_destroy(sfd_context *ctx);
const char *sfd_last_error(const sfd_context *ctx);
int sfd_abi_version(void);
/* Device memory of the context's GPU for inputs kept resident across calls,
 * such as the observed gathers of FWI iterations (no counterpart in the
 * reference): allocate, copy (kind 0 = host to device, 1 = device to host;
 * complete on return) and free. */
int sfd_device_alloc(sfd_context *ctx, size_t bytes, void **ptr);
int sfd_device_copy(sfd_context *ctx, void *dst, const void *src, size_t bytes, int kind);
int sfd_device_free(sfd_context *ctx, void *ptr);

/* One shot, end to end: the drop-in for Acoustic_1order_FD.cpp:237-393.
 *   vp, den  [nx*nz]   model (den NULL = 2400 kg/m^3 everywhere, :141)
 *   wavelet  [nt]      source samples (reference `signal`, :199-209); the
 *                      reference injects P[sx][sz] -= dt*den*vp*wavelet[k]
 *                      before the update of step k (:247-251)
 *   record   [nrec*nt] trace-major, like record.dat (:385-393); may be NULL
 *   snaps    [nsnap][popcount(snap_fields)][nx*nz], fields in P, Vx, Vz order
 *            (the reference writes Vx at k == 500 to Snap.dat, :343-355, and
 *            P every 10 steps to Pall.dat, :358-373); may be NULL */
int sfd_acoustic_forward(sfd_context *ctx, const sfd_acoustic_params *p,
                         const float *vp, const float *den, const float *wavelet,
                         float *record, float *snaps);

/* The same shot split in phases, so inputs stay resident in device memory:
 * create uploads the model and wavelet; run zeroes the wavefield and
 * propagates all nt steps on the device; fetch copies record and snapshots to
 * the host.  run may be repeated.  elapsed_ms (may be NULL) receives the
 * device time of the propagation measured with HIP events on the plan's
 * stream. */
int sfd_acoustic_plan_create(sfd_context *ctx, const sfd_acoustic_params *p,
                             const float *vp, const float *den, const float *wavelet,
                             sfd_acoustic_plan **plan);
int sfd_acoustic_plan_run(sfd_acoustic_plan *plan, float *elapsed_ms);
int sfd_acoustic_plan_fetch(sfd_acoustic_plan *plan, float *record, float *snaps);
int sfd_acoustic_plan_destroy(sfd_acoustic_plan *plan);

/* Step-level access (checkpoint / resume).  state is [3][nx*nz] (P, Vx, Vz);
 * pml is [4][nx*nz] (psi_px, psi_pz, psi_vx, psi_vz; C-PML only, NULL
 * otherwise).  step advances nsteps steps starting at step index k0 and
 * records into the plan's gather at columns k0..k0+nsteps-1. */
int sfd_acoustic_plan_set_state(sfd_acoustic_plan *plan, const float *state, const float *pml);
int sfd_acoustic_plan_get_state(sfd_acoustic_plan *plan, float *state, float *pml);
int sfd_acoustic_plan_step(sfd_acoustic_plan *plan, int k0, int nsteps);

/* Launch shape of the step kernel: waves per workgroup (1, 2 or 4) and rows
 * marched per workgroup.  0 = choose automatically. */
int sfd_acoustic_plan_set_tiling(sfd_acoustic_plan *plan, int waves, int chunk_rows);
/* Fast arithmetic (C-PML or sponge) runs by default as two tiled launches per
 * step (velocity, then pressure) instead of the fused march kernel: tile_rows
 * (8, 16 or 32) x 64 columns per workgroup of `waves` (4, 8 or 16; not every
 * pair is built); 0 = the default shape, tile_rows < 0 = the march kernel.
 * An explicit sfd_acoustic_plan_set_tiling also selects the march kernel.
 * Results are bitwise equal either way; the kernel may change between steps. */
int sfd_acoustic_plan_set_tiled(sfd_acoustic_plan *plan, int tile_rows, int waves);
/* Fast arithmetic as ONE launch per step (no counterpart in the reference),
 * on ping-pong buffers, so every field crosses memory about once per step
 * (28 B per cell with uniform density, against the two-launch step's 40 B).
 * Each workgroup of `waves` updates tile_rows x 64 cells:
 *   - tile_rows 16 or 32 (waves 4 or 8): the tiled step, which recomputes
 *     the new velocities its pressure stencils need around its tile;
 *   - tile_rows >= 64 (waves 4 or 8): the streaming step, which marches down
 *     its chunk of rows in blocks of 4 x waves rows, keeping the P and new Vx
 *     rows its stencils share in on-chip rings (4R - 2 extra P rows and 2R - 1
 *     extra new Vx rows per chunk).
 * 0 = 32 x 8.  Results are bitwise equal to the other kernels. */
int sfd_acoustic_plan_set_fused(sfd_acoustic_plan *plan, int tile_rows, int waves);

/* ---------------------------------------------------------------- batches
 * Many shots advanced together.  The reference runs its shots one after
 * another, one per MPI rank (forward/Q_forward/Q_multishot_fd_main_fsbc_zy_rotated.f90:124,
 * yang_version_2016_2_2/Iso_multishot_fd_main_fsbc_zy_m2.f90:71), each on its
 * own aperture-windowed grid (get_currshot_parameters, Q sub:288-428); a
 * batch holds such shots of different sizes and advances all of them with
 * one kernel launch per time step.  Every shot needs the same nt, radius,
 * boundary kind and arithmetic; batches record gathers (no snapshots). */
typedef struct sfd_acoustic_batch sfd_acoustic_batch;

enum sfd_batch_run {
    SFD_BATCH_GRAPH = 0,     /* one launch per step for all shots, the nt launches
                                captured once in a hipGraph and replayed */
    SFD_BATCH_LAUNCHES = 1,  /* one launch per step for all shots, no graph */
    SFD_BATCH_PER_SHOT = 2   /* the plain path: shot after shot, one launch per
                                shot and step, no graph, each shot at the
                                shape a plan of its own starts with (the
                                context's cached shape for its grid class, else
                                the library's default), not the batch's */
};

/* params[s], vp[s], den[s] (den may be NULL, or den[s] NULL), wavelet[s] as
 * in sfd_acoustic_forward, for shots s = 0 .. nshot-1. */
int sfd_acoustic_batch_create(sfd_context *ctx, int nshot, const sfd_acoustic_params *params,
                              const float *const *vp, const float *const *den,
                              const float *const *wavelet, sfd_acoustic_batch **batch);
/* Same tiling for every shot; invalidates a captured graph.  Explicit values
 * select the fused march kernel; (0, 0) = automatic: the shape the context's
 * autotuner cache holds for the batch's class, else the tiled step's default
 * (fast arithmetic) or the march kernel's. */
int sfd_acoustic_batch_set_tiling(sfd_acoustic_batch *batch, int waves, int chunk_rows);
/* Tiled step for every shot at tile_rows x 64 columns per workgroup of
 * `waves` (8x4, 16x4, 16x8, 32x4, 32x8 built); 0 = the default shape, tile_rows < 0
 * = the march kernel.  Invalidates a captured graph. */
int sfd_acoustic_batch_set_tiled(sfd_acoustic_batch *batch, int tile_rows, int waves);
/* One launch per step for every shot (as sfd_acoustic_plan_set_fused, the
 * same shapes).  Invalidates a captured graph. */
int sfd_acoustic_batch_set_fused(sfd_acoustic_batch *batch, int tile_rows, int waves);
/* Zeroes every wavefield and propagates all shots nt steps (how: enum
 * sfd_batch_run).  elapsed_ms: device time of the propagation. */
int sfd_acoustic_batch_run(sfd_acoustic_batch *batch, int how, float *elapsed_ms);
/* Gather of one shot, [nrec][nt] trace-major. */
int sfd_acoustic_batch_fetch(sfd_acoustic_batch *batch, int shot, float *record);
int sfd_acoustic_batch_destroy(sfd_acoustic_batch *batch);
/* Shot groups (no counterpart in the reference): a batch runs its shots in
 * groups of consecutive shots, each group for all nt steps before the next,
 * so that a group's fields stay in the GPU's 256 MB last-level cache (MALL)
 * across its steps.  bytes = the budget of one group's resident fields
 * (P, Vx, Vz, kf, a per-cell bf, C-PML memory fields); 0 = the library's
 * default (192 MiB), SIZE_MAX = every shot in one group.  Every group holds
 * at least one shot.  Results are bitwise the same for any grouping.
 * Invalidates a captured graph.  get_groups: the current number of groups. */
int sfd_acoustic_batch_set_group_bytes(sfd_acoustic_batch *batch, size_t bytes);
int sfd_acoustic_batch_get_groups(const sfd_acoustic_batch *batch, int *ngroups);

/* ---------------------------------------------------------------- autotuner
 * (no counterpart in the reference).  After a warm-up of up to 4 trials'
 * steps, times every built launch shape of the step (the two-launch and the
 * one-launch tiled steps' shapes in fast arithmetic, the fused march kernel
 * at 1, 2 and 4 waves) twice over
 * min(trial_steps, nt) steps from a zero state, keeps the fastest and records
 * it in the context's cache under the grid class:
 *   plan:  radius, boundary + arithmetic, nx and nz to the nearest power of 2
 *   batch: radius, boundary + arithmetic, shot count and total cells likewise.
 * Plans and batches created later in the context with the same class start
 * with the cached shape.  The state is reset (as by the run entry points);
 * results are bitwise the same for every shape.  best_us (may be NULL): device
 * time per step of the chosen shape. */
int sfd_acoustic_plan_autotune(sfd_acoustic_plan *plan, int trial_steps, float *best_us);
int sfd_acoustic_batch_autotune(sfd_acoustic_batch *batch, int trial_steps, float *best_us);
/* Current launch shape: kind 0 = tiled (a = tile rows, b = waves), kind 1 =
 * fused march kernel (a = waves, b = rows per chunk), kind 2 = one-launch
 * tiled step (a = tile rows, b = waves). */
int sfd_acoustic_plan_get_shape(const sfd_acoustic_plan *plan, int *kind, int *a, int *b);
int sfd_acoustic_batch_get_shape(const sfd_acoustic_batch *batch, int *kind, int *a, int *b);
/* The context's autotuner cache as a JSON file
 *   {"seisfd_tuning": 1, "entries": [{"arch", "class", "kind", "a", "b", "us_per_step"}, ...]}.
 * Saving merges into an existing file (entries of other GPUs are kept);
 * loading adds the entries measured on this GPU's architecture and rejects a
 * malformed file as a whole (SFD_EINVAL, nothing loaded). */
int sfd_context_save_tuning(sfd_context *ctx, const char *path);
int sfd_context_load_tuning(sfd_context *ctx, const char *path);
/* Number of grid classes in the context's cache. */
int sfd_context_tuning_size(const sfd_context *ctx);

/* ---------------------------------------------------------------- gradient
 * Velocity gradient by the discrete adjoint of the fast-arithmetic sponge or
 * C-PML step (no reference counterpart: the reference models data only).
 * The wavefield (with C-PML, its memory fields too) is checkpointed every
 * ckpt_every steps and recomputed one segment at a time during the reverse
 * sweep, so device memory grows with nt / ckpt_every rather than nt.  With
 * ckpt_every >= nt the forward keeps every step's divergence field itself
 * (nt fields per shot; 16 shots of 881 x 881 over 2500 steps: 124 GB) and
 * the reverse sweep recomputes nothing: one forward pass less, the same
 * gradient bit for bit.  Shots of a misfit gradient are processed in
 * groups that fit 80 % of the device memory free at the call.  The
 * C-PML profiles depend on max(vp); the gradient holds them fixed (the
 * derivative through the profiles' reference velocity is not included).
 * Every shot of one call needs the same boundary kind.
 *
 * sfd_acoustic_gradient: J = sum_{r,k} seed[r][k] record[r][k] (seed is
 *   [nrec][nt], the vector of a vector-Jacobian product); grad_vp [nx*nz]
 *   receives dJ/dvp, record (may be NULL) the forward gather.
 * sfd_acoustic_misfit_gradient: J = sum_s 1/2 |record_s - d_obs[s]|^2 over
 *   nshot shots on one model (vp, den of params[0]'s grid; wavelet[s] and
 *   the geometry of params[s] per shot); grad_vp receives dJ/dvp summed over
 *   shots, misfit (may be NULL) the value of J. */
int sfd_acoustic_gradient(sfd_context *ctx, const sfd_acoustic_params *p, const float *vp,
                          const float *den, const float *wavelet, const float *seed,
                          int ckpt_every, float *grad_vp, float *record);
int sfd_acoustic_misfit_gradient(sfd_context *ctx, int nshot, const sfd_acoustic_params *params,
                                 const float *vp, const float *den, const float *const *wavelet,
                                 const float *const *d_obs, int ckpt_every, float *grad_vp,
                                 double *misfit);

/* How the reverse sweep of sfd_acoustic_misfit_gradient_ex obtains the
 * forward wavefield. */
enum sfd_gradient_method {
    /* recompute each checkpoint segment forward, keeping its ckpt_every
       divergence fields (what sfd_acoustic_misfit_gradient does) */
    SFD_GRADIENT_RECOMPUTE = 0,
    /* boundary saving: the forward keeps the R-deep band just outside the box
       [max(w, R), n - max(w, R)) of both axes (w = sponge width; the C-PML
       width + 1 with C-PML) after every step; the reverse sweep runs the step
       backwards inside the box from the band and the checkpoints (reset at
       every segment end), with one divergence field per shot.  The absorbing
       layers cannot be run backwards, so grad_vp is zero outside the box
       (the source cell's term aside). */
    SFD_GRADIENT_BOUNDARY_SAVING = 1,
};
int sfd_acoustic_misfit_gradient_ex(sfd_context *ctx, int nshot, const sfd_acoustic_params *params,
                                    const float *vp, const float *den, const float *const *wavelet,
                                    const float *const *d_obs, int ckpt_every, int method,
                                    float *grad_vp, double *misfit);
/* As sfd_acoustic_misfit_gradient_ex (no counterpart in the reference), with
 * the observed gathers already in device memory: d_obs_dev[s] is a device
 * pointer to shot s's [nrec][nt] float gather, written before the call on
 * any stream and left unchanged during it.  FWI iterations keep their data
 * resident, so no host transfer of it enters the call.  (With host data,
 * sfd_acoustic_misfit_gradient_ex copies it on a second stream while the
 * forward runs.) */
int sfd_acoustic_misfit_gradient_dev(sfd_context *ctx, int nshot, const sfd_acoustic_params *params,
                                     const float *vp, const float *den, const float *const *wavelet,
                                     const float *const *d_obs_dev, int ckpt_every, int method,
                                     float *grad_vp, double *misfit);

/* Aperture window of one shot, the drop-in for get_currshot_parameters
 * (forward/Q_forward/Q_multishot_fd_subroutine_fsbc_zy_rotated.f90:315-373;
 * the same routine in yang_version_2016_2_2/Iso_multishot_fd_subroutine_fsbc_zy_m2_pub.f90:216-266).
 * model is [mnx][mnz] (z fastest).  Window column w (0 <= w < nxe) takes
 * model column clamp(x0 + w, 0, mnx-1); the window is then padded by edge
 * replication with pad_xlo / pad_xhi columns and pad_zlo / pad_zhi rows.
 * out is [nxe+pad_xlo+pad_xhi][mnz+pad_zlo+pad_zhi].  Host only. */
int sfd_window_model(const float *model, int mnx, int mnz, int x0, int nxe, int pad_xlo,
                     int pad_xhi, int pad_zlo, int pad_zhi, float *out);

/* Stability bound of the staggered leapfrog: the largest
 * vmax * dt * sqrt(1/dx^2 + 1/dz^2) the stencil allows, 1 / sum |C_m|.
 * Plans whose model exceeds it are rejected with SFD_ECFL. */
double sfd_acoustic_cfl_limit(int radius, const float *coef);


/* ---------------------------------------------------------------- visco-elastic
 * Rotated-staggered-grid visco-elastic propagator: the drop-in for the time
 * loop of forward/Q_forward, single_shot_modelling
 * (Q_multishot_fd_main_fsbc_zy_rotated.f90:169-484, step
 * extrapolation_one_step_10nd_absbc, Q_multishot_fd_subroutine_fsbc_zy_rotated.f90:695-839),
 * with L standard linear solids (the reference: L = 1) fitted by the
 * tau-method of pcalpas (Q sub:473-551).  vs = 0 everywhere is the
 * visco-acoustic limit (strxx = strzz = sigma, strxz = 0), run by one fused
 * kernel per step; any vs > 0 runs the visco-elastic kernels (two launches per
 * step).  Grid = the reference's padded shot grid nxea x nzea
 * (get_currshot_parameters, or sfd_window_model), [nx][nz] z fastest, 0-based. */
typedef struct sfd_visco_plan sfd_visco_plan;

typedef struct sfd_visco_params {
    int nx, nz;            /* padded shot grid (reference nxea, nzea) */
    int nt;                /* recorded samples (reference nt) */
    int delay;             /* steps run before the first recorded one (reference
                              decay, main:279-280); < 0 = nint(1.2 / (f0 dt)) */
    float dx, dz, dt;
    float f0;              /* Hz: the built-in wavelet and the tau-method band */
    int radius;            /* 5 (a100..a104, the reference) or 4 (a80..a83) */
    const float *coef;     /* a_0..a_{R-1}; NULL = the reference's table */
    int nmech;             /* L = 1 (reference), 2 or 3 */
    float q_scale;         /* Q multiplier before the fit; 0 = the reference's 0.9 (Q sub:525-526) */
    float f_low, f_high;   /* tau-method band in Hz; <= 0 = the reference's 2 Hz and 2.5 f0 */
    int sponge_xlo, sponge_xhi, sponge_zlo, sponge_zhi;  /* reference npxn, npxp, npzn, npzp:
                              strips of n+1 cells damped by cos(acos(R) d / n) (csponge, Q sub:555-605) */
    float sponge_r;        /* reference R (module constant 1e-3); 0 = 1e-3 */
    int sx, sz;            /* source cell (reference isxe - 1, isze - 1) */
    float src_sigma;       /* Gaussian source width in cells; 0 = the reference's 1.5 (addsrc) */
    int nrec;
    const int *rx, *rz;    /* receiver cells (reference: row itopo + nrz0 of every window column) */
    int nsnap;
    const int *snap_steps; /* ascending step indices in [0, delay + nt) after which fields are copied */
    int snap_fields;       /* mask of enum sfd_field: P = strzz (= sigma when vs = 0), VX, VZ, and
                              with vs > 0 also SXX, SZZ, SXZ */
    const float *qs;       /* [nx*nz] S-wave quality factor (reference qfs); NULL = qp */
} sfd_visco_params;

/* vp, qp [nx*nz] required; rho NULL = 2400 kg/m^3 everywhere (averaged over 4
 * cells as substagger does); vs NULL or all zero = visco-acoustic, otherwise
 * visco-elastic with mu = rho_avg vs^2; wavelet [delay + nt] or NULL = addsrc's Ricker
 * exp(-(pi f0 (it dt - 1.2/f0))^2)(1 - 2 (pi f0 (it dt - 1.2/f0))^2), it = 1..
 *   record [nrec][nt]: strzz at the receivers after steps delay .. delay+nt-1
 *          (reference record1, main:450-459; = record2 = strxx when vs = 0)
 *   snaps  [nsnap][popcount(snap_fields)][nx*nz], fields in P, vx, vz, sxx, szz, sxz order */
int sfd_visco_forward(sfd_context *ctx, const sfd_visco_params *p, const float *vp, const float *vs,
                      const float *rho, const float *qp, const float *wavelet, float *record, float *snaps);
int sfd_visco_plan_create(sfd_context *ctx, const sfd_visco_params *p, const float *vp, const float *vs,
                          const float *rho, const float *qp, const float *wavelet, sfd_visco_plan **plan);
/* Zeroes the fields and runs all delay + nt steps; elapsed_ms (may be NULL)
 * receives the device time measured with HIP events on the plan's stream. */
int sfd_visco_plan_run(sfd_visco_plan *plan, float *elapsed_ms);
int sfd_visco_plan_fetch(sfd_visco_plan *plan, float *record, float *snaps);
/* strxx at the receivers [nrec][nt] (reference record2). */
int sfd_visco_plan_fetch_xx(sfd_visco_plan *plan, float *record_xx);
/* vs = 0: explicit waves per workgroup (2 or 4) and rows marched per
 * workgroup select the fused march kernel; (0, 0) = the tiled step.  The
 * visco-elastic step has the tiled launches only (no effect). */
int sfd_visco_plan_set_tiling(sfd_visco_plan *plan, int waves, int chunk_rows);
/* The tiled step (two launches per step) at tile_rows x 64 columns per
 * workgroup of `waves` waves: 8 x 8 (the default; (0, 0)) or 16 x 8.
 * Results are bitwise the same at either shape. */
int sfd_visco_plan_set_tiled(sfd_visco_plan *plan, int tile_rows, int waves);
/* Autotuner (no counterpart in the reference), as sfd_acoustic_plan_autotune:
 * the tiled step at 8 x 8 and 16 x 8 against (vs = 0) the fused march kernel
 * at 2 and 4 waves, cached for the grid class (radius, L, vs = 0 or not, nx
 * and nz to the nearest power of 2); resets the state.  get_shape: kind 0 =
 * tiled (a = rows, b = waves), 1 = march (a = waves, b = rows per chunk). */
int sfd_visco_plan_autotune(sfd_visco_plan *plan, int trial_steps, float *best_us);
int sfd_visco_plan_get_shape(const sfd_visco_plan *plan, int *kind, int *a, int *b);
int sfd_visco_plan_destroy(sfd_visco_plan *plan);

/* Visco batches: nshot shots advanced together, one launch of each half-step
 * of the tiled step over the tiles of all of them per time step (the
 * reference runs its shots one after another, one per MPI rank, Q main:124).
 * params[s], vp[s], vs[s], rho[s], qp[s], wavelet[s] as in
 * sfd_visco_plan_create (vs, rho, wavelet may be NULL, or their entries);
 * each shot keeps its own grid (windows of any width).  Every shot needs the
 * same nt, delay, radius, L and the same kind (vs = 0 everywhere, or not).
 * Results are bitwise those of the shots' own plans. */
typedef struct sfd_visco_batch sfd_visco_batch;
int sfd_visco_batch_create(sfd_context *ctx, int nshot, const sfd_visco_params *params,
                           const float *const *vp, const float *const *vs, const float *const *rho,
                           const float *const *qp, const float *const *wavelet, sfd_visco_batch **batch);
/* Zeroes every wavefield and propagates all shots; elapsed_ms: device time. */
int sfd_visco_batch_run(sfd_visco_batch *batch, float *elapsed_ms);
/* Shot `shot`'s strzz gather [nrec][nt], its strxx gather (may be NULL) and
 * snapshots (may be NULL), as sfd_visco_plan_fetch / _fetch_xx. */
int sfd_visco_batch_fetch(sfd_visco_batch *batch, int shot, float *record, float *record_xx, float *snaps);
int sfd_visco_batch_destroy(sfd_visco_batch *batch);
/* Shot groups, as for acoustic batches (sfd_acoustic_batch_set_group_bytes):
 * the budget of one group's fields (state, model, strip parts); 0 = the
 * default (192 MiB), SIZE_MAX = one group; bitwise the same results. */
int sfd_visco_batch_set_group_bytes(sfd_visco_batch *batch, size_t bytes);
int sfd_visco_batch_get_groups(const sfd_visco_batch *batch, int *ngroups);

/* ---------------------------------------------------------------- elastic
 * Isotropic elastic standard-staggered-grid propagator with a free surface
 * over topography: the drop-in for the time loop of
 * forward/yang_version_2016_2_2, isoewm2dfsbc
 * (Iso_multishot_fd_subroutine_fsbc_zy_m2_pri.f90:57-230, called from
 * Iso_multishot_fd_main_fsbc_zy_m2.f90:193-201 in 'fsbc' mode), with its
 * set-up substagger (pub:323-547), csponge (pub:796-847) and the Gaussian
 * source window (pub:1470-1501).  10th-order x derivatives; z derivatives of
 * order 10 reduced to 8, 6, 4, 2 on the rows just below the surface; cosine
 * sponges on the left, right and bottom strips; tau_zz = 0 on and above the
 * surface, the tau_xz image and vz above it from tau_zz = 0.  Grid = the
 * reference's padded shot grid nxea x nzea (get_currshot_parameters, pub:188-313,
 * or sfd_window_model), [nx][nz] z fastest, 0-based. */
typedef struct sfd_elastic_plan sfd_elastic_plan;

enum sfd_elastic_mode {
    /* 'fsbc': free surface on row itopo of every column (isoewm2dfsbc) */
    SFD_ELASTIC_FREE_SURFACE = 0,
    /* 'abc' (isoewm2dabc, pri:232-386): no free surface; 10th-order stencils
     * from npzn rows above the surface, the z parts of those npzn + 1 rows
     * damped by the top cosine sponge (calpmlbclrtb, pub:851-892), the model
     * on and above the surface smoothed by the source's Gaussian window
     * (modipapml, pub:549-634), the source always into tau_xx */
    SFD_ELASTIC_ABSORBING_TOP = 1
};

typedef struct sfd_elastic_params {
    int nx, nz;            /* padded shot grid (reference nxea, nzea) */
    int nt;                /* recorded samples (reference nt) */
    int delay;             /* steps run before the first recorded one (reference
                              idecay = nint(decay/dt), pri:28-29); < 0 = nint(1.2 / (f0 dt)) */
    float dx, dz, dt;
    float f0;              /* Hz: the built-in wavelet */
    const int *itopo;      /* [nx] free-surface row of every column, 0-based (reference itopoe - 1,
                              itopoe = nint(topo/dz) + npzn, pub:270-272) */
    int sponge_xlo, sponge_xhi, sponge_zhi;  /* reference npxn, npxp, npzp: strips of n+1 cells
                              damped by cos(acos(R) d / n) (csponge); no top strip */
    float sponge_r;        /* reference R (module constant 1e-3); 0 = 1e-3 */
    int sx, sz;            /* source cell (reference isxe - 1, isze - 1) */
    int src_radius;        /* reference rgau: Gaussian window of 2 rgau - 1 cells, sigma 2; 0 = 50 */
    int nrec;
    const int *rx, *rz;    /* receiver cells (reference: row itopo + nrz0 of every window column) */
    int nsnap;
    const int *snap_steps; /* ascending step indices in [0, delay + nt) after which fields are copied */
    int snap_fields;       /* mask of enum sfd_field: VX, VZ, SXX, SZZ, SXZ */
    int mode;              /* enum sfd_elastic_mode */
    int sponge_zlo;        /* SFD_ELASTIC_ABSORBING_TOP: reference npzn (>= 1), the top sponge
                              of npzn + 1 rows ending on the surface row */
} sfd_elastic_params;

/* vp, vs, rho [nx*nz] required (vs = 0 cells are fluid); wavelet [delay + nt]
 * or NULL = addsrc's Ricker exp(-(pi f0 (it dt - delay dt))^2)(1 - 2 (...)^2),
 * it = 1.., injected as wavelet g dt / dz^2 into tau_zz and, when the source
 * lies below the surface, tau_xx.
 *   record_vx, record_vz [nrec][nt]: vx, vz at the receivers after steps
 *          delay .. delay+nt-1 (reference record1, record2, pri:199-210)
 *   snaps  [nsnap][popcount(snap_fields)][nx*nz], fields in vx, vz, sxx, szz, sxz order */
int sfd_elastic_forward(sfd_context *ctx, const sfd_elastic_params *p, const float *vp, const float *vs,
                        const float *rho, const float *wavelet, float *record_vx, float *record_vz, float *snaps);
int sfd_elastic_plan_create(sfd_context *ctx, const sfd_elastic_params *p, const float *vp, const float *vs,
                            const float *rho, const float *wavelet, sfd_elastic_plan **plan);
/* Zeroes the fields and runs all delay + nt steps; elapsed_ms (may be NULL)
 * receives the device time measured with HIP events on the plan's stream. */
int sfd_elastic_plan_run(sfd_elastic_plan *plan, float *elapsed_ms);
int sfd_elastic_plan_fetch(sfd_elastic_plan *plan, float *record_vx, float *record_vz, float *snaps);
/* Tile width in columns (8, 16 or 32) and waves per workgroup (4, 8 or 16;
 * not every pair is built); 0 = the default. */
int sfd_elastic_plan_set_tiling(sfd_elastic_plan *plan, int tile_cols, int waves);
/* Autotuner (no counterpart in the reference), as sfd_acoustic_plan_autotune:
 * times the eight built tilings over min(trial_steps, delay + nt) steps,
 * keeps the fastest and caches it for the grid class (mode, nx and nz to the
 * nearest power of 2); resets the state. */
int sfd_elastic_plan_autotune(sfd_elastic_plan *plan, int trial_steps, float *best_us);
int sfd_elastic_plan_get_shape(const sfd_elastic_plan *plan, int *tile_cols, int *waves);
int sfd_elastic_plan_destroy(sfd_elastic_plan *plan);

/* Several elastic shots stepped together (no counterpart in the reference:
 * its driver runs the shots one after another, distributing them over MPI
 * ranks, Iso_multishot_fd_main_fsbc_zy_m2.f90:71-81).  Shot i is what
 * sfd_elastic_plan_create(ctx, &params[i], vp[i], vs[i], rho[i], wavelet ? wavelet[i] : NULL)
 * makes; one launch per half-step covers the tiles of every shot, so a
 * survey of small shot grids fills the GPU.  Shots may differ in grid and
 * geometry but share the mode, nt and the delay.  Each shot's gathers and
 * snapshots are bitwise those of its own plan. */
typedef struct sfd_elastic_batch sfd_elastic_batch;
int sfd_elastic_batch_create(sfd_context *ctx, int nshot, const sfd_elastic_params *params,
                             const float *const *vp, const float *const *vs, const float *const *rho,
                             const float *const *wavelet, sfd_elastic_batch **batch);
/* Same tiling for every shot (as sfd_elastic_plan_set_tiling); 0 = the default. */
int sfd_elastic_batch_set_tiling(sfd_elastic_batch *batch, int tile_cols, int waves);
/* Autotuner, as sfd_elastic_plan_autotune, cached for the batch's class (mode,
 * shot count and cells per step to the nearest power of 2); resets the state. */
int sfd_elastic_batch_autotune(sfd_elastic_batch *batch, int trial_steps, float *best_us);
int sfd_elastic_batch_get_shape(const sfd_elastic_batch *batch, int *tile_cols, int *waves);
/* Shot groups, as for acoustic batches: the budget of one group's fields
 * (state, staggered model, strip parts); 0 = the default (256 MiB),
 * SIZE_MAX = one group; bitwise the same results. */
int sfd_elastic_batch_set_group_bytes(sfd_elastic_batch *batch, size_t bytes);
int sfd_elastic_batch_get_groups(const sfd_elastic_batch *batch, int *ngroups);
/* Zeroes every shot's fields and runs all delay + nt steps; elapsed_ms as
 * sfd_elastic_plan_run.  SFD_ENAN when a gather is not finite. */
int sfd_elastic_batch_run(sfd_elastic_batch *batch, float *elapsed_ms);
/* Shot `shot`'s outputs, as sfd_elastic_plan_fetch. */
int sfd_elastic_batch_fetch(sfd_elastic_batch *batch, int shot, float *record_vx, float *record_vz, float *snaps);
int sfd_elastic_batch_destroy(sfd_elastic_batch *batch);

/* ---------------------------------------------------------------- VTI
 * VTI qP-acoustic propagator (p and q fields, 2nd order in time, 10th-order
 * second derivatives, split PML on the four sides): the drop-in for the time
 * loop of de-multiple/convolution_srm/vti_qa_2d_c3_demo1, single_shot_modelling
 * (VTI_multishot_qacoustic_fd_main_c3.f90:429-520) with its per-step routines
 * extrapolation_one_step_1 / _2 (VTI_multishot_qacoustic_fd_subroutine_c3.f90:432-1258).
 * Grid = the reference's padded shot grid currshot_range_all x nz
 * (get_currshot_parameters, or sfd_vti_window), [nx][nz] z fastest, 0-based:
 * a Fortran caller passes its (nx, nz) arrays transposed.  A plan holds
 * nshot shots on grids of this one shape (the reference's shots all have
 * it), each with its own model, advanced together. */
typedef struct sfd_vti_plan sfd_vti_plan;

typedef struct sfd_vti_params {
    int nx, nz;            /* padded shot grid (reference currshot_range_all, nz) */
    int nt;                /* recorded samples (reference nt) */
    int delay;             /* steps run before the first recorded one; < 0 = the reference's
                              nint((nwt - 1) / 2), nwt = nint(2 * 0.07 / dt) + 1 (main:277, 469) */
    float dx, dz, dt;
    float f0;              /* Hz: the built-in wavelet */
    int pml_xlo, pml_xhi, pml_zlo, pml_zhi;  /* reference nx_bound_l, nx_bound_r, nz_bound_u, nz_bound_d */
    float pml_r;           /* reference R (module constant 1e-3); 0 = 1e-3 */
    const float *c2;       /* [5] second-derivative weights; NULL = coefficient_2nd(10) as the
                              reference solves it (fd_coe = fd_coe_pml_2nd) */
    const float *c1;       /* [5] first-derivative weights; NULL = coefficient_1st(9) (fd_coe_pml_1st) */
    int sx, sz;            /* source cell (reference currshot_nsx - 1, currshot_nsz - 1) */
    int nrec;
    const int *rx, *rz;    /* receiver cells, the same in every shot (reference: row
                              nrz0 + nz_bound_u, or row 6 when that is above it, of every window column) */
    int allow_eps_lt_delta;/* 0: cells with eps < delta are rejected (SFD_EINVAL): the
                              pseudo-acoustic system grows without bound there; nonzero: run them
                              anyway, as the reference does (a divergence then returns SFD_ENAN) */
} sfd_vti_params;

/* vz, eps, delta: [nshot][nx*nz].  wavelet [delay + nt] or NULL = the
 * reference's exp(-x^2)(1 - 2 x^2), x = pi f0 (it - 0.07/dt) dt for it = 1..nwt
 * (main:433-437), added to p at the source cell before each step.
 *   record [nshot][nrec][nt]: p at the receivers after steps delay .. delay+nt-1
 *   (reference record(iit, :), main:469-476). */
int sfd_vti_forward(sfd_context *ctx, const sfd_vti_params *p, int nshot, const float *vz, const float *eps,
                    const float *delta, const float *wavelet, float *record);
int sfd_vti_plan_create(sfd_context *ctx, const sfd_vti_params *p, int nshot, const float *vz, const float *eps,
                        const float *delta, const float *wavelet, sfd_vti_plan **plan);
/* Zeroes the fields and runs all delay + nt steps; elapsed_ms (may be NULL)
 * receives the device time measured with HIP events on the plan's stream. */
int sfd_vti_plan_run(sfd_vti_plan *plan, float *elapsed_ms);
/* record as above; state (may be NULL) [nshot][2][nx*nz]: p and q after the run. */
int sfd_vti_plan_fetch(sfd_vti_plan *plan, float *record, float *state);
/* Tile width in columns (8, 16, 32 or 64) and waves per workgroup (4, 8 or
 * 16; not every pair is built); 0 = the default. */
int sfd_vti_plan_set_tiling(sfd_vti_plan *plan, int tile_cols, int waves);
/* Autotuner (no counterpart in the reference), as sfd_elastic_plan_autotune:
 * cached for the grid class (nx, nz and nshot to the nearest power of 2);
 * resets the state. */
int sfd_vti_plan_autotune(sfd_vti_plan *plan, int trial_steps, float *best_us);
int sfd_vti_plan_get_shape(const sfd_vti_plan *plan, int *tile_cols, int *waves);
/* Shot groups (no counterpart in the reference), as for acoustic batches:
 * the plan's shots run in groups of consecutive shots, each group for all
 * its steps before the next, so that a group's fields (p, q at two levels,
 * vz, eps, delta, PML parts) stay in the 256 MB MALL.  bytes = the budget of
 * one group (0 = the default, 192 MiB; SIZE_MAX = every shot in one group);
 * results are bitwise the same for any grouping. */
int sfd_vti_plan_set_group_bytes(sfd_vti_plan *plan, size_t bytes);
int sfd_vti_plan_get_groups(const sfd_vti_plan *plan, int *ngroups);
int sfd_vti_plan_destroy(sfd_vti_plan *plan);
/* Host helpers (no device or context needed):
 * coefficient_2nd (kind 2) / coefficient_1st (kind 1) with k <= 8 weights, solved in
 * single precision as sub:1264-1472 does; the driver's wavelet (returns nwt, writes
 * min(nwt, nmax) samples); get_currshot_parameters (sub:268-338) plus the
 * isotropic rows 1..nsz0+nz_bound_u+6 (main:422-425): window column i (1-based)
 * takes model column clamp(i + xmin, 1, nxv), outputs
 * [range + nl + nr - 2][nzv + nu + nd - 2] (nr, nd >= 2). */
int sfd_vti_coefficients(int kind, int k, float *out);
int sfd_vti_wavelet(float dt, float f0, int nmax, float *w);
int sfd_vti_window(const float *vz, const float *eps, const float *delta, int nxv, int nzv, int xmin, int range,
                   int nl, int nr, int nu, int nd, int nsz0, float *ovz, float *oeps, float *odelta);

/* ---------------------------------------------------------------- SRM
 * Convolution surface-related multiple prediction, the consumer of the
 * modelled gathers: the drop-in for de-multiple/convolution_srm/convolution_srm.cpp:7-276
 * (its files become arrays).  Output trace (is, it) = the real part / ltt of
 * the inverse FFT of sum_k FFT(csg[is][k]) FFT(crg[is*ntrace + it*nshot + is + k])
 * / (ntrace + nshot) over k < ntrace with is + k < nshot, all traces
 * zero-padded to ltt = 2 lt, first lt samples kept (:107-108, 141-176, 201-252). */
typedef struct sfd_srm_plan sfd_srm_plan;

typedef struct sfd_srm_params {
    int nshot;             /* reference nshot */
    int ntrace;            /* traces per shot (reference nt) */
    int lt;                /* samples per trace */
    int ncrg;              /* traces in crg: at least ntrace (2 nshot - 1), what the reference reads */
} sfd_srm_params;

/* csg [nshot][ntrace][lt] (the reference's input file), crg [ncrg][lt]
 * (its inputcp file); out [nshot][ntrace][lt] (its output file). */
int sfd_srm_predict(sfd_context *ctx, const sfd_srm_params *p, const float *csg, const float *crg, float *out);
/* Traces the common-receiver array needs: ntrace (2 nshot - 1). */
int sfd_srm_crg_traces(int nshot, int ntrace);
/* Plan: uploads csg and crg once; run computes (device time in elapsed_ms,
 * may be NULL); fetch copies out [nshot][ntrace][lt]. */
int sfd_srm_plan_create(sfd_context *ctx, const sfd_srm_params *p, const float *csg, const float *crg,
                        sfd_srm_plan **plan);
int sfd_srm_plan_run(sfd_srm_plan *plan, float *elapsed_ms);
/* How the spectral sums are formed (no counterpart in the reference, which
 * loops): DIRECT = one workgroup per (frequency, shot) summing the trace
 * pairs; GEMM (ntrace = nshot only) = one batched complex matrix product per
 * frequency over the banded shot spectra (rocBLAS); AUTO = GEMM when
 * ntrace = nshot.  get_method returns the one a run uses. */
enum sfd_srm_method { SFD_SRM_AUTO = 0, SFD_SRM_DIRECT = 1, SFD_SRM_GEMM = 2 };
int sfd_srm_plan_set_method(sfd_srm_plan *plan, int method);
int sfd_srm_plan_get_method(const sfd_srm_plan *plan);
int sfd_srm_plan_fetch(sfd_srm_plan *plan, float *out);
int sfd_srm_plan_destroy(sfd_srm_plan *plan);

/* ---------------------------------------------------------------- files
 * The Fortran drivers' files on either side of the propagators (host code;
 * no device or context needed).  Parameter cards: a comment line before
 * every value line, strings as whole lines (trimmed), numbers read
 * list-directed or, where the reference writes '(F)' / '(I)', as Intel
 * Fortran's F15.7 / I12 fields (first 15 / 12 characters; an F field
 * without a decimal point is scaled by 1e-7).  err (may be NULL, 256 bytes)
 * receives the reason on SFD_EINVAL. */
typedef struct sfd_q_par {  /* Q_forward read_par, Q sub:121-234 */
    char vp_fn[256], vs_fn[256], rho_fn[256], qfp_fn[256], qfs_fn[256];
    char shot_fn1[256], shot_fn2[256], shot_fn3[256], shot_fn4[256];
    int nshot;
    float offset_min, offset_max, sx0, sz0, rz0, dsx;
    int nx, nz;
    float dx, dz;
    int nt;
    float dt, f0, ppxn, ppxp, ppzn, ppzp;
    int fdo2, fdo1;
} sfd_q_par;
int sfd_read_q_par(const char *path, sfd_q_par *par, char *err);

typedef struct sfd_iso_par {  /* yang_version read_par, pub:48-145 */
    char mode[256];           /* "fsbc" or "abc" */
    char vp_fn[256], vs_fn[256], rho_fn[256], topo_fn[256], snapx_fn[256], snapz_fn[256];
    char shot_fn1[256], shot_fn2[256], shot_fn3[256], shot_fn4[256];
    int nshot;
    float ofsmin, ofsmax, sx0, sz0, rz0, dsx;
    int nx, nz;
    float dx, dz;
    int nt;
    float dt, dtsnap, f0, decay, ppxn, ppxp, ppzn, ppzp;
    int rgau;
} sfd_iso_par;
int sfd_read_iso_par(const char *path, sfd_iso_par *par, char *err);

/* One shot's gather as an SU file (write_currshot_disk, Q sub:1476-1546 and
 * pub:1533-1590): per trace a 240-byte SU header (native byte order) then nt
 * floats; record [ntrace][nt].  Header: fldr = shot_no, trid = 1, sx =
 * int((src_col - 1) dx), ns = nt, dt = int(dt 1e6) us, per trace gx =
 * int((first_col - 1 + i - 1) dx) (i = 1..ntrace), offset = gx - sx and,
 * with with_cdp (Q_forward), cdp = nint((gx + sx) / dx) + 1; columns are
 * 1-based as in the Fortran. */
int sfd_write_su_shot(const char *path, const float *record, int ntrace, int nt, int shot_no, int src_col,
                      int first_col, float dx, float dt, int with_cdp);
/* merge_shot_files (Q sub:1550-1624, pub:1594-1684): the shot files, each
 * exactly ntrace traces of nt samples, concatenated in order into out_path;
 * remove_inputs deletes them afterwards (as the reference does). */
int sfd_merge_su(const char *const *shot_files, int nshot, int ntrace, int nt, const char *out_path,
                 int remove_inputs);

/* ---------------------------------------------------------------- allocation-free shot
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc): sfd_acoustic_forward in fast arithmetic, allocating
 * nothing.  Inputs and outputs are DEVICE arrays: d_vp [nx*nz], d_den
 * [nx*nz] or NULL (2400 kg/m^3), d_wavelet [nt], d_rx / d_rz [nrec]
 * receiver cells, d_record [nrec][nt]; p->rx / p->rz are not read and
 * p->nsnap must be 0.  The wavefield, folded model and tables live in one
 * static device array per GPU of sfd_noalloc_scratch_bytes() bytes (32 MiB),
 * held until the process ends: grids that need more are rejected with
 * SFD_EINVAL (about 1.7 M cells at radius 8).  Every call on a GPU shares
 * that array, so calls that overlap on one GPU race.  The gather is bitwise
 * the one sfd_acoustic_forward records (same step kernels).  Synchronous. */
int sfd_acoustic_forward_noalloc(sfd_context *ctx, const sfd_acoustic_params *p, const float *d_vp,
                                 const float *d_den, const float *d_wavelet, const int *d_rx, const int *d_rz,
                                 float *d_record);
size_t sfd_noalloc_scratch_bytes(void);

/* ---------------------------------------------------------------- measurement
 * Streaming-bandwidth probe (no counterpart in the reference): the rate a
 * hand-written kernel streams a device buffer of `bytes` bytes at (read: sum
 * every 16-byte vector; write: store a constant; copy: a second buffer of the
 * same size, bytes counted twice), the best of `reps` timed launches after one
 * warm-up over plain and nontemporal loads / stores, in GB/s (1e9 B/s).
 * Allocates and frees its buffers. */
enum sfd_bandwidth_kind { SFD_BW_READ = 0, SFD_BW_WRITE = 1, SFD_BW_COPY = 2 };
int sfd_bandwidth_probe(sfd_context *ctx, size_t bytes, int kind, int reps, double *gbps);

#ifdef __cplusplus
}
#endif
#endif
