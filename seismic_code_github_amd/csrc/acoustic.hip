// Acoustic shot plan: device buffers, launch dispatch and the C ABI entry
// points declared in include/seisfd.h (drop-in for the time loop of
// forward/Acoustic_FD_by_ss/src/Acoustic_1order_FD.cpp:237-393).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <tuple>

#include "acoustic_adjoint_cpml.h"
#include "acoustic_plan.h"
#include "acoustic_tiled.h"

namespace sfd {

// ---------------------------------------------------------------- helper kernels

// record[r][k] = rec_buf[k][slot[r]] : time-major gather buffer to the
// reference's trace-major record layout (:385-393).
__global__ void gather_transpose_kernel(const float* __restrict__ rec_buf, const int* __restrict__ slot,
                                        float* __restrict__ record, int nrec, int nt, int nslot) {
    __shared__ float tile[32][33];
    const int k0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
    for (int q = ty; q < 32; q += 8) {
        const int k = k0 + q, r = r0 + tx;
        tile[q][tx] = (k < nt && r < nrec) ? rec_buf[(size_t)k * nslot + slot[r]] : 0.f;
    }
    __syncthreads();
    for (int q = ty; q < 32; q += 8) {
        const int r = r0 + q, k = k0 + tx;
        if (r < nrec && k < nt) record[(size_t)r * nt + k] = tile[tx][q];
    }
}

// ---------------------------------------------------------------- launch table


template <int R, int NW>
StepKernel pick_mode(int mode) {
    switch (mode) {
    case MODE_FAST_SPONGE: return acoustic_step_kernel<R, NW, MODE_FAST_SPONGE, default_slots<R>()>;
    case MODE_FAST_CPML: return acoustic_step_kernel<R, NW, MODE_FAST_CPML, default_slots<R>()>;
    case MODE_REF_SPONGE: return acoustic_step_kernel<R, NW, MODE_REF_SPONGE, default_slots<R>()>;
    default: return nullptr;
    }
}

template <int R>
StepKernel pick_nw(int nw, int mode) {
    switch (nw) {
    case 1: return pick_mode<R, 1>(mode);
    case 2: return pick_mode<R, 2>(mode);
    case 4: return pick_mode<R, 4>(mode);
    default: return nullptr;
    }
}

int pad_rows(int radius) {
    switch (radius) {
    case 4: return kPadRows<4>;
    case 5: return kPadRows<5>;
    case 8: return kPadRows<8>;
    default: return 0;
    }
}

StepKernel pick_step_kernel(int radius, int nw, int mode) {
    switch (radius) {
    case 4: return pick_nw<4>(nw, mode);
    case 5: return pick_nw<5>(nw, mode);
    case 8: return pick_nw<8>(nw, mode);
    default: return nullptr;
    }
}


template <int R, int NW>
BatchKernel pick_batch_mode(int mode) {
    switch (mode) {
    case MODE_FAST_SPONGE: return acoustic_batch_step_kernel<R, NW, MODE_FAST_SPONGE, default_slots<R>()>;
    case MODE_FAST_CPML: return acoustic_batch_step_kernel<R, NW, MODE_FAST_CPML, default_slots<R>()>;
    case MODE_REF_SPONGE: return acoustic_batch_step_kernel<R, NW, MODE_REF_SPONGE, default_slots<R>()>;
    default: return nullptr;
    }
}

template <int R>
BatchKernel pick_batch_nw(int nw, int mode) {
    switch (nw) {
    case 1: return pick_batch_mode<R, 1>(mode);
    case 2: return pick_batch_mode<R, 2>(mode);
    case 4: return pick_batch_mode<R, 4>(mode);
    default: return nullptr;
    }
}

BatchKernel pick_batch_kernel(int radius, int nw, int mode) {
    switch (radius) {
    case 4: return pick_batch_nw<4>(nw, mode);
    case 5: return pick_batch_nw<5>(nw, mode);
    case 8: return pick_batch_nw<8>(nw, mode);
    default: return nullptr;
    }
}

}  // namespace sfd

using namespace sfd;


namespace sfd {

int validate(sfd_context* ctx, const sfd_acoustic_params* p, const float* vp, const float* wavelet) {
    if (!p || !vp || !wavelet) return fail(ctx, SFD_EINVAL, "null params, vp or wavelet");
    if (p->radius != 4 && p->radius != 5 && p->radius != 8)
        return fail(ctx, SFD_EINVAL, "radius must be 4, 5 or 8 (got %d)", p->radius);
    if (p->nx <= 2 * p->radius || p->nz <= 2 * p->radius)
        return fail(ctx, SFD_EINVAL, "grid %dx%d too small for radius %d", p->nx, p->nz, p->radius);
    if (p->nt < 0) return fail(ctx, SFD_EINVAL, "nt < 0");
    if (!(p->dx > 0) || !(p->dz > 0) || !(p->dt > 0)) return fail(ctx, SFD_EINVAL, "dx, dz, dt must be > 0");
    if (p->boundary != SFD_BOUNDARY_SPONGE && p->boundary != SFD_BOUNDARY_CPML)
        return fail(ctx, SFD_EINVAL, "unknown boundary kind %d", p->boundary);
    if (p->width < 0 || 2 * p->width >= p->nx || 2 * p->width >= p->nz)
        return fail(ctx, SFD_EINVAL, "boundary width %d does not fit the grid", p->width);
    if (p->boundary == SFD_BOUNDARY_CPML && !(p->pml_r > 0 && p->pml_r < 1))
        return fail(ctx, SFD_EINVAL, "C-PML reflection coefficient must be in (0,1)");
    if (p->arith != SFD_ARITH_FAST && p->arith != SFD_ARITH_REFERENCE)
        return fail(ctx, SFD_EINVAL, "unknown arithmetic %d", p->arith);
    if (p->arith == SFD_ARITH_REFERENCE && p->boundary != SFD_BOUNDARY_SPONGE)
        return fail(ctx, SFD_ENOTSUP, "reference arithmetic exists for the sponge boundary only");
    if (p->sx < 0 || p->sx >= p->nx || p->sz < 0 || p->sz >= p->nz)
        return fail(ctx, SFD_EINVAL, "source (%d,%d) outside the grid", p->sx, p->sz);
    if (p->nrec < 0 || (p->nrec > 0 && (!p->rx || !p->rz))) return fail(ctx, SFD_EINVAL, "bad receiver list");
    for (int r = 0; r < p->nrec; ++r)
        if (p->rx[r] < 0 || p->rx[r] >= p->nx || p->rz[r] < 0 || p->rz[r] >= p->nz)
            return fail(ctx, SFD_EINVAL, "receiver %d (%d,%d) outside the grid", r, p->rx[r], p->rz[r]);
    if (p->nsnap < 0 || (p->nsnap > 0 && !p->snap_steps)) return fail(ctx, SFD_EINVAL, "bad snapshot list");
    for (int s = 0; s < p->nsnap; ++s) {
        if (p->snap_steps[s] < 0 || p->snap_steps[s] >= p->nt || (s > 0 && p->snap_steps[s] <= p->snap_steps[s - 1]))
            return fail(ctx, SFD_EINVAL, "snapshot steps must be ascending and in [0, nt)");
    }
    if (p->nsnap > 0 && (p->snap_fields & 7) == 0) return fail(ctx, SFD_EINVAL, "snap_fields is empty");
    return SFD_OK;
}

// Sponge factors of the reference's table (:164-194) as two 1-D profiles.
// ax[i]: factor of x-edge row i, or 2.0 where row i is interior; az[j]: factor
// of z-edge column j, 1.0 inside.  The formula is evaluated in double exactly
// as the reference writes it and rounded to float once.
float sponge_factor(int d, int L) {
    return (float)(-0.38 * d * d / (1.0 * L * L) + 1);
}

void sponge_profiles(int nx, int nz, int L, std::vector<float>& ax, std::vector<float>& az) {
    ax.assign(nx, 2.0f);
    az.assign(nz, 1.0f);
    if (L <= 0) return;
    for (int i = 0; i < L; ++i) ax[i] = sponge_factor(L - i - 1, L);
    for (int i = nx - L; i < nx; ++i) ax[i] = sponge_factor(i - nx + L, L);
    for (int j = 0; j < L; ++j) az[j] = sponge_factor(L - j - 1, L);
    for (int j = nz - L; j < nz; ++j) az[j] = sponge_factor(j - nz + L, L);
}

// C-PML (Roden & Gedney 2000; Komatitsch & Martin 2007 with kappa = 1):
// d(x) = d0 xi^2, d0 = -3 vmax ln(R) / (2 w h), alpha = pi f0 (1 - xi),
// b = exp(-(d + alpha) dt), a = d (b - 1) / (d + alpha), xi = depth / w at
// x = i + shift cells; the absorbing cells are [0, w) and [n - w, n).
void cpml_profile(int n, int w, double h, double dt, double R, double vmax, double f0,
                  std::vector<float>& prof) {
    prof.assign(4 * (size_t)n, 0.f);
    for (int half = 0; half < 2; ++half) {
        for (int i = 0; i < n; ++i) {
            const double x = i + 0.5 * half;
            const double lo = w, hi = n - 1 - w;
            double depth = 0.0;
            if (x < lo) depth = lo - x;
            else if (x > hi) depth = x - hi;
            double b = 1.0, a = 0.0;
            if (w > 0 && depth > 0.0) {
                const double xi = std::min(depth / w, 1.0);
                const double d = -3.0 * vmax * std::log(R) / (2.0 * w * h) * xi * xi;
                const double alpha = M_PI * f0 * (1.0 - xi);
                b = std::exp(-(d + alpha) * dt);
                a = d * (b - 1.0) / (d + alpha);
            }
            prof[(size_t)(2 * half) * n + i] = (float)b;
            prof[(size_t)(2 * half + 1) * n + i] = (float)a;
        }
    }
}


template <int R, int TX, int NW>
bool tiled_mode(int mode, StepKernel* v, StepKernel* p) {
    switch (mode) {
#define SFD_TILED_MODE(M)                                \
    case M:                                              \
        *v = acoustic_tiled_velocity<R, TX, NW, M>;      \
        *p = acoustic_tiled_pressure<R, TX, NW, M>;      \
        return true;
        SFD_TILED_MODE(MODE_FAST_CPML)
        SFD_TILED_MODE(MODE_FAST_SPONGE)
#undef SFD_TILED_MODE
    default:
        return false;
    }
}

// Tiled shapes of single plans.  The default shape, measured on MI355X for
// the 2101 x 4101 C-PML grid (R = 8, us per step, both launches): 16 rows x 4
// waves 58.9, 32 x 8 60.4, 16 x 8 60.6, 8 x 4 62.9, 32 x 16 66.7, 8 x 8 78.3,
// 16 x 16 86.5 (the fused march kernel: 125.1).  The gradient's modes exist
// at the default shape.
template <int R>
bool pick_tiled_r(int tx, int nw, int mode, StepKernel* v, StepKernel* p) {
    if (tx == kTiledTX && nw == kTiledNW &&
        (mode == MODE_FWD_DIV || mode == MODE_ADJ || mode == MODE_FWD_DIV_CPML || mode == MODE_ADJ_CPML)) {
        if (mode == MODE_FWD_DIV) {
            *v = acoustic_tiled_velocity<R, kTiledTX, kTiledNW, MODE_FWD_DIV>;
            *p = acoustic_tiled_pressure<R, kTiledTX, kTiledNW, MODE_FWD_DIV>;
        } else if (mode == MODE_ADJ) {
            *v = acoustic_tiled_velocity<R, kTiledTX, kTiledNW, MODE_ADJ>;
            *p = acoustic_tiled_pressure<R, kTiledTX, kTiledNW, MODE_ADJ>;
        } else if (mode == MODE_FWD_DIV_CPML) {
            *v = acoustic_tiled_velocity<R, kTiledTX, kTiledNW, MODE_FWD_DIV_CPML>;
            *p = acoustic_tiled_pressure<R, kTiledTX, kTiledNW, MODE_FWD_DIV_CPML>;
        } else {
            // The C-PML adjoint's two launches (adjoint of the pressure half,
            // then of the velocity half) take the velocity / pressure slots.
            *v = acoustic_adjc_a1<R, kTiledTX, kTiledNW>;
            *p = acoustic_adjc_a2<R, kTiledTX, kTiledNW>;
        }
        return true;
    }
    switch (tx * 100 + nw) {
#define SFD_TILED(TX, NW) \
    case TX * 100 + NW: return tiled_mode<R, TX, NW>(mode, v, p);
        SFD_TILED(8, 4)
        SFD_TILED(8, 8)
        SFD_TILED(16, 4)
        SFD_TILED(16, 8)
        SFD_TILED(16, 16)
        SFD_TILED(32, 4)
        SFD_TILED(32, 8)
        SFD_TILED(32, 16)
#undef SFD_TILED
    default:
        return false;
    }
}

bool pick_tiled(int radius, int tx, int nw, int mode, StepKernel* v, StepKernel* p) {
    switch (radius) {
    case 4: return pick_tiled_r<4>(tx, nw, mode, v, p);
    case 5: return pick_tiled_r<5>(tx, nw, mode, v, p);
    case 8: return pick_tiled_r<8>(tx, nw, mode, v, p);
    default: return false;
    }
}

template <int R, int TX, int NW>
bool tiled_batch_mode(int mode, BatchKernel* v, BatchKernel* p) {
    switch (mode) {
#define SFD_TILED_BATCH(M)                                    \
    case M:                                                   \
        *v = acoustic_tiled_velocity_batch<R, TX, NW, M>;     \
        *p = acoustic_tiled_pressure_batch<R, TX, NW, M>;     \
        return true;
        SFD_TILED_BATCH(MODE_FAST_CPML)
        SFD_TILED_BATCH(MODE_FAST_SPONGE)
#undef SFD_TILED_BATCH
    default:
        return false;
    }
}

// Tiled shapes of batches (the autotuner's candidates); the gradient's modes
// at the default shape.
template <int R>
bool pick_tiled_batch_r(int tx, int nw, int mode, BatchKernel* v, BatchKernel* p) {
    if (tx == kTiledTX && nw == kTiledNW &&
        (mode == MODE_FWD_DIV || mode == MODE_ADJ || mode == MODE_REV || mode == MODE_FWD_DIV_CPML ||
         mode == MODE_ADJ_CPML)) {
        if (mode == MODE_FWD_DIV) {
            *v = acoustic_tiled_velocity_batch<R, kTiledTX, kTiledNW, MODE_FWD_DIV>;
            *p = acoustic_tiled_pressure_batch<R, kTiledTX, kTiledNW, MODE_FWD_DIV>;
        } else if (mode == MODE_FWD_DIV_CPML) {
            *v = acoustic_tiled_velocity_batch<R, kTiledTX, kTiledNW, MODE_FWD_DIV_CPML>;
            *p = acoustic_tiled_pressure_batch<R, kTiledTX, kTiledNW, MODE_FWD_DIV_CPML>;
        } else if (mode == MODE_ADJ_CPML) {
            *v = acoustic_adjc_a1_batch<R, kTiledTX, kTiledNW>;
            *p = acoustic_adjc_a2_batch<R, kTiledTX, kTiledNW>;
        } else if (mode == MODE_ADJ) {
            *v = acoustic_tiled_velocity_batch<R, kTiledTX, kTiledNW, MODE_ADJ>;
            *p = acoustic_tiled_pressure_batch<R, kTiledTX, kTiledNW, MODE_ADJ>;
        } else {
            *v = acoustic_tiled_velocity_batch<R, kTiledTX, kTiledNW, MODE_REV>;
            *p = acoustic_tiled_pressure_batch<R, kTiledTX, kTiledNW, MODE_REV>;
        }
        return true;
    }
    switch (tx * 100 + nw) {
    case 804: return tiled_batch_mode<R, 8, 4>(mode, v, p);
    case 1604: return tiled_batch_mode<R, 16, 4>(mode, v, p);
    case 1608: return tiled_batch_mode<R, 16, 8>(mode, v, p);
    case 3204: return tiled_batch_mode<R, 32, 4>(mode, v, p);
    case 3208: return tiled_batch_mode<R, 32, 8>(mode, v, p);
    default: return false;
    }
}

bool pick_tiled_batch(int radius, int tx, int nw, int mode, BatchKernel* v, BatchKernel* p) {
    switch (radius) {
    case 4: return pick_tiled_batch_r<4>(tx, nw, mode, v, p);
    case 5: return pick_tiled_batch_r<5>(tx, nw, mode, v, p);
    case 8: return pick_tiled_batch_r<8>(tx, nw, mode, v, p);
    default: return false;
    }
}

bool pick_fused_r4(int rows, int nw, int mode, bool ubf, StepKernel* k, BatchKernel* b);
bool pick_fused_r5(int rows, int nw, int mode, bool ubf, StepKernel* k, BatchKernel* b);
bool pick_fused_r8(int rows, int nw, int mode, bool ubf, StepKernel* k, BatchKernel* b);

bool pick_fused(int radius, int rows, int nw, int mode, bool ubf, StepKernel* k, BatchKernel* b) {
    switch (radius) {
    case 4: return pick_fused_r4(rows, nw, mode, ubf, k, b);
    case 5: return pick_fused_r5(rows, nw, mode, ubf, k, b);
    case 8: return pick_fused_r8(rows, nw, mode, ubf, k, b);
    default: return false;
    }
}

// Modes with a tiled step (the reference arithmetic stays on the march kernel).
bool tiled_mode_ok(int mode) { return mode == MODE_FAST_CPML || mode == MODE_FAST_SPONGE; }

// Chunks start at rows congruent to R modulo 2R (the kernel's queue slots are
// compile-time constants), so the row count of a chunk is a multiple of 2R.
int round_chunk(int ch, int radius) {
    const int q = 2 * radius;
    return std::max(q, (ch + q - 1) / q * q);
}

void auto_tiling(const sfd_acoustic_plan* pl, int* nw, int* chunk) {
    // Workgroups of 2 waves; the row chunk is chosen so the grid holds about
    // 3072 waves (12 per CU), and never shorter than 48 rows: the queue fill
    // costs about 4R marched rows per chunk.  (Measured on MI355X for the
    // 2101 x 4101 grid: 2 waves x 48-64 rows is the fastest shape.)
    const int upd_z = pl->nz - 2 * pl->radius, upd_x = pl->nx - 2 * pl->radius;
    const int w = upd_z >= 128 ? 2 : 1;
    const long waves_z = (upd_z + 63) / 64;
    int ch = (int)std::max<long>(48, (long)upd_x * waves_z / 3072);
    ch = std::min(ch, std::max(upd_x, 1));
    *nw = w;
    *chunk = round_chunk(ch, pl->radius);
}

// Watchdog after a run: the gather and the final pressure field.
int check_plan_finite(sfd_acoustic_plan* pl) {
    const int rc = check_finite(pl->ctx, pl->rec_buf.as<float>(), pl->rec_buf.bytes / 4, "the gather");
    if (rc) return rc;
    return check_finite(pl->ctx, pl->grid(pl->p[pl->cur]), (size_t)pl->nx * pl->nz, "the pressure field");
}

int zero_state(sfd_acoustic_plan* pl) {
    hipStream_t s = pl->ctx->stream;
    for (int b = 0; b < 2; ++b) {
        SFD_HIP(pl->ctx, hipMemsetAsync(pl->p[b].ptr, 0, pl->p[b].bytes, s));
        SFD_HIP(pl->ctx, hipMemsetAsync(pl->vx[b].ptr, 0, pl->vx[b].bytes, s));
        SFD_HIP(pl->ctx, hipMemsetAsync(pl->vz[b].ptr, 0, pl->vz[b].bytes, s));
        for (Field* d : {&pl->psi_px[b], &pl->psi_vx[b], &pl->psi_pz[b], &pl->psi_vz[b]})
            if (d->ptr) SFD_HIP(pl->ctx, hipMemsetAsync(d->ptr, 0, d->bytes, s));
    }
    if (pl->rec_buf.ptr) SFD_HIP(pl->ctx, hipMemsetAsync(pl->rec_buf.ptr, 0, pl->rec_buf.bytes, s));
    pl->cur = 0;
    pl->v_raw = false;
    return SFD_OK;
}

// Step-kernel arguments of a plan whose current time level is in buffers
// `cur` (inputs) and cur ^ 1 (outputs) of the given state fields.
AcousticStepArgs make_args(const sfd_acoustic_plan* pl, const Field* P, const Field* VX, const Field* VZ, int cur,
                           int k) {
    AcousticStepArgs a{};
    a.arena = pl->arena.as<float>();
    a.arena_bytes = (uint32_t)pl->arena.bytes;
    a.nx = pl->nx;
    a.nz = pl->nz;
    a.chunk = pl->chunk;
    a.k = k;
    a.kf = pl->grid(pl->kf);
    a.bf = pl->grid(pl->bf);
    a.bf_const = pl->bf_const;
    a.ax = pl->ax.as<float>() + pl->pad;
    a.az = pl->az.as<float>();
    a.pml_x = pl->pml_x.as<float>();
    a.pml_z = pl->pml_z.as<float>();
    a.pml_xw = pl->pml_xw;
    a.pml_zw = pl->pml_zw;
    a.sponge_w = pl->mode == MODE_FAST_CPML ? 0 : pl->width;
    a.sx = pl->sx;
    a.sz = pl->sz;
    a.amp = pl->amp.as<float>();
    a.rec_row = pl->rec_row.as<int4>() + pl->pad;
    a.rec_more = pl->rec_more.as<int4>();
    a.rec_buf = pl->rec_buf.as<float>();
    a.nslot = pl->nslot;
    const bool ref = pl->mode == MODE_REF_SPONGE;
    for (int m = 0; m < 8; ++m) a.cx[m] = ref ? pl->coef[m] : pl->coef[m] / pl->dx;
    a.zx = pl->dx / pl->dz;
    a.dtdx = pl->dt / pl->dx;
    a.dtdz = pl->dt / pl->dz;
    a.inv_dx = 1.0 / (double)pl->dx;
    a.inv_dz = 1.0 / (double)pl->dz;
    const int TZ = 64 * pl->nw;
    a.ntz = (pl->nz - 2 * pl->radius + TZ - 1) / TZ;
    a.nchunk = (pl->nx - 2 * pl->radius + pl->chunk - 1) / pl->chunk;
    const int c = cur, o = cur ^ 1;
    a.p_in = pl->grid(P[c]);
    a.vx_in = pl->grid(VX[c]);
    a.vz_in = pl->grid(VZ[c]);
    a.p_out = pl->grid(P[o]);
    a.vx_out = pl->grid(VX[o]);
    a.vz_out = pl->grid(VZ[o]);
    a.psi_px_in = pl->psi_px[c].as<float>();
    a.psi_px_out = pl->psi_px[o].as<float>();
    a.psi_vx_in = pl->psi_vx[c].as<float>();
    a.psi_vx_out = pl->psi_vx[o].as<float>();
    a.psi_pz_in = pl->psi_pz[c].as<float>();
    a.psi_pz_out = pl->psi_pz[o].as<float>();
    a.psi_vz_in = pl->psi_vz[c].as<float>();
    a.psi_vz_out = pl->psi_vz[o].as<float>();
    a.div = pl->grid_period_div();
    a.div_stride = pl->div_stride();
    a.div_period = pl->div_fields > 0 ? pl->div_fields : 1;
    a.grad = pl->grad_period > 0 ? pl->grid(pl->grad) : nullptr;
    a.src_adj = pl->src_adj.as<float>();
    // Box of the reverse reconstruction.
    const int bx = rev_margin(pl);
    a.rev_x0 = bx;
    a.rev_x1 = pl->nx - bx;
    a.rev_z0 = bx;
    a.rev_z1 = pl->nz - bx;
    return a;
}

int rev_margin(const sfd_acoustic_plan* pl) {
    return std::max(pl->mode == MODE_FAST_CPML ? pl->width + 1 : pl->width, pl->radius);
}

AcousticStepArgs make_args(const sfd_acoustic_plan* pl, int cur, int k) {
    return make_args(pl, pl->p, pl->vx, pl->vz, cur, k);
}

// Autotuner grid class of a plan: radius, boundary and arithmetic, nx and nz
// to the nearest power of 2.
const char* mode_name(int mode) {
    switch (mode) {
    case MODE_FAST_CPML: return "cpml";
    case MODE_FAST_SPONGE: return "sponge";
    case MODE_REF_SPONGE: return "sponge-ref";
    default: return "other";
    }
}


std::string plan_class(const sfd_acoustic_plan* pl) {
    char buf[96];
    snprintf(buf, sizeof buf, "acoustic:R%d:%s:x%d:z%d", pl->radius, mode_name(pl->mode), log2_round(pl->nx),
             log2_round(pl->nz));
    return buf;
}

// Sets a plan's launch shape: kind 0 = tiled (a rows, b waves), 1 = march (a
// waves, b chunk rows), 2 = one-launch tiled (a rows per wave, b waves).
int apply_shape(sfd_acoustic_plan* pl, const TunedShape& t) {
    if (t.kind == 2) return sfd_acoustic_plan_set_fused(pl, t.a, t.b);
    return t.kind == 0 ? sfd_acoustic_plan_set_tiled(pl, t.a, t.b) : sfd_acoustic_plan_set_tiling(pl, t.a, t.b);
}

// A plan's launch-shape fields, saved and put back around a run at another
// shape.
struct PlanShape {
    int nw, chunk, ttx, tnw, ftx, fnw;
    StepKernel kernel, tvel, tpres, fstep;
};

PlanShape save_shape(const sfd_acoustic_plan* pl) {
    return {pl->nw, pl->chunk, pl->ttx, pl->tnw, pl->ftx, pl->fnw, pl->kernel, pl->tvel, pl->tpres, pl->fstep};
}

void restore_shape(sfd_acoustic_plan* pl, const PlanShape& s) {
    pl->nw = s.nw;
    pl->chunk = s.chunk;
    pl->ttx = s.ttx;
    pl->tnw = s.tnw;
    pl->ftx = s.ftx;
    pl->fnw = s.fnw;
    pl->kernel = s.kernel;
    pl->tvel = s.tvel;
    pl->tpres = s.tpres;
    pl->fstep = s.fstep;
}

// The shape a plan created on its own starts with: the context's cached shape
// for its grid class, else the tiled step's default (fast arithmetic) or the
// march kernel's automatic tiling.
int apply_plain_shape(sfd_acoustic_plan* pl) {
    const auto tuned = pl->ctx->tuning.find(plan_class(pl));
    if (tuned != pl->ctx->tuning.end() && apply_shape(pl, tuned->second) == SFD_OK) return SFD_OK;
    return sfd_acoustic_plan_set_tiling(pl, 0, 0);
}

float sponge_at(const sfd_acoustic_plan* pl, int i, int j) {
    if (pl->mode == MODE_FAST_CPML || pl->width <= 0) return 1.0f;
    const int L = pl->width;
    if (i < L) return sponge_factor(L - i - 1, L);
    if (i >= pl->nx - L) return sponge_factor(i - pl->nx + L, L);
    if (j < L) return sponge_factor(L - j - 1, L);
    if (j >= pl->nz - L) return sponge_factor(j - pl->nz + L, L);
    return 1.0f;
}

// Velocities of a tiled sponge run are stored unabsorbed; this writes them
// absorbed (dst = src * the cell's sponge factor, dst == src allowed), as the
// fused kernel stores them.  Fields are nx x nz from row 0; cells outside
// the updated region hold zero either way.
__global__ void absorb_velocity_kernel(float* dst, const float* src, const float* __restrict__ ax,
                                       const float* __restrict__ az, int sw, int nx, int nz) {
    const size_t n = (size_t)nx * nz;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(c / nz), j = (int)(c - (size_t)i * nz);
        dst[c] = src[c] * tiled_sponge(i, nx, sw, ax[i], az[j]);
    }
}

int launch_absorb(const sfd_acoustic_plan* pl, float* dst, const float* src) {
    const size_t n = (size_t)pl->nx * pl->nz;
    const int bs = 256, nb = (int)std::min<size_t>((n + bs - 1) / bs, 4096);
    hipLaunchKernelGGL(absorb_velocity_kernel, dim3(nb), dim3(bs), 0, pl->ctx->stream, dst, src,
                       pl->ax.as<float>() + pl->pad, pl->az.as<float>(), pl->width, pl->nx, pl->nz);
    SFD_HIP(pl->ctx, hipGetLastError());
    return SFD_OK;
}

void tiled_args(const sfd_acoustic_plan* pl, AcousticStepArgs& a, bool v_raw) {
    // In place: within each launch a cell's outputs are read by no other
    // workgroup (velocities read P's halo, pressure reads the velocities'
    // halo), so one set of buffers suffices and the step's working set is
    // half the ping-pong one (it then fits the 256 MB MALL at 2101 x 4101).
    a.p_out = const_cast<float*>(a.p_in);
    a.vx_out = const_cast<float*>(a.vx_in);
    a.vz_out = const_cast<float*>(a.vz_in);
    a.psi_px_out = const_cast<float*>(a.psi_px_in);
    a.psi_vx_out = const_cast<float*>(a.psi_vx_in);
    a.psi_pz_out = const_cast<float*>(a.psi_pz_in);
    a.psi_vz_out = const_cast<float*>(a.psi_vz_in);
    a.ntz = (pl->nz - 2 * pl->radius + kTiledTZ - 1) / kTiledTZ;
    a.nchunk = (pl->nx - 2 * pl->radius + pl->ttx - 1) / pl->ttx;
    a.v_raw = v_raw ? 1 : 0;
}

void fused_args(const sfd_acoustic_plan* pl, AcousticStepArgs& a, bool v_raw) {
    a.chunk = pl->ftx;
    a.ntz = (pl->nz - 2 * pl->radius + kTiledTZ - 1) / kTiledTZ;
    a.nchunk = (pl->nx - 2 * pl->radius + pl->ftx - 1) / pl->ftx;
    a.v_raw = v_raw ? 1 : 0;
}

int launch_tiled(const sfd_acoustic_plan* pl, StepKernel v, StepKernel p, const AcousticStepArgs& a) {
    const dim3 grid(a.ntz * a.nchunk), block(64 * pl->tnw);
    hipLaunchKernelGGL(v, grid, block, 0, pl->ctx->stream, a);
    SFD_HIP(pl->ctx, hipGetLastError());
    hipLaunchKernelGGL(p, grid, block, 0, pl->ctx->stream, a);
    SFD_HIP(pl->ctx, hipGetLastError());
    return SFD_OK;
}

// Enqueues steps k0 .. k0+nsteps-1 on the context stream.
int enqueue_steps(sfd_acoustic_plan* pl, int k0, int nsteps) {
    sfd_context* ctx = pl->ctx;
    const size_t nbytes = (size_t)pl->nx * pl->nz * sizeof(float);
    size_t snap_next = std::lower_bound(pl->snap_steps.begin(), pl->snap_steps.end(), k0) - pl->snap_steps.begin();
    const bool tiled = pl->ttx > 0, fused = !tiled && pl->ftx > 0;
    if (!tiled && !fused && pl->v_raw && nsteps > 0) {
        // The march kernel reads absorbed velocities.
        for (Field* f : {&pl->vx[pl->cur], &pl->vz[pl->cur]}) {
            const int rc = launch_absorb(pl, pl->grid(*f), pl->grid(*f));
            if (rc) return rc;
        }
        pl->v_raw = false;
    }
    for (int s = 0; s < nsteps; ++s) {
        const int k = k0 + s;
        AcousticStepArgs a = make_args(pl, pl->cur, k);
        if (tiled) {
            tiled_args(pl, a, pl->v_raw);
            const int rc = launch_tiled(pl, pl->tvel, pl->tpres, a);
            if (rc) return rc;
            pl->v_raw = pl->mode == MODE_FAST_SPONGE;
        } else if (fused) {
            fused_args(pl, a, pl->v_raw);
            hipLaunchKernelGGL(pl->fstep, dim3(a.ntz * a.nchunk), dim3(64 * pl->fnw), 0, ctx->stream, a);
            SFD_HIP(ctx, hipGetLastError());
            pl->cur ^= 1;
            pl->v_raw = pl->mode == MODE_FAST_SPONGE;
        } else {
            hipLaunchKernelGGL(pl->kernel, dim3(a.ntz * a.nchunk), dim3(64 * pl->nw), 0, ctx->stream, a);
            SFD_HIP(ctx, hipGetLastError());
            pl->cur ^= 1;
        }
        const int o = pl->cur;
        while (snap_next < pl->snap_steps.size() && pl->snap_steps[snap_next] == k) {
            float* dst = pl->snaps.as<float>() + snap_next * pl->nsnap_fields * (size_t)pl->nx * pl->nz;
            const float* src[3] = {pl->grid(pl->p[o]), pl->grid(pl->vx[o]), pl->grid(pl->vz[o])};
            for (int f = 0; f < 3; ++f) {
                if (!(pl->snap_fields & (1 << f))) continue;
                if (f > 0 && pl->v_raw) {
                    const int rc = launch_absorb(pl, dst, src[f]);
                    if (rc) return rc;
                } else {
                    SFD_HIP(ctx, hipMemcpyAsync(dst, src[f], nbytes, hipMemcpyDeviceToDevice, ctx->stream));
                }
                dst += (size_t)pl->nx * pl->nz;
            }
            ++snap_next;
        }
    }
    return SFD_OK;
}

}  // namespace sfd

extern "C" {

int sfd_acoustic_plan_create(sfd_context* ctx, const sfd_acoustic_params* p, const float* vp,
                             const float* den, const float* wavelet, sfd_acoustic_plan** out) {
    return plan_create(ctx, p, vp, den, wavelet, 0, out);
}

}  // extern "C"

int sfd::plan_create(sfd_context* ctx, const sfd_acoustic_params* p, const float* vp, const float* den,
                     const float* wavelet, int grad_period, sfd_acoustic_plan** out, int div_fields,
                     PlanModelCache* mc) {
    if (!ctx || !out) return SFD_EINVAL;
    *out = nullptr;
    int rc = validate(ctx, p, vp, wavelet);
    if (rc) return rc;
    SFD_HIP(ctx, hipSetDevice(ctx->device));

    sfd_acoustic_plan* pl = new (std::nothrow) sfd_acoustic_plan;
    if (!pl) return fail(ctx, SFD_ENOMEM, "host allocation failed");
    pl->ctx = ctx;
    pl->grad_period = grad_period;
    pl->div_fields = grad_period > 0 ? (div_fields > 0 ? div_fields : grad_period) : 0;
    pl->nx = p->nx;
    pl->nz = p->nz;
    pl->nt = p->nt;
    pl->dx = p->dx;
    pl->dz = p->dz;
    pl->dt = p->dt;
    pl->radius = p->radius;
    pl->width = p->width;
    pl->sx = p->sx;
    pl->sz = p->sz;
    pl->mode = (p->boundary == SFD_BOUNDARY_CPML) ? MODE_FAST_CPML
               : (p->arith == SFD_ARITH_REFERENCE) ? MODE_REF_SPONGE : MODE_FAST_SPONGE;
    if (p->coef) {
        for (int m = 0; m < p->radius; ++m) pl->coef[m] = p->coef[m];
    } else {
        default_coefs(p->radius, pl->coef);
    }
    auto bail = [&](int code) {
        delete pl;
        return code;
    };

    const int nx = p->nx, nz = p->nz;
    const size_t n = (size_t)nx * nz;
    // The model's fields from an earlier plan of the same call, if it had them.
    const bool shared = mc && mc->valid && mc->vp == vp && mc->den == den && mc->nx == nx && mc->nz == nz &&
                        mc->radius == p->radius && mc->mode == pl->mode && mc->dt == p->dt;

    // Stability: vmax dt sqrt(1/dx^2 + 1/dz^2) sum|C| <= 1.
    float vmax = shared ? mc->vmax : 0.f;
    for (size_t c = 0; !shared && c < n; ++c) {
        if (!(vp[c] > 0.f) || !std::isfinite(vp[c]))
            return bail(fail(ctx, SFD_EINVAL, "vp must be finite and > 0 (cell %zu)", c));
        vmax = std::max(vmax, vp[c]);
        if (den && (!(den[c] > 0.f) || !std::isfinite(den[c])))
            return bail(fail(ctx, SFD_EINVAL, "den must be finite and > 0 (cell %zu)", c));
    }
    const double courant = vmax * (double)p->dt * std::sqrt(1.0 / ((double)p->dx * p->dx) + 1.0 / ((double)p->dz * p->dz));
    const double limit = sfd_acoustic_cfl_limit(p->radius, pl->coef);
    if (courant > limit)
        return bail(fail(ctx, SFD_ECFL, "unstable: vmax*dt*sqrt(1/dx^2+1/dz^2) = %.4f exceeds %.4f", courant, limit));

    // Per-cell coefficients.  FAST: kf = rho vp^2 dt, bf = dt/rho.  REF: kf is
    // the reference's float product DEN*Vp*Vp*dt (:293) and bf holds DEN.
    std::vector<float> hk(shared ? 0 : n), hb(shared ? 0 : n);
    const bool ref = pl->mode == MODE_REF_SPONGE;
    // In fast arithmetic both are zero outside the updated cells [R, N-R)^2,
    // so the kernel's reads there yield zero velocities without masking.
    const int R = p->radius;
    for (int i = 0; !shared && i < nx; ++i) {
        for (int jj = 0; jj < nz; ++jj) {
            const size_t c = (size_t)i * nz + jj;
            const float d = den ? den[c] : 2400.0f;
            const bool upd = i >= R && i < nx - R && jj >= R && jj < nz - R;
            hk[c] = (ref || upd) ? d * vp[c] * vp[c] * p->dt : 0.f;
            hb[c] = ref ? d : upd ? (float)((double)p->dt / (double)d) : 0.f;
        }
    }
    // Uniform density: every updated cell's bf is the same number.
    if (shared) {
        pl->bf_const = mc->bf_const;
    } else if (!ref) {
        bool uniform = true;
        const float d0 = den ? den[0] : 2400.0f;
        for (size_t c = 1; den && uniform && c < n; ++c) uniform = den[c] == d0;
        if (uniform) pl->bf_const = (float)((double)p->dt / (double)d0);
    }
    // Source amplitude per step, the reference's float product dt*DEN*Vp*signal (:249-250).
    std::vector<float> hamp(std::max(p->nt, 1), 0.f);
    {
        const size_t s = (size_t)p->sx * nz + p->sz;
        const float d = den ? den[s] : 2400.0f;
        for (int k = 0; k < p->nt; ++k) hamp[k] = p->dt * d * vp[s] * wavelet[k];
    }
    // Receivers: one slot per distinct cell; runs of consecutive z cells in a
    // row become spans (jlo, jhi, slot).  A row's first span goes to its
    // entry in rec_row, further spans to rec_more, chained.
    pl->nrec = p->nrec;
    pl->pad = pad_rows(p->radius);
    const int pad = pl->pad;
    std::vector<int> hslot;
    std::vector<int4> hrow, hmore;
    receiver_spans(nx, pad, p->nrec, p->rx, p->rz, hslot, hrow, hmore, &pl->nslot);
    // Absorbing boundary tables.
    std::vector<float> hax, haz, hpx, hpz;
    sponge_profiles(nx, nz, p->boundary == SFD_BOUNDARY_SPONGE ? p->width : 0, hax, haz);
    if (p->boundary == SFD_BOUNDARY_CPML) {
        cpml_profile(nx, p->width, p->dx, p->dt, p->pml_r, vmax, p->pml_f0, hpx);
        cpml_profile(nz, p->width, p->dz, p->dt, p->pml_r, vmax, p->pml_f0, hpz);
        pl->pml_xw = p->width + 1;
        pl->pml_zw = p->width + 1;
    }

    hipStream_t s = ctx->stream;
    hipError_t e = hipSuccess;
    const size_t npad = (size_t)(nx + 2 * pad) * nz;
    const bool cpml = pl->mode == MODE_FAST_CPML;
    const size_t xs = cpml ? (size_t)2 * pl->pml_xw * nz * 4 : 0, zs = cpml ? (size_t)2 * pl->pml_zw * nx * 4 : 0;
    for (int pass = 0; pass < 2; ++pass) {
        Carver cv;
        if (pass == 1) cv.base = pl->arena.as<char>();
        for (int b = 0; b < 2; ++b) {
            pl->p[b] = cv.take(npad * 4);
            pl->vx[b] = cv.take(npad * 4);
            pl->vz[b] = cv.take(npad * 4);
        }
        pl->kf = cv.take(npad * 4);
        pl->bf = cv.take(npad * 4);
        for (int b = 0; b < 2; ++b) {
            pl->psi_px[b] = cv.take(xs);
            pl->psi_vx[b] = cv.take(xs);
            pl->psi_pz[b] = cv.take(zs);
            pl->psi_vz[b] = cv.take(zs);
        }
        if (grad_period > 0) {
            for (int b = 0; b < 2; ++b) {
                pl->ap[b] = cv.take(npad * 4);
                pl->avx[b] = cv.take(npad * 4);
                pl->avz[b] = cv.take(npad * 4);
            }
            pl->grad = cv.take(npad * 4);
        }
        if (pass == 0) {
            if (cv.used >= (size_t)1 << 31)
                return bail(fail(ctx, SFD_ENOTSUP, "grid needs %zu bytes of fields; the step kernel addresses < 2 GiB",
                                 cv.used));
            e = pl->arena.alloc(cv.used);
            if (e != hipSuccess) return bail(hip_check(ctx, e, "arena allocation"));
        }
    }
    if (grad_period > 0) {
        // The divergence ring is an allocation of its own: the kernels form
        // each step's field base in 64-bit arithmetic, so the ring may pass
        // 2 GiB (nt fields when the gradient keeps every step's divergence).
        const size_t db = npad * 4 * (size_t)pl->div_fields;
        e = pl->div_mem.alloc(db);
        if (e != hipSuccess) return bail(hip_check(ctx, e, "divergence allocation"));
        pl->div = Field{pl->div_mem.ptr, db};
    }
    // Per-row arrays are padded with kPadRows rows on both sides (zeros; the
    // sponge row table with 2.0, "not an x-edge row").
    std::vector<float> hk_pad, hb_pad;
    if (shared) {
        if (e == hipSuccess) e = hipMemcpyAsync(pl->kf.ptr, mc->kf, pl->kf.bytes, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(pl->bf.ptr, mc->bf, pl->bf.bytes, hipMemcpyDeviceToDevice, s);
    } else {
        hk_pad.assign((size_t)(nx + 2 * pad) * nz, 0.f);
        hb_pad.assign(hk_pad.size(), 0.f);
        std::copy(hk.begin(), hk.end(), hk_pad.begin() + (size_t)pad * nz);
        std::copy(hb.begin(), hb.end(), hb_pad.begin() + (size_t)pad * nz);
        if (e == hipSuccess) e = hipMemcpyAsync(pl->kf.ptr, hk_pad.data(), pl->kf.bytes, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(pl->bf.ptr, hb_pad.data(), pl->bf.bytes, hipMemcpyHostToDevice, s);
    }
    std::vector<float> hax_pad(nx + 2 * pad, 2.0f);
    std::copy(hax.begin(), hax.end(), hax_pad.begin() + pad);
    if (e == hipSuccess) e = upload(pl->ax, hax_pad.data(), hax_pad.size(), s);
    if (e == hipSuccess) e = upload(pl->az, haz.data(), haz.size(), s);
    if (e == hipSuccess && !hpx.empty()) e = upload(pl->pml_x, hpx.data(), hpx.size(), s);
    if (e == hipSuccess && !hpz.empty()) e = upload(pl->pml_z, hpz.data(), hpz.size(), s);
    if (e == hipSuccess) e = upload(pl->amp, hamp.data(), hamp.size(), s);
    if (e == hipSuccess) e = upload(pl->rec_row, hrow.data(), hrow.size(), s);
    if (e == hipSuccess && !hmore.empty()) e = upload(pl->rec_more, hmore.data(), hmore.size(), s);
    if (e == hipSuccess && p->nrec > 0) e = upload(pl->rec_slot, hslot.data(), hslot.size(), s);
    if (e == hipSuccess && p->nrec > 0) e = pl->rec_buf.alloc((size_t)pl->nslot * std::max(p->nt, 1) * 4);
    if (e == hipSuccess && p->nrec > 0) e = pl->record.alloc((size_t)p->nrec * std::max(p->nt, 1) * 4);
    pl->nsnap = p->nsnap;
    pl->snap_fields = p->snap_fields & 7;
    pl->nsnap_fields = ((pl->snap_fields >> 0) & 1) + ((pl->snap_fields >> 1) & 1) + ((pl->snap_fields >> 2) & 1);
    pl->snap_steps.assign(p->snap_steps, p->snap_steps + p->nsnap);
    if (e == hipSuccess && p->nsnap > 0) e = pl->snaps.alloc((size_t)p->nsnap * pl->nsnap_fields * n * 4);
    if (e != hipSuccess) return bail(hip_check(ctx, e, "plan allocation"));

    auto_tiling(pl, &pl->nw, &pl->chunk);
    pl->kernel = pick_step_kernel(pl->radius, pl->nw, pl->mode);
    if (!pl->kernel) return bail(fail(ctx, SFD_ENOTSUP, "no kernel for this configuration"));
    if (tiled_mode_ok(pl->mode) && pick_tiled(pl->radius, kTiledTX, kTiledNW, pl->mode, &pl->tvel, &pl->tpres)) {
        pl->ttx = kTiledTX;
        pl->tnw = kTiledNW;
    }
    // A shape the context's autotuner measured for this grid class wins.
    const auto tuned = ctx->tuning.find(plan_class(pl));
    if (tuned != ctx->tuning.end()) (void)apply_shape(pl, tuned->second);
    rc = zero_state(pl);
    if (rc) return bail(rc);
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return bail(hip_check(ctx, e, "plan setup"));
    if (mc && !shared) {
        mc->valid = true;
        mc->vp = vp;
        mc->den = den;
        mc->nx = nx;
        mc->nz = nz;
        mc->radius = p->radius;
        mc->mode = pl->mode;
        mc->dt = p->dt;
        mc->vmax = vmax;
        mc->bf_const = pl->bf_const;
        mc->kf = pl->kf.ptr;
        mc->bf = pl->bf.ptr;
    }
    *out = pl;
    return SFD_OK;
}

extern "C" {

int sfd_acoustic_plan_set_tiling(sfd_acoustic_plan* pl, int waves, int chunk_rows) {
    if (!pl) return SFD_EINVAL;
    int nw = waves, ch = chunk_rows;
    int anw, ach;
    auto_tiling(pl, &anw, &ach);
    if (nw == 0) nw = anw;
    if (ch == 0) ch = ach;
    if (ch < 1) return fail(pl->ctx, SFD_EINVAL, "chunk_rows must be >= 1");
    ch = round_chunk(ch, pl->radius);
    StepKernel k = pick_step_kernel(pl->radius, nw, pl->mode);
    if (!k) return fail(pl->ctx, SFD_EINVAL, "waves must be 1, 2 or 4 (got %d)", waves);
    pl->nw = nw;
    pl->chunk = ch;
    pl->kernel = k;
    // An explicit march tiling selects the march kernel; automatic restores
    // the tiled step where it applies.
    pl->ttx = pl->tnw = 0;
    pl->ftx = pl->fnw = 0;
    if (waves == 0 && chunk_rows == 0 && tiled_mode_ok(pl->mode) &&
        pick_tiled(pl->radius, kTiledTX, kTiledNW, pl->mode, &pl->tvel, &pl->tpres)) {
        pl->ttx = kTiledTX;
        pl->tnw = kTiledNW;
    }
    return SFD_OK;
}

int sfd_acoustic_plan_set_tiled(sfd_acoustic_plan* pl, int tile_rows, int waves) {
    if (!pl) return SFD_EINVAL;
    if (tile_rows < 0) {
        pl->ttx = pl->tnw = 0;
        pl->ftx = pl->fnw = 0;
        return SFD_OK;
    }
    if (!tiled_mode_ok(pl->mode)) return fail(pl->ctx, SFD_ENOTSUP, "the tiled step exists in fast arithmetic only");
    const int tx = tile_rows ? tile_rows : kTiledTX, nw = waves ? waves : kTiledNW;
    StepKernel v, p;
    if (!pick_tiled(pl->radius, tx, nw, pl->mode, &v, &p))
        return fail(pl->ctx, SFD_EINVAL, "no tiled step for %d rows x %d waves", tx, nw);
    pl->ttx = tx;
    pl->tnw = nw;
    pl->tvel = v;
    pl->tpres = p;
    pl->ftx = pl->fnw = 0;
    return SFD_OK;
}

int sfd_acoustic_plan_set_fused(sfd_acoustic_plan* pl, int rows, int waves) {
    if (!pl) return SFD_EINVAL;
    if (!tiled_mode_ok(pl->mode))
        return fail(pl->ctx, SFD_ENOTSUP, "the one-launch step exists in fast arithmetic only");
    const int tr = rows ? rows : kFusedShapes[0][0], nw = waves ? waves : kFusedShapes[0][1];
    StepKernel k;
    if (!pick_fused(pl->radius, tr, nw, pl->mode, pl->bf_const != 0.f, &k, nullptr))
        return fail(pl->ctx, SFD_EINVAL, "no one-launch step for %d rows x %d waves", tr, nw);
    pl->ftx = tr;
    pl->fnw = nw;
    pl->fstep = k;
    pl->ttx = pl->tnw = 0;
    return SFD_OK;
}

int sfd_acoustic_plan_run(sfd_acoustic_plan* pl, float* elapsed_ms) {
    if (!pl) return SFD_EINVAL;
    sfd_context* ctx = pl->ctx;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    int rc = zero_state(pl);
    if (rc) return rc;
    SFD_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    rc = enqueue_steps(pl, 0, pl->nt);
    if (rc) return rc;
    SFD_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    SFD_HIP(ctx, hipEventSynchronize(ctx->ev1));
    if (elapsed_ms) SFD_HIP(ctx, hipEventElapsedTime(elapsed_ms, ctx->ev0, ctx->ev1));
    return check_plan_finite(pl);
}

int sfd_acoustic_plan_step(sfd_acoustic_plan* pl, int k0, int nsteps) {
    if (!pl) return SFD_EINVAL;
    if (k0 < 0 || nsteps < 0 || k0 + nsteps > pl->nt)
        return fail(pl->ctx, SFD_EINVAL, "steps [%d, %d) outside [0, nt=%d)", k0, k0 + nsteps, pl->nt);
    SFD_HIP(pl->ctx, hipSetDevice(pl->ctx->device));
    int rc = enqueue_steps(pl, k0, nsteps);
    if (rc) return rc;
    SFD_HIP(pl->ctx, hipStreamSynchronize(pl->ctx->stream));
    return SFD_OK;
}

int sfd_acoustic_plan_fetch(sfd_acoustic_plan* pl, float* record, float* snaps) {
    if (!pl) return SFD_EINVAL;
    sfd_context* ctx = pl->ctx;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    if (record && pl->nrec > 0 && pl->nt > 0) {
        const dim3 grid((pl->nt + 31) / 32, (pl->nrec + 31) / 32), block(32, 8);
        hipLaunchKernelGGL(gather_transpose_kernel, grid, block, 0, ctx->stream, pl->rec_buf.as<float>(),
                           pl->rec_slot.as<int>(), pl->record.as<float>(), pl->nrec, pl->nt, pl->nslot);
        SFD_HIP(ctx, hipGetLastError());
        SFD_HIP(ctx, hipMemcpyAsync(record, pl->record.ptr, (size_t)pl->nrec * pl->nt * 4, hipMemcpyDeviceToHost,
                                    ctx->stream));
    }
    if (snaps && pl->nsnap > 0)
        SFD_HIP(ctx, hipMemcpyAsync(snaps, pl->snaps.ptr, pl->snaps.bytes, hipMemcpyDeviceToHost, ctx->stream));
    SFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SFD_OK;
}

int sfd_acoustic_plan_set_state(sfd_acoustic_plan* pl, const float* state, const float* pml) {
    if (!pl || !state) return SFD_EINVAL;
    sfd_context* ctx = pl->ctx;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)pl->nx * pl->nz;
    hipStream_t s = ctx->stream;
    // Both ping-pong buffers: cells outside the updated region are never
    // written by a step, and the one-launch step reads them from either.
    for (int b = 0; b < 2; ++b) {
        SFD_HIP(ctx, hipMemcpyAsync(pl->grid(pl->p[b]), state, n * 4, hipMemcpyHostToDevice, s));
        SFD_HIP(ctx, hipMemcpyAsync(pl->grid(pl->vx[b]), state + n, n * 4, hipMemcpyHostToDevice, s));
        SFD_HIP(ctx, hipMemcpyAsync(pl->grid(pl->vz[b]), state + 2 * n, n * 4, hipMemcpyHostToDevice, s));
    }
    pl->v_raw = false;  // the given velocities are absorbed ones
    if (pl->mode == MODE_FAST_CPML) {
        if (!pml) return fail(ctx, SFD_EINVAL, "C-PML plans need the pml state");
        // Full-grid psi fields -> strips.
        const int nx = pl->nx, nz = pl->nz, xw = pl->pml_xw, zw = pl->pml_zw;
        std::vector<float> xs((size_t)2 * xw * nz), zs((size_t)2 * zw * nx);
        const float* f[4] = {pml, pml + n, pml + 2 * n, pml + 3 * n};  // px, pz, vx, vz
        Field* xb[2] = {&pl->psi_px[pl->cur], &pl->psi_vx[pl->cur]};
        const float* xf[2] = {f[0], f[2]};
        for (int q = 0; q < 2; ++q) {
            for (int r = 0; r < 2 * xw; ++r) {
                const int i = r < xw ? r : nx - xw + (r - xw);
                std::memcpy(&xs[(size_t)r * nz], xf[q] + (size_t)i * nz, nz * 4);
            }
            SFD_HIP(ctx, hipMemcpyAsync(xb[q]->ptr, xs.data(), xs.size() * 4, hipMemcpyHostToDevice, s));
            SFD_HIP(ctx, hipStreamSynchronize(s));
        }
        Field* zb[2] = {&pl->psi_pz[pl->cur], &pl->psi_vz[pl->cur]};
        const float* zf[2] = {f[1], f[3]};
        for (int q = 0; q < 2; ++q) {
            for (int i = 0; i < nx; ++i)
                for (int c = 0; c < 2 * zw; ++c) {
                    const int jj = c < zw ? c : nz - zw + (c - zw);
                    zs[(size_t)i * 2 * zw + c] = zf[q][(size_t)i * nz + jj];
                }
            SFD_HIP(ctx, hipMemcpyAsync(zb[q]->ptr, zs.data(), zs.size() * 4, hipMemcpyHostToDevice, s));
            SFD_HIP(ctx, hipStreamSynchronize(s));
        }
    }
    SFD_HIP(ctx, hipStreamSynchronize(s));
    return SFD_OK;
}

int sfd_acoustic_plan_get_state(sfd_acoustic_plan* pl, float* state, float* pml) {
    if (!pl || !state) return SFD_EINVAL;
    sfd_context* ctx = pl->ctx;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)pl->nx * pl->nz;
    hipStream_t s = ctx->stream;
    SFD_HIP(ctx, hipMemcpyAsync(state, pl->grid(pl->p[pl->cur]), n * 4, hipMemcpyDeviceToHost, s));
    SFD_HIP(ctx, hipMemcpyAsync(state + n, pl->grid(pl->vx[pl->cur]), n * 4, hipMemcpyDeviceToHost, s));
    SFD_HIP(ctx, hipMemcpyAsync(state + 2 * n, pl->grid(pl->vz[pl->cur]), n * 4, hipMemcpyDeviceToHost, s));
    SFD_HIP(ctx, hipStreamSynchronize(s));
    if (pl->v_raw) {
        // Stored unabsorbed by the tiled sponge step: apply the factor as the
        // fused kernel does when it stores them.
        for (int i = 0; i < pl->nx; ++i)
            for (int jj = 0; jj < pl->nz; ++jj) {
                const float f = sponge_at(pl, i, jj);
                state[n + (size_t)i * pl->nz + jj] *= f;
                state[2 * n + (size_t)i * pl->nz + jj] *= f;
            }
    }
    if (pml && pl->mode == MODE_FAST_CPML) {
        const int nx = pl->nx, nz = pl->nz, xw = pl->pml_xw, zw = pl->pml_zw;
        std::fill(pml, pml + 4 * n, 0.f);
        std::vector<float> xs((size_t)2 * xw * nz), zs((size_t)2 * zw * nx);
        Field* xb[2] = {&pl->psi_px[pl->cur], &pl->psi_vx[pl->cur]};
        float* xf[2] = {pml, pml + 2 * n};
        for (int q = 0; q < 2; ++q) {
            SFD_HIP(ctx, hipMemcpyAsync(xs.data(), xb[q]->ptr, xs.size() * 4, hipMemcpyDeviceToHost, s));
            SFD_HIP(ctx, hipStreamSynchronize(s));
            for (int r = 0; r < 2 * xw; ++r) {
                const int i = r < xw ? r : nx - xw + (r - xw);
                std::memcpy(xf[q] + (size_t)i * nz, &xs[(size_t)r * nz], nz * 4);
            }
        }
        Field* zb[2] = {&pl->psi_pz[pl->cur], &pl->psi_vz[pl->cur]};
        float* zf[2] = {pml + n, pml + 3 * n};
        for (int q = 0; q < 2; ++q) {
            SFD_HIP(ctx, hipMemcpyAsync(zs.data(), zb[q]->ptr, zs.size() * 4, hipMemcpyDeviceToHost, s));
            SFD_HIP(ctx, hipStreamSynchronize(s));
            for (int i = 0; i < nx; ++i)
                for (int c = 0; c < 2 * zw; ++c) {
                    const int jj = c < zw ? c : nz - zw + (c - zw);
                    zf[q][(size_t)i * nz + jj] = zs[(size_t)i * 2 * zw + c];
                }
        }
    }
    return SFD_OK;
}

int sfd_acoustic_plan_destroy(sfd_acoustic_plan* pl) {
    if (!pl) return SFD_EINVAL;
    (void)hipSetDevice(pl->ctx->device);
    (void)hipStreamSynchronize(pl->ctx->stream);
    delete pl;
    return SFD_OK;
}

int sfd_acoustic_forward(sfd_context* ctx, const sfd_acoustic_params* p, const float* vp, const float* den,
                         const float* wavelet, float* record, float* snaps) {
    sfd_acoustic_plan* pl = nullptr;
    int rc = sfd_acoustic_plan_create(ctx, p, vp, den, wavelet, &pl);
    if (rc) return rc;
    rc = sfd_acoustic_plan_run(pl, nullptr);
    if (!rc) rc = sfd_acoustic_plan_fetch(pl, record, snaps);
    sfd_acoustic_plan_destroy(pl);
    return rc;
}

}  // extern "C"

// ------------------------------------------------------------------ batches

struct sfd_acoustic_batch {
    sfd_context* ctx = nullptr;
    std::vector<sfd_acoustic_plan*> plans;
    int nt = 0, radius = 0, mode = 0, nw = 0;
    int ntile = 0;
    DeviceBuffer shot_args, tile_shot, tile_base, tile_map;
    BatchKernel kernel = nullptr;
    // Fast arithmetic: two tiled launches per step, in place, at ttx rows x
    // tnw waves per tile.
    bool tiled = false;
    int ttx = 0, tnw = 0;
    BatchKernel tvel = nullptr, tpres = nullptr;
    // Or the one-launch tiled step (ping-pong): tile rows, waves.
    bool fused = false;
    int ftx = 0, fnw = 0;
    BatchKernel fstep = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // Groups of consecutive shots run one after another, each for all its
    // steps, so that a group's fields stay in the 256 MB MALL across steps:
    // shots [s0, s1), their argument blocks from args_off (parity-major) and
    // tiles from tile_off in the tables.  group_bytes: the budget of one
    // group's resident fields (0 = the default, SIZE_MAX = one group).
    struct Group {
        int s0, s1, args_off, tile_off, ntile;
    };
    std::vector<Group> groups;
    size_t group_bytes = 0;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~sfd_acoustic_batch() {
        drop_graph();
        for (auto* p : plans) sfd_acoustic_plan_destroy(p);
    }
};

namespace {

// Sets a batch's launch shape (kinds as apply_shape).
int apply_batch_shape(sfd_acoustic_batch* b, const TunedShape& t) {
    if (t.kind == 2) return sfd_acoustic_batch_set_fused(b, t.a, t.b);
    return t.kind == 0 ? sfd_acoustic_batch_set_tiled(b, t.a, t.b) : sfd_acoustic_batch_set_tiling(b, t.a, t.b);
}

// Autotuner grid class of a batch: radius, boundary and arithmetic, the shot
// count and the total cells to the nearest power of 2.
std::string batch_class(const sfd_acoustic_batch* b) {
    double cells = 0;
    for (auto* pl : b->plans) cells += (double)pl->nx * pl->nz;
    char buf[96];
    snprintf(buf, sizeof buf, "batch:R%d:%s:n%d:c%d", b->radius, mode_name(b->mode),
             log2_round((double)b->plans.size()), log2_round(cells));
    return buf;
}

// Bytes a shot's step keeps resident when it runs in place: P, Vx, Vz, kf
// (and bf unless the density is uniform) and the C-PML memory fields.
size_t resident_bytes(const sfd_acoustic_plan* pl) {
    size_t n = (size_t)pl->nx * pl->nz * (pl->bf_const != 0.f ? 16 : 20);
    for (const Field* f : {&pl->psi_px[0], &pl->psi_vx[0], &pl->psi_pz[0], &pl->psi_vz[0]}) n += f->bytes;
    return n;
}

// Default budget of one group's resident fields: measured on MI355X (256 MB
// MALL) with the 64 windowed shots of 1101 x 1100-2600 cells.
constexpr size_t kBatchGroupBytes = (size_t)192 << 20;

// Tile tables and per-shot argument blocks (both ping-pong orders; the
// two-launch tiled step runs in place and needs one), group by group.
int build_batch_tables(sfd_acoustic_batch* b) {
    const int n = (int)b->plans.size();
    const size_t budget = b->group_bytes ? b->group_bytes : kBatchGroupBytes;
    b->groups.clear();
    for (int s = 0; s < n;) {
        size_t used = 0;
        int e = s;
        do used += resident_bytes(b->plans[e++]);
        while (e < n && used + resident_bytes(b->plans[e]) <= budget);
        b->groups.push_back({s, e, 0, 0, 0});
        s = e;
    }
    std::vector<AcousticStepArgs> args(2 * (size_t)n);
    std::vector<int> base(n), tshot;
    std::vector<int2> tmap;
    int t = 0;
    for (auto& g : b->groups) {
        const int ng = g.s1 - g.s0;
        g.args_off = 2 * g.s0;
        g.tile_off = t;
        int tl = 0;
        for (int s = g.s0; s < g.s1; ++s) {
            sfd_acoustic_plan* pl = b->plans[s];
            const int ls = s - g.s0;
            AcousticStepArgs* ga = args.data() + g.args_off;
            for (int par = 0; par < 2; ++par) ga[(size_t)par * ng + ls] = make_args(pl, par, 0);
            // Batches start from zero fields, so the tiled sponge steps may
            // read their velocities as unabsorbed from the first step on.
            if (b->tiled) tiled_args(pl, ga[ls], b->mode == MODE_FAST_SPONGE);
            if (b->fused)
                for (int par = 0; par < 2; ++par) fused_args(pl, ga[(size_t)par * ng + ls], b->mode == MODE_FAST_SPONGE);
            const int nt = ga[ls].ntz * ga[ls].nchunk;
            base[s] = tl;  // tile_base and tile_shot hold group-local values
            tshot.insert(tshot.end(), nt, ls);
            for (int q = 0; q < nt; ++q) tmap.push_back(make_int2(ls, q));
            tl += nt;
        }
        g.ntile = tl;
        t += tl;
    }
    b->ntile = t;
    hipStream_t st = b->ctx->stream;
    hipError_t e = upload(b->shot_args, args.data(), args.size(), st);
    if (e == hipSuccess) e = upload(b->tile_shot, tshot.data(), tshot.size(), st);
    if (e == hipSuccess) e = upload(b->tile_base, base.data(), base.size(), st);
    if (e == hipSuccess) e = upload(b->tile_map, tmap.data(), tmap.size(), st);
    if (e != hipSuccess) return hip_check(b->ctx, e, "batch tables");
    b->drop_graph();
    return SFD_OK;
}

// The first nsteps steps of every shot, group after group.
int enqueue_batch_steps(sfd_acoustic_batch* b, int nsteps) {
    for (const auto& g : b->groups) {
        BatchStepArgs a{};
        a.shots = b->shot_args.as<AcousticStepArgs>() + g.args_off;
        a.tile_shot = b->tile_shot.as<int>() + g.tile_off;
        a.tile_base = b->tile_base.as<int>() + g.s0;
        a.tile_map = b->tile_map.as<int2>() + g.tile_off;
        a.nshot = g.s1 - g.s0;
        a.ntile = g.ntile;
        for (int k = 0; k < nsteps; ++k) {
            a.k = k;
            a.parity = k & 1;
            if (b->tiled) {
                const dim3 block(64 * b->tnw);
                hipLaunchKernelGGL(b->tvel, dim3(g.ntile), block, 0, b->ctx->stream, a);
                SFD_HIP(b->ctx, hipGetLastError());
                hipLaunchKernelGGL(b->tpres, dim3(g.ntile), block, 0, b->ctx->stream, a);
            } else if (b->fused) {
                hipLaunchKernelGGL(b->fstep, dim3(g.ntile), dim3(64 * b->fnw), 0, b->ctx->stream, a);
            } else {
                hipLaunchKernelGGL(b->kernel, dim3(g.ntile), dim3(64 * b->nw), 0, b->ctx->stream, a);
            }
            SFD_HIP(b->ctx, hipGetLastError());
        }
    }
    return SFD_OK;
}

}  // namespace

extern "C" {

int sfd_acoustic_batch_create(sfd_context* ctx, int nshot, const sfd_acoustic_params* params,
                              const float* const* vp, const float* const* den, const float* const* wavelet,
                              sfd_acoustic_batch** out) {
    if (!ctx || !out) return SFD_EINVAL;
    *out = nullptr;
    if (nshot <= 0 || !params || !vp || !wavelet) return fail(ctx, SFD_EINVAL, "empty batch");
    for (int s = 0; s < nshot; ++s) {
        if (params[s].nt != params[0].nt || params[s].radius != params[0].radius ||
            params[s].boundary != params[0].boundary || params[s].arith != params[0].arith)
            return fail(ctx, SFD_EINVAL, "shot %d: every shot of a batch needs the same nt, radius, boundary kind "
                                         "and arithmetic", s);
        if (params[s].nsnap != 0) return fail(ctx, SFD_ENOTSUP, "shot %d: batches record gathers only", s);
    }
    sfd_acoustic_batch* b = new (std::nothrow) sfd_acoustic_batch;
    if (!b) return fail(ctx, SFD_ENOMEM, "host allocation failed");
    b->ctx = ctx;
    for (int s = 0; s < nshot; ++s) {
        sfd_acoustic_plan* pl = nullptr;
        const int rc = sfd_acoustic_plan_create(ctx, &params[s], vp[s], den ? den[s] : nullptr, wavelet[s], &pl);
        if (rc) {
            delete b;
            return rc;
        }
        b->plans.push_back(pl);
    }
    b->nt = params[0].nt;
    b->radius = params[0].radius;
    b->mode = b->plans[0]->mode;
    const int rc = sfd_acoustic_batch_set_tiling(b, 0, 0);
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return SFD_OK;
}

int sfd_acoustic_batch_set_tiled(sfd_acoustic_batch* b, int tile_rows, int waves) {
    if (!b) return SFD_EINVAL;
    if (tile_rows < 0) return sfd_acoustic_batch_set_tiling(b, 2, 0);
    if (!tiled_mode_ok(b->mode)) return fail(b->ctx, SFD_ENOTSUP, "the tiled step exists in fast arithmetic only");
    const int tx = tile_rows ? tile_rows : kTiledTX, nw = waves ? waves : kTiledNW;
    BatchKernel v, p;
    if (!pick_tiled_batch(b->radius, tx, nw, b->mode, &v, &p))
        return fail(b->ctx, SFD_EINVAL, "no batched tiled step for %d rows x %d waves", tx, nw);
    for (auto* pl : b->plans) {
        const int rc = sfd_acoustic_plan_set_tiled(pl, tx, nw);
        if (rc) return rc;
    }
    b->tiled = true;
    b->fused = false;
    b->ttx = tx;
    b->tnw = nw;
    b->tvel = v;
    b->tpres = p;
    return build_batch_tables(b);
}

int sfd_acoustic_batch_set_fused(sfd_acoustic_batch* b, int rows, int waves) {
    if (!b) return SFD_EINVAL;
    if (!tiled_mode_ok(b->mode))
        return fail(b->ctx, SFD_ENOTSUP, "the one-launch step exists in fast arithmetic only");
    const int tr = rows ? rows : kFusedShapes[0][0], nw = waves ? waves : kFusedShapes[0][1];
    BatchKernel k;
    bool ubf = true;  // every shot has uniform density
    for (auto* pl : b->plans) ubf = ubf && pl->bf_const != 0.f;
    if (!pick_fused(b->radius, tr, nw, b->mode, ubf, nullptr, &k))
        return fail(b->ctx, SFD_EINVAL, "no batched one-launch step for %d rows x %d waves", tr, nw);
    for (auto* pl : b->plans) {
        const int rc = sfd_acoustic_plan_set_fused(pl, tr, nw);
        if (rc) return rc;
    }
    b->tiled = false;
    b->fused = true;
    b->ftx = tr;
    b->fnw = nw;
    b->fstep = k;
    return build_batch_tables(b);
}

int sfd_acoustic_batch_set_group_bytes(sfd_acoustic_batch* b, size_t bytes) {
    if (!b) return SFD_EINVAL;
    b->group_bytes = bytes;
    return build_batch_tables(b);
}

int sfd_acoustic_batch_get_groups(const sfd_acoustic_batch* b, int* ngroups) {
    if (!b || !ngroups) return SFD_EINVAL;
    *ngroups = (int)b->groups.size();
    return SFD_OK;
}

int sfd_acoustic_batch_set_tiling(sfd_acoustic_batch* b, int waves, int chunk_rows) {
    if (!b) return SFD_EINVAL;
    // Automatic: the shape the context's autotuner cache holds for this
    // batch's class, else the tiled step's default shape (fast arithmetic)
    // or the march kernel's default tiling.
    if (waves == 0 && chunk_rows == 0) {
        const auto it = b->ctx->tuning.find(batch_class(b));
        if (it != b->ctx->tuning.end()) {
            const int rc = apply_batch_shape(b, it->second);
            if (rc == SFD_OK) return rc;
        }
        if (tiled_mode_ok(b->mode)) return sfd_acoustic_batch_set_tiled(b, kTiledTX, kTiledNW);
    }
    b->tiled = false;
    b->fused = false;
    // Batches hold far more cells than one shot, so the default chunk is
    // longer (less queue fill per cell) while the grid still fills the GPU.
    const int nw = waves ? waves : 2;
    long cells = 0;
    for (auto* pl : b->plans) cells += (long)pl->nx * pl->nz;
    const int ch = chunk_rows ? chunk_rows : (cells > 64L * 1000 * 1000 ? 128 : 0);
    for (auto* pl : b->plans) {
        const int rc = sfd_acoustic_plan_set_tiling(pl, nw, ch);
        if (rc) return rc;
    }
    b->nw = nw;
    b->kernel = pick_batch_kernel(b->radius, nw, b->mode);
    if (!b->kernel) return fail(b->ctx, SFD_EINVAL, "waves must be 1, 2 or 4 (got %d)", waves);
    return build_batch_tables(b);
}

int sfd_acoustic_batch_run(sfd_acoustic_batch* b, int how, float* elapsed_ms) {
    if (!b) return SFD_EINVAL;
    sfd_context* ctx = b->ctx;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    if (how != SFD_BATCH_GRAPH && how != SFD_BATCH_LAUNCHES && how != SFD_BATCH_PER_SHOT)
        return fail(ctx, SFD_EINVAL, "unknown batch run mode %d", how);
    for (auto* pl : b->plans) {
        const int rc = zero_state(pl);  // outside any graph: no memset nodes
        if (rc) return rc;
    }
    if (how == SFD_BATCH_GRAPH && !b->exec) {
        SFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
        SFD_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_batch_steps(b, b->nt);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        SFD_HIP(ctx, e);
        b->graph = g;
        SFD_HIP(ctx, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    }
    SFD_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (how == SFD_BATCH_GRAPH) {
        SFD_HIP(ctx, hipGraphLaunch(b->exec, ctx->stream));
    } else if (how == SFD_BATCH_LAUNCHES) {
        const int rc = enqueue_batch_steps(b, b->nt);
        if (rc) return rc;
    } else {
        // The plain path: every shot at the shape a plan of its own starts
        // with (the context's cached shape for its grid class, else the
        // tiled default), not the batch's shape; the batch's shape is put
        // back afterwards (host fields only, so nothing is timed for it).
        for (auto* pl : b->plans) {
            const PlanShape keep = save_shape(pl);
            int rc = apply_plain_shape(pl);
            if (!rc) rc = enqueue_steps(pl, 0, pl->nt);
            restore_shape(pl, keep);
            if (rc) return rc;
        }
    }
    SFD_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    SFD_HIP(ctx, hipEventSynchronize(ctx->ev1));
    if (elapsed_ms) SFD_HIP(ctx, hipEventElapsedTime(elapsed_ms, ctx->ev0, ctx->ev1));
    if (how != SFD_BATCH_PER_SHOT) {
        for (auto* pl : b->plans) {
            if (b->tiled || b->fused) pl->v_raw = b->nt > 0 && b->mode == MODE_FAST_SPONGE;
            if (!b->tiled) pl->cur = b->nt & 1;
        }
    }
    for (auto* pl : b->plans) {
        const int rc = check_plan_finite(pl);
        if (rc) return rc;
    }
    return SFD_OK;
}

int sfd_acoustic_batch_fetch(sfd_acoustic_batch* b, int shot, float* record) {
    if (!b) return SFD_EINVAL;
    if (shot < 0 || shot >= (int)b->plans.size()) return fail(b->ctx, SFD_EINVAL, "shot %d out of range", shot);
    return sfd_acoustic_plan_fetch(b->plans[shot], record, nullptr);
}

int sfd_acoustic_batch_destroy(sfd_acoustic_batch* b) {
    if (!b) return SFD_EINVAL;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
    return SFD_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ autotuner

namespace {

// Candidates: every built tiled shape, the one-launch tiled shapes and
// streaming chunks (fast arithmetic) and the march kernel at 1, 2 and 4 waves
// with its automatic chunk.
std::vector<TunedShape> shape_candidates(int mode, bool batch) {
    std::vector<TunedShape> c;
    if (tiled_mode_ok(mode)) {
        static const int single[][2] = {{8, 4}, {8, 8}, {16, 4}, {16, 8}, {16, 16}, {32, 4}, {32, 8}, {32, 16}};
        static const int batched[][2] = {{8, 4}, {16, 4}, {16, 8}, {32, 4}, {32, 8}};
        if (batch)
            for (auto& t : batched) c.push_back({0, t[0], t[1], 0.f});
        else
            for (auto& t : single) c.push_back({0, t[0], t[1], 0.f});
        for (auto& t : kFusedShapes) c.push_back({2, t[0], t[1], 0.f});
        for (auto& t : kStreamShapes) c.push_back({2, t[0], t[1], 0.f});
    }
    for (int w : {1, 2, 4}) c.push_back({1, w, 0, 0.f});
    return c;
}

}  // namespace

extern "C" {

int sfd_acoustic_plan_autotune(sfd_acoustic_plan* pl, int trial_steps, float* best_us) {
    if (!pl) return SFD_EINVAL;
    sfd_context* ctx = pl->ctx;
    if (trial_steps < 1) return fail(ctx, SFD_EINVAL, "trial_steps must be >= 1");
    const int n = std::min(trial_steps, pl->nt);
    if (n < 1) return fail(ctx, SFD_EINVAL, "the plan has no time steps to time");
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    auto run = [&](int m) { return m < 0 ? zero_state(pl) : enqueue_steps(pl, 0, m); };
    // The GPU first runs a few trials' worth of steps (clocks settle), then
    // every candidate is timed twice, in two passes, keeping its better time.
    float warm = 0.f;
    int rc = time_steps(ctx, std::min(pl->nt, 4 * n), run, &warm);
    if (rc) return rc;
    std::vector<TunedShape> cand = shape_candidates(pl->mode, false);
    for (auto& t : cand) t.us = 3.4e38f;
    for (int pass = 0; pass < 2; ++pass)
        for (TunedShape& t : cand) {
            rc = apply_shape(pl, t);
            if (rc) return rc;
            if (t.kind == 1) t.b = pl->chunk;
            float us = 0.f;
            rc = time_steps(ctx, n, run, &us);
            if (rc) return rc;
            t.us = std::min(t.us, us);
        }
    TunedShape best = cand[0];
    for (const TunedShape& t : cand)
        if (t.us < best.us) best = t;
    rc = apply_shape(pl, best);
    if (!rc) rc = zero_state(pl);
    if (rc) return rc;
    SFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->tuning[plan_class(pl)] = best;
    if (best_us) *best_us = best.us;
    return SFD_OK;
}

int sfd_acoustic_batch_autotune(sfd_acoustic_batch* b, int trial_steps, float* best_us) {
    if (!b) return SFD_EINVAL;
    sfd_context* ctx = b->ctx;
    if (trial_steps < 1) return fail(ctx, SFD_EINVAL, "trial_steps must be >= 1");
    const int n = std::min(trial_steps, b->nt);
    if (n < 1) return fail(ctx, SFD_EINVAL, "the batch has no time steps to time");
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    auto run = [&](int m) -> int {
        if (m >= 0) return enqueue_batch_steps(b, m);
        for (auto* pl : b->plans) {
            const int rc = zero_state(pl);
            if (rc) return rc;
        }
        return SFD_OK;
    };
    float warm = 0.f;
    int rc = time_steps(ctx, std::min(b->nt, 4 * n), run, &warm);
    if (rc) return rc;
    std::vector<TunedShape> cand = shape_candidates(b->mode, true);
    for (auto& t : cand) t.us = 3.4e38f;
    for (int pass = 0; pass < 2; ++pass)
        for (TunedShape& t : cand) {
            rc = apply_batch_shape(b, t);
            if (rc) return rc;
            if (t.kind == 1) t.b = b->plans[0]->chunk;
            float us = 0.f;
            rc = time_steps(ctx, n, run, &us);
            if (rc) return rc;
            t.us = std::min(t.us, us);
        }
    TunedShape best = cand[0];
    for (const TunedShape& t : cand)
        if (t.us < best.us) best = t;
    rc = apply_batch_shape(b, best);
    if (!rc) rc = run(-1);
    if (rc) return rc;
    SFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->tuning[batch_class(b)] = best;
    if (best_us) *best_us = best.us;
    return SFD_OK;
}

int sfd_acoustic_plan_get_shape(const sfd_acoustic_plan* pl, int* kind, int* a, int* b) {
    if (!pl || !kind || !a || !b) return SFD_EINVAL;
    *kind = pl->ttx > 0 ? 0 : pl->ftx > 0 ? 2 : 1;
    *a = pl->ttx > 0 ? pl->ttx : pl->ftx > 0 ? pl->ftx : pl->nw;
    *b = pl->ttx > 0 ? pl->tnw : pl->ftx > 0 ? pl->fnw : pl->chunk;
    return SFD_OK;
}

int sfd_acoustic_batch_get_shape(const sfd_acoustic_batch* bt, int* kind, int* a, int* b) {
    if (!bt || !kind || !a || !b) return SFD_EINVAL;
    *kind = bt->tiled ? 0 : bt->fused ? 2 : 1;
    *a = bt->tiled ? bt->ttx : bt->fused ? bt->ftx : bt->nw;
    *b = bt->tiled ? bt->tnw : bt->fused ? bt->fnw : bt->plans[0]->chunk;
    return SFD_OK;
}

}  // extern "C"
