// Velocity gradient of the acoustic propagator by the discrete adjoint
// (no counterpart in the reference, which has no adjoint or inversion code).
//
// For J = sum_{r,k} seed[r][k] * record[r][k] (the L2 misfit is the case
// seed = record - observed) the adjoint of the fast-arithmetic sponge step is
// the same staggered step applied to the scaled adjoint fields
// (K lambda_P, -b lambda_V), K = rho vp^2 dt, b = dt / rho.  Per shot:
//   1. forward over all nt steps, saving the wavefield every `period` steps;
//   2. seed (residual) per receiver cell, time-major, on the device;
//   3. backward, segment by segment from the last: restore the segment's
//      checkpoint, recompute its steps storing div_k = D-Vx' + D-Vz', then run
//      the adjoint steps k = end-1 .. start, each accumulating
//      g += P_adj_in * div_k and followed by the injection of seed_{k-1};
//   4. dJ/dvp = -2 g / vp, plus the source term -sum_k adj_src_k w_k / vp_s^2
//      (the injected amplitude dt rho_s vp_s w_k depends on vp at the source).
// Memory: one segment of divergences plus nt/period checkpoints per shot.
// The forward's last segment stores its own divergences, so step 3
// recomputes every segment but the last; with period >= nt (nt divergence
// fields per shot) it recomputes nothing.
// Every step is the two-launch tiled step (acoustic_tiled.h) in place on the
// first buffer of each field; the velocities, checkpoints included, keep the
// tiled sponge step's unabsorbed values.
//
// Boundary saving (SFD_GRADIENT_BOUNDARY_SAVING, shot groups): instead of
// recomputing each segment forward and keeping its divergences, the forward
// saves after every step the P, Vx, Vz values of the R-deep band just outside
// the box [max(w, R), n - max(w, R)) on both axes (w = sponge width), where
// the sponge factor is 1.  The backward sweep then runs the forward step in
// reverse inside the box (MODE_REV: P_k = P_k+1 + kf div_k, V_k = V_k+1 +
// bf D(P_k - source), the band restored from the saves before each step, the
// state reset from the checkpoint at every segment end), so one divergence
// field per shot suffices.  The sponge is dissipative and cannot be run
// backwards, so the gradient exists inside the box only (zero in the sponge).
#include <algorithm>
#include <cmath>
#include <new>

#include "acoustic_plan.h"

using namespace sfd;

namespace sfd {

// P[cell[s]] += w[s] * res[k * nslot + s]: the seed of step k at the
// receiver cells, absorbed (w = sponge factor * K of the cell).
__global__ void adjoint_inject_kernel(float* __restrict__ p, const int* __restrict__ cell,
                                      const float* __restrict__ w, const float* __restrict__ res, int nslot,
                                      int k) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nslot) p[cell[s]] += w[s] * res[(size_t)k * nslot + s];
}

namespace {

struct GradShot {
    sfd_acoustic_plan* pl = nullptr;
    int nseg = 0;
    std::vector<int> slot_cell;
};

int check_gradient_params(sfd_context* ctx, const sfd_acoustic_params* p) {
    if ((p->boundary != SFD_BOUNDARY_SPONGE && p->boundary != SFD_BOUNDARY_CPML) || p->arith != SFD_ARITH_FAST)
        return fail(ctx, SFD_ENOTSUP, "the gradient exists for the sponge and C-PML boundaries in fast arithmetic");
    if (p->nt < 1) return fail(ctx, SFD_EINVAL, "nt < 1");
    const int R = p->radius;
    if (p->sx < R || p->sx >= p->nx - R || p->sz < R || p->sz >= p->nz - R)
        return fail(ctx, SFD_EINVAL, "the gradient needs the source inside the updated cells");
    return SFD_OK;
}

bool is_cpml(const sfd_acoustic_plan* pl) { return pl->mode == MODE_FAST_CPML; }

// The C-PML memory fields of a plan's state (buffer 0): x strips, z strips.
int psi_fields(sfd_acoustic_plan* pl, Field* f[4], int buf = 0) {
    if (!is_cpml(pl)) return 0;
    f[0] = &pl->psi_px[buf];
    f[1] = &pl->psi_vx[buf];
    f[2] = &pl->psi_pz[buf];
    f[3] = &pl->psi_vz[buf];
    return 4;
}

// Floats of one checkpoint: P, Vx, Vz and, with C-PML, the memory fields.
size_t ckpt_floats(sfd_acoustic_plan* pl) {
    size_t n = 3 * (size_t)pl->nx * pl->nz;
    Field* f[4];
    for (int q = 0, m = psi_fields(pl, f); q < m; ++q) n += f[q]->bytes / 4;
    return n;
}

// Copies the state (buffer 0) to or from checkpoint `seg` (1-based).
int ckpt_copy(sfd_acoustic_plan* pl, int seg, bool save) {
    sfd_context* ctx = pl->ctx;
    const size_t n = (size_t)pl->nx * pl->nz;
    float* ck = pl->ckpt.as<float>() + (size_t)(seg - 1) * ckpt_floats(pl);
    Field* f[3] = {&pl->p[0], &pl->vx[0], &pl->vz[0]};
    for (int q = 0; q < 3; ++q) {
        float* dst = save ? ck + q * n : pl->grid(*f[q]);
        const float* src = save ? pl->grid(*f[q]) : ck + q * n;
        SFD_HIP(ctx, hipMemcpyAsync(dst, src, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
    }
    ck += 3 * n;
    Field* m[4];
    for (int q = 0, c = psi_fields(pl, m); q < c; ++q) {
        if (save)
            SFD_HIP(ctx, hipMemcpyAsync(ck, m[q]->ptr, m[q]->bytes, hipMemcpyDeviceToDevice, ctx->stream));
        else
            SFD_HIP(ctx, hipMemcpyAsync(m[q]->ptr, ck, m[q]->bytes, hipMemcpyDeviceToDevice, ctx->stream));
        ck += m[q]->bytes / 4;
    }
    return SFD_OK;
}

// Zeroes the adjoint state: P, Vx, Vz and (C-PML) the adjoint strip fields,
// which live in the second copies of the plan's strip fields.
int zero_adjoint(sfd_acoustic_plan* pl) {
    sfd_context* ctx = pl->ctx;
    hipStream_t s = ctx->stream;
    for (Field* f : {&pl->ap[0], &pl->avx[0], &pl->avz[0]}) SFD_HIP(ctx, hipMemsetAsync(f->ptr, 0, f->bytes, s));
    Field* m[4];
    for (int q = 0, c = psi_fields(pl, m, 1); q < c; ++q) SFD_HIP(ctx, hipMemsetAsync(m[q]->ptr, 0, m[q]->bytes, s));
    return SFD_OK;
}

// Arguments of an adjoint step (in place on the adjoint fields).
AcousticStepArgs adjoint_args(const sfd_acoustic_plan* pl, int k) {
    AcousticStepArgs a = make_args(pl, pl->ap, pl->avx, pl->avz, 0, k);
    tiled_args(pl, a, true);
    if (pl->mode == MODE_FAST_CPML) {
        a.psi_px_in = a.psi_px_out = pl->psi_px[1].as<float>();
        a.psi_vx_in = a.psi_vx_out = pl->psi_vx[1].as<float>();
        a.psi_pz_in = a.psi_pz_out = pl->psi_pz[1].as<float>();
        a.psi_vz_in = a.psi_vz_out = pl->psi_vz[1].as<float>();
    }
    return a;
}

// Creates the plan with its gradient workspace and the receiver-cell tables.
int grad_setup(sfd_context* ctx, const sfd_acoustic_params* p, const float* vp, const float* den,
               const float* wavelet, int period, GradShot* g, int method = SFD_GRADIENT_RECOMPUTE,
               PlanModelCache* mc = nullptr) {
    sfd_acoustic_params q = *p;
    q.nsnap = 0;
    q.snap_steps = nullptr;
    const bool bsave = method == SFD_GRADIENT_BOUNDARY_SAVING;
    int rc = plan_create(ctx, &q, vp, den, wavelet, period, &g->pl, bsave ? 1 : 0, mc);
    if (rc) return rc;
    sfd_acoustic_plan* pl = g->pl;
    // Forward, recomputation and adjoint all run the tiled step at the
    // default shape (the velocity fields keep the tiled step's unabsorbed
    // values, checkpoints included).
    rc = sfd_acoustic_plan_set_tiled(pl, kTiledTX, kTiledNW);
    if (rc) return rc;
    g->nseg = (pl->nt + period - 1) / period;
    // slot -> cell (row-major over the padded state field) and weight.
    std::vector<int> cell(std::max(pl->nslot, 1), 0);
    std::vector<float> w(std::max(pl->nslot, 1), 0.f);
    std::vector<int> slot_of(p->nrec);
    {
        // Slots were assigned to distinct cells in (x, z) order.
        std::vector<std::pair<long, int>> cells(p->nrec);
        for (int r = 0; r < p->nrec; ++r) cells[r] = {(long)p->rx[r] * pl->nz + p->rz[r], r};
        std::sort(cells.begin(), cells.end());
        int slot = -1;
        long prev = -1;
        for (auto& [c, r] : cells) {
            if (c != prev) ++slot;
            prev = c;
            slot_of[r] = slot;
            const int i = (int)(c / pl->nz), j = (int)(c % pl->nz);
            const bool upd = i >= p->radius && i < pl->nx - p->radius && j >= p->radius && j < pl->nz - p->radius;
            const float d = den ? den[c] : 2400.0f;
            cell[slot] = (int)((long)(i + pl->pad) * pl->nz + j);
            // The sponge adjoint runs on K-scaled fields (K = rho vp^2 dt,
            // times the cell's sponge factor); the C-PML adjoint unscaled.
            if (!upd) w[slot] = 0.f;
            else w[slot] = is_cpml(pl) ? 1.f : sponge_at(pl, i, j) * (d * vp[c] * vp[c] * p->dt);
        }
    }
    g->slot_cell = slot_of;
    hipStream_t s = ctx->stream;
    hipError_t e = upload(pl->inj_cell, cell.data(), cell.size(), s);
    if (e == hipSuccess) e = upload(pl->inj_w, w.data(), w.size(), s);
    if (e == hipSuccess) e = pl->res.alloc((size_t)std::max(pl->nslot, 1) * pl->nt * 4);
    if (e == hipSuccess) e = pl->src_adj.alloc((size_t)pl->nt * 4);
    if (e == hipSuccess) e = pl->ckpt.alloc((size_t)std::max(g->nseg - 1, 1) * ckpt_floats(pl) * 4);
    if (e == hipSuccess && bsave) {
        // Band cells: within R of the box along x (same columns) or along z
        // (same rows), the reach of the stencils of the box's cells.
        const int b = rev_margin(pl), R = pl->radius;
        const int x0 = b, x1 = pl->nx - b, z0 = b, z1 = pl->nz - b;
        if (x0 >= x1 || z0 >= z1) return fail(ctx, SFD_EINVAL, "grid too small for boundary saving");
        std::vector<int> band;
        for (int i = x0 - R; i < x1 + R; ++i)
            for (int j = z0 - R; j < z1 + R; ++j) {
                const bool xin = i >= x0 && i < x1, zin = j >= z0 && j < z1;
                if (xin != zin) band.push_back(i * pl->nz + j);
            }
        pl->nband = (int)band.size();
        e = upload(pl->band_cell, band.data(), band.size(), s);
        if (e == hipSuccess) e = pl->bands.alloc((size_t)pl->nt * 3 * band.size() * 4);
    }
    if (e != hipSuccess) return hip_check(ctx, e, "gradient workspace");
    return SFD_OK;
}

// Forward over all steps, saving the state before every segment but the first.
// The last segment runs the recomputation's kernels, which also store each of
// its steps' divergences (the same state and records bit for bit), so the
// backward sweep starts with that segment's adjoint and recomputes only the
// others; with one segment (checkpoint interval >= nt) it recomputes nothing.
int grad_forward(GradShot* g) {
    sfd_acoustic_plan* pl = g->pl;
    int rc = zero_state(pl);
    if (rc) return rc;
    const int T = pl->grad_period;
    const bool cpml = is_cpml(pl);
    StepKernel dv, dp;
    if (!pick_tiled(pl->radius, pl->ttx, pl->tnw, cpml ? MODE_FWD_DIV_CPML : MODE_FWD_DIV, &dv, &dp))
        return fail(pl->ctx, SFD_ENOTSUP, "no gradient kernel for this radius / tiling");
    for (int seg = 0; seg < g->nseg; ++seg) {
        const int k0 = seg * T, n = std::min(T, pl->nt - k0);
        if (seg > 0) {
            rc = ckpt_copy(pl, seg, true);  // the tiled step runs in place: the state is in buffer 0
            if (rc) return rc;
        }
        if (seg < g->nseg - 1) {
            rc = enqueue_steps(pl, k0, n);
            if (rc) return rc;
            continue;
        }
        for (int k = k0; k < k0 + n; ++k) {
            AcousticStepArgs a = make_args(pl, 0, k);
            tiled_args(pl, a, true);
            rc = launch_tiled(pl, dv, dp, a);
            if (rc) return rc;
        }
        pl->v_raw = !cpml;
    }
    return SFD_OK;
}

// Time-major per-slot seed from the trace-major per-receiver seed.
int grad_upload_seed(GradShot* g, const float* seed) {
    sfd_acoustic_plan* pl = g->pl;
    const int nt = pl->nt, ns = std::max(pl->nslot, 1);
    std::vector<float> res((size_t)ns * nt, 0.f);
    for (int r = 0; r < (int)g->slot_cell.size(); ++r)
        for (int k = 0; k < nt; ++k) res[(size_t)k * ns + g->slot_cell[r]] += seed[(size_t)r * nt + k];
    SFD_HIP(pl->ctx, hipMemcpyAsync(pl->res.ptr, res.data(), res.size() * 4, hipMemcpyHostToDevice,
                                    pl->ctx->stream));
    SFD_HIP(pl->ctx, hipStreamSynchronize(pl->ctx->stream));
    return SFD_OK;
}

int launch_inject(sfd_acoustic_plan* pl, float* p_grid, int k) {
    if (pl->nslot <= 0) return SFD_OK;
    const int bs = 256;
    hipLaunchKernelGGL(adjoint_inject_kernel, dim3((pl->nslot + bs - 1) / bs), dim3(bs), 0, pl->ctx->stream,
                       p_grid - (size_t)pl->pad * pl->nz, pl->inj_cell.as<int>(), pl->inj_w.as<float>(),
                       pl->res.as<float>(), pl->nslot, k);
    SFD_HIP(pl->ctx, hipGetLastError());
    return SFD_OK;
}

int grad_backward(GradShot* g) {
    sfd_acoustic_plan* pl = g->pl;
    sfd_context* ctx = pl->ctx;
    hipStream_t s = ctx->stream;
    const int T = pl->grad_period, nt = pl->nt;
    int rc = zero_adjoint(pl);
    if (rc) return rc;
    SFD_HIP(ctx, hipMemsetAsync(pl->grad.ptr, 0, pl->grad.bytes, s));
    SFD_HIP(ctx, hipMemsetAsync(pl->src_adj.ptr, 0, pl->src_adj.bytes, s));
    StepKernel dv, dp, av, ap;
    const bool cpml = is_cpml(pl);
    if (!pick_tiled(pl->radius, pl->ttx, pl->tnw, cpml ? MODE_FWD_DIV_CPML : MODE_FWD_DIV, &dv, &dp) ||
        !pick_tiled(pl->radius, pl->ttx, pl->tnw, cpml ? MODE_ADJ_CPML : MODE_ADJ, &av, &ap))
        return fail(ctx, SFD_ENOTSUP, "no gradient kernel for this radius / tiling");
    // Every step runs in place on buffer 0 of the state and adjoint fields.
    rc = launch_inject(pl, pl->grid(pl->ap[0]), nt - 1);
    if (rc) return rc;
    for (int seg = g->nseg - 1; seg >= 0; --seg) {
        const int k0 = seg * T, k1 = std::min(k0 + T, nt);
        // Restore the forward state of step k0 and recompute the divergences
        // (the last segment's: the forward stored them).
        pl->cur = 0;
        if (seg < g->nseg - 1) {
            if (seg == 0) {
                for (Field* f : {&pl->p[0], &pl->vx[0], &pl->vz[0]})
                    SFD_HIP(ctx, hipMemsetAsync(f->ptr, 0, f->bytes, s));
                Field* m[4];
                for (int q = 0, c = psi_fields(pl, m); q < c; ++q)
                    SFD_HIP(ctx, hipMemsetAsync(m[q]->ptr, 0, m[q]->bytes, s));
            } else {
                rc = ckpt_copy(pl, seg, false);
                if (rc) return rc;
            }
            for (int k = k0; k < k1; ++k) {
                AcousticStepArgs a = make_args(pl, 0, k);
                tiled_args(pl, a, true);
                rc = launch_tiled(pl, dv, dp, a);
                if (rc) return rc;
            }
        }
        // Adjoint steps of the segment, in reverse.
        for (int k = k1 - 1; k >= k0; --k) {
            const AcousticStepArgs a = adjoint_args(pl, k);
            rc = launch_tiled(pl, av, ap, a);
            if (rc) return rc;
            if (k >= 1) {
                rc = launch_inject(pl, pl->grid(pl->ap[0]), k - 1);
                if (rc) return rc;
            }
        }
    }
    pl->v_raw = !cpml;
    return SFD_OK;
}

// dJ/dvp from the accumulator and the source adjoint; adds into grad_vp.
// The source's share of dJ/dvp: the injected amplitude dt rho_s vp_s w_k
// depends on vp at the source cell.
void grad_source_term(const sfd_acoustic_plan* pl, const std::vector<float>& sa, const float* vp,
                      const float* wavelet, float* grad_vp) {
    const size_t sc = (size_t)pl->sx * pl->nz + pl->sz;
    double t = 0.0;
    for (int k = 0; k < pl->nt; ++k) t += (double)sa[k] * wavelet[k];
    grad_vp[sc] += (float)(-t / ((double)vp[sc] * vp[sc]));
}

int grad_finish(GradShot* g, const float* vp, const float* wavelet, float* grad_vp) {
    sfd_acoustic_plan* pl = g->pl;
    sfd_context* ctx = pl->ctx;
    const size_t n = (size_t)pl->nx * pl->nz;
    std::vector<float> acc(n), sa(pl->nt);
    SFD_HIP(ctx, hipMemcpyAsync(acc.data(), pl->grid(pl->grad), n * 4, hipMemcpyDeviceToHost, ctx->stream));
    SFD_HIP(ctx, hipMemcpyAsync(sa.data(), pl->src_adj.ptr, (size_t)pl->nt * 4, hipMemcpyDeviceToHost, ctx->stream));
    SFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t c = 0; c < n; ++c) grad_vp[c] += -2.0f * acc[c] / vp[c];
    grad_source_term(pl, sa, vp, wavelet, grad_vp);
    return SFD_OK;
}

int grad_one(sfd_context* ctx, const sfd_acoustic_params* p, const float* vp, const float* den, const float* wavelet,
             const float* seed, const float* d_obs, int period, float* grad_vp, float* record, double* misfit) {
    int rc = check_gradient_params(ctx, p);
    if (rc) return rc;
    if (period < 1) return fail(ctx, SFD_EINVAL, "checkpoint interval must be >= 1");
    GradShot g;
    rc = grad_setup(ctx, p, vp, den, wavelet, std::min(period, p->nt), &g);
    if (!rc) rc = grad_forward(&g);
    std::vector<float> rec((size_t)p->nrec * p->nt), sd;
    if (!rc) rc = sfd_acoustic_plan_fetch(g.pl, rec.data(), nullptr);
    if (!rc) {
        if (d_obs) {
            sd.resize(rec.size());
            double m = 0.0;
            for (size_t q = 0; q < rec.size(); ++q) {
                sd[q] = rec[q] - d_obs[q];
                m += 0.5 * (double)sd[q] * sd[q];
            }
            if (misfit) *misfit += m;
            seed = sd.data();
        }
        if (record) std::copy(rec.begin(), rec.end(), record);
        rc = grad_upload_seed(&g, seed);
    }
    if (!rc) rc = grad_backward(&g);
    if (!rc) rc = grad_finish(&g, vp, wavelet, grad_vp);
    if (g.pl) sfd_acoustic_plan_destroy(g.pl);
    return rc;
}

// ------------------------------------------------------------ shot groups
// Shots of one gradient that share a model grid advance together: one tiled
// step (a velocity and a pressure launch) per time step for all of them in
// the forward, the recomputation and the adjoint sweep, and one launch per
// step for all their residual injections.

// One receiver slot of one shot: p[shot][cell] += w * res[shot][k * nslot + slot].
struct InjEntry {
    int shot, cell;
    float w;
    int slot;
};

__global__ void adjoint_inject_group_kernel(float* const* __restrict__ p, const InjEntry* __restrict__ e, int n,
                                            const float* const* __restrict__ res, const int* __restrict__ nslot,
                                            int k) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const InjEntry t = e[q];
    p[t.shot][t.cell] += t.w * res[t.shot][(size_t)k * nslot[t.shot] + t.slot];
}

// Residual of one shot on the device: res[k][slot] = sum over the slot's
// receivers r (ascending) of record(slot, k) - obs[r][k] (obs trace-major,
// the gathers' layout), the seed of the adjoint; partial[block] = that
// block's share of sum 0.5 (record - obs)^2 (a fixed tree, so J does not
// depend on scheduling).  rec_buf is the time-major gather buffer.
__global__ __launch_bounds__(256) void residual_kernel(const float* __restrict__ rec_buf, int nslot, int nt,
                                                       const float* __restrict__ obs,
                                                       const int* __restrict__ slot_start,
                                                       const int* __restrict__ slot_recv, float* __restrict__ res,
                                                       double* __restrict__ partial) {
    __shared__ double red[256];
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (q < (size_t)nslot * nt) {
        const int slot = (int)(q / nt), k = (int)(q - (size_t)slot * nt);
        const float v = rec_buf[(size_t)k * nslot + slot];
        float acc = 0.f;
        for (int p = slot_start[slot]; p < slot_start[slot + 1]; ++p) {
            const float sd = v - obs[(size_t)slot_recv[p] * nt + k];
            acc += sd;
            m += 0.5 * (double)sd * sd;
        }
        res[(size_t)k * nslot + slot] = acc;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// out[c] = sum over the group's shots (in order) of -2 acc_s[c] / vp[c].
__global__ void grad_sum_kernel(float* __restrict__ out, const float* const* __restrict__ acc, int nshot,
                                const float* __restrict__ vp, size_t n) {
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        float g = 0.f;
        for (int s = 0; s < nshot; ++s) g += -2.0f * acc[s][c] / vp[c];
        out[c] = g;
    }
}

// Boundary saving: one band cell of one shot.
struct BandEntry {
    int shot, cell, idx;  // cell: offset from row 0 of the shot's fields; idx: its band slot
};

// The shot's band of step k: band[k][field][idx] = field[cell], fields P, Vx, Vz.
__global__ void band_save_kernel(const BandEntry* __restrict__ e, int n, float* const* __restrict__ fields,
                                 float* const* __restrict__ bands, const int* __restrict__ nband, int k) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const BandEntry t = e[q];
    const size_t nb = (size_t)nband[t.shot];
    float* b = bands[t.shot] + (size_t)k * 3 * nb + t.idx;
    for (int f = 0; f < 3; ++f) b[f * nb] = fields[3 * t.shot + f][t.cell];
}

// Before the reverse step k: the velocities of the band from save kv (the
// state after step k) and P from save kp (the state before step k; kp < 0 =
// the zero initial state).
__global__ void band_restore_kernel(const BandEntry* __restrict__ e, int n, float* const* __restrict__ fields,
                                    float* const* __restrict__ bands, const int* __restrict__ nband, int kv,
                                    int kp) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const BandEntry t = e[q];
    const size_t nb = (size_t)nband[t.shot];
    const float* bv = bands[t.shot] + (size_t)kv * 3 * nb + t.idx;
    fields[3 * t.shot][t.cell] = kp < 0 ? 0.f : bands[t.shot][(size_t)kp * 3 * nb + t.idx];
    fields[3 * t.shot + 1][t.cell] = bv[nb];
    fields[3 * t.shot + 2][t.cell] = bv[2 * nb];
}

struct GradGroup {
    sfd_context* ctx = nullptr;
    std::vector<GradShot> shots;
    int nt = 0, period = 0, nseg = 0, ntile = 0, ninj = 0, nbandall = 0;
    bool bsave = false;
    // Tiled step kernels in launch order: (velocity, pressure) of the
    // forward, the recomputation and the adjoint, (pressure, velocity) of
    // the reverse step; all steps run in place on buffer 0.
    BatchKernel kfwd[2] = {}, kdiv[2] = {}, kadj[2] = {}, krev[2] = {};
    DeviceBuffer fwd_args, adj_args, tile_map, inj, inj_p, inj_res, inj_nslot;
    DeviceBuffer band_e, band_fields, band_ptr, band_n;
    ~GradGroup() {
        for (auto& g : shots)
            if (g.pl) sfd_acoustic_plan_destroy(g.pl);
    }
    int nshot() const { return (int)shots.size(); }
    BatchStepArgs args(const DeviceBuffer& blocks, int k) const {
        BatchStepArgs a{};
        a.shots = blocks.as<AcousticStepArgs>();
        a.tile_map = tile_map.as<int2>();
        a.nshot = nshot();
        a.ntile = ntile;
        a.k = k;
        return a;
    }
};

int group_tables(GradGroup* G) {
    sfd_context* ctx = G->ctx;
    const int n = G->nshot();
    std::vector<AcousticStepArgs> fa(n), aa(n);
    std::vector<float*> pp(n);
    std::vector<const float*> rp(n);
    std::vector<int> ns(n);
    std::vector<InjEntry> inj;
    // The shots share the model grid, vp and den; with equal dt their
    // per-cell coefficients kf and bf are equal too, and every step reads
    // the first shot's copy (one copy in the caches for the whole group).
    bool share = true;
    for (int s = 1; s < n; ++s) share = share && G->shots[s].pl->dt == G->shots[0].pl->dt;
    for (int s = 0; s < n; ++s) {
        sfd_acoustic_plan* pl = G->shots[s].pl;
        fa[s] = make_args(pl, 0, 0);
        tiled_args(pl, fa[s], true);
        aa[s] = adjoint_args(pl, 0);
        if (share && s > 0) {
            fa[s].kf = aa[s].kf = fa[0].kf;
            fa[s].bf = aa[s].bf = fa[0].bf;
        }
        pp[s] = pl->grid(pl->ap[0]) - (size_t)pl->pad * pl->nz;
        rp[s] = pl->res.as<float>();
        ns[s] = std::max(pl->nslot, 1);
        std::vector<int> cell(pl->nslot);
        std::vector<float> w(pl->nslot);
        if (pl->nslot > 0) {
            SFD_HIP(ctx, hipMemcpy(cell.data(), pl->inj_cell.ptr, pl->nslot * sizeof(int), hipMemcpyDeviceToHost));
            SFD_HIP(ctx, hipMemcpy(w.data(), pl->inj_w.ptr, pl->nslot * sizeof(float), hipMemcpyDeviceToHost));
        }
        for (int q = 0; q < pl->nslot; ++q) inj.push_back({s, cell[q], w[q], q});
    }
    // Every shot has the same tiles (same grid); tiles go shot after shot.
    // (Interleaving the shots' tiles, so that each kf / bf tile is read
    // for all shots in turn, measured 2 % slower on MI355X for 16 shots of
    // 881 x 881: it loses more halo reuse between a shot's neighbouring
    // tiles than it gains.)
    const int per = fa[0].ntz * fa[0].nchunk;
    std::vector<int2> tmap((size_t)n * per);
    for (int g = 0; g < n * per; ++g) tmap[g] = make_int2(g / per, g % per);
    G->ntile = n * per;
    G->ninj = (int)inj.size();
    hipStream_t st = ctx->stream;
    hipError_t e = upload(G->fwd_args, fa.data(), fa.size(), st);
    if (e == hipSuccess) e = upload(G->adj_args, aa.data(), aa.size(), st);
    if (e == hipSuccess) e = upload(G->tile_map, tmap.data(), tmap.size(), st);
    if (e == hipSuccess) e = upload(G->inj_p, pp.data(), pp.size(), st);
    if (e == hipSuccess) e = upload(G->inj_res, rp.data(), rp.size(), st);
    if (e == hipSuccess) e = upload(G->inj_nslot, ns.data(), ns.size(), st);
    if (e == hipSuccess && !inj.empty()) e = upload(G->inj, inj.data(), inj.size(), st);
    if (e == hipSuccess && G->bsave) {
        std::vector<BandEntry> be;
        std::vector<float*> fp(3 * (size_t)n), bp(n);
        std::vector<int> nb(n);
        for (int s = 0; s < n; ++s) {
            sfd_acoustic_plan* pl = G->shots[s].pl;
            std::vector<int> cell(pl->nband);
            e = hipMemcpy(cell.data(), pl->band_cell.ptr, cell.size() * sizeof(int), hipMemcpyDeviceToHost);
            if (e != hipSuccess) break;
            for (int q = 0; q < pl->nband; ++q) be.push_back({s, cell[q], q});
            fp[3 * s] = pl->grid(pl->p[0]);
            fp[3 * s + 1] = pl->grid(pl->vx[0]);
            fp[3 * s + 2] = pl->grid(pl->vz[0]);
            bp[s] = pl->bands.as<float>();
            nb[s] = pl->nband;
        }
        G->nbandall = (int)be.size();
        if (e == hipSuccess) e = upload(G->band_e, be.data(), be.size(), st);
        if (e == hipSuccess) e = upload(G->band_fields, fp.data(), fp.size(), st);
        if (e == hipSuccess) e = upload(G->band_ptr, bp.data(), bp.size(), st);
        if (e == hipSuccess) e = upload(G->band_n, nb.data(), nb.size(), st);
    }
    if (e != hipSuccess) return hip_check(ctx, e, "gradient group tables");
    return SFD_OK;
}

int group_band(GradGroup* G, bool save, int k) {
    const int bs = 256, nb = (G->nbandall + bs - 1) / bs;
    if (save)
        hipLaunchKernelGGL(band_save_kernel, dim3(nb), dim3(bs), 0, G->ctx->stream, G->band_e.as<BandEntry>(),
                           G->nbandall, G->band_fields.as<float*>(), G->band_ptr.as<float*>(), G->band_n.as<int>(), k);
    else
        hipLaunchKernelGGL(band_restore_kernel, dim3(nb), dim3(bs), 0, G->ctx->stream, G->band_e.as<BandEntry>(),
                           G->nbandall, G->band_fields.as<float*>(), G->band_ptr.as<float*>(), G->band_n.as<int>(), k,
                           k - 1);
    SFD_HIP(G->ctx, hipGetLastError());
    return SFD_OK;
}

int group_inject(GradGroup* G, int k) {
    if (G->ninj == 0) return SFD_OK;
    const int bs = 256;
    hipLaunchKernelGGL(adjoint_inject_group_kernel, dim3((G->ninj + bs - 1) / bs), dim3(bs), 0, G->ctx->stream,
                       G->inj_p.as<float*>(), G->inj.as<InjEntry>(), G->ninj,
                       G->inj_res.as<const float*>(), G->inj_nslot.as<int>(), k);
    SFD_HIP(G->ctx, hipGetLastError());
    return SFD_OK;
}

int group_launch(GradGroup* G, const BatchKernel kern[2], const DeviceBuffer& blocks, int k) {
    const BatchStepArgs a = G->args(blocks, k);
    for (int q = 0; q < 2; ++q) {
        hipLaunchKernelGGL(kern[q], dim3(G->ntile), dim3(64 * kTiledNW), 0, G->ctx->stream, a);
        SFD_HIP(G->ctx, hipGetLastError());
    }
    return SFD_OK;
}

// Copies the current state of every shot to (or from) its checkpoint `seg`.
int group_checkpoint(GradGroup* G, int seg, bool save) {
    for (auto& g : G->shots) {
        const int rc = ckpt_copy(g.pl, seg, save);
        if (rc) return rc;
    }
    return SFD_OK;
}

// Misfit gradient of up to G->shots.size() shots on one model; adds into
// grad_vp and *misfit.
int group_run(sfd_context* ctx, int nshot, const sfd_acoustic_params* params, const float* vp, const float* den,
              const float* const* wavelet, const float* const* d_obs, bool obs_dev, int period, int method,
              float* grad_vp, double* misfit) {
    GradGroup G;
    G.ctx = ctx;
    G.nt = params[0].nt;
    G.period = std::min(period, G.nt);
    G.bsave = method == SFD_GRADIENT_BOUNDARY_SAVING;
    G.shots.resize(nshot);
    PlanModelCache mc;  // the shots share vp and den: one computation of kf and bf
    for (int s = 0; s < nshot; ++s) {
        int rc = check_gradient_params(ctx, &params[s]);
        if (!rc && params[s].nt != G.nt) rc = fail(ctx, SFD_EINVAL, "shot %d: every shot needs the same nt", s);
        if (!rc && params[s].radius != params[0].radius)
            rc = fail(ctx, SFD_EINVAL, "shot %d: every shot needs the same radius", s);
        if (!rc) rc = grad_setup(ctx, &params[s], vp, den, wavelet[s], G.period, &G.shots[s], method, &mc);
        if (rc) return rc;
    }
    const int R = params[0].radius;
    for (int s = 1; s < nshot; ++s)
        if (params[s].boundary != params[0].boundary)
            return fail(ctx, SFD_EINVAL, "shot %d: every shot of a gradient needs the same boundary kind", s);
    const bool cpml = is_cpml(G.shots[0].pl);
    if (!pick_tiled_batch(R, kTiledTX, kTiledNW, cpml ? MODE_FAST_CPML : MODE_FAST_SPONGE, &G.kfwd[0], &G.kfwd[1]) ||
        !pick_tiled_batch(R, kTiledTX, kTiledNW, cpml ? MODE_FWD_DIV_CPML : MODE_FWD_DIV, &G.kdiv[0], &G.kdiv[1]) ||
        !pick_tiled_batch(R, kTiledTX, kTiledNW, cpml ? MODE_ADJ_CPML : MODE_ADJ, &G.kadj[0], &G.kadj[1]) ||
        !pick_tiled_batch(R, kTiledTX, kTiledNW, MODE_REV, &G.krev[1], &G.krev[0]))
        return fail(ctx, SFD_ENOTSUP, "no gradient kernel for radius %d", R);
    G.nseg = (G.nt + G.period - 1) / G.period;
    int rc = group_tables(&G);
    if (rc) return rc;
    hipStream_t s = ctx->stream;

    // Residual tables of every shot (receivers of each slot, in receiver
    // order), uploaded before the forward is queued; one block of misfit
    // partials per 256 residual cells.
    DeviceBuffer tables, partial, obs_all;
    std::vector<size_t> t_st(nshot), t_rv(nshot), poff(nshot + 1, 0), ooff(nshot + 1, 0);
    {
        std::vector<int> tab;
        for (int q = 0; q < nshot; ++q) {
            GradShot& g = G.shots[q];
            const int nrec = params[q].nrec, nslot = g.pl->nslot;
            const bool any = nrec > 0 && nslot > 0;
            ooff[q + 1] = ooff[q] + (size_t)nrec * G.nt;
            poff[q + 1] = poff[q] + (any ? ((size_t)nslot * G.nt + 255) / 256 : 0);
            t_st[q] = tab.size();
            if (!any) continue;
            std::vector<int> st(nslot + 1, 0), rv(nrec);
            for (int r = 0; r < nrec; ++r) ++st[g.slot_cell[r] + 1];
            for (int k = 0; k < nslot; ++k) st[k + 1] += st[k];
            std::vector<int> fill(st.begin(), st.end() - 1);
            for (int r = 0; r < nrec; ++r) rv[fill[g.slot_cell[r]]++] = r;
            tab.insert(tab.end(), st.begin(), st.end());
            t_rv[q] = tab.size();
            tab.insert(tab.end(), rv.begin(), rv.end());
        }
        hipError_t e = upload(tables, tab.data(), tab.size(), s);
        if (e == hipSuccess) e = partial.alloc(std::max<size_t>(poff[nshot], 1) * 8);
        if (e == hipSuccess && !obs_dev) e = obs_all.alloc(std::max<size_t>(ooff[nshot], 1) * 4);
        if (e != hipSuccess) return hip_check(ctx, e, "residual buffers");
    }

    // 1. forward, checkpointing before every segment but the first.  Without
    // boundary saving the last segment runs the recomputation's kernels,
    // which store its steps' divergences as well (with one segment: every
    // step's), so step 3 recomputes only the other segments.
    const int kdiv0 = G.bsave ? G.nt : (G.nseg - 1) * G.period;
    for (auto& g : G.shots) {
        rc = zero_state(g.pl);
        if (rc) return rc;
    }
    for (int k = 0; k < G.nt; ++k) {
        if (k > 0 && k % G.period == 0) {
            rc = group_checkpoint(&G, k / G.period, true);
            if (rc) return rc;
        }
        rc = group_launch(&G, k >= kdiv0 ? G.kdiv : G.kfwd, G.fwd_args, k);
        if (!rc && G.bsave) rc = group_band(&G, true, k);
        if (rc) return rc;
    }
    for (auto& g : G.shots) g.pl->v_raw = !cpml;
    // 2. residuals and seeds on the device from the gather buffers.  Host
    // data is copied on a second stream while the forward runs; device data
    // is read in place.  The misfit partials come back with the gradient.
    if (!obs_dev && ooff[nshot] > 0) {
        hipStream_t side = nullptr;
        hipEvent_t done = nullptr;
        hipError_t e = hipStreamCreateWithFlags(&side, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&done, hipEventDisableTiming);
        for (int q = 0; q < nshot && e == hipSuccess; ++q)
            if (ooff[q + 1] > ooff[q])
                e = hipMemcpyAsync(obs_all.as<float>() + ooff[q], d_obs[q], (ooff[q + 1] - ooff[q]) * 4,
                                   hipMemcpyHostToDevice, side);
        if (e == hipSuccess) e = hipEventRecord(done, side);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, done, 0);
        if (done) (void)hipEventDestroy(done);
        if (side) (void)hipStreamDestroy(side);
        if (e != hipSuccess) return hip_check(ctx, e, "observed data transfer");
    }
    for (int q = 0; q < nshot; ++q) {
        sfd_acoustic_plan* pl = G.shots[q].pl;
        const int nslot = pl->nslot;
        if (poff[q + 1] == poff[q]) continue;
        const float* o = obs_dev ? d_obs[q] : obs_all.as<float>() + ooff[q];
        hipLaunchKernelGGL(residual_kernel, dim3((unsigned)(poff[q + 1] - poff[q])), dim3(256), 0, s,
                           pl->rec_buf.as<float>(), nslot, G.nt, o, tables.as<int>() + t_st[q],
                           tables.as<int>() + t_rv[q], pl->res.as<float>(), partial.as<double>() + poff[q]);
        SFD_HIP(ctx, hipGetLastError());
    }
    // 3. backward: per segment, restore, recompute with divergences, adjoint.
    for (auto& g : G.shots) {
        sfd_acoustic_plan* pl = g.pl;
        rc = zero_adjoint(pl);
        if (rc) return rc;
        SFD_HIP(ctx, hipMemsetAsync(pl->grad.ptr, 0, pl->grad.bytes, s));
        SFD_HIP(ctx, hipMemsetAsync(pl->src_adj.ptr, 0, pl->src_adj.bytes, s));
        if (G.bsave) SFD_HIP(ctx, hipMemsetAsync(pl->div.ptr, 0, pl->div.bytes, s));
    }
    rc = group_inject(&G, G.nt - 1);
    if (rc) return rc;
    if (G.bsave) {
        // One sweep from the final state: each step reconstructs the forward
        // state before it inside the box (resetting to the checkpoint at a
        // segment end), then runs the adjoint step with its divergence.
        for (int k = G.nt - 1; k >= 0; --k) {
            if ((k + 1) % G.period == 0 && k + 1 < G.nt) {
                rc = group_checkpoint(&G, (k + 1) / G.period, false);
                if (rc) return rc;
            }
            rc = group_band(&G, false, k);
            if (!rc) rc = group_launch(&G, G.krev, G.fwd_args, k);
            if (!rc) rc = group_launch(&G, G.kadj, G.adj_args, k);
            if (!rc && k >= 1) rc = group_inject(&G, k - 1);
            if (rc) return rc;
        }
    }
    for (int seg = G.bsave ? -1 : G.nseg - 1; seg >= 0; --seg) {
        const int k0 = seg * G.period, k1 = std::min(k0 + G.period, G.nt);
        if (k0 < kdiv0) {
            if (seg == 0) {
                for (auto& g : G.shots) {
                    for (Field* f : {&g.pl->p[0], &g.pl->vx[0], &g.pl->vz[0]})
                        SFD_HIP(ctx, hipMemsetAsync(f->ptr, 0, f->bytes, s));
                    Field* m[4];
                    for (int q = 0, c = psi_fields(g.pl, m); q < c; ++q)
                        SFD_HIP(ctx, hipMemsetAsync(m[q]->ptr, 0, m[q]->bytes, s));
                }
            } else {
                rc = group_checkpoint(&G, seg, false);
                if (rc) return rc;
            }
            for (int k = k0; k < k1; ++k) {
                rc = group_launch(&G, G.kdiv, G.fwd_args, k);
                if (rc) return rc;
            }
        }
        for (int k = k1 - 1; k >= k0; --k) {
            rc = group_launch(&G, G.kadj, G.adj_args, k);
            if (rc) return rc;
            if (k >= 1) {
                rc = group_inject(&G, k - 1);
                if (rc) return rc;
            }
        }
    }
    // 4. dJ/dvp summed over the group's shots on the device, then the source
    // terms of each shot; the misfit from the partials (per shot in block
    // order, then shot after shot).
    if (misfit && poff[nshot] > 0) {
        std::vector<double> parts(poff[nshot]);
        SFD_HIP(ctx, hipMemcpyAsync(parts.data(), partial.ptr, poff[nshot] * 8, hipMemcpyDeviceToHost, s));
        SFD_HIP(ctx, hipStreamSynchronize(s));
        for (int q = 0; q < nshot; ++q) {
            double m = 0.0;
            for (size_t b = poff[q]; b < poff[q + 1]; ++b) m += parts[b];
            *misfit += m;
        }
    }
    {
        sfd_acoustic_plan* p0 = G.shots[0].pl;
        const size_t n = (size_t)p0->nx * p0->nz;
        DeviceBuffer dvp, sum, accp;
        std::vector<const float*> acc(nshot);
        for (int q = 0; q < nshot; ++q) acc[q] = G.shots[q].pl->grid(G.shots[q].pl->grad);
        hipError_t e = upload(dvp, vp, n, s);
        if (e == hipSuccess) e = upload(accp, acc.data(), acc.size(), s);
        if (e == hipSuccess) e = sum.alloc(n * 4);
        if (e != hipSuccess) return hip_check(ctx, e, "gradient sum");
        hipLaunchKernelGGL(grad_sum_kernel, dim3(2048), dim3(256), 0, s, sum.as<float>(), accp.as<const float*>(), nshot,
                           dvp.as<float>(), n);
        SFD_HIP(ctx, hipGetLastError());
        std::vector<float> g(n);
        SFD_HIP(ctx, hipMemcpyAsync(g.data(), sum.ptr, n * 4, hipMemcpyDeviceToHost, s));
        SFD_HIP(ctx, hipStreamSynchronize(s));
        for (size_t c = 0; c < n; ++c) grad_vp[c] += g[c];
        for (int q = 0; q < nshot; ++q) {
            sfd_acoustic_plan* pl = G.shots[q].pl;
            std::vector<float> sa(pl->nt);
            SFD_HIP(ctx, hipMemcpyAsync(sa.data(), pl->src_adj.ptr, (size_t)pl->nt * 4, hipMemcpyDeviceToHost, s));
            SFD_HIP(ctx, hipStreamSynchronize(s));
            grad_source_term(pl, sa, vp, wavelet[q], grad_vp);
        }
    }
    return SFD_OK;
}

// Watchdog of a host result: SFD_ENAN when any value is Inf or NaN.
int host_finite(sfd_context* ctx, const float* a, size_t n, const char* what) {
    for (size_t c = 0; c < n; ++c)
        if (!std::isfinite(a[c]))
            return fail(ctx, SFD_ENAN, "%s holds non-finite values: the run diverged or an input is not finite", what);
    return SFD_OK;
}

// The misfit gradient behind both entry points (observed data on the host
// or, obs_dev, in device memory).
int misfit_gradient(sfd_context* ctx, int nshot, const sfd_acoustic_params* params, const float* vp, const float* den,
                    const float* const* wavelet, const float* const* d_obs, bool obs_dev, int ckpt_every, int method,
                    float* grad_vp, double* misfit) {
    if (!ctx || nshot < 1 || !params || !vp || !wavelet || !d_obs || !grad_vp) return SFD_EINVAL;
    if (method != SFD_GRADIENT_RECOMPUTE && method != SFD_GRADIENT_BOUNDARY_SAVING)
        return fail(ctx, SFD_EINVAL, "unknown gradient method %d", method);
    for (int s = 1; s < nshot; ++s)
        if (params[s].nx != params[0].nx || params[s].nz != params[0].nz)
            return fail(ctx, SFD_EINVAL, "shot %d: every shot of a gradient shares the model grid", s);
    if (ckpt_every < 1) return fail(ctx, SFD_EINVAL, "checkpoint interval must be >= 1");
    std::fill(grad_vp, grad_vp + (size_t)params[0].nx * params[0].nz, 0.f);
    if (misfit) *misfit = 0.0;
    // Groups sized so that their workspaces (state, adjoint state, the
    // divergences of one segment or, with boundary saving, one divergence
    // field and the band of every step, and the checkpoints of every shot)
    // stay below 80 % of the device memory free at the call (64 GiB where the
    // query fails).  A checkpoint interval >= nt keeps all nt divergences of
    // every shot: 16 shots of 881 x 881 over 2500 steps take 124 GB.
    size_t free_b = 0, total_b = 0;
    const double budget =
        hipMemGetInfo(&free_b, &total_b) == hipSuccess ? 0.8 * (double)free_b : 64.0 * (1ull << 30);
    const double field = (double)params[0].nx * params[0].nz * 4.0;
    const int period = std::min(ckpt_every, std::max(params[0].nt, 1));
    const bool bsave = method == SFD_GRADIENT_BOUNDARY_SAVING;
    const double band = 4.0 * params[0].radius * (params[0].nx + params[0].nz) * 4.0 * 3.0 * params[0].nt;
    const double per_shot = field * (16.0 + (bsave ? 1 : period)) * 1.2 +
                            field * 3.0 * ((params[0].nt + period - 1) / period) + (bsave ? band : 0.0);
    const int gmax = (int)std::max(1.0, std::min((double)nshot, budget / per_shot));
    for (int s0 = 0; s0 < nshot; s0 += gmax) {
        const int n = std::min(gmax, nshot - s0);
        const int rc = group_run(ctx, n, params + s0, vp, den, wavelet + s0, d_obs + s0, obs_dev, period, method,
                                 grad_vp, misfit);
        if (rc) return rc;
    }
    if (misfit && !std::isfinite(*misfit)) return fail(ctx, SFD_ENAN, "the misfit is not finite: the run diverged");
    return host_finite(ctx, grad_vp, (size_t)params[0].nx * params[0].nz, "the gradient");
}

}  // namespace
}  // namespace sfd

extern "C" {

int sfd_acoustic_gradient(sfd_context* ctx, const sfd_acoustic_params* p, const float* vp, const float* den,
                          const float* wavelet, const float* seed, int ckpt_every, float* grad_vp, float* record) {
    if (!ctx || !p || !vp || !wavelet || !seed || !grad_vp) return SFD_EINVAL;
    std::fill(grad_vp, grad_vp + (size_t)p->nx * p->nz, 0.f);
    const int rc = grad_one(ctx, p, vp, den, wavelet, seed, nullptr, ckpt_every, grad_vp, record, nullptr);
    if (rc) return rc;
    return host_finite(ctx, grad_vp, (size_t)p->nx * p->nz, "the gradient");
}

int sfd_acoustic_misfit_gradient(sfd_context* ctx, int nshot, const sfd_acoustic_params* params, const float* vp,
                                 const float* den, const float* const* wavelet, const float* const* d_obs,
                                 int ckpt_every, float* grad_vp, double* misfit) {
    return sfd_acoustic_misfit_gradient_ex(ctx, nshot, params, vp, den, wavelet, d_obs, ckpt_every,
                                           SFD_GRADIENT_RECOMPUTE, grad_vp, misfit);
}

int sfd_acoustic_misfit_gradient_ex(sfd_context* ctx, int nshot, const sfd_acoustic_params* params, const float* vp,
                                    const float* den, const float* const* wavelet, const float* const* d_obs,
                                    int ckpt_every, int method, float* grad_vp, double* misfit) {
    return sfd::misfit_gradient(ctx, nshot, params, vp, den, wavelet, d_obs, false, ckpt_every, method, grad_vp,
                                misfit);
}

int sfd_acoustic_misfit_gradient_dev(sfd_context* ctx, int nshot, const sfd_acoustic_params* params, const float* vp,
                                     const float* den, const float* const* wavelet, const float* const* d_obs_dev,
                                     int ckpt_every, int method, float* grad_vp, double* misfit) {
    return sfd::misfit_gradient(ctx, nshot, params, vp, den, wavelet, d_obs_dev, true, ckpt_every, method, grad_vp,
                                misfit);
}

}  // extern "C"
