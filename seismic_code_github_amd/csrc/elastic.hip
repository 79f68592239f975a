// Isotropic elastic free-surface shot plan and its C ABI entry points
// (include/seisfd.h): the drop-in for the time loop of
// forward/yang_version_2016_2_2, isoewm2dfsbc
// (Iso_multishot_fd_subroutine_fsbc_zy_m2_pri.f90:57-230).  Per-shot set-up
// follows the reference's own routines:
//   substagger            (pub:323-547)   staggered moduli and buoyancies
//   csponge               (pub:796-847)   cosine sponge tables
//   get_gauss_filter_coe  (pub:1470-1501) Gaussian source window
// and each step is elastic_velocity_kernel + elastic_stress_kernel
// (elastic_kernel.h).
#include <algorithm>
#include <climits>
#include <cmath>
#include <new>
#include <vector>

#include "common.h"
#include "elastic_kernel.h"

struct sfd_elastic_plan {
    sfd_context* ctx = nullptr;
    int nx = 0, nz = 0, nt = 0, delay = 0, nsteps = 0;
    int pitch = 0, ncols = 0;  // padded field: ncols columns of pitch floats
    int nrec = 0, nsnap = 0, snap_fields = 0, nsnap_fields = 0;
    std::vector<int> snap_steps;
    sfd::ElasticStepArgs args{};
    int tx = 0, nw = 0;
    void (*vel)(sfd::ElasticStepArgs) = nullptr;
    void (*str)(sfd::ElasticStepArgs) = nullptr;
    sfd::DeviceBuffer arena;
    sfd::Field vx, vz, sxx, szz, sxz, parts, ghost, b2e, b1e, lm, ls, mu;
    sfd::DeviceBuffer itopo, kst, srow, fxn, fxh, ratio, fzn, fzh, tzn, tzh, g, amp, rx, rz, rec_vx, rec_vz, snaps;
    bool abc = false;

    size_t field_floats() const { return (size_t)ncols * pitch; }
    float* cell0(const sfd::Field& f) const { return f.as<float>() + (size_t)sfd::kElPad * pitch + sfd::kElPad; }
};

// Several shots stepped together: each shot keeps its own plan (fields,
// model, gathers); one launch per half-step covers the tiles of every shot.
struct sfd_elastic_batch {
    sfd_context* ctx = nullptr;
    std::vector<sfd_elastic_plan*> plans;
    bool abc = false;
    int nsteps = 0, delay = 0;
    int tx = 0, nw = 0, ntile = 0;
    void (*vel)(sfd::ElasticBatchArgs) = nullptr;
    void (*str)(sfd::ElasticBatchArgs) = nullptr;
    sfd::DeviceBuffer shot_args, tile_map;
    // Groups of consecutive shots run one after another, each for all its
    // steps, so that a group's fields stay in the 256 MB MALL: shots
    // [s0, s1), tiles from tile_off (shot indices group-local).
    struct Group {
        int s0, s1, tile_off, ntile;
    };
    std::vector<Group> groups;
    size_t group_bytes = 0;  // budget of one group's fields; 0 = the default
    ~sfd_elastic_batch() {
        for (auto* pl : plans) sfd_elastic_plan_destroy(pl);
    }
};

namespace sfd {
namespace {

constexpr float kPi = 3.1415926f;  // the reference's module constant (module:4)

// substagger (pub:323-547) in the reference's single-precision arithmetic:
// mu = rho vs^2, lambda = rho vp^2 - 2 mu; busg1e = z-average of the inverse
// 4-cell density average, busg2e = x-average of 1/rho, musge = 4-cell
// harmonic average of the z-averaged mu (0 next to a fluid cell), lambsg =
// x-average of lambda, lambmusg = x-average of lambda + 2 mu; the last row /
// column copies its neighbour or keeps the unaveraged value as the reference
// does; finally every row above a column's surface repeats the surface row.
void stagger(int nx, int nz, const float* vp, const float* vs, const float* rho, const int* itopo,
             std::vector<float> out[5]) {
    const size_t n = (size_t)nx * nz;
    auto at = [nz](std::vector<float>& v, int i, int k) -> float& { return v[(size_t)i * nz + k]; };
    std::vector<float> mu(n), lamb(n), musg(n), mge(n), lsg(n), lmsg(n), b1(n), b2(n), b1e(n), b2e(n);
    for (size_t c = 0; c < n; ++c) {
        mu[c] = rho[c] * vs[c] * vs[c];
        lamb[c] = rho[c] * vp[c] * vp[c] - 2.0f * mu[c];
    }
    for (int i = 0; i < nx - 1; ++i)
        for (int k = 0; k < nz - 1; ++k) {
            const float* r = rho + (size_t)i * nz + k;
            at(b1, i, k) = 0.25f * (((r[0] + r[1]) + r[nz + 1]) + r[nz]);
        }
    for (int i = 0; i < nx; ++i) at(b1, i, nz - 1) = at(b1, i, nz - 2);
    for (int k = 0; k < nz; ++k) at(b1, nx - 1, k) = at(b1, nx - 2, k);
    for (size_t c = 0; c < n; ++c) b1[c] = 1.0f / b1[c];
    for (int i = 0; i < nx; ++i) {
        for (int k = 1; k < nz; ++k) at(b1e, i, k) = 0.5f * (at(b1, i, k) + at(b1, i, k - 1));
        at(b1e, i, 0) = at(b1, i, 0);
    }
    for (size_t c = 0; c < n; ++c) b2[c] = 1.0f / rho[c];
    for (int i = 0; i < nx - 1; ++i)
        for (int k = 0; k < nz; ++k) at(b2e, i, k) = 0.5f * (at(b2, i, k) + at(b2, i + 1, k));
    for (int k = 0; k < nz; ++k) at(b2e, nx - 1, k) = at(b2, nx - 1, k);
    for (int i = 0; i < nx; ++i) {
        for (int k = 0; k < nz - 1; ++k) at(musg, i, k) = 0.5f * (at(mu, i, k + 1) + at(mu, i, k));
        at(musg, i, nz - 1) = at(mu, i, nz - 1);
    }
    for (int i = 0; i < nx - 1; ++i)
        for (int k = 1; k < nz; ++k) {
            const float m0 = at(musg, i, k), m1 = at(musg, i + 1, k), m2 = at(musg, i + 1, k - 1),
                        m3 = at(musg, i, k - 1);
            at(mge, i, k) = (m0 != 0 && m3 != 0 && m1 != 0 && m2 != 0)
                                ? 4.0f / (((1.0f / m0 + 1.0f / m1) + 1.0f / m2) + 1.0f / m3)
                                : 0.0f;
        }
    for (int i = 0; i < nx; ++i) at(mge, i, 0) = at(musg, i, 0);
    for (int k = 0; k < nz; ++k) at(mge, nx - 1, k) = at(musg, nx - 1, k);
    for (int k = 0; k < nz; ++k) {
        for (int i = 0; i < nx - 1; ++i) at(lsg, i, k) = 0.5f * (at(lamb, i + 1, k) + at(lamb, i, k));
        at(lsg, nx - 1, k) = at(lamb, nx - 1, k);
    }
    for (int i = 0; i < nx - 1; ++i)
        for (int k = 0; k < nz; ++k)
            at(lmsg, i, k) = 0.5f * (((at(lamb, i, k) + 2 * at(mu, i, k)) + at(lamb, i + 1, k)) + 2 * at(mu, i + 1, k));
    for (int k = 0; k < nz; ++k) at(lmsg, nx - 1, k) = at(lamb, nx - 1, k) + 2 * at(mu, nx - 1, k);
    std::vector<float>* keep[5] = {&b2e, &b1e, &lmsg, &lsg, &mge};
    for (int q = 0; q < 5; ++q) {
        std::vector<float>& v = *keep[q];
        for (int i = 0; i < nx; ++i)
            for (int k = itopo[i]; k >= 1; --k) at(v, i, k - 1) = at(v, i, k);
        out[q].swap(v);
    }
}

// csponge (pub:796-847) for one side of np + 1 cells, outermost first: node
// table cos(acos(R) (np - i + 1) h / (np h)) and the table shifted half a cell
// inwards (lo side, last entry 1) or outwards (hi side).
void csponge(int np, float h, float Rc, bool lo, std::vector<float>& node, std::vector<float>& half) {
    node.assign(np + 1, 1.0f);
    half.assign(np + 1, 1.0f);
    if (np <= 0) return;
    const float fac = std::acos(Rc) / ((float)np * h);
    for (int i = 1; i <= np + 1; ++i) {
        node[i - 1] = std::cos(fac * (float)(np - i + 1) * h);
        half[i - 1] = lo ? std::cos(fac * ((float)(np - i + 1) * h - h / 2))
                         : std::cos(fac * ((float)(np - i + 1) * h + h / 2));
    }
    if (lo) half[np] = 1.0f;
}

// gauss_filter_2d (pub:592-634), as one thread runs it: rows 0 .. itopo(i)
// of every column replaced in place, column after column, by the window sum
// (weights g, window 2 rgau - 1 clipped to the grid, not renormalised) of the
// field as it stands, so smoothed cells feed later ones.  (The reference runs
// the columns under OpenMP, which races on shared neighbours; the serial order
// is the deterministic reading.)
void gauss_smooth(int nx, int nz, int rg, const int* itopo, const std::vector<float>& g, std::vector<float>& u) {
    const int ng = 2 * rg - 1;
    for (int i = 0; i < nx; ++i)
        for (int k = 0; k <= itopo[i]; ++k) {
            float sum = 0.f;
            for (int n = std::max(0, i - (rg - 1)); n <= std::min(nx - 1, i + (rg - 1)); ++n) {
                const int idx2 = n - (i - (rg - 1));
                for (int m = std::max(0, k - (rg - 1)); m <= std::min(nz - 1, k + (rg - 1)); ++m)
                    sum += u[(size_t)n * nz + m] * g[(size_t)idx2 * ng + (m - (k - (rg - 1)))];
            }
            u[(size_t)i * nz + k] = sum;
        }
}

int validate(sfd_context* ctx, const sfd_elastic_params* p, const float* vp, const float* vs, const float* rho) {
    if (!p || !vp || !vs || !rho || !p->itopo) return fail(ctx, SFD_EINVAL, "null params, vp, vs, rho or itopo");
    if (p->nx < 12 || p->nz < 12) return fail(ctx, SFD_EINVAL, "grid %dx%d too small (12x12 at least)", p->nx, p->nz);
    if (p->nt < 0) return fail(ctx, SFD_EINVAL, "nt < 0");
    if (!(p->dx > 0) || !(p->dz > 0) || !(p->dt > 0)) return fail(ctx, SFD_EINVAL, "dx, dz, dt must be > 0");
    if (!(p->f0 > 0)) return fail(ctx, SFD_EINVAL, "f0 must be > 0");
    if (p->sponge_xlo < 0 || p->sponge_xhi < 0 || p->sponge_zhi < 0)
        return fail(ctx, SFD_EINVAL, "sponge widths must be >= 0");
    if (p->mode != SFD_ELASTIC_FREE_SURFACE && p->mode != SFD_ELASTIC_ABSORBING_TOP)
        return fail(ctx, SFD_EINVAL, "mode must be SFD_ELASTIC_FREE_SURFACE or SFD_ELASTIC_ABSORBING_TOP");
    if (p->mode == SFD_ELASTIC_ABSORBING_TOP && p->sponge_zlo < 1)
        return fail(ctx, SFD_EINVAL, "the absorbing top needs sponge_zlo >= 1 (reference npzn)");
    if (p->sponge_xlo + p->sponge_xhi + 2 > p->nx)
        return fail(ctx, SFD_EINVAL, "left and right sponge strips overlap in %d columns", p->nx);
    const int kb0 = p->nz - 1 - p->sponge_zhi;
    if (kb0 < 1) return fail(ctx, SFD_EINVAL, "bottom sponge covers the grid");
    for (int i = 0; i < p->nx; ++i)
        if (p->itopo[i] < 0 || p->itopo[i] + 6 > kb0)
            return fail(ctx, SFD_EINVAL,
                        "surface row %d of column %d outside [0, %d]: the surface and the 5 rows of reduced "
                        "order below it must lie above the bottom sponge",
                        p->itopo[i], i, kb0 - 6);
    if (p->sx < 0 || p->sx >= p->nx || p->sz < 0 || p->sz >= p->nz)
        return fail(ctx, SFD_EINVAL, "source (%d,%d) outside the grid", p->sx, p->sz);
    if (p->src_radius < 0) return fail(ctx, SFD_EINVAL, "src_radius must be >= 0");
    if (p->nrec < 0 || (p->nrec > 0 && (!p->rx || !p->rz))) return fail(ctx, SFD_EINVAL, "bad receiver list");
    for (int r = 0; r < p->nrec; ++r)
        if (p->rx[r] < 0 || p->rx[r] >= p->nx || p->rz[r] < 0 || p->rz[r] >= p->nz)
            return fail(ctx, SFD_EINVAL, "receiver %d (%d,%d) outside the grid", r, p->rx[r], p->rz[r]);
    if (p->nsnap < 0 || (p->nsnap > 0 && !p->snap_steps)) return fail(ctx, SFD_EINVAL, "bad snapshot list");
    if (p->nsnap > 0 && (p->snap_fields & 62) == 0) return fail(ctx, SFD_EINVAL, "snap_fields is empty");
    return SFD_OK;
}

using ElasticKernel = void (*)(ElasticStepArgs);

// Default tiling, measured on MI355X for the 903 x 603 shot grid (us per
// step, both kernels): 8 x 8 24.2, 8 x 4 25.9, 16 x 8 26.0, 16 x 16 29.5,
// 32 x 16 32.3, 16 x 4 36.8, 32 x 8 38.5.  One column per wave and small
// tiles win: the step is latency-bound and more workgroups hide it.
constexpr int kElDefaultTX = 8, kElDefaultNW = 8;

template <int TX, int NW>
void el_pair(bool abc, ElasticKernel* v, ElasticKernel* s) {
    *v = abc ? elastic_velocity_kernel<TX, NW, true> : elastic_velocity_kernel<TX, NW, false>;
    *s = abc ? elastic_stress_kernel<TX, NW, true> : elastic_stress_kernel<TX, NW, false>;
}

// The instantiated tilings: tile columns x waves per workgroup.
bool el_pick(int tx, int nw, bool abc, ElasticKernel* v, ElasticKernel* s) {
    switch (tx * 100 + nw) {
    case 804: el_pair<8, 4>(abc, v, s); return true;
    case 808: el_pair<8, 8>(abc, v, s); return true;
    case 1604: el_pair<16, 4>(abc, v, s); return true;
    case 1608: el_pair<16, 8>(abc, v, s); return true;
    case 1616: el_pair<16, 16>(abc, v, s); return true;
    case 3204: el_pair<32, 4>(abc, v, s); return true;
    case 3208: el_pair<32, 8>(abc, v, s); return true;
    case 3216: el_pair<32, 16>(abc, v, s); return true;
    default: return false;
    }
}

using ElasticBatchKernel = void (*)(ElasticBatchArgs);

template <int TX, int NW>
void el_batch_pair(bool abc, ElasticBatchKernel* v, ElasticBatchKernel* s) {
    *v = abc ? elastic_velocity_batch<TX, NW, true> : elastic_velocity_batch<TX, NW, false>;
    *s = abc ? elastic_stress_batch<TX, NW, true> : elastic_stress_batch<TX, NW, false>;
}

bool el_batch_pick(int tx, int nw, bool abc, ElasticBatchKernel* v, ElasticBatchKernel* s) {
    switch (tx * 100 + nw) {
    case 804: el_batch_pair<8, 4>(abc, v, s); return true;
    case 808: el_batch_pair<8, 8>(abc, v, s); return true;
    case 1604: el_batch_pair<16, 4>(abc, v, s); return true;
    case 1608: el_batch_pair<16, 8>(abc, v, s); return true;
    case 1616: el_batch_pair<16, 16>(abc, v, s); return true;
    case 3204: el_batch_pair<32, 4>(abc, v, s); return true;
    case 3208: el_batch_pair<32, 8>(abc, v, s); return true;
    case 3216: el_batch_pair<32, 16>(abc, v, s); return true;
    default: return false;
    }
}

// Tilings the autotuners try (tile columns, waves).
constexpr int kElShapes[][2] = {{8, 4}, {8, 8}, {16, 4}, {16, 8}, {16, 16}, {32, 4}, {32, 8}, {32, 16}};

int el_set_tiling(sfd_elastic_plan* pl, int tx, int nw) {
    ElasticKernel v, s;
    if (!el_pick(tx, nw, pl->abc, &v, &s))
        return fail(pl->ctx, SFD_EINVAL, "no elastic kernel for %d tile columns x %d waves", tx, nw);
    pl->tx = tx;
    pl->nw = nw;
    pl->vel = v;
    pl->str = s;
    pl->args.ntz = (pl->nz + kElTZ - 1) / kElTZ;
    pl->args.ntiles = pl->args.ntz * ((pl->nx + tx - 1) / tx);
    return SFD_OK;
}

// Autotuner grid class: mode, nx and nz to the nearest power of 2.
std::string elastic_class(const sfd_elastic_plan* pl) {
    char buf[96];
    snprintf(buf, sizeof buf, "elastic:%s:x%d:z%d", pl->abc ? "abc" : "fsbc", log2_round(pl->nx), log2_round(pl->nz));
    return buf;
}

int elastic_zero(sfd_elastic_plan* pl) {
    sfd_context* ctx = pl->ctx;
    // Wavefields, strip parts and the pending surface values lie at the front
    // of the arena, before the model fields.
    const size_t state_bytes = (size_t)((const char*)pl->b2e.ptr - pl->arena.as<char>());
    SFD_HIP(ctx, hipMemsetAsync(pl->arena.ptr, 0, state_bytes, ctx->stream));
    return SFD_OK;
}

// Snapshots the plan takes after step s (the next one is snap_steps[*next]).
int elastic_snaps(sfd_elastic_plan* pl, int s, size_t* next) {
    sfd_context* ctx = pl->ctx;
    const size_t cell_bytes = (size_t)pl->nz * sizeof(float);
    const ElasticStepArgs& a = pl->args;
    while (*next < pl->snap_steps.size() && pl->snap_steps[*next] == s) {
        float* dst = pl->snaps.as<float>() + *next * pl->nsnap_fields * (size_t)pl->nx * pl->nz;
        const float* src[5] = {a.vx, a.vz, a.sxx, a.szz, a.sxz};
        for (int f = 0; f < 5; ++f) {
            if (!(pl->snap_fields & (2 << f))) continue;
            SFD_HIP(ctx, hipMemcpy2DAsync(dst, cell_bytes, src[f], (size_t)pl->pitch * sizeof(float), cell_bytes,
                                          pl->nx, hipMemcpyDeviceToDevice, ctx->stream));
            dst += (size_t)pl->nx * pl->nz;
        }
        ++*next;
    }
    return SFD_OK;
}

// Enqueues the first nsteps steps (all of them by default).
int elastic_enqueue(sfd_elastic_plan* pl, int nsteps = -1) {
    sfd_context* ctx = pl->ctx;
    const dim3 grid(pl->args.ntiles), block(64 * pl->nw);
    ElasticStepArgs a = pl->args;
    size_t snap_next = 0;
    const int ns = nsteps < 0 ? pl->nsteps : std::min(nsteps, pl->nsteps);
    for (int s = 0; s < ns; ++s) {
        a.step = s;
        a.rec_col = (s >= pl->delay && pl->nrec > 0) ? s - pl->delay : -1;
        hipLaunchKernelGGL(pl->vel, grid, block, 0, ctx->stream, a);
        SFD_HIP(ctx, hipGetLastError());
        hipLaunchKernelGGL(pl->str, grid, block, 0, ctx->stream, a);
        SFD_HIP(ctx, hipGetLastError());
        const int rc = elastic_snaps(pl, s, &snap_next);
        if (rc) return rc;
    }
    return SFD_OK;
}

// Default budget of one shot group's fields (state, staggered model, strip
// parts, the pending vz row): measured on MI355X with the parameter file's
// 32-shot survey (per step: one group 516, 160 MiB 428, 192 MiB 422, 256 MiB
// 408 us).
constexpr size_t kElGroupBytes = (size_t)256 << 20;

// Batch tables: every shot's argument block (at tile_cols x waves) and the
// (group-local shot, tile) of every tile, group after group.
int el_batch_set_tiling(sfd_elastic_batch* b, int tx, int nw) {
    ElasticBatchKernel v, s;
    if (!el_batch_pick(tx, nw, b->abc, &v, &s))
        return fail(b->ctx, SFD_EINVAL, "no elastic kernel for %d tile columns x %d waves", tx, nw);
    const int n = (int)b->plans.size();
    const size_t budget = b->group_bytes ? b->group_bytes : kElGroupBytes;
    b->groups.clear();
    for (int s0 = 0; s0 < n;) {
        size_t used = 0;
        int e = s0;
        do used += b->plans[e++]->arena.bytes;
        while (e < n && used + b->plans[e]->arena.bytes <= budget);
        b->groups.push_back({s0, e, 0, 0});
        s0 = e;
    }
    std::vector<ElasticStepArgs> args;
    std::vector<int2> map;
    for (auto& g : b->groups) {
        g.tile_off = (int)map.size();
        for (int i = g.s0; i < g.s1; ++i) {
            const int rc = el_set_tiling(b->plans[i], tx, nw);
            if (rc) return rc;
            args.push_back(b->plans[i]->args);
            for (int t = 0; t < b->plans[i]->args.ntiles; ++t) map.push_back(make_int2(i - g.s0, t));
        }
        g.ntile = (int)map.size() - g.tile_off;
    }
    hipStream_t st = b->ctx->stream;
    hipError_t e = upload(b->shot_args, args.data(), args.size(), st);
    if (e == hipSuccess) e = upload(b->tile_map, map.data(), map.size(), st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_check(b->ctx, e, "elastic batch tables");
    b->tx = tx;
    b->nw = nw;
    b->vel = v;
    b->str = s;
    b->ntile = (int)map.size();
    return SFD_OK;
}

int el_batch_zero(sfd_elastic_batch* b) {
    for (auto* pl : b->plans) {
        const int rc = elastic_zero(pl);
        if (rc) return rc;
    }
    return SFD_OK;
}

// Enqueues the first nsteps steps of every shot (all of them by default),
// shot group after shot group.
int el_batch_enqueue(sfd_elastic_batch* b, int nsteps = -1) {
    sfd_context* ctx = b->ctx;
    const dim3 block(64 * b->nw);
    const int ns = nsteps < 0 ? b->nsteps : std::min(nsteps, b->nsteps);
    for (const auto& g : b->groups) {
        const dim3 grid(g.ntile);
        ElasticBatchArgs a{};
        a.shots = b->shot_args.as<ElasticStepArgs>() + g.s0;
        a.tile_map = b->tile_map.as<int2>() + g.tile_off;
        a.ntile = g.ntile;
        std::vector<size_t> snap_next(g.s1 - g.s0, 0);
        for (int s = 0; s < ns; ++s) {
            a.step = s;
            a.rec_col = s >= b->delay ? s - b->delay : -1;
            hipLaunchKernelGGL(b->vel, grid, block, 0, ctx->stream, a);
            SFD_HIP(ctx, hipGetLastError());
            hipLaunchKernelGGL(b->str, grid, block, 0, ctx->stream, a);
            SFD_HIP(ctx, hipGetLastError());
            for (int i = g.s0; i < g.s1; ++i) {
                const int rc = elastic_snaps(b->plans[i], s, &snap_next[i - g.s0]);
                if (rc) return rc;
            }
        }
    }
    return SFD_OK;
}

// Autotuner grid class of a batch: mode, shot count and cells per step to
// the nearest power of 2.
std::string elastic_batch_class(const sfd_elastic_batch* b) {
    size_t cells = 0;
    for (auto* pl : b->plans) cells += (size_t)pl->nx * pl->nz;
    char buf[96];
    snprintf(buf, sizeof buf, "elastic-batch:%s:n%d:c%d", b->abc ? "abc" : "fsbc", log2_round((double)b->plans.size()),
             log2_round((double)cells));
    return buf;
}

}  // namespace
}  // namespace sfd

using namespace sfd;

extern "C" {

int sfd_elastic_plan_create(sfd_context* ctx, const sfd_elastic_params* p, const float* vp, const float* vs,
                            const float* rho, const float* wavelet, sfd_elastic_plan** out) {
    if (!ctx || !out) return SFD_EINVAL;
    *out = nullptr;
    int rc = validate(ctx, p, vp, vs, rho);
    if (rc) return rc;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    const int nx = p->nx, nz = p->nz;
    const size_t n = (size_t)nx * nz;
    float vmax = 0.f;
    for (size_t c = 0; c < n; ++c) {
        if (!(vp[c] > 0.f) || !std::isfinite(vp[c])) return fail(ctx, SFD_EINVAL, "vp must be finite and > 0 (cell %zu)", c);
        if (!(vs[c] >= 0.f) || !std::isfinite(vs[c])) return fail(ctx, SFD_EINVAL, "vs must be finite and >= 0 (cell %zu)", c);
        if (!(rho[c] > 0.f) || !std::isfinite(rho[c])) return fail(ctx, SFD_EINVAL, "rho must be finite and > 0 (cell %zu)", c);
        vmax = std::max(vmax, vp[c]);
    }
    float a10[5], a8[4];
    default_coefs(5, a10);
    default_coefs(4, a8);
    // Leapfrog stability of the staggered system: the derivative symbols are
    // at most 2 sum|a| / h, so vp dt sum|a| sqrt(1/dx^2 + 1/dz^2) <= 1.
    double sa = 0;
    for (float c : a10) sa += std::fabs((double)c);
    const double courant = vmax * (double)p->dt * std::sqrt(1.0 / ((double)p->dx * p->dx) + 1.0 / ((double)p->dz * p->dz));
    if (courant * sa > 1.0)
        return fail(ctx, SFD_ECFL, "unstable: vmax*dt*sqrt(1/dx^2+1/dz^2) = %.4f exceeds %.4f", courant, 1.0 / sa);

    sfd_elastic_plan* pl = new (std::nothrow) sfd_elastic_plan;
    if (!pl) return fail(ctx, SFD_ENOMEM, "host allocation failed");
    auto bail = [&](int code) {
        delete pl;
        return code;
    };
    pl->ctx = ctx;
    pl->nx = nx;
    pl->nz = nz;
    pl->nt = p->nt;
    // The reference runs nt + idecay steps, idecay = nint(decay / dt), and
    // records from step idecay on (pri:117, 199-210); its parameter file
    // recommends decay = 1.2 / f0.
    pl->delay = p->delay >= 0 ? p->delay : (int)std::lround(1.2f / (p->f0 * p->dt));
    pl->nsteps = pl->delay + p->nt;
    for (int s = 0; s < p->nsnap; ++s)
        if (p->snap_steps[s] < 0 || p->snap_steps[s] >= pl->nsteps || (s > 0 && p->snap_steps[s] <= p->snap_steps[s - 1]))
            return bail(fail(ctx, SFD_EINVAL, "snapshot steps must be ascending and in [0, delay + nt)"));
    pl->pitch = (nz + kElTZ - 1) / kElTZ * kElTZ + 2 * kElPad;
    // Columns -kElPad .. nx + kElTXMax + kElPad - 1: the widest tile's last
    // halo column stays inside.
    pl->ncols = nx + kElTXMax + 2 * kElPad;

    pl->abc = p->mode == SFD_ELASTIC_ABSORBING_TOP;
    const int npzn = pl->abc ? p->sponge_zlo : 0;

    // ---- Gaussian window (get_gauss_filter_coe(gc2d, 2.0, 2 rgau - 1, ...),
    // normalised by its float sum in the reference's order) ----
    const int rg = p->src_radius > 0 ? p->src_radius : 50;
    const int ng = 2 * rg - 1;
    std::vector<float> g2((size_t)ng * ng);
    {
        const float sigma2 = 2.0f * 2.0f;
        const int cen = (ng + 1) / 2;
        float sum = 0.f;
        for (int j = 1; j <= ng; ++j)
            for (int i = 1; i <= ng; ++i) {
                const float dis = (float)((j - cen) * (j - cen) + (i - cen) * (i - cen));
                g2[(size_t)(j - 1) * ng + (i - 1)] = 1.0f / (2.0f * kPi * sigma2) * std::exp(-dis / (2.0f * sigma2));
            }
        for (int j = 1; j <= ng; ++j)
            for (int i = 1; i <= ng; ++i) sum += g2[(size_t)(j - 1) * ng + (i - 1)];
        for (float& v : g2) v = v / sum;
    }

    // ---- model: substagger, and in 'abc' mode modipapml (pub:549-590), the
    // Gaussian smoothing of the parameters on and above the surface ----
    std::vector<float> st[5];
    stagger(nx, nz, vp, vs, rho, p->itopo, st);
    if (pl->abc)
        for (int q = 0; q < 5; ++q) gauss_smooth(nx, nz, rg, p->itopo, g2, st[q]);

    // ---- sponge tables (csponge) per column and row ----
    const int npxn = p->sponge_xlo, npxp = p->sponge_xhi, npzp = p->sponge_zhi;
    const float Rc = p->sponge_r > 0 ? p->sponge_r : 1e-3f;
    std::vector<float> xn, xn1, xp, xp2, zp, zp2;
    csponge(npxn, p->dx, Rc, true, xn, xn1);
    csponge(npxp, p->dx, Rc, false, xp, xp2);
    csponge(npzp, p->dz, Rc, false, zp, zp2);
    std::vector<float> tn, tn1;
    csponge(npzn, p->dz, Rc, true, tn, tn1);  // 'abc' top tables spgzn, spg1zn
    // First updated row of a column: its surface row ('fsbc'), or npzn rows
    // above it ('abc'; row -1 of the reference's split parts is never summed).
    std::vector<int> hkst(nx);
    for (int i = 0; i < nx; ++i) hkst[i] = std::max(0, p->itopo[i] - npzn);
    std::vector<int> hsrow(nx, INT_MAX);
    std::vector<float> hfxn(nx, 1.f), hfxh(nx, 1.f), hfzn(nz, 1.f), hfzh(nz, 1.f), hratio(nx);
    for (int i = 0; i <= npxn; ++i) {
        hsrow[i] = hkst[i];
        hfxn[i] = xn[i];
        hfxh[i] = xn1[i];
    }
    // The right strip's damped rows start at the first updated row of the
    // column counted the same distance from the LEFT edge (pub:917-922, 871-876).
    for (int i = 0; i <= npxp; ++i) {
        hsrow[nx - 1 - i] = hkst[i];
        hfxn[nx - 1 - i] = xp[i];
        hfxh[nx - 1 - i] = xp2[i];
    }
    for (int i = 0; i <= npzp; ++i) {
        hfzn[nz - 1 - i] = zp[i];
        hfzh[nz - 1 - i] = zp2[i];
    }
    for (int i = 0; i < nx; ++i) {
        const size_t c = (size_t)i * nz + p->itopo[i];
        hratio[i] = st[3][c] / st[2][c];  // lambsg / lambmusg at the surface (calfsbcv)
    }

    // ---- source window: addsrc places gc2d centred on the source cell,
    // clipped by the grid ----
    {
        auto weight = [&](int x, int z) -> float {
            const int idx2 = x - p->sx + rg, idx1 = z - p->sz + rg;
            if (idx1 < 1 || idx1 > ng || idx2 < 1 || idx2 > ng) return 0.f;
            return g2[(size_t)(idx2 - 1) * ng + (idx1 - 1)];
        };
        int x0 = nx, x1 = -1, z0 = nz, z1 = -1;
        for (int x = std::max(0, p->sx - rg); x <= std::min(nx - 1, p->sx + rg); ++x)
            for (int z = std::max(0, p->sz - rg); z <= std::min(nz - 1, p->sz + rg); ++z)
                if (weight(x, z) != 0.f) {
                    x0 = std::min(x0, x);
                    x1 = std::max(x1, x);
                    z0 = std::min(z0, z);
                    z1 = std::max(z1, z);
                }
        if (x1 < 0) x0 = z0 = 0, x1 = z1 = -1;
        pl->args.sx0 = x0;
        pl->args.sx1 = x1 + 1;
        pl->args.sz0 = z0;
        pl->args.sz1 = z1 + 1;
        std::vector<float> hg(std::max<size_t>(1, (size_t)(x1 + 1 - x0) * (z1 + 1 - z0)), 0.f);
        for (int x = x0; x <= x1; ++x)
            for (int z = z0; z <= z1; ++z) hg[(size_t)(x - x0) * (z1 + 1 - z0) + (z - z0)] = weight(x, z);
        hipError_t e = upload(pl->g, hg.data(), hg.size(), ctx->stream);
        if (e != hipSuccess) return bail(hip_check(ctx, e, "source window"));
    }
    // Wavelet of step s (Fortran it = s + 1): the given samples or addsrc's
    // exp(-(pi f0 (it dt - idecay dt))^2) (1 - 2 (pi f0 (it dt - idecay dt))^2).
    std::vector<float> hamp(std::max(pl->nsteps, 1), 0.f);
    for (int s = 0; s < pl->nsteps; ++s) {
        if (wavelet) {
            hamp[s] = wavelet[s];
        } else {
            const float x = kPi * p->f0 * ((float)(s + 1) * p->dt - (float)pl->delay * p->dt);
            hamp[s] = std::exp(-(x * x)) * (1.0f - 2.0f * (x * x));
        }
    }

    // ---- device arena: state (5 fields, strip parts, pending surface values), then model ----
    const int nl = npxn + 1, nr = npxp + 1, nb = npzp + 1, kb0 = nz - nb;
    const size_t topbase = (size_t)(nl + nr) * nz + (size_t)(nx - nl - nr) * nb;
    const size_t nstrip = topbase + (pl->abc ? (size_t)(nx - nl - nr) * (npzn + 1) : 0);
    if (nstrip >= (size_t)INT_MAX / 10) return bail(fail(ctx, SFD_ENOTSUP, "sponge strips too large"));
    const size_t fb = pl->field_floats() * sizeof(float);
    hipStream_t sm = ctx->stream;
    hipError_t e = hipSuccess;
    for (int pass = 0; pass < 2; ++pass) {
        Carver cv;
        if (pass == 1) cv.base = pl->arena.as<char>();
        pl->vx = cv.take(fb);
        pl->vz = cv.take(fb);
        pl->sxx = cv.take(fb);
        pl->szz = cv.take(fb);
        pl->sxz = cv.take(fb);
        pl->parts = cv.take(10 * nstrip * sizeof(float));
        pl->ghost = cv.take((size_t)nx * sizeof(float));
        pl->b2e = cv.take(fb);
        pl->b1e = cv.take(fb);
        pl->lm = cv.take(fb);
        pl->ls = cv.take(fb);
        pl->mu = cv.take(fb);
        if (pass == 0) {
            e = pl->arena.alloc(cv.used);
            if (e != hipSuccess) return bail(hip_check(ctx, e, "arena allocation"));
        }
    }
    {
        std::vector<float> padded(pl->field_floats(), 0.f);
        const Field* dst[5] = {&pl->b2e, &pl->b1e, &pl->lm, &pl->ls, &pl->mu};
        for (int q = 0; q < 5 && e == hipSuccess; ++q) {
            for (int i = 0; i < nx; ++i)
                std::copy(st[q].begin() + (size_t)i * nz, st[q].begin() + (size_t)(i + 1) * nz,
                          padded.begin() + (size_t)(i + kElPad) * pl->pitch + kElPad);
            e = hipMemcpyAsync(dst[q]->ptr, padded.data(), fb, hipMemcpyHostToDevice, sm);
            if (e == hipSuccess) e = hipStreamSynchronize(sm);
        }
    }
    if (e == hipSuccess) e = upload(pl->itopo, p->itopo, nx, sm);
    if (e == hipSuccess) e = upload(pl->kst, hkst.data(), nx, sm);
    if (e == hipSuccess) e = upload(pl->srow, hsrow.data(), nx, sm);
    if (e == hipSuccess) e = upload(pl->tzn, tn.data(), tn.size(), sm);
    if (e == hipSuccess) e = upload(pl->tzh, tn1.data(), tn1.size(), sm);
    if (e == hipSuccess) e = upload(pl->fxn, hfxn.data(), nx, sm);
    if (e == hipSuccess) e = upload(pl->fxh, hfxh.data(), nx, sm);
    if (e == hipSuccess) e = upload(pl->ratio, hratio.data(), nx, sm);
    if (e == hipSuccess) e = upload(pl->fzn, hfzn.data(), nz, sm);
    if (e == hipSuccess) e = upload(pl->fzh, hfzh.data(), nz, sm);
    if (e == hipSuccess) e = upload(pl->amp, hamp.data(), hamp.size(), sm);
    pl->nrec = p->nrec;
    if (e == hipSuccess && p->nrec > 0) e = upload(pl->rx, p->rx, p->nrec, sm);
    if (e == hipSuccess && p->nrec > 0) e = upload(pl->rz, p->rz, p->nrec, sm);
    if (e == hipSuccess && p->nrec > 0) e = pl->rec_vx.alloc((size_t)p->nrec * std::max(p->nt, 1) * 4);
    if (e == hipSuccess && p->nrec > 0) e = pl->rec_vz.alloc((size_t)p->nrec * std::max(p->nt, 1) * 4);
    pl->nsnap = p->nsnap;
    pl->snap_fields = p->snap_fields & 62;
    pl->nsnap_fields = __builtin_popcount(pl->snap_fields);
    pl->snap_steps.assign(p->snap_steps, p->snap_steps + p->nsnap);
    if (e == hipSuccess && p->nsnap > 0) e = pl->snaps.alloc((size_t)p->nsnap * pl->nsnap_fields * n * 4);
    if (e != hipSuccess) return bail(hip_check(ctx, e, "plan allocation"));

    ElasticStepArgs& a = pl->args;
    a.nx = nx;
    a.nz = nz;
    a.pitch = pl->pitch;
    a.vx = pl->cell0(pl->vx);
    a.vz = pl->cell0(pl->vz);
    a.sxx = pl->cell0(pl->sxx);
    a.szz = pl->cell0(pl->szz);
    a.sxz = pl->cell0(pl->sxz);
    a.b2e = pl->cell0(pl->b2e);
    a.b1e = pl->cell0(pl->b1e);
    a.lm = pl->cell0(pl->lm);
    a.ls = pl->cell0(pl->ls);
    a.mu = pl->cell0(pl->mu);
    a.itopo = pl->itopo.as<int>();
    a.kst = pl->kst.as<int>();
    a.srow = pl->srow.as<int>();
    a.tzn = pl->tzn.as<float>();
    a.tzh = pl->tzh.as<float>();
    a.npzn = npzn;
    a.topbase = (int)topbase;
    a.fxn = pl->fxn.as<float>();
    a.fxh = pl->fxh.as<float>();
    a.ratio = pl->ratio.as<float>();
    a.fzn = pl->fzn.as<float>();
    a.fzh = pl->fzh.as<float>();
    a.ghost = pl->ghost.as<float>();
    a.parts = pl->parts.as<float>();
    a.nstrip = (int)nstrip;
    a.nl = nl;
    a.nr = nr;
    a.kb0 = kb0;
    a.nb = nb;
    a.cx = p->dt / p->dx;
    a.cz = p->dt / p->dz;
    for (int m = 0; m < 5; ++m) a.a10[m] = a10[m];
    for (int m = 0; m < 4; ++m) a.a8[m] = a8[m];
    // a60..a62, a40, a41 (module:12-15)
    a.a6[0] = 75.0f / 64.0f;
    a.a6[1] = -25.0f / 384.0f;
    a.a6[2] = 3.0f / 640.0f;
    a.a4[0] = 9.0f / 8.0f;
    a.a4[1] = -1.0f / 24.0f;
    // "if source at free surface, don't increment strxxx" (pri:163-167)
    // ('abc' mode always increments strxxx, pri:330)
    a.src_xx = (pl->abc || p->sz > p->itopo[p->sx]) ? 1 : 0;
    a.g = pl->g.as<float>();
    a.amp = pl->amp.as<float>();
    a.dt = p->dt;
    a.dz2 = p->dz * p->dz;
    a.nrec = p->nrec;
    a.nrt = std::max(p->nt, 1);
    a.rx = pl->rx.as<int>();
    a.rz = pl->rz.as<int>();
    a.rec_vx = pl->rec_vx.as<float>();
    a.rec_vz = pl->rec_vz.as<float>();

    rc = el_set_tiling(pl, kElDefaultTX, kElDefaultNW);
    if (rc) return bail(rc);
    // A shape the context's autotuner measured for this grid class wins.
    const auto tuned = ctx->tuning.find(elastic_class(pl));
    if (tuned != ctx->tuning.end()) (void)el_set_tiling(pl, tuned->second.a, tuned->second.b);
    rc = elastic_zero(pl);
    if (rc) return bail(rc);
    e = hipStreamSynchronize(sm);
    if (e != hipSuccess) return bail(hip_check(ctx, e, "plan setup"));
    *out = pl;
    return SFD_OK;
}

int sfd_elastic_plan_set_tiling(sfd_elastic_plan* pl, int tile_cols, int waves) {
    if (!pl) return SFD_EINVAL;
    return el_set_tiling(pl, tile_cols ? tile_cols : kElDefaultTX, waves ? waves : kElDefaultNW);
}

int sfd_elastic_plan_autotune(sfd_elastic_plan* pl, int trial_steps, float* best_us) {
    if (!pl) return SFD_EINVAL;
    sfd_context* ctx = pl->ctx;
    if (trial_steps < 1) return fail(ctx, SFD_EINVAL, "trial_steps must be >= 1");
    const int n = std::min(trial_steps, pl->nsteps);
    if (n < 1) return fail(ctx, SFD_EINVAL, "the plan has no time steps to time");
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    auto run = [&](int m) { return m < 0 ? elastic_zero(pl) : elastic_enqueue(pl, m); };
    float warm = 0.f;
    int rc = time_steps(ctx, std::min(pl->nsteps, 4 * n), run, &warm);
    if (rc) return rc;
    const auto& shapes = kElShapes;
    constexpr int nshape = (int)(sizeof(kElShapes) / sizeof(kElShapes[0]));
    float us[nshape];
    for (float& u : us) u = 3.4e38f;
    for (int pass = 0; pass < 2; ++pass)
        for (int q = 0; q < nshape; ++q) {
            rc = el_set_tiling(pl, shapes[q][0], shapes[q][1]);
            float t = 0.f;
            if (!rc) rc = time_steps(ctx, n, run, &t);
            if (rc) return rc;
            us[q] = std::min(us[q], t);
        }
    int best = 0;
    for (int q = 1; q < nshape; ++q)
        if (us[q] < us[best]) best = q;
    rc = el_set_tiling(pl, shapes[best][0], shapes[best][1]);
    if (!rc) rc = elastic_zero(pl);
    if (rc) return rc;
    SFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    TunedShape t;
    t.kind = 0;
    t.a = shapes[best][0];
    t.b = shapes[best][1];
    t.us = us[best];
    ctx->tuning[elastic_class(pl)] = t;
    if (best_us) *best_us = us[best];
    return SFD_OK;
}

int sfd_elastic_plan_get_shape(const sfd_elastic_plan* pl, int* tile_cols, int* waves) {
    if (!pl || !tile_cols || !waves) return SFD_EINVAL;
    *tile_cols = pl->tx;
    *waves = pl->nw;
    return SFD_OK;
}

int sfd_elastic_plan_run(sfd_elastic_plan* pl, float* elapsed_ms) {
    if (!pl) return SFD_EINVAL;
    sfd_context* ctx = pl->ctx;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    int rc = elastic_zero(pl);
    if (rc) return rc;
    SFD_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    rc = elastic_enqueue(pl);
    if (rc) return rc;
    SFD_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    SFD_HIP(ctx, hipEventSynchronize(ctx->ev1));
    if (elapsed_ms) SFD_HIP(ctx, hipEventElapsedTime(elapsed_ms, ctx->ev0, ctx->ev1));
    rc = check_finite(ctx, pl->rec_vx.as<float>(), pl->rec_vx.bytes / 4, "the vx gather");
    if (rc) return rc;
    return check_finite(ctx, pl->rec_vz.as<float>(), pl->rec_vz.bytes / 4, "the vz gather");
}

int sfd_elastic_plan_fetch(sfd_elastic_plan* pl, float* record_vx, float* record_vz, float* snaps) {
    if (!pl) return SFD_EINVAL;
    sfd_context* ctx = pl->ctx;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    const size_t rb = (size_t)pl->nrec * pl->nt * sizeof(float);
    if (record_vx && rb)
        SFD_HIP(ctx, hipMemcpyAsync(record_vx, pl->rec_vx.ptr, rb, hipMemcpyDeviceToHost, ctx->stream));
    if (record_vz && rb)
        SFD_HIP(ctx, hipMemcpyAsync(record_vz, pl->rec_vz.ptr, rb, hipMemcpyDeviceToHost, ctx->stream));
    if (snaps && pl->nsnap > 0)
        SFD_HIP(ctx, hipMemcpyAsync(snaps, pl->snaps.ptr, pl->snaps.bytes, hipMemcpyDeviceToHost, ctx->stream));
    SFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SFD_OK;
}

int sfd_elastic_plan_destroy(sfd_elastic_plan* pl) {
    if (!pl) return SFD_EINVAL;
    (void)hipSetDevice(pl->ctx->device);
    (void)hipStreamSynchronize(pl->ctx->stream);
    delete pl;
    return SFD_OK;
}

int sfd_elastic_batch_create(sfd_context* ctx, int nshot, const sfd_elastic_params* params, const float* const* vp,
                             const float* const* vs, const float* const* rho, const float* const* wavelet,
                             sfd_elastic_batch** out) {
    if (!ctx || !out) return SFD_EINVAL;
    *out = nullptr;
    if (nshot < 1 || !params || !vp || !vs || !rho) return fail(ctx, SFD_EINVAL, "a batch needs nshot >= 1 shots");
    sfd_elastic_batch* b = new (std::nothrow) sfd_elastic_batch;
    if (!b) return fail(ctx, SFD_ENOMEM, "host allocation failed");
    b->ctx = ctx;
    for (int i = 0; i < nshot; ++i) {
        sfd_elastic_plan* pl = nullptr;
        const int rc = sfd_elastic_plan_create(ctx, &params[i], vp[i], vs[i], rho[i], wavelet ? wavelet[i] : nullptr,
                                               &pl);
        if (rc) {
            delete b;
            return rc;
        }
        b->plans.push_back(pl);
        if (i == 0) {
            b->abc = pl->abc;
            b->nsteps = pl->nsteps;
            b->delay = pl->delay;
        } else if (pl->abc != b->abc || pl->nsteps != b->nsteps || pl->delay != b->delay) {
            delete b;
            return fail(ctx, SFD_EINVAL, "shot %d: the shots of a batch share the mode, nt and the delay", i);
        }
    }
    int tx = kElDefaultTX, nw = kElDefaultNW;
    const auto tuned = ctx->tuning.find(elastic_batch_class(b));
    if (tuned != ctx->tuning.end()) {
        tx = tuned->second.a;
        nw = tuned->second.b;
    }
    const int rc = el_batch_set_tiling(b, tx, nw);
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return SFD_OK;
}

int sfd_elastic_batch_set_tiling(sfd_elastic_batch* b, int tile_cols, int waves) {
    if (!b) return SFD_EINVAL;
    SFD_HIP(b->ctx, hipSetDevice(b->ctx->device));
    return el_batch_set_tiling(b, tile_cols ? tile_cols : kElDefaultTX, waves ? waves : kElDefaultNW);
}

int sfd_elastic_batch_autotune(sfd_elastic_batch* b, int trial_steps, float* best_us) {
    if (!b) return SFD_EINVAL;
    sfd_context* ctx = b->ctx;
    if (trial_steps < 1) return fail(ctx, SFD_EINVAL, "trial_steps must be >= 1");
    const int n = std::min(trial_steps, b->nsteps);
    if (n < 1) return fail(ctx, SFD_EINVAL, "the batch has no time steps to time");
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    auto run = [&](int m) { return m < 0 ? el_batch_zero(b) : el_batch_enqueue(b, m); };
    float warm = 0.f;
    int rc = time_steps(ctx, std::min(b->nsteps, 2 * n), run, &warm);
    if (rc) return rc;
    constexpr int ns = (int)(sizeof(kElShapes) / sizeof(kElShapes[0]));
    float us[ns];
    for (float& u : us) u = 3.4e38f;
    for (int pass = 0; pass < 2; ++pass)
        for (int q = 0; q < ns; ++q) {
            rc = el_batch_set_tiling(b, kElShapes[q][0], kElShapes[q][1]);
            float t = 0.f;
            if (!rc) rc = time_steps(ctx, n, run, &t);
            if (rc) return rc;
            us[q] = std::min(us[q], t);
        }
    int best = 0;
    for (int q = 1; q < ns; ++q)
        if (us[q] < us[best]) best = q;
    rc = el_batch_set_tiling(b, kElShapes[best][0], kElShapes[best][1]);
    if (!rc) rc = el_batch_zero(b);
    if (rc) return rc;
    SFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    TunedShape t;
    t.kind = 0;
    t.a = kElShapes[best][0];
    t.b = kElShapes[best][1];
    t.us = us[best];
    ctx->tuning[elastic_batch_class(b)] = t;
    if (best_us) *best_us = us[best];
    return SFD_OK;
}

int sfd_elastic_batch_set_group_bytes(sfd_elastic_batch* b, size_t bytes) {
    if (!b) return SFD_EINVAL;
    b->group_bytes = bytes;
    return el_batch_set_tiling(b, b->tx, b->nw);
}

int sfd_elastic_batch_get_groups(const sfd_elastic_batch* b, int* ngroups) {
    if (!b || !ngroups) return SFD_EINVAL;
    *ngroups = (int)b->groups.size();
    return SFD_OK;
}

int sfd_elastic_batch_get_shape(const sfd_elastic_batch* b, int* tile_cols, int* waves) {
    if (!b || !tile_cols || !waves) return SFD_EINVAL;
    *tile_cols = b->tx;
    *waves = b->nw;
    return SFD_OK;
}

int sfd_elastic_batch_run(sfd_elastic_batch* b, float* elapsed_ms) {
    if (!b) return SFD_EINVAL;
    sfd_context* ctx = b->ctx;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    int rc = el_batch_zero(b);
    if (rc) return rc;
    SFD_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    rc = el_batch_enqueue(b);
    if (rc) return rc;
    SFD_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    SFD_HIP(ctx, hipEventSynchronize(ctx->ev1));
    if (elapsed_ms) SFD_HIP(ctx, hipEventElapsedTime(elapsed_ms, ctx->ev0, ctx->ev1));
    for (auto* pl : b->plans) {
        rc = check_finite(ctx, pl->rec_vx.as<float>(), pl->rec_vx.bytes / 4, "a vx gather of the batch");
        if (!rc) rc = check_finite(ctx, pl->rec_vz.as<float>(), pl->rec_vz.bytes / 4, "a vz gather of the batch");
        if (rc) return rc;
    }
    return SFD_OK;
}

int sfd_elastic_batch_fetch(sfd_elastic_batch* b, int shot, float* record_vx, float* record_vz, float* snaps) {
    if (!b) return SFD_EINVAL;
    if (shot < 0 || shot >= (int)b->plans.size())
        return fail(b->ctx, SFD_EINVAL, "shot %d outside the batch (%zu shots)", shot, b->plans.size());
    return sfd_elastic_plan_fetch(b->plans[shot], record_vx, record_vz, snaps);
}

int sfd_elastic_batch_destroy(sfd_elastic_batch* b) {
    if (!b) return SFD_EINVAL;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
    return SFD_OK;
}

int sfd_elastic_forward(sfd_context* ctx, const sfd_elastic_params* p, const float* vp, const float* vs,
                        const float* rho, const float* wavelet, float* record_vx, float* record_vz, float* snaps) {
    sfd_elastic_plan* pl = nullptr;
    int rc = sfd_elastic_plan_create(ctx, p, vp, vs, rho, wavelet, &pl);
    if (rc) return rc;
    rc = sfd_elastic_plan_run(pl, nullptr);
    if (!rc) rc = sfd_elastic_plan_fetch(pl, record_vx, record_vz, snaps);
    sfd_elastic_plan_destroy(pl);
    return rc;
}

}  // extern "C"
