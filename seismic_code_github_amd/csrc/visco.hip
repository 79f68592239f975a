// Visco-acoustic rotated-staggered-grid shot plan and its C ABI entry points
// (include/seisfd.h): the drop-in for the time loop of forward/Q_forward,
// single_shot_modelling (Q_multishot_fd_main_fsbc_zy_rotated.f90:169-484) with
// vs = 0.  Per-shot set-up follows the reference's own routines:
//   substagger  (Q sub:434-469)  density average and buoyancy
//   pcalpas     (Q sub:473-551)  moduli and the tau-method relaxation times
//   csponge     (Q sub:555-605)  cosine sponge tables
//   addsrc + get_gauss_filter_coe (Q sub:1340-1444)  Gaussian source footprint
// and the step is visco_step_kernel (visco_kernel.h).
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "common.h"
#include "visco_el_kernel.h"
#include "visco_kernel.h"

struct sfd_visco_plan {
    sfd_context* ctx = nullptr;
    int nx = 0, nz = 0, nt = 0, delay = 0, nsteps = 0, radius = 0, nmech = 0;
    int nrec = 0, nslot = 0, nsnap = 0, snap_fields = 0, nsnap_fields = 0;
    std::vector<int> snap_steps;
    int pad = 0;
    int nw = 0, chunk = 0;
    int cur = 0;
    int xlo = 0, xhi = 0, zlo = 0, zhi = 0;
    int sx0 = 0, sx1 = 0, sz0 = 0, sz1 = 0;
    float hx = 0, hz = 0, dt = 0;
    float a[5] = {};
    float c[sfd::kViscoMaxMech] = {};
    sfd::DeviceBuffer arena;
    sfd::Field s[2], vx[2], vz[2], r, b, ku, kr, xs, zs;
    sfd::DeviceBuffer dvx, dsx, dvz, dsz, g, amp, rec_row, rec_more, rec_buf, rec_slot, record, snaps;
    void (*kernel)(sfd::ViscoStepArgs) = nullptr;
    // Two launches per step on padded fields (visco_el_kernel.h): always for
    // vs > 0 (elastic), by default for vs = 0 (ve_on; an explicit march
    // tiling switches the vs = 0 plan to the fused march kernel).
    bool elastic = false, ve_on = false;
    sfd::ViscoElArgs ve{};
    sfd::DeviceBuffer ve_arena, ve_xs, ve_zs, ve_rx, ve_rz, rec_zz, rec_xx;
    int ve_pitch = 0, ve_rows = 0;
    int ve_tx = 8, ve_nw = 8;  // tile rows and waves of the two launches
    void (*ve_vel)(sfd::ViscoElArgs) = nullptr;
    void (*ve_str)(sfd::ViscoElArgs) = nullptr;

    float* grid(const sfd::Field& f) const { return f.as<float>() + (size_t)pad * nz; }
    size_t field_floats() const { return (size_t)(nx + 2 * pad) * nz; }
};

// Several shots of the same length stepped together (the reference's shot
// loop, Q main:124, on one GPU): each shot keeps its own plan (fields, model,
// gathers); one launch per half-step covers the tiles of every shot.
struct sfd_visco_batch {
    sfd_context* ctx = nullptr;
    std::vector<sfd_visco_plan*> plans;
    int nsteps = 0, delay = 0, nt = 0, ntile = 0;
    void (*vel)(sfd::ViscoBatchArgs) = nullptr;
    void (*str)(sfd::ViscoBatchArgs) = nullptr;
    sfd::DeviceBuffer shot_args, tile_map;
    // Groups of consecutive shots run one after another, each for all its
    // steps, so that a group's fields stay in the 256 MB MALL: shots
    // [s0, s1), tiles from tile_off (shot indices group-local).
    struct Group {
        int s0, s1, tile_off, ntile;
    };
    std::vector<Group> groups;
    size_t group_bytes = 0;  // budget of one group's fields; 0 = the default
    ~sfd_visco_batch() {
        for (auto* pl : plans) sfd_visco_plan_destroy(pl);
    }
};

namespace sfd {

__global__ void visco_gather_kernel(const float* __restrict__ buf, const int* __restrict__ slot, float* record,
                                    int nrec, int nslot, int nt) {
    const int r = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrec && k < nt) record[(size_t)r * nt + k] = buf[(size_t)k * nslot + slot[r]];
}

namespace {

template <int R, int L>
ViscoKernel visco_pick_nw(int nw) {
    switch (nw) {
    case 2: return visco_step_kernel<R, L, 2, visco_slots<R>()>;
    case 4: return visco_step_kernel<R, L, 4, visco_slots<R>()>;
    default: return nullptr;
    }
}

template <int R>
ViscoKernel visco_pick_l(int L, int nw) {
    switch (L) {
    case 1: return visco_pick_nw<R, 1>(nw);
    case 2: return visco_pick_nw<R, 2>(nw);
    case 3: return visco_pick_nw<R, 3>(nw);
    default: return nullptr;
    }
}

ViscoKernel visco_pick(int R, int L, int nw) {
    switch (R) {
    case 4: return visco_pick_l<4>(L, nw);
    case 5: return visco_pick_l<5>(L, nw);
    default: return nullptr;
    }
}

// Tile shapes of the two launches (rows x 64 columns, waves): 8 x 8 (one
// row per wave, the default and the batches' shape) and 16 x 8.
constexpr int kVeTX = 8, kVeNW = 8;
constexpr int kVeTXMax = 16;  // the padded fields' extra rows cover the tallest tile
constexpr int kVeShapes[][2] = {{8, 8}, {16, 8}};

template <int R, int TX, int NW>
bool ve_pick_shape(int L, bool elastic, void (**v)(ViscoElArgs), void (**st)(ViscoElArgs)) {
    *v = elastic ? visco_el_velocity_kernel<R, TX, NW> : visco_ac_velocity_kernel<R, TX, NW>;
    switch (L) {
    case 1: *st = elastic ? visco_el_stress_kernel<R, 1, TX, NW> : visco_ac_stress_kernel<R, 1, TX, NW>; return true;
    case 2: *st = elastic ? visco_el_stress_kernel<R, 2, TX, NW> : visco_ac_stress_kernel<R, 2, TX, NW>; return true;
    case 3: *st = elastic ? visco_el_stress_kernel<R, 3, TX, NW> : visco_ac_stress_kernel<R, 3, TX, NW>; return true;
    default: return false;
    }
}

template <int R>
bool ve_pick_r(int L, bool elastic, int tx, int nw, void (**v)(ViscoElArgs), void (**st)(ViscoElArgs)) {
    if (tx == 8 && nw == 8) return ve_pick_shape<R, 8, 8>(L, elastic, v, st);
    if (tx == 16 && nw == 8) return ve_pick_shape<R, 16, 8>(L, elastic, v, st);
    return false;
}

// The two tiled launches of the visco-elastic (elastic) or visco-acoustic
// step at tile shape tx x nw.
bool ve_pick(int R, int L, bool elastic, int tx, int nw, void (**v)(ViscoElArgs), void (**st)(ViscoElArgs)) {
    if (R == 4) return ve_pick_r<4>(L, elastic, tx, nw, v, st);
    if (R == 5) return ve_pick_r<5>(L, elastic, tx, nw, v, st);
    return false;
}

using ViscoBatchKernel = void (*)(ViscoBatchArgs);

template <int R, int L>
void ve_batch_pair(bool elastic, ViscoBatchKernel* v, ViscoBatchKernel* st) {
    *v = elastic ? visco_el_velocity_batch<R, kVeTX, kVeNW> : visco_ac_velocity_batch<R, kVeTX, kVeNW>;
    *st = elastic ? visco_el_stress_batch<R, L, kVeTX, kVeNW> : visco_ac_stress_batch<R, L, kVeTX, kVeNW>;
}

template <int R>
bool ve_batch_pick_r(int L, bool elastic, ViscoBatchKernel* v, ViscoBatchKernel* st) {
    switch (L) {
    case 1: ve_batch_pair<R, 1>(elastic, v, st); return true;
    case 2: ve_batch_pair<R, 2>(elastic, v, st); return true;
    case 3: ve_batch_pair<R, 3>(elastic, v, st); return true;
    default: return false;
    }
}

// The batched forms of ve_pick's launches.
bool ve_batch_pick(int R, int L, bool elastic, ViscoBatchKernel* v, ViscoBatchKernel* st) {
    if (R == 4) return ve_batch_pick_r<4>(L, elastic, v, st);
    if (R == 5) return ve_batch_pick_r<5>(L, elastic, v, st);
    return false;
}

int visco_pad_rows(int R) { return R == 4 ? kViscoPadRows<4> : kViscoPadRows<5>; }

constexpr float kPi = 3.1415926f;  // the reference's module constant (Q sub:4)

// pcalpas (Q sub:506-542): relaxation times of L mechanisms over the band
// [w_lo, w_hi] (rad/s) and the tau of quality factor q.  One mechanism:
// tau_sigma = 2 / (w_hi - w_lo) and tau = (1/q) [ln(1 + w^2 ts^2)] /
// [atan(w ts) - w ts / (1 + w^2 ts^2)], both brackets taken between the band
// edges.  L > 1: tau_sigma_l = 1 / w_l, w_l log-spaced over the band, and the
// least-squares tau of Blanch, Robertsson & Symes (1995):
// tau = (1/q) sum_l I0_l / (sum_l I1_l + 2 sum_{k<l} I2_kl).
void relaxation_times(int L, double wlo, double whi, double* ts) {
    if (L == 1) {
        ts[0] = 2.0 / (whi - wlo);
        return;
    }
    for (int l = 0; l < L; ++l) ts[l] = 1.0 / (wlo * std::pow(whi / wlo, (double)l / (L - 1)));
}

double tau_of_q(int L, double wlo, double whi, const double* ts, double q) {
    double s0 = 0, s1 = 0, s2 = 0;
    auto i0 = [](double w, double t) { return std::log(1 + w * w * t * t) / (2 * t); };
    auto i1 = [](double w, double t) { return (std::atan(w * t) - w * t / (1 + w * w * t * t)) / (2 * t); };
    for (int l = 0; l < L; ++l) {
        s0 += i0(whi, ts[l]) - i0(wlo, ts[l]);
        s1 += i1(whi, ts[l]) - i1(wlo, ts[l]);
        for (int k = 0; k < l; ++k) {
            const double tk = ts[k], tl = ts[l];
            auto i2 = [&](double w) { return std::atan(w * tk) / tk - std::atan(w * tl) / tl; };
            s2 += tk * tl / (tl * tl - tk * tk) * (i2(whi) - i2(wlo));
        }
    }
    return s0 / (s1 + 2 * s2) / q;
}

// csponge (Q sub:555-605) for one side of np + 1 cells, outermost first:
// spg(i) = cos(acos(R) (np - i + 1) / np) at the nodes and, shifted by half a
// cell inwards (lo side: spg1, with spg1(np+1) = 1) or outwards (hi side:
// spg2), at the stress positions.
void csponge_side(int np, float h, float Rc, bool lo, std::vector<float>& node, std::vector<float>& half) {
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

int validate(sfd_context* ctx, const sfd_visco_params* p, const float* vp, const float* qp) {
    if (!p || !vp || !qp) return fail(ctx, SFD_EINVAL, "null params, vp or qp");
    if (p->radius != 4 && p->radius != 5) return fail(ctx, SFD_EINVAL, "radius must be 4 or 5 (got %d)", p->radius);
    if (p->nmech < 1 || p->nmech > kViscoMaxMech)
        return fail(ctx, SFD_EINVAL, "nmech must be 1..%d (got %d)", kViscoMaxMech, p->nmech);
    if (p->nx < 2 || p->nz < 2) return fail(ctx, SFD_EINVAL, "grid %dx%d too small", p->nx, p->nz);
    if (p->nt < 0) return fail(ctx, SFD_EINVAL, "nt < 0");
    if (!(p->dx > 0) || !(p->dz > 0) || !(p->dt > 0)) return fail(ctx, SFD_EINVAL, "dx, dz, dt must be > 0");
    if (!(p->f0 > 0)) return fail(ctx, SFD_EINVAL, "f0 must be > 0");
    const int xs[4] = {p->sponge_xlo, p->sponge_xhi, p->sponge_zlo, p->sponge_zhi};
    for (int v : xs)
        if (v < 0) return fail(ctx, SFD_EINVAL, "sponge widths must be >= 0");
    if (p->sponge_xlo + p->sponge_xhi + 2 > p->nx || p->sponge_zlo + p->sponge_zhi + 2 > p->nz)
        return fail(ctx, SFD_EINVAL, "sponge strips overlap in a %dx%d grid", p->nx, p->nz);
    if (p->sx < 0 || p->sx >= p->nx || p->sz < 0 || p->sz >= p->nz)
        return fail(ctx, SFD_EINVAL, "source (%d,%d) outside the grid", p->sx, p->sz);
    if (p->nrec < 0 || (p->nrec > 0 && (!p->rx || !p->rz))) return fail(ctx, SFD_EINVAL, "bad receiver list");
    for (int r = 0; r < p->nrec; ++r)
        if (p->rx[r] < 0 || p->rx[r] >= p->nx || p->rz[r] < 0 || p->rz[r] >= p->nz)
            return fail(ctx, SFD_EINVAL, "receiver %d (%d,%d) outside the grid", r, p->rx[r], p->rz[r]);
    if (p->nsnap < 0 || (p->nsnap > 0 && !p->snap_steps)) return fail(ctx, SFD_EINVAL, "bad snapshot list");
    if (p->nsnap > 0 && (p->snap_fields & 7) == 0) return fail(ctx, SFD_EINVAL, "snap_fields is empty");
    return SFD_OK;
}

// Tiling: workgroups of 2 waves over z tiles of 128 - (4R - 2) output
// columns, x chunks sized so the grid holds about 2600 workgroups (10 per
// CU), never shorter than 8 rows.  Measured on MI355X for the 2081 x 1081
// grid: 2 waves x 8-16 rows runs 1.6x faster than 4 waves x 32 rows; the
// march is latency-bound, so parallel workgroups pay for the longer fill.
void visco_auto_tiling(const sfd_visco_plan* pl, int* nw, int* chunk) {
    const int R = pl->radius;
    const int w = 2;
    const int ow = 64 * w - (4 * R - 2);
    const long ntz = (pl->nz + ow - 1) / ow;
    const long target = 2600;
    int ch = (int)std::max<long>(8, (long)pl->nx * ntz / target);
    *nw = w;
    *chunk = std::min(ch, pl->nx);
}

ViscoStepArgs make_visco_args(const sfd_visco_plan* pl, int k) {
    ViscoStepArgs a{};
    const int cur = pl->cur, nxt = cur ^ 1;
    a.arena = pl->arena.as<float>();
    a.arena_bytes = (uint32_t)pl->arena.bytes;
    a.nx = pl->nx;
    a.nz = pl->nz;
    const int ow = 64 * pl->nw - (4 * pl->radius - 2);
    a.ntz = (pl->nz + ow - 1) / ow;
    a.chunk = pl->chunk;
    a.nchunk = (pl->nx + pl->chunk - 1) / pl->chunk;
    a.k = k;
    a.rec_col = (k >= pl->delay && pl->nrec > 0) ? k - pl->delay : -1;
    a.s_in = pl->grid(pl->s[cur]);
    a.s_out = pl->grid(pl->s[nxt]);
    a.vx_in = pl->grid(pl->vx[cur]);
    a.vx_out = pl->grid(pl->vx[nxt]);
    a.vz_in = pl->grid(pl->vz[cur]);
    a.vz_out = pl->grid(pl->vz[nxt]);
    a.r = pl->grid(pl->r);
    a.r_stride = pl->field_floats();
    a.b = pl->grid(pl->b);
    a.ku = pl->grid(pl->ku);
    a.kr = pl->grid(pl->kr);
    a.xlo = pl->xlo;
    a.xhi = pl->xhi;
    a.zlo = pl->zlo;
    a.zhi = pl->zhi;
    a.xs = pl->xs.as<float>();
    a.zs = pl->zs.as<float>();
    a.dvx = pl->dvx.as<float>();
    a.dsx = pl->dsx.as<float>();
    a.dvz = pl->dvz.as<float>();
    a.dsz = pl->dsz.as<float>();
    a.hx = pl->hx;
    a.hz = pl->hz;
    for (int m = 0; m < 5; ++m) a.a[m] = pl->a[m];
    a.dt = pl->dt;
    for (int l = 0; l < kViscoMaxMech; ++l) a.c[l] = pl->c[l];
    a.sx0 = pl->sx0;
    a.sx1 = pl->sx1;
    a.sz0 = pl->sz0;
    a.sz1 = pl->sz1;
    a.g = pl->g.as<float>();
    a.amp = pl->amp.as<float>();
    a.rec_row = pl->rec_row.as<int4>() + pl->pad;
    a.rec_more = pl->rec_more.as<int4>();
    a.rec_buf = pl->rec_buf.as<float>();
    a.nslot = pl->nslot;
    return a;
}

int visco_zero(sfd_visco_plan* pl) {
    sfd_context* ctx = pl->ctx;
    hipStream_t s = ctx->stream;
    // Wavefields, memory variables and strip parts lie at the front of the
    // arena, before the model fields.
    const char* base = pl->arena.as<char>();
    const size_t state_bytes = (size_t)((const char*)pl->b.ptr - base);
    SFD_HIP(ctx, hipMemsetAsync(pl->arena.ptr, 0, state_bytes, s));
    pl->cur = 0;
    return SFD_OK;
}

// Padded fields, strips, receiver tables and kernel arguments of the tiled
// two-launch step: vs > 0 (elastic: vx, vz, sxx, szz, sxz, 3L memory fields;
// model b, kup, kus, krp, krs) or vs = 0 (vx, vz, sigma, L memory fields;
// model b, ku, kr).  The sponge tables (x tables with pl->pad rows of
// padding), amp and the source footprint are on the device already.
int ve_setup(sfd_visco_plan* pl, const sfd_visco_params* p, bool elastic, const std::vector<float>& hb,
             const std::vector<float>& hku, const std::vector<float>& hkr, const std::vector<float>* hkus,
             const std::vector<float>* hkrs) {
    sfd_context* ctx = pl->ctx;
    const int nx = pl->nx, nz = pl->nz, L = pl->nmech, R = pl->radius;
    pl->ve_tx = kVeTX;
    pl->ve_nw = kVeNW;
    if (!ve_pick(R, L, elastic, pl->ve_tx, pl->ve_nw, &pl->ve_vel, &pl->ve_str))
        return fail(ctx, SFD_ENOTSUP, "no tiled visco kernel");
    pl->ve_pitch = (nz + kVeTZ - 1) / kVeTZ * kVeTZ + 2 * kVePad;
    pl->ve_rows = nx + kVeTXMax + 2 * kVePad;
    const size_t fe = (size_t)pl->ve_pitch * pl->ve_rows;  // floats per padded field
    const int nstate = elastic ? 5 + 3 * L : 3 + L, nmodel = elastic ? 5 : 3, nparts = elastic ? 5 + 3 * L : 2 + L;
    const int xw = pl->xlo + pl->xhi, zw = pl->zlo + pl->zhi;
    hipError_t e = pl->ve_arena.alloc((size_t)(nstate + nmodel) * fe * sizeof(float));
    if (e == hipSuccess) e = pl->ve_xs.alloc((size_t)nparts * xw * nz * sizeof(float));
    if (e == hipSuccess) e = pl->ve_zs.alloc((size_t)nparts * nx * zw * sizeof(float));
    if (e != hipSuccess) return hip_check(ctx, e, "tiled visco allocation");
    float* base = pl->ve_arena.as<float>();
    auto cell0 = [&](int q) { return base + (size_t)q * fe + (size_t)kVePad * pl->ve_pitch + kVePad; };
    ViscoElArgs& a = pl->ve;
    a = ViscoElArgs{};
    a.nx = nx;
    a.nz = nz;
    a.pitch = pl->ve_pitch;
    a.ntz = (nz + kVeTZ - 1) / kVeTZ;
    a.ntiles = a.ntz * ((nx + pl->ve_tx - 1) / pl->ve_tx);
    a.vx = cell0(0);
    a.vz = cell0(1);
    std::vector<const std::vector<float>*> model;
    if (elastic) {
        a.sxx = cell0(2);
        a.szz = cell0(3);
        a.sxz = cell0(4);
        a.r = cell0(5);
        a.kup = cell0(nstate + 1);
        a.kus = cell0(nstate + 2);
        a.krp = cell0(nstate + 3);
        a.krs = cell0(nstate + 4);
        model = {&hb, &hku, hkus, &hkr, hkrs};
    } else {
        a.sxx = a.szz = cell0(2);
        a.r = cell0(3);
        a.kup = cell0(nstate + 1);
        a.krp = cell0(nstate + 2);
        model = {&hb, &hku, &hkr};
    }
    a.b = cell0(nstate);
    a.rstride = fe;
    a.hx = pl->hx;
    a.hz = pl->hz;
    a.dt = pl->dt;
    for (int m = 0; m < 5; ++m) a.a[m] = m < R ? pl->a[m] : 0.f;
    for (int l = 0; l < kVeMaxMech; ++l) a.c[l] = l < L ? pl->c[l] : 0.f;
    a.xlo = pl->xlo;
    a.xhi = pl->xhi;
    a.zlo = pl->zlo;
    a.zhi = pl->zhi;
    a.xs = pl->ve_xs.as<float>();
    a.zs = pl->ve_zs.as<float>();
    a.sx0 = pl->sx0;
    a.sx1 = pl->sx1;
    a.sz0 = pl->sz0;
    a.sz1 = pl->sz1;
    a.g = pl->g.as<float>();
    a.nrec = p->nrec;
    a.nrt = std::max(p->nt, 1);
    hipStream_t sm = ctx->stream;
    // Model fields into their padded layout.
    std::vector<float> padded(fe, 0.f);
    for (int q = 0; q < nmodel && e == hipSuccess; ++q) {
        for (int i = 0; i < nx; ++i)
            std::copy(model[q]->begin() + (size_t)i * nz, model[q]->begin() + (size_t)(i + 1) * nz,
                      padded.begin() + (size_t)(i + kVePad) * pl->ve_pitch + kVePad);
        e = hipMemcpyAsync(base + (size_t)(nstate + q) * fe, padded.data(), fe * sizeof(float),
                           hipMemcpyHostToDevice, sm);
        if (e == hipSuccess) e = hipStreamSynchronize(sm);
    }
    if (e == hipSuccess && p->nrec > 0) e = upload(pl->ve_rx, p->rx, p->nrec, sm);
    if (e == hipSuccess && p->nrec > 0) e = upload(pl->ve_rz, p->rz, p->nrec, sm);
    if (e == hipSuccess && p->nrec > 0) e = pl->rec_zz.alloc((size_t)p->nrec * a.nrt * 4);
    if (e == hipSuccess && p->nrec > 0 && elastic) e = pl->rec_xx.alloc((size_t)p->nrec * a.nrt * 4);
    if (e != hipSuccess) return hip_check(ctx, e, "tiled visco set-up");
    a.dvx = pl->dvx.as<float>() + pl->pad;
    a.dsx = pl->dsx.as<float>() + pl->pad;
    a.dvz = pl->dvz.as<float>();
    a.dsz = pl->dsz.as<float>();
    a.amp = pl->amp.as<float>();
    a.rx = pl->ve_rx.as<int>();
    a.rz = pl->ve_rz.as<int>();
    a.rec_zz = pl->rec_zz.as<float>();
    a.rec_xx = elastic ? pl->rec_xx.as<float>() : nullptr;
    pl->ve_on = true;
    return SFD_OK;
}

// Autotuner grid class: radius, mechanisms, vs = 0 or not, nx and nz to the
// nearest power of 2.
std::string visco_class(const sfd_visco_plan* pl) {
    char buf[96];
    snprintf(buf, sizeof buf, "visco:R%d:L%d:%s:x%d:z%d", pl->radius, pl->nmech, pl->elastic ? "elastic" : "acoustic",
             log2_round(pl->nx), log2_round(pl->nz));
    return buf;
}

// Switches the plan to the two tiled launches at tile_rows x waves.
int ve_set_shape(sfd_visco_plan* pl, int tile_rows, int waves) {
    void (*v)(ViscoElArgs) = nullptr;
    void (*st)(ViscoElArgs) = nullptr;
    if (!ve_pick(pl->radius, pl->nmech, pl->elastic, tile_rows, waves, &v, &st))
        return fail(pl->ctx, SFD_EINVAL, "no tiled visco step of %d rows x %d waves (8 x 8 and 16 x 8 are built)",
                    tile_rows, waves);
    pl->ve_tx = tile_rows;
    pl->ve_nw = waves;
    pl->ve_vel = v;
    pl->ve_str = st;
    pl->ve.ntiles = pl->ve.ntz * ((pl->nx + tile_rows - 1) / tile_rows);
    pl->ve_on = true;
    return SFD_OK;
}

int ve_zero(sfd_visco_plan* pl) {
    sfd_context* ctx = pl->ctx;
    // State (vx, vz, the stresses, the memory fields) lies at the front of
    // the arena, before the model fields.
    const size_t state_bytes = (size_t)((const char*)pl->ve.b - pl->ve_arena.as<char>()) -
                               ((size_t)kVePad * pl->ve_pitch + kVePad) * sizeof(float);
    SFD_HIP(ctx, hipMemsetAsync(pl->ve_arena.ptr, 0, state_bytes, ctx->stream));
    if (pl->ve_xs.bytes) SFD_HIP(ctx, hipMemsetAsync(pl->ve_xs.ptr, 0, pl->ve_xs.bytes, ctx->stream));
    if (pl->ve_zs.bytes) SFD_HIP(ctx, hipMemsetAsync(pl->ve_zs.ptr, 0, pl->ve_zs.bytes, ctx->stream));
    return SFD_OK;
}

// Snapshots of the tiled step's fields after step k.
int ve_snaps(sfd_visco_plan* pl, int k, size_t* snap_next) {
    sfd_context* ctx = pl->ctx;
    const ViscoElArgs& a = pl->ve;
    const size_t nb = (size_t)pl->nz * sizeof(float), pb = (size_t)pl->ve_pitch * sizeof(float);
    while (*snap_next < pl->snap_steps.size() && pl->snap_steps[*snap_next] == k) {
        float* dst = pl->snaps.as<float>() + *snap_next * pl->nsnap_fields * (size_t)pl->nx * pl->nz;
        const float* src[6] = {a.szz, a.vx, a.vz, a.sxx, a.szz, a.sxz};  // sxz only with vs > 0
        for (int f = 0; f < 6; ++f) {
            if (!(pl->snap_fields & (1 << f))) continue;
            SFD_HIP(ctx, hipMemcpy2DAsync(dst, nb, src[f], pb, nb, pl->nx, hipMemcpyDeviceToDevice, ctx->stream));
            dst += (size_t)pl->nx * pl->nz;
        }
        ++*snap_next;
    }
    return SFD_OK;
}

// The gather of the last of ns steps (the velocity launch of the next step
// gathers every other).
int ve_last_gather(sfd_visco_plan* pl, int ns) {
    if (pl->nrec > 0 && pl->nt > 0 && ns - 1 >= pl->delay) {
        ViscoElArgs a = pl->ve;
        a.rec_col = ns - 1 - pl->delay;
        hipLaunchKernelGGL(visco_el_gather_kernel, dim3((pl->nrec + 255) / 256), dim3(256), 0, pl->ctx->stream, a);
        SFD_HIP(pl->ctx, hipGetLastError());
    }
    return SFD_OK;
}

// Enqueues the first nsteps steps (all of them by default).
int ve_enqueue(sfd_visco_plan* pl, int nsteps = -1) {
    sfd_context* ctx = pl->ctx;
    ViscoElArgs a = pl->ve;
    const dim3 grid(a.ntiles), block(64 * pl->ve_nw);
    const bool rec = pl->nrec > 0 && pl->nt > 0;
    size_t snap_next = 0;
    const int ns = nsteps < 0 ? pl->nsteps : std::min(nsteps, pl->nsteps);
    for (int k = 0; k < ns; ++k) {
        a.step = k;
        // the velocity launch of step k gathers the stresses of step k - 1
        a.rec_col = (rec && k - 1 >= pl->delay) ? k - 1 - pl->delay : -1;
        hipLaunchKernelGGL(pl->ve_vel, grid, block, 0, ctx->stream, a);
        SFD_HIP(ctx, hipGetLastError());
        hipLaunchKernelGGL(pl->ve_str, grid, block, 0, ctx->stream, a);
        SFD_HIP(ctx, hipGetLastError());
        const int rc = ve_snaps(pl, k, &snap_next);
        if (rc) return rc;
    }
    return ve_last_gather(pl, ns);
}

// A batch's steps, shot group after shot group: one launch of each
// half-step over the group's tiles.
int ve_batch_enqueue(sfd_visco_batch* b, int nsteps = -1) {
    sfd_context* ctx = b->ctx;
    const dim3 block(64 * kVeNW);
    const int ns = nsteps < 0 ? b->nsteps : std::min(nsteps, b->nsteps);
    for (const auto& g : b->groups) {
        const dim3 grid(g.ntile);
        ViscoBatchArgs a{};
        a.shots = b->shot_args.as<ViscoElArgs>() + g.s0;
        a.tile_map = b->tile_map.as<int2>() + g.tile_off;
        a.ntile = g.ntile;
        std::vector<size_t> snap_next(g.s1 - g.s0, 0);
        for (int k = 0; k < ns; ++k) {
            a.step = k;
            a.rec_col = (b->nt > 0 && k - 1 >= b->delay) ? k - 1 - b->delay : -1;
            hipLaunchKernelGGL(b->vel, grid, block, 0, ctx->stream, a);
            SFD_HIP(ctx, hipGetLastError());
            hipLaunchKernelGGL(b->str, grid, block, 0, ctx->stream, a);
            SFD_HIP(ctx, hipGetLastError());
            for (int i = g.s0; i < g.s1; ++i) {
                const int rc = ve_snaps(b->plans[i], k, &snap_next[i - g.s0]);
                if (rc) return rc;
            }
        }
        for (int i = g.s0; i < g.s1; ++i) {
            const int rc = ve_last_gather(b->plans[i], ns);
            if (rc) return rc;
        }
    }
    return SFD_OK;
}

// Default budget of one shot group's fields (state, model, strip parts):
// measured on MI355X with the Q_forward card's survey.
constexpr size_t kVeGroupBytes = (size_t)192 << 20;

// Shot groups and the tile tables (group-local shot indices).
int ve_batch_tables(sfd_visco_batch* b) {
    const int n = (int)b->plans.size();
    const size_t budget = b->group_bytes ? b->group_bytes : kVeGroupBytes;
    auto bytes = [&](int i) {
        const sfd_visco_plan* pl = b->plans[i];
        return pl->ve_arena.bytes + pl->ve_xs.bytes + pl->ve_zs.bytes;
    };
    b->groups.clear();
    for (int s0 = 0; s0 < n;) {
        size_t used = 0;
        int e = s0;
        do used += bytes(e++);
        while (e < n && used + bytes(e) <= budget);
        b->groups.push_back({s0, e, 0, 0});
        s0 = e;
    }
    std::vector<ViscoElArgs> args;
    std::vector<int2> map;
    for (auto& g : b->groups) {
        g.tile_off = (int)map.size();
        for (int i = g.s0; i < g.s1; ++i) {
            args.push_back(b->plans[i]->ve);
            for (int t = 0; t < b->plans[i]->ve.ntiles; ++t) map.push_back(make_int2(i - g.s0, t));
        }
        g.ntile = (int)map.size() - g.tile_off;
    }
    hipStream_t st = b->ctx->stream;
    hipError_t e = upload(b->shot_args, args.data(), args.size(), st);
    if (e == hipSuccess) e = upload(b->tile_map, map.data(), map.size(), st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_check(b->ctx, e, "visco batch tables");
    b->ntile = (int)map.size();
    return SFD_OK;
}

int visco_enqueue(sfd_visco_plan* pl, int k0, int nsteps) {
    sfd_context* ctx = pl->ctx;
    const int ow = 64 * pl->nw - (4 * pl->radius - 2);
    const int ntz = (pl->nz + ow - 1) / ow, nchunk = (pl->nx + pl->chunk - 1) / pl->chunk;
    const dim3 grid(ntz * nchunk), block(64 * pl->nw);
    const size_t nb = (size_t)pl->nx * pl->nz * sizeof(float);
    size_t snap_next = std::lower_bound(pl->snap_steps.begin(), pl->snap_steps.end(), k0) - pl->snap_steps.begin();
    for (int s = 0; s < nsteps; ++s) {
        const int k = k0 + s;
        const ViscoStepArgs a = make_visco_args(pl, k);
        hipLaunchKernelGGL(pl->kernel, grid, block, 0, ctx->stream, a);
        SFD_HIP(ctx, hipGetLastError());
        pl->cur ^= 1;
        while (snap_next < pl->snap_steps.size() && pl->snap_steps[snap_next] == k) {
            float* dst = pl->snaps.as<float>() + snap_next * pl->nsnap_fields * (size_t)pl->nx * pl->nz;
            const float* src[3] = {pl->grid(pl->s[pl->cur]), pl->grid(pl->vx[pl->cur]), pl->grid(pl->vz[pl->cur])};
            for (int f = 0; f < 3; ++f) {
                if (!(pl->snap_fields & (1 << f))) continue;
                SFD_HIP(ctx, hipMemcpyAsync(dst, src[f], nb, hipMemcpyDeviceToDevice, ctx->stream));
                dst += (size_t)pl->nx * pl->nz;
            }
            ++snap_next;
        }
    }
    return SFD_OK;
}

}  // namespace
}  // namespace sfd

using namespace sfd;

extern "C" {

int sfd_visco_plan_create(sfd_context* ctx, const sfd_visco_params* p, const float* vp, const float* vs,
                          const float* rho, const float* qp, const float* wavelet, sfd_visco_plan** out) {
    if (!ctx || !out) return SFD_EINVAL;
    *out = nullptr;
    int rc = validate(ctx, p, vp, qp);
    if (rc) return rc;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    const int nx = p->nx, nz = p->nz, R = p->radius, L = p->nmech;
    const size_t n = (size_t)nx * nz;
    // Any vs > 0 selects the visco-elastic kernels; vs = 0 everywhere is the
    // visco-acoustic limit (sxx = szz, sxz = 0): two tiled launches per step,
    // or the fused march kernel on request (sfd_visco_plan_set_tiling).
    bool elastic = false;
    if (vs) {
        for (size_t c = 0; c < n; ++c) {
            if (!(vs[c] >= 0.f) || !std::isfinite(vs[c])) return fail(ctx, SFD_EINVAL, "vs must be finite and >= 0 (cell %zu)", c);
            if (vs[c] != 0.0f) elastic = true;
        }
    }
    if (p->qs)
        for (size_t c = 0; c < n; ++c)
            if (!(p->qs[c] > 0.f) || !std::isfinite(p->qs[c])) return fail(ctx, SFD_EINVAL, "qs must be finite and > 0 (cell %zu)", c);
    float vmax = 0.f;
    for (size_t c = 0; c < n; ++c) {
        if (!(vp[c] > 0.f) || !std::isfinite(vp[c])) return fail(ctx, SFD_EINVAL, "vp must be finite and > 0 (cell %zu)", c);
        if (!(qp[c] > 0.f) || !std::isfinite(qp[c])) return fail(ctx, SFD_EINVAL, "qp must be finite and > 0 (cell %zu)", c);
        if (rho && (!(rho[c] > 0.f) || !std::isfinite(rho[c])))
            return fail(ctx, SFD_EINVAL, "rho must be finite and > 0 (cell %zu)", c);
        vmax = std::max(vmax, vp[c]);
    }

    sfd_visco_plan* pl = new (std::nothrow) sfd_visco_plan;
    if (!pl) return fail(ctx, SFD_ENOMEM, "host allocation failed");
    auto bail = [&](int code) {
        delete pl;
        return code;
    };
    pl->ctx = ctx;
    pl->nx = nx;
    pl->nz = nz;
    pl->nt = p->nt;
    pl->radius = R;
    pl->nmech = L;
    pl->dt = p->dt;
    pl->hx = 0.5f / p->dx;
    pl->hz = 0.5f / p->dz;
    if (p->coef) {
        for (int m = 0; m < R; ++m) pl->a[m] = p->coef[m];
    } else {
        default_coefs(R, pl->a);
    }
    // Recording starts after `delay` steps: the reference runs nt + decay
    // steps, decay = nint(nwt / 2), nwt = 2 nint(1.2 / (f0 dt)) (main:279-280).
    pl->delay = p->delay >= 0 ? p->delay : (int)std::lround(1.2f / (p->f0 * p->dt));
    pl->nsteps = pl->delay + p->nt;
    for (int s = 0; s < p->nsnap; ++s)
        if (p->snap_steps[s] < 0 || p->snap_steps[s] >= pl->nsteps || (s > 0 && p->snap_steps[s] <= p->snap_steps[s - 1]))
            return bail(fail(ctx, SFD_EINVAL, "snapshot steps must be ascending and in [0, delay + nt)"));

    // Stability of the leapfrog with the rotated-grid operator: the symbol of
    // Dx^2 + Dz^2 is at most (2 sum|a| / h)^2 with h = min(dx, dz), so
    // vmax dt / h <= 1 / sum|a|.
    double sa = 0;
    for (int m = 0; m < R; ++m) sa += std::fabs((double)pl->a[m]);
    const double courant = vmax * (double)p->dt / std::min((double)p->dx, (double)p->dz);
    if (courant * sa > 1.0)
        return bail(fail(ctx, SFD_ECFL, "unstable: vmax*dt/min(dx,dz) = %.4f exceeds %.4f", courant, 1.0 / sa));

    // ---- model (substagger, pcalpas) ----
    std::vector<float> rb(n);
    for (size_t c = 0; c < n; ++c) rb[c] = rho ? rho[c] : 2400.0f;
    for (int i = 0; i < nx - 1; ++i)
        for (int k = 0; k < nz - 1; ++k) {
            float* r0 = &rb[(size_t)i * nz + k];
            *r0 = 0.25f * (((r0[0] + r0[1]) + r0[nz + 1]) + r0[nz]);
        }
    for (int i = 0; i < nx; ++i) rb[(size_t)i * nz + nz - 1] = rb[(size_t)i * nz + nz - 2];
    for (int k = 0; k < nz; ++k) rb[(size_t)(nx - 1) * nz + k] = rb[(size_t)(nx - 2) * nz + k];
    const double wlo = 2.0 * kPi * (p->f_low > 0 ? p->f_low : 2.0f);
    const double whi = 2.0 * kPi * (p->f_high > 0 ? p->f_high : 2.5f * p->f0);
    if (!(whi > wlo)) return bail(fail(ctx, SFD_EINVAL, "tau-method band: f_high must exceed f_low"));
    const double qscale = p->q_scale > 0 ? p->q_scale : 0.9;
    double ts[kViscoMaxMech];
    relaxation_times(L, wlo, whi, ts);
    for (int l = 0; l < L; ++l) {
        pl->c[l] = (float)((double)p->dt / ts[l]);
        if (pl->c[l] >= 1.0f)
            return bail(fail(ctx, SFD_ECFL, "dt / tau_sigma = %.3f >= 1: the memory-variable update is unstable", pl->c[l]));
    }
    // tau is proportional to 1/q: evaluate the band integrals once.
    const double tau1 = tau_of_q(L, wlo, whi, ts, 1.0);
    std::vector<float> hb(n), hku(n), hkr(n), hkus, hkrs;
    if (elastic) {
        hkus.resize(n);
        hkrs.resize(n);
    }
    for (size_t c = 0; c < n; ++c) {
        const double m = (double)rb[c] * (double)vp[c] * (double)vp[c];
        const double tau = tau1 / ((double)qp[c] * qscale);
        hb[c] = (float)((double)p->dt / rb[c]);
        hku[c] = (float)((double)p->dt * m * (1.0 + L * tau));
        hkr[c] = (float)(m * tau);
        if (elastic) {
            // mu = rho_avg vs^2 with the S-wave tau from qs (qp where qs is not given)
            const double mu = (double)rb[c] * (double)vs[c] * (double)vs[c];
            const double taus = tau1 / ((double)(p->qs ? p->qs[c] : qp[c]) * qscale);
            hkus[c] = (float)((double)p->dt * mu * (1.0 + L * taus));
            hkrs[c] = (float)(mu * taus);
        }
    }

    // ---- sponge profiles (csponge) ----
    pl->xlo = p->sponge_xlo + 1;
    pl->xhi = p->sponge_xhi + 1;
    pl->zlo = p->sponge_zlo + 1;
    pl->zhi = p->sponge_zhi + 1;
    const float Rc = p->sponge_r > 0 ? p->sponge_r : 1e-3f;
    std::vector<float> xn, xn1, xp, xp2, zn, zn1, zp, zp2;
    csponge_side(p->sponge_xlo, p->dx, Rc, true, xn, xn1);
    csponge_side(p->sponge_xhi, p->dx, Rc, false, xp, xp2);
    csponge_side(p->sponge_zlo, p->dz, Rc, true, zn, zn1);
    csponge_side(p->sponge_zhi, p->dz, Rc, false, zp, zp2);
    const int pad = visco_pad_rows(R);
    pl->pad = pad;
    std::vector<float> hdvx(nx + 2 * pad, 1.0f), hdsx(nx + 2 * pad, 1.0f), hdvz(nz, 1.0f), hdsz(nz, 1.0f);
    for (int i = 0; i < pl->xlo; ++i) {
        hdvx[pad + i] = xn[i];
        hdsx[pad + i] = xn1[i];
    }
    for (int i = 0; i < pl->xhi; ++i) {
        hdvx[pad + nx - 1 - i] = xp[i];
        hdsx[pad + nx - 1 - i] = xp2[i];
    }
    for (int k = 0; k < pl->zlo; ++k) {
        hdvz[k] = zn[k];
        hdsz[k] = zn1[k];
    }
    for (int k = 0; k < pl->zhi; ++k) {
        hdvz[nz - 1 - k] = zp[k];
        hdsz[nz - 1 - k] = zp2[k];
    }

    // ---- Gaussian source footprint (get_gauss_filter_coe over the whole
    // grid, n1 = nz, n2 = nx, normalised by its float sum; placed by addsrc
    // so that weight (idx1, idx2) lands on cell (sz - (hw1-1) + idx1 - 1,
    // sx - (hw2-1) + idx2 - 1)) ----
    {
        const int n1 = nz, n2 = nx;
        const float sigma = p->src_sigma > 0 ? p->src_sigma : 1.5f, sigma2 = sigma * sigma;
        const int ixcen = (n1 + 1) / 2, izcen = (n2 + 1) / 2;
        std::vector<float> g2((size_t)n1 * n2);
        float sum = 0.f;
        for (int j = 1; j <= n2; ++j)
            for (int i = 1; i <= n1; ++i) {
                const float dis = (float)((j - izcen) * (j - izcen) + (i - ixcen) * (i - ixcen));
                g2[(size_t)(j - 1) * n1 + (i - 1)] = 1.0f / (2.0f * kPi * sigma2) * std::exp(-dis / (2.0f * sigma2));
            }
        for (size_t c = 0; c < g2.size(); ++c) sum += g2[c];
        const int hw1 = (int)std::lround((n1 + 1) / 2.0), hw2 = (int)std::lround((n2 + 1) / 2.0);
        // Bounding box (grid cells) of the nonzero weights.
        int x0 = nx, x1 = -1, z0 = nz, z1 = -1;
        auto weight = [&](int x, int z) -> float {
            const int idx2 = x - p->sx + hw2, idx1 = z - p->sz + hw1;
            if (idx1 < 1 || idx1 > n1 || idx2 < 1 || idx2 > n2) return 0.f;
            return g2[(size_t)(idx2 - 1) * n1 + (idx1 - 1)] / sum;
        };
        for (int x = 0; x < nx; ++x)
            for (int z = 0; z < nz; ++z)
                if (weight(x, z) != 0.f) {
                    x0 = std::min(x0, x);
                    x1 = std::max(x1, x);
                    z0 = std::min(z0, z);
                    z1 = std::max(z1, z);
                }
        if (x1 < 0) {
            x0 = z0 = 0;
            x1 = z1 = -1;
        }
        pl->sx0 = x0;
        pl->sx1 = x1 + 1;
        pl->sz0 = z0;
        pl->sz1 = z1 + 1;
        std::vector<float> hg(std::max<size_t>(1, (size_t)(pl->sx1 - pl->sx0) * (pl->sz1 - pl->sz0)), 0.f);
        for (int x = pl->sx0; x < pl->sx1; ++x)
            for (int z = pl->sz0; z < pl->sz1; ++z)
                hg[(size_t)(x - pl->sx0) * (pl->sz1 - pl->sz0) + (z - pl->sz0)] = weight(x, z);
        hipError_t e = upload(pl->g, hg.data(), hg.size(), ctx->stream);
        if (e != hipSuccess) return bail(hip_check(ctx, e, "source footprint"));
    }
    // Wavelet per step k (Fortran step it = k + 1): the given samples or
    // addsrc's exp(-(pi f0 (it dt - 1.2/f0))^2) (1 - 2 (pi f0 (it dt - 1.2/f0))^2),
    // times dt / dz^2.
    std::vector<float> hamp(std::max(pl->nsteps, 1), 0.f);
    for (int k = 0; k < pl->nsteps; ++k) {
        float w;
        if (wavelet) {
            w = wavelet[k];
        } else {
            const float x = kPi * p->f0 * ((float)(k + 1) * p->dt - 1.2f / p->f0);
            w = std::exp(-(x * x)) * (1.0f - 2.0f * (x * x));
        }
        hamp[k] = w * p->dt / (p->dz * p->dz);
    }

    if (elastic) {
        pl->elastic = true;
        pl->nrec = p->nrec;
        pl->nsnap = p->nsnap;
        pl->snap_fields = p->snap_fields & 63;
        pl->nsnap_fields = __builtin_popcount(pl->snap_fields);
        pl->snap_steps.assign(p->snap_steps, p->snap_steps + p->nsnap);
        hipStream_t sm = ctx->stream;
        hipError_t e2 = upload(pl->dvx, hdvx.data(), hdvx.size(), sm);
        if (e2 == hipSuccess) e2 = upload(pl->dsx, hdsx.data(), hdsx.size(), sm);
        if (e2 == hipSuccess) e2 = upload(pl->dvz, hdvz.data(), hdvz.size(), sm);
        if (e2 == hipSuccess) e2 = upload(pl->dsz, hdsz.data(), hdsz.size(), sm);
        if (e2 == hipSuccess) e2 = upload(pl->amp, hamp.data(), hamp.size(), sm);
        if (e2 == hipSuccess && p->nsnap > 0) e2 = pl->snaps.alloc((size_t)p->nsnap * pl->nsnap_fields * n * 4);
        if (e2 != hipSuccess) return bail(hip_check(ctx, e2, "visco-elastic set-up"));
        rc = ve_setup(pl, p, true, hb, hku, hkr, &hkus, &hkrs);
        if (rc) return bail(rc);
        // A tile shape the context's autotuner measured for this grid class.
        const auto tuned = ctx->tuning.find(visco_class(pl));
        if (tuned != ctx->tuning.end() && tuned->second.kind == 0)
            (void)ve_set_shape(pl, tuned->second.a, tuned->second.b);
        rc = ve_zero(pl);
        if (rc) return bail(rc);
        e2 = hipStreamSynchronize(sm);
        if (e2 != hipSuccess) return bail(hip_check(ctx, e2, "plan setup"));
        *out = pl;
        return SFD_OK;
    }

    // ---- receivers ----
    pl->nrec = p->nrec;
    std::vector<int> hslot;
    std::vector<int4> hrow, hmore;
    receiver_spans(nx, pad, p->nrec, p->rx, p->rz, hslot, hrow, hmore, &pl->nslot);

    // ---- device arena: state (s, vx, vz ping-pong, r_l), then strips, then model ----
    hipStream_t st = ctx->stream;
    hipError_t e = hipSuccess;
    const size_t fb = pl->field_floats() * 4;
    const size_t xsb = (size_t)(1 + L) * (pl->xlo + pl->xhi) * nz * 4;
    const size_t zsb = (size_t)(1 + L) * nx * (pl->zlo + pl->zhi) * 4;
    for (int pass = 0; pass < 2; ++pass) {
        Carver cv;
        if (pass == 1) cv.base = pl->arena.as<char>();
        for (int b2 = 0; b2 < 2; ++b2) {
            pl->s[b2] = cv.take(fb);
            pl->vx[b2] = cv.take(fb);
            pl->vz[b2] = cv.take(fb);
        }
        pl->r = cv.take(fb * L);
        pl->xs = cv.take(xsb);
        pl->zs = cv.take(zsb);
        pl->b = cv.take(fb);
        pl->ku = cv.take(fb);
        pl->kr = cv.take(fb);
        if (pass == 0) {
            if (cv.used >= (size_t)1 << 31)
                return bail(fail(ctx, SFD_ENOTSUP, "grid needs %zu bytes of fields; the step kernel addresses < 2 GiB", cv.used));
            e = pl->arena.alloc(cv.used);
            if (e != hipSuccess) return bail(hip_check(ctx, e, "arena allocation"));
        }
    }
    // r_l fields must be field_floats() apart: the carver keeps one block of L.
    auto put_model = [&](const Field& f, const std::vector<float>& h) -> hipError_t {
        std::vector<float> padded(pl->field_floats(), 0.f);
        std::copy(h.begin(), h.end(), padded.begin() + (size_t)pad * nz);
        hipError_t e2 = hipMemcpyAsync(f.ptr, padded.data(), f.bytes, hipMemcpyHostToDevice, st);
        if (e2 == hipSuccess) e2 = hipStreamSynchronize(st);
        return e2;
    };
    if (e == hipSuccess) e = put_model(pl->b, hb);
    if (e == hipSuccess) e = put_model(pl->ku, hku);
    if (e == hipSuccess) e = put_model(pl->kr, hkr);
    if (e == hipSuccess) e = upload(pl->dvx, hdvx.data(), hdvx.size(), st);
    if (e == hipSuccess) e = upload(pl->dsx, hdsx.data(), hdsx.size(), st);
    if (e == hipSuccess) e = upload(pl->dvz, hdvz.data(), hdvz.size(), st);
    if (e == hipSuccess) e = upload(pl->dsz, hdsz.data(), hdsz.size(), st);
    if (e == hipSuccess) e = upload(pl->amp, hamp.data(), hamp.size(), st);
    if (e == hipSuccess) e = upload(pl->rec_row, hrow.data(), hrow.size(), st);
    if (e == hipSuccess && !hmore.empty()) e = upload(pl->rec_more, hmore.data(), hmore.size(), st);
    if (e == hipSuccess && p->nrec > 0) e = upload(pl->rec_slot, hslot.data(), hslot.size(), st);
    if (e == hipSuccess && p->nrec > 0) e = pl->rec_buf.alloc((size_t)pl->nslot * std::max(p->nt, 1) * 4);
    if (e == hipSuccess && p->nrec > 0) e = pl->record.alloc((size_t)p->nrec * std::max(p->nt, 1) * 4);
    pl->nsnap = p->nsnap;
    pl->snap_fields = p->snap_fields & 7;
    pl->nsnap_fields = (pl->snap_fields & 1) + ((pl->snap_fields >> 1) & 1) + ((pl->snap_fields >> 2) & 1);
    pl->snap_steps.assign(p->snap_steps, p->snap_steps + p->nsnap);
    if (e == hipSuccess && p->nsnap > 0) e = pl->snaps.alloc((size_t)p->nsnap * pl->nsnap_fields * n * 4);
    if (e != hipSuccess) return bail(hip_check(ctx, e, "plan allocation"));

    visco_auto_tiling(pl, &pl->nw, &pl->chunk);
    pl->kernel = visco_pick(R, L, pl->nw);
    if (!pl->kernel) return bail(fail(ctx, SFD_ENOTSUP, "no kernel for this configuration"));
    // The tiled two-launch step runs by default.
    rc = ve_setup(pl, p, false, hb, hku, hkr, nullptr, nullptr);
    if (rc) return bail(rc);
    // A shape the context's autotuner measured for this grid class: a tile
    // shape of the tiled step, or a march tiling.
    const auto tuned = ctx->tuning.find(visco_class(pl));
    if (tuned != ctx->tuning.end() && tuned->second.kind == 0)
        (void)ve_set_shape(pl, tuned->second.a, tuned->second.b);
    if (tuned != ctx->tuning.end() && tuned->second.kind == 1)
        (void)sfd_visco_plan_set_tiling(pl, tuned->second.a, tuned->second.b);
    rc = ve_zero(pl);
    if (!rc) rc = visco_zero(pl);
    if (rc) return bail(rc);
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return bail(hip_check(ctx, e, "plan setup"));
    *out = pl;
    return SFD_OK;
}

int sfd_visco_plan_set_tiled(sfd_visco_plan* pl, int tile_rows, int waves) {
    if (!pl) return SFD_EINVAL;
    if (tile_rows == 0 && waves == 0) {
        tile_rows = kVeTX;
        waves = kVeNW;
    }
    return ve_set_shape(pl, tile_rows, waves);
}

int sfd_visco_plan_set_tiling(sfd_visco_plan* pl, int waves, int chunk_rows) {
    if (!pl) return SFD_EINVAL;
    if (pl->elastic) return SFD_OK;  // the visco-elastic step has the tiled launches only
    // vs = 0: automatic = the tiled step; an explicit tiling selects the
    // fused march kernel.
    pl->ve_on = waves == 0 && chunk_rows == 0;
    int nw = pl->nw, ch = pl->chunk;
    if (waves != 0 || chunk_rows != 0) {
        visco_auto_tiling(pl, &nw, &ch);
        if (waves) nw = waves;
        if (chunk_rows) ch = chunk_rows;
    } else {
        visco_auto_tiling(pl, &nw, &ch);
    }
    if (nw != 2 && nw != 4) return fail(pl->ctx, SFD_EINVAL, "waves must be 2 or 4 (got %d)", nw);
    if (ch < 1) return fail(pl->ctx, SFD_EINVAL, "chunk_rows must be >= 1");
    auto k = visco_pick(pl->radius, pl->nmech, nw);
    if (!k) return fail(pl->ctx, SFD_ENOTSUP, "no kernel for this tiling");
    pl->nw = nw;
    pl->chunk = std::min(ch, pl->nx);
    pl->kernel = k;
    return SFD_OK;
}

int sfd_visco_plan_autotune(sfd_visco_plan* pl, int trial_steps, float* best_us) {
    if (!pl) return SFD_EINVAL;
    sfd_context* ctx = pl->ctx;
    if (trial_steps < 1) return fail(ctx, SFD_EINVAL, "trial_steps must be >= 1");
    const int n = std::min(trial_steps, pl->nsteps);
    if (n < 1) return fail(ctx, SFD_EINVAL, "the plan has no time steps to time");
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    auto run = [&](int m) {
        if (m < 0) return pl->ve_on ? ve_zero(pl) : visco_zero(pl);
        return pl->ve_on ? ve_enqueue(pl, m) : visco_enqueue(pl, 0, m);
    };
    // Candidates: the tiled step at each built tile shape, and with vs = 0
    // the march kernel at 2 and 4 waves with its automatic chunk.
    std::vector<TunedShape> cand;
    for (auto& sh : kVeShapes) {
        TunedShape t;
        t.kind = 0;
        t.a = sh[0];
        t.b = sh[1];
        cand.push_back(t);
    }
    if (!pl->elastic)
        for (int w : {2, 4}) {
            TunedShape t;
            t.kind = 1;
            t.a = w;
            cand.push_back(t);
        }
    auto apply = [&](const TunedShape& t) {
        return t.kind == 0 ? ve_set_shape(pl, t.a, t.b) : sfd_visco_plan_set_tiling(pl, t.a, t.b);
    };
    float warm = 0.f;
    int rc = time_steps(ctx, std::min(pl->nsteps, 4 * n), run, &warm);
    if (rc) return rc;
    for (auto& t : cand) t.us = 3.4e38f;
    for (int pass = 0; pass < 2; ++pass)
        for (TunedShape& t : cand) {
            rc = apply(t);
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
    rc = apply(best);
    if (!rc) rc = pl->ve_on ? ve_zero(pl) : visco_zero(pl);
    if (rc) return rc;
    SFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->tuning[visco_class(pl)] = best;
    if (best_us) *best_us = best.us;
    return SFD_OK;
}

int sfd_visco_plan_get_shape(const sfd_visco_plan* pl, int* kind, int* a, int* b) {
    if (!pl || !kind || !a || !b) return SFD_EINVAL;
    *kind = pl->ve_on ? 0 : 1;
    *a = pl->ve_on ? pl->ve_tx : pl->nw;
    *b = pl->ve_on ? pl->ve_nw : pl->chunk;
    return SFD_OK;
}

int sfd_visco_plan_run(sfd_visco_plan* pl, float* elapsed_ms) {
    if (!pl) return SFD_EINVAL;
    sfd_context* ctx = pl->ctx;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    int rc = pl->ve_on ? ve_zero(pl) : visco_zero(pl);
    if (rc) return rc;
    SFD_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    rc = pl->ve_on ? ve_enqueue(pl) : visco_enqueue(pl, 0, pl->nsteps);
    if (rc) return rc;
    SFD_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    SFD_HIP(ctx, hipEventSynchronize(ctx->ev1));
    if (elapsed_ms) SFD_HIP(ctx, hipEventElapsedTime(elapsed_ms, ctx->ev0, ctx->ev1));
    const sfd::DeviceBuffer& rec = pl->ve_on ? pl->rec_zz : pl->rec_buf;
    return check_finite(ctx, rec.as<float>(), rec.bytes / 4, "the gather");
}

int sfd_visco_plan_fetch(sfd_visco_plan* pl, float* record, float* snaps) {
    if (!pl) return SFD_EINVAL;
    sfd_context* ctx = pl->ctx;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    if (pl->ve_on) {
        if (record && pl->nrec > 0 && pl->nt > 0)
            SFD_HIP(ctx, hipMemcpyAsync(record, pl->rec_zz.ptr, (size_t)pl->nrec * pl->nt * 4, hipMemcpyDeviceToHost,
                                        ctx->stream));
        if (snaps && pl->nsnap > 0)
            SFD_HIP(ctx, hipMemcpyAsync(snaps, pl->snaps.ptr, pl->snaps.bytes, hipMemcpyDeviceToHost, ctx->stream));
        SFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return SFD_OK;
    }
    if (record && pl->nrec > 0 && pl->nt > 0) {
        const dim3 block(256), grid((pl->nt + 255) / 256, pl->nrec);
        hipLaunchKernelGGL(visco_gather_kernel, grid, block, 0, ctx->stream, pl->rec_buf.as<float>(),
                           pl->rec_slot.as<int>(), pl->record.as<float>(), pl->nrec, pl->nslot, pl->nt);
        SFD_HIP(ctx, hipGetLastError());
        SFD_HIP(ctx, hipMemcpyAsync(record, pl->record.ptr, (size_t)pl->nrec * pl->nt * 4, hipMemcpyDeviceToHost,
                                    ctx->stream));
    }
    if (snaps && pl->nsnap > 0)
        SFD_HIP(ctx, hipMemcpyAsync(snaps, pl->snaps.ptr, pl->snaps.bytes, hipMemcpyDeviceToHost, ctx->stream));
    SFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SFD_OK;
}

int sfd_visco_plan_fetch_xx(sfd_visco_plan* pl, float* record_xx) {
    if (!pl) return SFD_EINVAL;
    if (!pl->elastic) return sfd_visco_plan_fetch(pl, record_xx, nullptr);  // sxx = szz when vs = 0
    sfd_context* ctx = pl->ctx;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    if (record_xx && pl->nrec > 0 && pl->nt > 0)
        SFD_HIP(ctx, hipMemcpyAsync(record_xx, pl->rec_xx.ptr, (size_t)pl->nrec * pl->nt * 4, hipMemcpyDeviceToHost,
                                    ctx->stream));
    SFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SFD_OK;
}

int sfd_visco_plan_destroy(sfd_visco_plan* pl) {
    if (!pl) return SFD_EINVAL;
    (void)hipSetDevice(pl->ctx->device);
    (void)hipStreamSynchronize(pl->ctx->stream);
    delete pl;
    return SFD_OK;
}

int sfd_visco_batch_create(sfd_context* ctx, int nshot, const sfd_visco_params* params, const float* const* vp,
                           const float* const* vs, const float* const* rho, const float* const* qp,
                           const float* const* wavelet, sfd_visco_batch** out) {
    if (!ctx || !out) return SFD_EINVAL;
    *out = nullptr;
    if (nshot < 1 || !params || !vp || !qp) return fail(ctx, SFD_EINVAL, "a batch needs nshot >= 1 shots");
    sfd_visco_batch* b = new (std::nothrow) sfd_visco_batch;
    if (!b) return fail(ctx, SFD_ENOMEM, "host allocation failed");
    b->ctx = ctx;
    for (int i = 0; i < nshot; ++i) {
        sfd_visco_plan* pl = nullptr;
        int rc = sfd_visco_plan_create(ctx, &params[i], vp[i], vs ? vs[i] : nullptr, rho ? rho[i] : nullptr, qp[i],
                                       wavelet ? wavelet[i] : nullptr, &pl);
        if (rc) {
            delete b;
            return rc;
        }
        b->plans.push_back(pl);
        const sfd_visco_plan* p0 = b->plans[0];
        if (pl->nsteps != p0->nsteps || pl->delay != p0->delay || pl->nt != p0->nt || pl->radius != p0->radius ||
            pl->nmech != p0->nmech || pl->elastic != p0->elastic) {
            delete b;
            return fail(ctx, SFD_EINVAL,
                        "shot %d: the shots of a batch share nt, the delay, the radius, L and whether vs > 0", i);
        }
        // The batch runs the tiled step at its one tile shape (a cached march
        // tiling or tile shape does not apply).
        rc = ve_set_shape(pl, kVeTX, kVeNW);
        if (rc) {
            delete b;
            return rc;
        }
    }
    const sfd_visco_plan* p0 = b->plans[0];
    b->nsteps = p0->nsteps;
    b->delay = p0->delay;
    b->nt = p0->nt;
    if (!ve_batch_pick(p0->radius, p0->nmech, p0->elastic, &b->vel, &b->str)) {
        delete b;
        return fail(ctx, SFD_ENOTSUP, "no batched visco kernel for R = %d, L = %d", p0->radius, p0->nmech);
    }
    const int rc = ve_batch_tables(b);
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return SFD_OK;
}

int sfd_visco_batch_set_group_bytes(sfd_visco_batch* b, size_t bytes) {
    if (!b) return SFD_EINVAL;
    b->group_bytes = bytes;
    return ve_batch_tables(b);
}

int sfd_visco_batch_get_groups(const sfd_visco_batch* b, int* ngroups) {
    if (!b || !ngroups) return SFD_EINVAL;
    *ngroups = (int)b->groups.size();
    return SFD_OK;
}

int sfd_visco_batch_run(sfd_visco_batch* b, float* elapsed_ms) {
    if (!b) return SFD_EINVAL;
    sfd_context* ctx = b->ctx;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    for (auto* pl : b->plans) {
        const int rc = ve_zero(pl);
        if (rc) return rc;
    }
    SFD_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    int rc = ve_batch_enqueue(b);
    if (rc) return rc;
    SFD_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    SFD_HIP(ctx, hipEventSynchronize(ctx->ev1));
    if (elapsed_ms) SFD_HIP(ctx, hipEventElapsedTime(elapsed_ms, ctx->ev0, ctx->ev1));
    for (auto* pl : b->plans) {
        rc = check_finite(ctx, pl->rec_zz.as<float>(), pl->rec_zz.bytes / 4, "a gather of the batch");
        if (rc) return rc;
    }
    return SFD_OK;
}

int sfd_visco_batch_fetch(sfd_visco_batch* b, int shot, float* record, float* record_xx, float* snaps) {
    if (!b) return SFD_EINVAL;
    if (shot < 0 || shot >= (int)b->plans.size())
        return fail(b->ctx, SFD_EINVAL, "shot %d outside the batch (%zu shots)", shot, b->plans.size());
    int rc = sfd_visco_plan_fetch(b->plans[shot], record, snaps);
    if (!rc && record_xx) rc = sfd_visco_plan_fetch_xx(b->plans[shot], record_xx);
    return rc;
}

int sfd_visco_batch_destroy(sfd_visco_batch* b) {
    if (!b) return SFD_EINVAL;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
    return SFD_OK;
}

int sfd_visco_forward(sfd_context* ctx, const sfd_visco_params* p, const float* vp, const float* vs,
                      const float* rho, const float* qp, const float* wavelet, float* record, float* snaps) {
    sfd_visco_plan* pl = nullptr;
    int rc = sfd_visco_plan_create(ctx, p, vp, vs, rho, qp, wavelet, &pl);
    if (rc) return rc;
    rc = sfd_visco_plan_run(pl, nullptr);
    if (!rc) rc = sfd_visco_plan_fetch(pl, record, snaps);
    sfd_visco_plan_destroy(pl);
    return rc;
}

}  // extern "C"
