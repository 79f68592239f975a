// VTI qP-acoustic propagator: the p / q scheme of
// de-multiple/convolution_srm/vti_qa_2d_c3_demo1 (extrapolation_one_step_1 and
// _2, VTI_multishot_qacoustic_fd_subroutine_c3.f90:432-1258, time loop of
// single_shot_modelling, VTI_multishot_qacoustic_fd_main_c3.f90:429-520).
//
//   p' = 2p - p_ + dt^2 vz^2 ((1 + 2 delta) Dxx (p + q) + Dzz p)
//   q' = 2q - q_ + dt^2 vz^2 2 (eps - delta) Dxx (p + q)
//
// 10th-order second derivatives (5 weights), split PML: in the left / right
// strips (all rows) p = U1 + U2 + U3 and q = Q1 + Q2 are advanced part by
// part with the x damping d(vx) and its derivative, in the top / bottom
// strips (columns between them) p = U1 + U2 + U3 with the z damping d(vz)
// while q keeps the interior formula.
//
// MI355X mapping.  Both routines read only the current level, so one launch
// per step advances p, q and every PML part.  Every x operator acts on
// s = p + q and every z operator on p, so a workgroup stages s over its tile
// and 5-column x halo and p over its tile and 5-row z halo in LDS, once per
// step.  Tiles are 64 rows (one per lane) x TX columns, each wave owning
// TX / NW whole columns, so a wave is either all x-strip or all not; tiles
// are dealt to the XCDs in bands of whole columns (halo re-reads hit the
// XCD's own L2).  A batch of shots on grids of one shape runs as one launch
// per step (blockIdx.y = shot).
//
// The source sample of step k is added to p at the source cell as p is
// staged, and the previous step's sample is added to the old level read
// there, so the fields in memory are the reference's arrays after each step
// and receivers read them directly (the next launch records the previous
// step's output).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "common.h"
#include "device_util.h"

namespace sfd {
namespace {

constexpr int kVTZ = 64;     // tile rows (one per lane)
constexpr int kVH = 5;       // stencil half-width
constexpr int kVPad = 8;     // zero cells around the grid in every field
constexpr int kVTXMax = 64;  // widest tile; the field's extra columns cover its staging

struct VtiArgs {
    int nx, nz, pitch, ntz, ntiles;
    long long fs;                 // floats between two shots' copies of a field
    const float *uc, *qc;         // current level (cell (0, 0) of shot 0)
    float *un, *qn;               // old level in, new level out
    const float *vz, *ep, *de;    // model
    float *xs, *zs;               // PML parts: [shot][12][nxs], [shot][7][nzs]
    int nxs, nzs;
    int nl, nr, nu, nd, par;      // strip widths; par: slot of a two-level part holding the old level
    float dx, dz, dt, dt2, idx2, idz2, lnr;
    float c2[5], c1[5];
    int sx, sz;
    float w_cur, w_prev;          // source samples of this and the previous step
    int nrec, nt, sample;         // sample < 0: nothing to record
    const int *rx, *rz;
    float* rec;                   // [shot][nrec][nt]
    int shot0;                    // shot of blockIdx.y = 0 (shot groups)
};

// p of the previous step's output at the receivers.
__device__ __forceinline__ void vti_record(const VtiArgs& a, int shot, int tid, int nthr) {
    if (a.sample < 0) return;
    const float* u = a.uc + shot * a.fs;
    for (int r = tid; r < a.nrec; r += nthr)
        a.rec[((size_t)shot * a.nrec + r) * a.nt + a.sample] = u[(ptrdiff_t)a.rx[r] * a.pitch + a.rz[r]];
}

__global__ void vti_gather_kernel(VtiArgs a) { vti_record(a, a.shot0 + blockIdx.y, threadIdx.x, blockDim.x); }

// The split-PML recursions (sub:651-663 and 1141-1153), slots [old, cur] of
// P1 (slot base 0) and Pt (slot base 2), P2 single-level (slot 4); the old
// slot takes the new value.  S2 = the second x (z) derivative, S1 the first;
// returns P1 + P2.  Operand order as the reference writes it.  P views one
// shot's parts, slot q at byte offset q * nb, this lane's cell at byte vs.
__device__ __forceinline__ float vti_pml(const Buf& P, uint32_t nb, int vs, int o, int c, float k, float S2, float S1,
                                         float d, float dd, float h, float dt, float dt2) {
    const float h2 = h * h;
    const float p1 = (P.ld(vs, (uint32_t)c * nb) * (2.f * h2 + 2.f * d * dt * h2 - d * d * h2 * dt2) -
                      h2 * P.ld(vs, (uint32_t)o * nb) + k * dt2 * (S2 * h2)) /
                     (h2 + 2.f * d * dt * h2);
    const float pt = (P.ld(vs, (uint32_t)(2 + c) * nb) * (2.f * h + 2.f * d * dt * h - dt2 * h * (d * d)) -
                      h * P.ld(vs, (uint32_t)(2 + o) * nb) - k * dt2 * dd * (S1 * h)) /
                     (h + 2.f * d * dt * h);
    const float p2 = P.ld(vs, 4 * nb) * (1.f - d * dt) + pt * dt;
    P.st(p1, vs, (uint32_t)o * nb);
    P.st(pt, vs, (uint32_t)(2 + o) * nb);
    P.st(p2, vs, 4 * nb);
    return p1 + p2;
}

// Damping at signed distance a into a strip of thickness L (sub:645-649).
__device__ __forceinline__ void vti_damp(float v, float a, float L, float lnr, float* d, float* dd) {
    const float g = 3.f * v / (2.f * L);
    *d = g * ((a / L) * (a / L)) * lnr;
    *dd = g * (2.f * a / (L * L)) * lnr;
}

// Fields are read and written through raw buffer views of the shot's copy
// from its padded origin (cell (-kVPad, -kVPad)): a lane's offset is its row
// (z), a column's offset a wave-uniform scalar, so a cell costs no 64-bit
// address arithmetic.  Own-cell loads of columns past nx and rows past nz
// read the padding and are not used.
template <int TX, int NW>
__global__ __launch_bounds__(64 * NW) void vti_step_kernel(VtiArgs a) {
    constexpr int WX = TX + 2 * kVH, HZ = kVTZ + 2 * kVH, T = 64 * NW;
    __shared__ float s_s[WX * kVTZ];  // p + q: tile columns and the x halo
    __shared__ float s_u[TX * HZ];    // p: tile columns, rows with the z halo
    __shared__ float s_q[TX * kVTZ];  // q: tile
    const int shot = a.shot0 + blockIdx.y;
    if (blockIdx.x == 0) vti_record(a, shot, threadIdx.x, T);
    const int t = xcd_tile(blockIdx.x, a.ntiles);
    const int z0 = (t % a.ntz) * kVTZ, x0 = (t / a.ntz) * TX;
    const int tz = threadIdx.x & 63, vl = tz * 4;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int k = z0 + tz;
    const int n_x = a.nx;
    const bool inz = k >= a.nu && k < a.nz - a.nd;
    const long long pad0 = shot * a.fs - ((long long)kVPad * a.pitch + kVPad);
    const Buf UC = field_buf(a.uc + pad0), QC = field_buf(a.qc + pad0);
    const Buf UN = field_buf(a.un + pad0), QN = field_buf(a.qn + pad0);
    const uint32_t cb = (uint32_t)a.pitch * 4u;  // bytes per column
    auto colo = [&](int col) { return (uint32_t)(col + kVPad) * cb + (uint32_t)(z0 + kVPad) * 4u; };  // (col, z0)
    // The cells' own old level and model are loaded first, so they are in
    // flight together with the staging loads (the old p only where the
    // interior formula reads it, the old q only outside the x strips).
    constexpr int J = TX / NW;
    float r_uo[J], r_qo[J], r_vz[J], r_ep[J], r_de[J];
    {
        const Buf VZ = field_buf(a.vz + pad0), EP = field_buf(a.ep + pad0), DE = field_buf(a.de + pad0);
        const int vu = inz ? vl : kOOB;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int col = x0 + wv + j * NW;
            const uint32_t o = colo(col);
            r_vz[j] = VZ.ld(vl, o);
            r_ep[j] = EP.ld(vl, o);
            r_de[j] = DE.ld(vl, o);
            const bool inx = col >= a.nl && col < n_x - a.nr;
            r_qo[j] = inx ? QN.ld(vl, o) : 0.f;
            r_uo[j] = inx ? UN.ld(vu, o) : 0.f;
        }
    }
    {
        // s = p + q over WX columns x 64 rows, whole columns per wave; q and
        // p of the tile's own columns from the same loads; p's 5-row z halo
        // of the own columns (10 cells a column) spread over all threads.
        constexpr int QS = (WX + NW - 1) / NW;
        constexpr int NH = TX * 2 * kVH, QH = (NH + T - 1) / T;
        float vu[QS], vq[QS], vh[QH];
        int hl[QH];
        const uint32_t cs0 = colo(x0 - kVH);
#pragma unroll
        for (int q = 0; q < QS; ++q) {
            const int c = wv + q * NW;
            vu[q] = vq[q] = 0.f;
            if (QS * NW == WX || c < WX) {
                vu[q] = UC.ld(vl, cs0 + (uint32_t)c * cb);
                vq[q] = QC.ld(vl, cs0 + (uint32_t)c * cb);
            }
        }
        const uint32_t h0 = (uint32_t)(x0 + kVPad) * cb + (uint32_t)(z0 + kVPad - kVH) * 4u;  // (x0, z0 - 5)
#pragma unroll
        for (int q = 0; q < QH; ++q) {
            const int e = threadIdx.x + q * T;
            const int c = e / (2 * kVH), r = e - c * (2 * kVH);
            const int rr = r < kVH ? r : r + kVTZ;  // row - (z0 - 5)
            hl[q] = e < NH ? c * HZ + rr : -1;
            float u = e < NH ? UC.ld((int)((uint32_t)c * cb) + rr * 4, h0) : 0.f;
            if (e < NH && x0 + c == a.sx && z0 - kVH + rr == a.sz) u += a.w_cur;
            vh[q] = u;
        }
#pragma unroll
        for (int q = 0; q < QS; ++q) {
            const int c = wv + q * NW;
            if (QS * NW == WX || c < WX) {
                float u = vu[q];
                if (x0 - kVH + c == a.sx && k == a.sz) u += a.w_cur;
                s_s[c * kVTZ + tz] = u + vq[q];
                if (c >= kVH && c < kVH + TX) {
                    s_q[(c - kVH) * kVTZ + tz] = vq[q];
                    s_u[(c - kVH) * HZ + kVH + tz] = u;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < QH; ++q)
            if (hl[q] >= 0) s_u[hl[q]] = vh[q];
    }
    __syncthreads();
    if (k >= a.nz) return;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int cl = wv + j * NW, col = x0 + cl;
        if (col >= n_x) break;
        const uint32_t o = colo(col);
        const float* ps = s_s + (cl + kVH) * kVTZ + tz;
        const float* pu = s_u + cl * HZ + tz + kVH;
        const float s0 = ps[0], u0 = pu[0], q0 = s_q[cl * kVTZ + tz];
        float dxx = 0.f, dzz = 0.f;  // Dxx (p + q), Dzz p
#pragma unroll
        for (int m = 1; m <= 5; ++m) {
            dxx += a.c2[m - 1] * (ps[m * kVTZ] + ps[-m * kVTZ] - 2.f * s0);
            dzz += a.c2[m - 1] * (pu[m] + pu[-m] - 2.f * u0);
        }
        dxx *= a.idx2;
        dzz *= a.idz2;
        const float vz = r_vz[j], ep = r_ep[j], de = r_de[j];
        const float t = a.dt2 * (vz * vz);
        const bool inx = col >= a.nl && col < n_x - a.nr;
        const int po = a.par, pc = 1 - a.par;
        if (inx) {
            QN.st(2.f * q0 - r_qo[j] + t * (2.f * (ep - de) * dxx), vl, o);
            if (inz) {
                float uo = r_uo[j];
                if (col == a.sx && k == a.sz) uo += a.w_prev;
                UN.st(2.f * u0 - uo + t * ((1.f + 2.f * de) * dxx + dzz), vl, o);
            } else {
                // top / bottom strip (sub:785-853, 869-947): z-split p
                float s1 = 0.f;
#pragma unroll
                for (int m = 1; m <= 5; ++m) s1 += a.c1[m - 1] * (pu[m] - pu[-m]);
                s1 /= a.dx;  // the reference divides the z derivative by dx (sub:796)
                const bool top = k < a.nu;
                const float az = top ? (float)(k - a.nu) * a.dz : (float)(k + 1 - (a.nz - a.nd)) * a.dz;
                const float L = (float)(top ? a.nu : a.nd) * a.dz;
                const uint32_t nb = (uint32_t)a.nzs * 4u;
                const Buf P = field_buf(a.zs + (size_t)shot * 7 * a.nzs);
                const int vs = ((col - a.nl) * (a.nu + a.nd) + (top ? k : a.nu + k - (a.nz - a.nd))) * 4;
                float d, dd;
                vti_damp(vz, az, L, a.lnr, &d, &dd);
                const float uz = vti_pml(P, nb, vs, po, pc, vz * vz, dzz, s1, d, dd, a.dz, a.dt, a.dt2);
                const float vn2 = (vz * sqrtf(1.f + 2.f * de)) * (vz * sqrtf(1.f + 2.f * de));
                const float u3 =
                    2.f * P.ld(vs, (uint32_t)(5 + pc) * nb) - P.ld(vs, (uint32_t)(5 + po) * nb) + vn2 * a.dt2 * dxx;
                P.st(u3, vs, (uint32_t)(5 + po) * nb);
                UN.st(uz + u3, vl, o);
            }
        } else {
            // left / right strip (sub:561-667, 669-781; q: 1072-1256)
            float s1 = 0.f;
#pragma unroll
            for (int m = 1; m <= 5; ++m) s1 += a.c1[m - 1] * (ps[m * kVTZ] - ps[-m * kVTZ]);
            s1 /= a.dx;
            const bool left = col < a.nl;
            const float ax = left ? (float)(col - a.nl) * a.dx : (float)(col + 1 - (n_x - a.nr)) * a.dx;
            const float L = (float)(left ? a.nl : a.nr) * a.dx;
            const uint32_t nb = (uint32_t)a.nxs * 4u;
            const float* xp = a.xs + (size_t)shot * 12 * a.nxs;
            const Buf P = field_buf(xp);
            const int vs = ((left ? col : a.nl + col - (n_x - a.nr)) * a.nz + k) * 4;
            const float vn = vz * sqrtf(1.f + 2.f * de), vx = vz * sqrtf(1.f + 2.f * ep);
            float d, dd;
            vti_damp(vx, ax, L, a.lnr, &d, &dd);
            const float ux = vti_pml(P, nb, vs, po, pc, vn * vn, dxx, s1, d, dd, a.dx, a.dt, a.dt2);
            const float u3 = 2.f * P.ld(vs, (uint32_t)(5 + pc) * nb) - P.ld(vs, (uint32_t)(5 + po) * nb) +
                             vz * vz * a.dt2 * (dzz * (a.dz * a.dz)) / (a.dz * a.dz);
            P.st(u3, vs, (uint32_t)(5 + po) * nb);
            UN.st(ux + u3, vl, o);
            QN.st(vti_pml(field_buf(xp + 7 * (size_t)a.nxs), nb, vs, po, pc, vx * vx - vn * vn, dxx, s1, d, dd, a.dx,
                          a.dt, a.dt2),
                  vl, o);
        }
    }
}

using VtiKernel = void (*)(VtiArgs);

bool vti_pick(int tx, int nw, VtiKernel* k) {
    switch (tx * 100 + nw) {
    case 804: *k = vti_step_kernel<8, 4>; return true;
    case 808: *k = vti_step_kernel<8, 8>; return true;
    case 1604: *k = vti_step_kernel<16, 4>; return true;
    case 1608: *k = vti_step_kernel<16, 8>; return true;
    case 3208: *k = vti_step_kernel<32, 8>; return true;
    case 3216: *k = vti_step_kernel<32, 16>; return true;
    case 6408: *k = vti_step_kernel<64, 8>; return true;
    case 6416: *k = vti_step_kernel<64, 16>; return true;
    default: return false;
    }
}
constexpr int kVtiShapes[][2] = {{8, 4}, {8, 8}, {16, 4}, {16, 8}, {32, 8}, {32, 16}, {64, 8}, {64, 16}};
constexpr int kVtiDefaultTX = 64, kVtiDefaultNW = 16;  // the fastest on the card survey (912 us per step of 300 shots)

// coefficient_2nd / coefficient_1st (sub:1264-1472): the k x k system
// a(i, j) = i^e / e!, e = 2j (kind 2) or 2j - 1 (kind 1), inverted in single
// precision by k Sherman-Morrison updates of the identity; weight i = half
// the (1, i) entry of the inverse.  Sequential sums, no contraction, so the
// weights are those of the reference's single-precision solve.
int vti_coefficients(int kind, int k, float* out) {
#pragma clang fp contract(off)
    if ((kind != 1 && kind != 2) || k < 1 || k > 8 || !out) return SFD_EINVAL;
    float a[8][8], inv[8][8], tmp[8], col[8], f3[8][8];
    for (int i = 1; i <= k; ++i)
        for (int j = 1; j <= k; ++j) {
            const int e = kind == 2 ? 2 * j : 2 * j - 1;
            float fact = 1.f;
            for (int m = 1; m <= e; ++m) fact = fact * (float)m;
            a[i - 1][j - 1] = powf((float)i, (float)e) / fact;
        }
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) inv[i][j] = i == j ? 1.f : 0.f;
    for (int i = 0; i < k; ++i) {
        for (int j = 0; j < k; ++j) tmp[j] = a[i][j];
        tmp[i] = tmp[i] - 1.f;
        float s = 0.f;
        for (int j = 0; j < k; ++j) s = s + tmp[j] * inv[j][i];
        const float f1 = s + 1.f;
        for (int j = 0; j < k; ++j) col[j] = inv[j][i];
        for (int r = 0; r < k; ++r)
            for (int c = 0; c < k; ++c) {
                float v = 0.f;
                for (int m = 0; m < k; ++m) v = v + (col[r] * tmp[m]) * inv[m][c];
                f3[r][c] = v / f1;
            }
        for (int r = 0; r < k; ++r)
            for (int c = 0; c < k; ++c) inv[r][c] = inv[r][c] - f3[r][c];
    }
    for (int i = 0; i < k; ++i) {
        float v = 0.f;
        for (int m = 0; m < k; ++m) v = v + inv[0][m] * (m == i ? 0.5f : 0.f);
        out[i] = v;
    }
    return SFD_OK;
}

// The driver's wavelet (main:433-437), single precision with pi = 3.1415926.
int vti_wavelet(float dt, float f0, int nmax, float* w) {
#pragma clang fp contract(off)
    const float pi = 3.1415926f;
    const int nwt = (int)lroundf(2.0f * 0.07f / dt) + 1;
    for (int it = 1; it <= nwt && it <= nmax; ++it) {
        const float x = pi * f0 * ((float)it - 0.07f / dt) * dt;
        w[it - 1] = expf(-(x * x)) * (1.f - 2.f * (x * x));
    }
    return nwt;
}

}  // namespace
}  // namespace sfd

using namespace sfd;

struct sfd_vti_plan {
    sfd_context* ctx = nullptr;
    int nx = 0, nz = 0, nshot = 0, nsteps = 0, delay = 0, nt = 0, nrec = 0;
    int pitch = 0, ncol = 0, tx = kVtiDefaultTX, nw = kVtiDefaultNW;
    size_t fs = 0;
    DeviceBuffer arena, rx, rz, rec;
    float* u[2] = {nullptr, nullptr};
    float* q[2] = {nullptr, nullptr};
    size_t state_bytes = 0;
    std::vector<float> wav;  // [nsteps + 1], zero past the wavelet
    // Shots run in groups of gshot consecutive shots, each group for all its
    // steps before the next, so a group's fields stay in the 256 MB MALL.
    int gshot = 0;
    size_t group_bytes = 0;  // budget of one group's fields; 0 = the default
    VtiArgs args{};
    VtiKernel kern = nullptr;
};

namespace sfd {
namespace {

int vti_set_tiling(sfd_vti_plan* pl, int tx, int nw) {
    VtiKernel k;
    if (!vti_pick(tx, nw, &k)) return fail(pl->ctx, SFD_EINVAL, "no VTI kernel for %d tile columns x %d waves", tx, nw);
    pl->tx = tx;
    pl->nw = nw;
    pl->kern = k;
    pl->args.ntz = (pl->nz + kVTZ - 1) / kVTZ;
    pl->args.ntiles = pl->args.ntz * ((pl->nx + tx - 1) / tx);
    return SFD_OK;
}

std::string vti_class(const sfd_vti_plan* pl) {
    char buf[96];
    snprintf(buf, sizeof buf, "vti:x%d:z%d:s%d", log2_round(pl->nx), log2_round(pl->nz), log2_round(pl->nshot));
    return buf;
}

int vti_zero(sfd_vti_plan* pl) {
    SFD_HIP(pl->ctx, hipMemsetAsync(pl->arena.ptr, 0, pl->state_bytes, pl->ctx->stream));
    return SFD_OK;
}

// Default budget of one shot group's fields (p, q at two levels, vz, eps,
// delta and the PML parts): measured on MI355X with the VTI card's survey.
constexpr size_t kVtiGroupBytes = (size_t)192 << 20;

// Shots per group for the plan's budget (at least one).
void vti_set_groups(sfd_vti_plan* pl) {
    const size_t per = pl->fs * 7 * sizeof(float) + (12 * (size_t)pl->args.nxs + 7 * (size_t)pl->args.nzs) * sizeof(float);
    const size_t budget = pl->group_bytes ? pl->group_bytes : kVtiGroupBytes;
    pl->gshot = (int)std::max<size_t>(1, std::min<size_t>(pl->nshot, budget / std::max<size_t>(per, 1)));
}

// Enqueues the first nsteps steps (all of them by default) from a zero
// state, shot group after shot group.
int vti_enqueue(sfd_vti_plan* pl, int nsteps = -1) {
    sfd_context* ctx = pl->ctx;
    const int ns = nsteps < 0 ? pl->nsteps : std::min(nsteps, pl->nsteps);
    const dim3 block(64 * pl->nw);
    for (int g0 = 0; g0 < pl->nshot; g0 += pl->gshot) {
        const int gn = std::min(pl->gshot, pl->nshot - g0);
        const dim3 grid(pl->args.ntiles, gn);
        VtiArgs a = pl->args;
        a.shot0 = g0;
        int cur = 1;
        for (int s = 0; s < ns; ++s) {
            const int old = 1 - cur;
            a.uc = pl->u[cur];
            a.qc = pl->q[cur];
            a.un = pl->u[old];
            a.qn = pl->q[old];
            a.par = s & 1;
            a.w_cur = pl->wav[s];
            a.w_prev = s > 0 ? pl->wav[s - 1] : 0.f;
            a.sample = (s - 1 >= pl->delay && s - 1 - pl->delay < pl->nt) ? s - 1 - pl->delay : -1;
            hipLaunchKernelGGL(pl->kern, grid, block, 0, ctx->stream, a);
            SFD_HIP(ctx, hipGetLastError());
            cur = old;
        }
        if (ns > 0 && ns - 1 >= pl->delay && pl->nrec > 0) {
            a.uc = pl->u[cur];
            a.sample = ns - 1 - pl->delay < pl->nt ? ns - 1 - pl->delay : -1;
            hipLaunchKernelGGL(vti_gather_kernel, dim3(1, gn), dim3(256), 0, ctx->stream, a);
            SFD_HIP(ctx, hipGetLastError());
        }
    }
    return SFD_OK;
}

}  // namespace
}  // namespace sfd

extern "C" {

int sfd_vti_coefficients(int kind, int k, float* out) { return vti_coefficients(kind, k, out); }

int sfd_vti_wavelet(float dt, float f0, int nmax, float* w) {
    if (!(dt > 0.f) || nmax < 0 || (nmax > 0 && !w)) return SFD_EINVAL;
    return vti_wavelet(dt, f0, nmax, w);
}

int sfd_vti_window(const float* vz, const float* eps, const float* delta, int nxv, int nzv, int xmin, int range,
                   int nl, int nr, int nu, int nd, int nsz0, float* ovz, float* oeps, float* odelta) {
    if (!vz || !eps || !delta || !ovz || !oeps || !odelta || nxv < 1 || nzv < 1 || range < 1 || nl < 0 || nr < 2 ||
        nu < 0 || nd < 2)
        return SFD_EINVAL;
    const int nx = range + nl + nr - 2, nz = nzv + nu + nd - 2;
    const size_t n = (size_t)nx * nz;
    std::fill(ovz, ovz + n, 0.f);
    std::fill(oeps, oeps + n, 0.f);
    std::fill(odelta, odelta + n, 0.f);
    // window column i (1-based) <- model column i + xmin, clamped (sub:291-299)
    for (int i = 0; i < range; ++i) {
        const int m = std::min(std::max(i + 1 + xmin, 1), nxv) - 1;
        const size_t g = (size_t)(i + nl) * nz + nu, s = (size_t)m * nzv;
        std::copy(vz + s, vz + s + nzv, ovz + g);
        std::copy(eps + s, eps + s + nzv, oeps + g);
        std::copy(delta + s, delta + s + nzv, odelta + g);
    }
    // vz (only) edge-replicated: rows above and below, then columns (sub:303-333)
    for (int i = 0; i < nx; ++i) {
        float* c = ovz + (size_t)i * nz;
        for (int j = 0; j < nu; ++j) c[j] = c[nu];
        for (int j = nu + nzv; j < nz; ++j) c[j] = c[nu + nzv - 1];
    }
    for (int i = 0; i < nl; ++i) std::copy(ovz + (size_t)nl * nz, ovz + (size_t)(nl + 1) * nz, ovz + (size_t)i * nz);
    for (int i = nl + range; i < nx; ++i)
        std::copy(ovz + (size_t)(nl + range - 1) * nz, ovz + (size_t)(nl + range) * nz, ovz + (size_t)i * nz);
    // isotropic rows 1 .. nsz0 + nu + 6 (main:422-425)
    const int top = std::min(nsz0 + nu + 6, nz);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < top; ++j) oeps[(size_t)i * nz + j] = odelta[(size_t)i * nz + j] = 0.f;
    return SFD_OK;
}

int sfd_vti_plan_create(sfd_context* ctx, const sfd_vti_params* p, int nshot, const float* vz, const float* eps,
                        const float* delta, const float* wavelet, sfd_vti_plan** out) {
    if (!ctx || !out) return SFD_EINVAL;
    *out = nullptr;
    if (!p || !vz || !eps || !delta) return fail(ctx, SFD_EINVAL, "params, vz, eps and delta are required");
    if (nshot < 1 || nshot > 65535) return fail(ctx, SFD_EINVAL, "nshot must be in [1, 65535]");
    if (p->nx < 1 || p->nz < 1 || p->nt < 0) return fail(ctx, SFD_EINVAL, "bad grid or nt");
    if (!(p->dx > 0.f) || !(p->dz > 0.f) || !(p->dt > 0.f)) return fail(ctx, SFD_EINVAL, "dx, dz, dt must be > 0");
    if (p->pml_xlo < 0 || p->pml_xhi < 0 || p->pml_zlo < 0 || p->pml_zhi < 0 || p->pml_xlo + p->pml_xhi > p->nx ||
        p->pml_zlo + p->pml_zhi > p->nz)
        return fail(ctx, SFD_EINVAL, "PML widths must be >= 0 and fit the grid");
    if (p->sx < 0 || p->sx >= p->nx || p->sz < 0 || p->sz >= p->nz) return fail(ctx, SFD_EINVAL, "source outside the grid");
    if (p->nrec < 0 || (p->nrec > 0 && (!p->rx || !p->rz))) return fail(ctx, SFD_EINVAL, "bad receivers");
    for (int r = 0; r < p->nrec; ++r)
        if (p->rx[r] < 0 || p->rx[r] >= p->nx || p->rz[r] < 0 || p->rz[r] >= p->nz)
            return fail(ctx, SFD_EINVAL, "receiver %d outside the grid", r);
    float c2[5], c1[5];
    if (p->c2) std::memcpy(c2, p->c2, sizeof c2);
    else vti_coefficients(2, 5, c2);
    if (p->c1) std::memcpy(c1, p->c1, sizeof c1);
    else vti_coefficients(1, 5, c1);
    const int nwt = vti_wavelet(p->dt, p->f0, 0, nullptr);
    const int delay = p->delay >= 0 ? p->delay : (int)lroundf((float)(nwt - 1) / 2.0f);
    const int nsteps = delay + p->nt;
    const size_t n = (size_t)p->nx * p->nz, ntot = n * (size_t)nshot;
    // Stability of the second-order recursion: dt^2 vz^2 ((1 + 2 max(eps, delta, 0)) K / dx^2 + K / dz^2) <= 4,
    // K the largest magnitude of the second-derivative symbol sum c_m (2 cos(m w) - 2).
    double K = 0;
    for (int q = 0; q <= 512; ++q) {
        const double w = M_PI * q / 512.0;
        double s = 0;
        for (int m = 0; m < 5; ++m) s += c2[m] * (2 * std::cos((m + 1) * w) - 2);
        K = std::max(K, std::fabs(s));
    }
    double worst = 0;
    size_t nunstable = 0, first_unstable = 0;
    for (size_t c = 0; c < ntot; ++c) {
        if (eps[c] < delta[c] && nunstable++ == 0) first_unstable = c;
        if (!(vz[c] > 0.f) || !std::isfinite(vz[c]) || !std::isfinite(eps[c]) || !std::isfinite(delta[c]) ||
            eps[c] <= -0.5f || delta[c] <= -0.5f)
            return fail(ctx, SFD_EINVAL, "vz must be finite and > 0, eps and delta finite and > -1/2 (cell %zu)", c);
        const double an = 1 + 2 * std::max({(double)eps[c], (double)delta[c], 0.0});
        worst = std::max(worst, (double)vz[c] * vz[c] * (an * K / ((double)p->dx * p->dx) + K / ((double)p->dz * p->dz)));
    }
    if (nunstable && !p->allow_eps_lt_delta) {
        const size_t sh = first_unstable / n, c = first_unstable % n;
        return fail(ctx, SFD_EINVAL,
                    "eps < delta in %zu cells (first: shot %zu, x %zu, z %zu): the pseudo-acoustic VTI system "
                    "is unstable there; set allow_eps_lt_delta to run it anyway",
                    nunstable, sh, c / p->nz, c % p->nz);
    }
    if ((double)p->dt * p->dt * worst > 4.0)
        return fail(ctx, SFD_ECFL, "unstable: dt^2 max(vz^2 ((1+2 max(eps,delta)) K/dx^2 + K/dz^2)) = %.4f exceeds 4",
                    (double)p->dt * p->dt * worst);
    SFD_HIP(ctx, hipSetDevice(ctx->device));

    sfd_vti_plan* pl = new (std::nothrow) sfd_vti_plan;
    if (!pl) return fail(ctx, SFD_ENOMEM, "host allocation failed");
    auto bail = [&](int code) {
        delete pl;
        return code;
    };
    pl->ctx = ctx;
    pl->nx = p->nx;
    pl->nz = p->nz;
    pl->nshot = nshot;
    pl->nsteps = nsteps;
    pl->delay = delay;
    pl->nt = p->nt;
    pl->nrec = p->nrec;
    pl->pitch = (p->nz + 2 * kVPad + 63) / 64 * 64;
    // columns: pad, grid, pad, the widest tile's overhang and one more for the
    // last staged column's rows past the pitch
    pl->ncol = p->nx + 2 * kVPad + kVTXMax + 2;
    pl->fs = (size_t)pl->ncol * pl->pitch;
    const int nl = p->pml_xlo, nr = p->pml_xhi, nu = p->pml_zlo, nd = p->pml_zhi;
    const size_t nxs = (size_t)(nl + nr) * p->nz, nzs = (size_t)(p->nx - nl - nr) * (nu + nd);

    Carver cv;
    Field fu[2], fq[2], fx, fz, fm[3];
    size_t total = 0;
    for (int pass = 0; pass < 2; ++pass) {
        cv.used = 0;
        for (auto& f : fu) f = cv.take(pl->fs * nshot * sizeof(float));
        for (auto& f : fq) f = cv.take(pl->fs * nshot * sizeof(float));
        fx = cv.take(12 * nxs * nshot * sizeof(float));
        fz = cv.take(7 * nzs * nshot * sizeof(float));
        pl->state_bytes = (cv.used + 255) / 256 * 256;
        for (auto& f : fm) f = cv.take(pl->fs * nshot * sizeof(float));
        if (pass == 0) {
            total = cv.used;
            hipError_t e = pl->arena.alloc(total);
            if (e != hipSuccess) return bail(hip_check(ctx, e, "VTI plan arena"));
            cv.base = pl->arena.as<char>();
        }
    }
    const size_t off = (size_t)kVPad * pl->pitch + kVPad;  // cell (0, 0)
    for (int l = 0; l < 2; ++l) {
        pl->u[l] = fu[l].as<float>() + off;
        pl->q[l] = fq[l].as<float>() + off;
    }
    hipStream_t sm = ctx->stream;
    // model fields: zero border, then each shot's [nx][nz] into its padded copy
    SFD_HIP(ctx, hipMemsetAsync(fm[0].ptr, 0, (char*)pl->arena.ptr + total - (char*)fm[0].ptr, sm));
    const float* src[3] = {vz, eps, delta};
    for (int f = 0; f < 3; ++f)
        for (int s = 0; s < nshot; ++s) {
            hipError_t e = hipMemcpy2DAsync(fm[f].as<float>() + off + s * pl->fs, (size_t)pl->pitch * sizeof(float),
                                            src[f] + (size_t)s * n, (size_t)p->nz * sizeof(float),
                                            (size_t)p->nz * sizeof(float), p->nx, hipMemcpyHostToDevice, sm);
            if (e != hipSuccess) return bail(hip_check(ctx, e, "VTI model upload"));
        }
    hipError_t e = hipStreamSynchronize(sm);
    if (e != hipSuccess) return bail(hip_check(ctx, e, "VTI model upload"));
    if (p->nrec > 0) {
        e = upload(pl->rx, p->rx, p->nrec, sm);
        if (e == hipSuccess) e = upload(pl->rz, p->rz, p->nrec, sm);
        if (e == hipSuccess) e = pl->rec.alloc((size_t)nshot * p->nrec * std::max(p->nt, 1) * sizeof(float));
        if (e == hipSuccess) e = hipMemsetAsync(pl->rec.ptr, 0, pl->rec.bytes, sm);
        if (e != hipSuccess) return bail(hip_check(ctx, e, "VTI receivers"));
    }
    pl->wav.assign((size_t)nsteps + 1, 0.f);
    if (wavelet) std::copy(wavelet, wavelet + nsteps, pl->wav.begin());
    else vti_wavelet(p->dt, p->f0, nsteps, pl->wav.data());

    VtiArgs& a = pl->args;
    a.nx = p->nx;
    a.nz = p->nz;
    a.pitch = pl->pitch;
    a.fs = (long long)pl->fs;
    a.vz = fm[0].as<float>() + off;
    a.ep = fm[1].as<float>() + off;
    a.de = fm[2].as<float>() + off;
    a.xs = fx.as<float>();
    a.zs = fz.as<float>();
    a.nxs = (int)nxs;
    a.nzs = (int)nzs;
    a.nl = nl;
    a.nr = nr;
    a.nu = nu;
    a.nd = nd;
    a.dx = p->dx;
    a.dz = p->dz;
    a.dt = p->dt;
    a.dt2 = p->dt * p->dt;
    a.idx2 = 1.f / (p->dx * p->dx);
    a.idz2 = 1.f / (p->dz * p->dz);
    const float R = p->pml_r > 0.f ? p->pml_r : 1e-3f;
    a.lnr = logf(1.f / R);
    std::memcpy(a.c2, c2, sizeof c2);
    std::memcpy(a.c1, c1, sizeof c1);
    a.sx = p->sx;
    a.sz = p->sz;
    a.nrec = p->nrec;
    a.nt = std::max(p->nt, 1);
    a.rx = pl->rx.as<int>();
    a.rz = pl->rz.as<int>();
    a.rec = pl->rec.as<float>();
    vti_set_groups(pl);
    int rc = vti_set_tiling(pl, kVtiDefaultTX, kVtiDefaultNW);
    if (rc) return bail(rc);
    const auto tuned = ctx->tuning.find(vti_class(pl));
    if (tuned != ctx->tuning.end()) (void)vti_set_tiling(pl, tuned->second.a, tuned->second.b);
    rc = vti_zero(pl);
    if (rc) return bail(rc);
    e = hipStreamSynchronize(sm);
    if (e != hipSuccess) return bail(hip_check(ctx, e, "VTI plan setup"));
    *out = pl;
    return SFD_OK;
}

int sfd_vti_plan_set_tiling(sfd_vti_plan* pl, int tile_cols, int waves) {
    if (!pl) return SFD_EINVAL;
    return vti_set_tiling(pl, tile_cols ? tile_cols : kVtiDefaultTX, waves ? waves : kVtiDefaultNW);
}

int sfd_vti_plan_set_group_bytes(sfd_vti_plan* pl, size_t bytes) {
    if (!pl) return SFD_EINVAL;
    pl->group_bytes = bytes;
    vti_set_groups(pl);
    return SFD_OK;
}

int sfd_vti_plan_get_groups(const sfd_vti_plan* pl, int* ngroups) {
    if (!pl || !ngroups) return SFD_EINVAL;
    *ngroups = (pl->nshot + pl->gshot - 1) / pl->gshot;
    return SFD_OK;
}

int sfd_vti_plan_get_shape(const sfd_vti_plan* pl, int* tile_cols, int* waves) {
    if (!pl || !tile_cols || !waves) return SFD_EINVAL;
    *tile_cols = pl->tx;
    *waves = pl->nw;
    return SFD_OK;
}

int sfd_vti_plan_autotune(sfd_vti_plan* pl, int trial_steps, float* best_us) {
    if (!pl) return SFD_EINVAL;
    sfd_context* ctx = pl->ctx;
    if (trial_steps < 1) return fail(ctx, SFD_EINVAL, "trial_steps must be >= 1");
    const int n = std::min(trial_steps, pl->nsteps);
    if (n < 1) return fail(ctx, SFD_EINVAL, "the plan has no time steps to time");
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    auto run = [&](int m) { return m < 0 ? vti_zero(pl) : vti_enqueue(pl, m); };
    float warm = 0.f;
    int rc = time_steps(ctx, std::min(pl->nsteps, 4 * n), run, &warm);
    if (rc) return rc;
    constexpr int NS = sizeof(kVtiShapes) / sizeof(kVtiShapes[0]);
    float us[NS];
    for (float& u : us) u = 3.4e38f;
    for (int pass = 0; pass < 2; ++pass)
        for (int q = 0; q < NS; ++q) {
            rc = vti_set_tiling(pl, kVtiShapes[q][0], kVtiShapes[q][1]);
            float t = 0.f;
            if (!rc) rc = time_steps(ctx, n, run, &t);
            if (rc) return rc;
            us[q] = std::min(us[q], t);
        }
    int best = 0;
    for (int q = 1; q < NS; ++q)
        if (us[q] < us[best]) best = q;
    rc = vti_set_tiling(pl, kVtiShapes[best][0], kVtiShapes[best][1]);
    if (!rc) rc = vti_zero(pl);
    if (rc) return rc;
    SFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    TunedShape t;
    t.kind = 0;
    t.a = kVtiShapes[best][0];
    t.b = kVtiShapes[best][1];
    t.us = us[best];
    ctx->tuning[vti_class(pl)] = t;
    if (best_us) *best_us = us[best];
    return SFD_OK;
}

int sfd_vti_plan_run(sfd_vti_plan* pl, float* elapsed_ms) {
    if (!pl) return SFD_EINVAL;
    sfd_context* ctx = pl->ctx;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    int rc = vti_zero(pl);
    if (rc) return rc;
    SFD_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    rc = vti_enqueue(pl);
    if (rc) return rc;
    SFD_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    SFD_HIP(ctx, hipEventSynchronize(ctx->ev1));
    if (elapsed_ms) SFD_HIP(ctx, hipEventElapsedTime(elapsed_ms, ctx->ev0, ctx->ev1));
    rc = check_finite(ctx, pl->rec.as<float>(), pl->rec.bytes / 4, "the gathers");
    if (rc) return rc;
    // the level written last (see sfd_vti_plan_fetch), every shot's p
    const int cur = (pl->nsteps % 2 == 0) ? 1 : 0;
    const size_t off = (size_t)kVPad * pl->pitch + kVPad;  // cell (0, 0) within a shot's field
    return check_finite(ctx, pl->u[cur] - off, pl->fs * pl->nshot, "the p field");
}

int sfd_vti_plan_fetch(sfd_vti_plan* pl, float* record, float* state) {
    if (!pl) return SFD_EINVAL;
    sfd_context* ctx = pl->ctx;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    if (record && pl->nrec > 0 && pl->nt > 0)
        SFD_HIP(ctx, hipMemcpyAsync(record, pl->rec.ptr, (size_t)pl->nshot * pl->nrec * pl->nt * sizeof(float),
                                    hipMemcpyDeviceToHost, ctx->stream));
    if (state) {
        // the level written last: after nsteps steps, u[nsteps even ? 1 : 0]
        const int cur = (pl->nsteps % 2 == 0) ? 1 : 0;
        const size_t n = (size_t)pl->nx * pl->nz;
        for (int s = 0; s < pl->nshot; ++s)
            for (int f = 0; f < 2; ++f) {
                const float* src = (f == 0 ? pl->u[cur] : pl->q[cur]) + s * pl->fs;
                SFD_HIP(ctx, hipMemcpy2DAsync(state + ((size_t)s * 2 + f) * n, (size_t)pl->nz * sizeof(float), src,
                                              (size_t)pl->pitch * sizeof(float), (size_t)pl->nz * sizeof(float), pl->nx,
                                              hipMemcpyDeviceToHost, ctx->stream));
            }
    }
    SFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SFD_OK;
}

int sfd_vti_plan_destroy(sfd_vti_plan* pl) {
    if (!pl) return SFD_EINVAL;
    (void)hipSetDevice(pl->ctx->device);
    (void)hipStreamSynchronize(pl->ctx->stream);
    delete pl;
    return SFD_OK;
}

int sfd_vti_forward(sfd_context* ctx, const sfd_vti_params* p, int nshot, const float* vz, const float* eps,
                    const float* delta, const float* wavelet, float* record) {
    sfd_vti_plan* pl = nullptr;
    int rc = sfd_vti_plan_create(ctx, p, nshot, vz, eps, delta, wavelet, &pl);
    if (rc) return rc;
    rc = sfd_vti_plan_run(pl, nullptr);
    if (!rc) rc = sfd_vti_plan_fetch(pl, record, nullptr);
    sfd_vti_plan_destroy(pl);
    return rc;
}

}  // extern "C"
