// Fused time-step kernel of the visco-acoustic rotated-staggered-grid (RSG)
// propagator.  Reference: forward/Q_forward/Q_multishot_fd_subroutine_fsbc_zy_rotated.f90,
// extrapolation_one_step_10nd_absbc (:695-839) in the limit vs = 0, where
// strxx and strzz receive identical updates (one field sigma) and strxz, vxz,
// vzx stay zero; with L standard linear solids (the reference: L = 1).
//
// One launch advances every cell of the grid one step:
//   velocity   vx += b (0.5/dx) (dd + da),  vz += b (0.5/dz) (dd - da)      (:747-773)
//              dd, da = sums of a_n over the two diagonals of sigma (calderxabs/
//              calderzabs xflag = 0, :843-966); x part damped in the left/right
//              strips (vx), z part in the top/bottom strips (vz)      (calpmlbclrtb)
//   stress     div = Dx'(vx) + Dz'(vz) (xflag = 1, :779-782)
//              sigma += ku div + dt sum_l r_l                          (calpstrxx, :993-1030)
//   memory     r_l += -(dt / tau_sigma_l) (kr div + r_l)              (calprxx, :1135-1172)
//   sponge     the x and z parts of sigma and r_l (split fields, :786-836) are
//              kept only in the strips where they are damped; elsewhere the
//              unsplit sums are stored
//   source     Gaussian footprint added to sigma (and its x part)      (addsrc, :1340-1392)
//   receivers  the new sigma at the receiver cells                     (main :450-459)
//
// Layout: fields are [nx][nz], z contiguous.  A workgroup of TZ = 64*NW lanes
// (one z column each) marches through a chunk of x rows.  Three LDS rings of
// S = 2R+2 rows hold sigma (old), vx and vz (new); iteration i loads sigma row
// i+2R+1, computes v at row i+R+1 and sigma at row i, then one barrier.  Tiles
// overlap by 4R-2 columns so that every lane loads exactly its own column: the
// middle TZ-4R+2 lanes produce output.  All cells are updated; reads outside
// the grid are zero (the Fortran's zero halo, -4:nz+5).
#pragma once
#include <hip/hip_runtime.h>

#include "device_util.h"

namespace sfd {

constexpr int kViscoMaxMech = 3;

// Zero rows before and after every per-row field (prefetch and queue fill run
// past both ends of a chunk without clamping).
template <int R>
constexpr int kViscoPadRows = 4 * R + 12;

struct ViscoStepArgs {
    // Every field below lies in one device allocation starting at `arena`.
    const float* arena;
    uint32_t arena_bytes;
    int nx, nz;
    int chunk, ntz, nchunk;
    int k;        // step index (0-based): source sample
    int rec_col;  // gather column of this step, < 0 = not recorded
    const float* s_in;
    float* s_out;
    const float* vx_in;
    float* vx_out;
    const float* vz_in;
    float* vz_out;
    float* r;             // r_1 .. r_L, updated in place, r_stride floats apart
    size_t r_stride;
    const float* b;       // dt / rho_avg
    const float* ku;      // dt * M * (1 + L tau)   (unrelaxed modulus * dt)
    const float* kr;      // M * tau
    // Sponge strips: x strips are rows [0, xlo) and [nx - xhi, nx), z strips
    // columns [0, zlo) and [nz - zhi, nz).  xs = [1+L][xlo+xhi][nz] holds the
    // x parts of sigma and r_l there; zs = [1+L][nx][zlo+zhi] the z parts.
    int xlo, xhi, zlo, zhi;
    float* xs;
    float* zs;
    // Damping per row (x parts) and per column (z parts), 1 outside the
    // strips: velocity tables (spgx*, spgz*) and stress tables (spg1*, spg2*).
    // The row tables dvx, dsx start kViscoPadRows rows before row 0 (padding
    // value 1); the column tables dvz, dsz start at column 0.
    const float* dvx;
    const float* dsx;
    const float* dvz;
    const float* dsz;
    float hx, hz;  // 0.5 / dx, 0.5 / dz
    float a[5];    // a_0 .. a_{R-1}
    float dt;
    float c[kViscoMaxMech];  // dt / tau_sigma_l
    // Source footprint: weights g over rows [sx0, sx1) x columns [sz0, sz1);
    // amp[k] = wavelet(k) dt / dz^2.
    int sx0, sx1, sz0, sz1;
    const float* g;
    const float* amp;
    // Receivers, as in the acoustic kernel: rec_row[i] = (jlo, jhi, slot,
    // more), samples to rec_buf[rec_col * nslot + slot + j - jlo].
    const int4* rec_row;
    const int4* rec_more;
    float* rec_buf;
    int nslot;
};

template <int R, int NW>
struct ViscoLds {
    static constexpr int TZ = 64 * NW, S = 2 * R + 2, W = TZ + 2 * R;
    float s[S][W];
    float vx[S][W];
    float vz[S][W];
};

template <int L>
struct ViscoLoads {
    float sf;              // sigma_in(f), f = i + 2R + 1
    float vx, vz, b;       // at row g = i + R + 1
    float ku, kr;          // at row i
    float r[L];            // r_l at row i
    int rjlo, rjhi, rslot, rmore;
};

template <int R, int L, int NW, int NS>
__global__ __launch_bounds__(64 * NW) void visco_step_kernel(const ViscoStepArgs a) {
    constexpr int TZ = 64 * NW;
    constexpr int S = 2 * R + 2;
    constexpr int PD = NS - 1;
    constexpr int PAD = kViscoPadRows<R>;
    static_assert(S % NS == 0, "prefetch slots must divide the ring length");
    static_assert(4 * R + PD <= PAD, "padding too small for the prefetch distance");
    static_assert(L >= 1 && L <= kViscoMaxMech, "mechanism count");
    __shared__ ViscoLds<R, NW> lds;

    const int nx = a.nx, nz = a.nz;
    const int tile = xcd_tile((int)blockIdx.x, a.ntz * a.nchunk);
    const int lane = threadIdx.x;
    constexpr int OW = TZ - (4 * R - 2);  // output columns per tile
    const int c0 = (tile % a.ntz) * OW - (2 * R - 1);
    const int c = c0 + lane;
    const int i0 = (tile / a.ntz) * a.chunk;
    const int i1 = min(i0 + a.chunk, nx);
    if (i0 >= i1) return;
    const bool cin = c >= 0 && c < nz;
    const bool out = cin && lane >= 2 * R - 1 && lane < TZ - 2 * R + 1;

    const uint32_t rowb = (uint32_t)nz * 4u;
    const Buf AR(a.arena, a.arena_bytes);
    auto fo = [&](const float* p) -> uint32_t { return (uint32_t)((const char*)p - (const char*)a.arena); };
    const uint32_t padb = (uint32_t)PAD * rowb;
    const uint32_t o_sin = fo(a.s_in) - padb, o_sout = fo(a.s_out) - padb;
    const uint32_t o_vxin = fo(a.vx_in) - padb, o_vxout = fo(a.vx_out) - padb;
    const uint32_t o_vzin = fo(a.vz_in) - padb, o_vzout = fo(a.vz_out) - padb;
    const uint32_t o_r = fo(a.r) - padb, r_strb = (uint32_t)(a.r_stride * 4u);
    const uint32_t o_b = fo(a.b) - padb, o_ku = fo(a.ku) - padb, o_kr = fo(a.kr) - padb;
    auto so = [&](int r) -> uint32_t { return (uint32_t)(r + PAD) * rowb; };
    const Buf RR(reinterpret_cast<const int*>(a.rec_row) - 4 * PAD, (uint32_t)(nx + 2 * PAD) * 16u);

    const int vo_c = cin ? 4 * c : kOOB;   // lanes inside the grid
    const int vo_o = out ? 4 * c : kOOB;   // output lanes

    float co[R];
#pragma unroll
    for (int m = 0; m < R; ++m) co[m] = in_vgpr(a.a[m]);
    float cl[L];
#pragma unroll
    for (int l = 0; l < L; ++l) cl[l] = a.c[l];
    const float hx = a.hx, hz = a.hz, dt = a.dt;
    const float dvz_c = cin ? a.dvz[c] : 1.0f;

    // z strip of this lane's column, and whether any output lane has one.
    const int zw = a.zlo + a.zhi;
    int zc = -1;
    if (cin && c < a.zlo) zc = c;
    else if (cin && c >= nz - a.zhi) zc = c - (nz - a.zhi) + a.zlo;
    const bool ztile = (c0 + 2 * R - 1 < a.zlo) || (c0 + TZ - 2 * R + 1 > nz - a.zhi);
    const int vo_z = (out && zc >= 0) ? 4 * zc : kOOB;
    const float dsz_c = (zc >= 0) ? a.dsz[c] : 1.0f;
    const uint32_t o_zs = fo(a.zs), zs_fb = (uint32_t)nx * zw * 4u;
    const int xw = a.xlo + a.xhi;
    const uint32_t o_xs = fo(a.xs), xs_fb = (uint32_t)xw * rowb;
    // Source box columns of this lane.
    const int sg_w = a.sz1 - a.sz0;
    const bool src_col = out && c >= a.sz0 && c < a.sz1;
    const float amp = ro(a.amp, a.k);
    __builtin_amdgcn_s_waitcnt(0);

    const int i_first = i0 - 4 * R;
    auto load_row = [&](int i, ViscoLoads<L>& Ld) {
        const int f = i + 2 * R + 1, g = i + R + 1;
        const int ic = min(max(i, i0), i1 - 1);  // r_l is updated in place: stay in the chunk
        Ld.sf = AR.ld(vo_c, o_sin + so(f));
        Ld.vx = AR.ld(vo_c, o_vxin + so(g));
        Ld.vz = AR.ld(vo_c, o_vzin + so(g));
        Ld.b = AR.ld(vo_c, o_b + so(g));
        Ld.ku = AR.ld(vo_o, o_ku + so(i));
        Ld.kr = AR.ld(vo_o, o_kr + so(i));
#pragma unroll
        for (int l = 0; l < L; ++l) Ld.r[l] = AR.ld(vo_o, o_r + l * r_strb + so(ic));
        const auto rv = __builtin_amdgcn_raw_buffer_load_b128(RR.r, 0, (int)((uint32_t)(ic + PAD) * 16u), 0);
        Ld.rjlo = (int)rv[0];
        Ld.rjhi = (int)rv[1];
        Ld.rslot = (int)rv[2];
        Ld.rmore = (int)rv[3];
    };

    ViscoLoads<L> ld[NS];
    const int base0 = floor_div(i_first, S) * S;
    for (int base = base0 - S; base < i1; base += S) {
        static_for<S>([&](auto U_) {
            constexpr int U = decltype(U_)::value;
            const int i = base + U;
            load_row(i + PD, ld[(U + PD) % NS]);
            if (i < i_first || i >= i1) return;
            const ViscoLoads<L>& Ld = ld[U % NS];
            constexpr int s_i = U % S;                  // ring slot of row i
            constexpr int s_f = (U + 2 * R + 1) % S;    // row f = i + 2R + 1
            constexpr int s_g = (U + R + 1) % S;        // row g = i + R + 1
            const int g = i + R + 1;

            const float s_old = lds.s[s_i][R + lane];
            lds.s[s_f][R + lane] = Ld.sf;

            // ---- v(g) from sigma rows g-R .. g+R-1 (calderxabs/zabs, xflag 0) ----
            if (i >= i0 - 2 * R) {
                float dd = 0.f, da = 0.f;
                static_for<R>([&](auto N_) {
                    constexpr int n = decltype(N_)::value;
                    constexpr int sp = (U + R + 1 + n) % S;  // row g + n
                    constexpr int sm = (U + R - n) % S;      // row g - n - 1
                    const float p1 = lds.s[sp][R + lane + n], p2 = lds.s[sm][R + lane - n - 1];
                    const float p3 = lds.s[sp][R + lane - n - 1], p4 = lds.s[sm][R + lane + n];
                    dd = __builtin_fmaf(co[n], p1 - p2, dd);
                    da = __builtin_fmaf(co[n], p3 - p4, da);
                });
                const bool gin = cin && g >= 0 && g < nx;
                const float bx = Ld.b;
                float vxn = ro(a.dvx, g + PAD) * __builtin_fmaf(bx * hx, dd + da, Ld.vx);
                float vzn = dvz_c * __builtin_fmaf(bx * hz, dd - da, Ld.vz);
                vxn = gin ? vxn : 0.f;
                vzn = gin ? vzn : 0.f;
                lds.vx[s_g][R + lane] = vxn;
                lds.vz[s_g][R + lane] = vzn;
                if (g >= i0 && g < i1) {
                    AR.st(vxn, vo_o, o_vxout + so(g));
                    AR.st(vzn, vo_o, o_vzout + so(g));
                }
            }

            // ---- sigma(i), r_l(i) from v rows i-R+1 .. i+R (xflag 1) ----
            if (i >= i0) {
                float dd1 = 0.f, da1 = 0.f, dd4 = 0.f, da4 = 0.f;
                static_for<R>([&](auto N_) {
                    constexpr int n = decltype(N_)::value;
                    constexpr int vp = (U + 1 + n) % S;   // row i + 1 + n
                    constexpr int vm = (U - n + S) % S;   // row i - n
                    const float x1 = lds.vx[vp][R + lane + 1 + n], x2 = lds.vx[vm][R + lane - n];
                    const float x3 = lds.vx[vp][R + lane - n], x4 = lds.vx[vm][R + lane + 1 + n];
                    const float z1 = lds.vz[vp][R + lane + 1 + n], z2 = lds.vz[vm][R + lane - n];
                    const float z3 = lds.vz[vp][R + lane - n], z4 = lds.vz[vm][R + lane + 1 + n];
                    dd1 = __builtin_fmaf(co[n], x1 - x2, dd1);
                    da1 = __builtin_fmaf(co[n], x3 - x4, da1);
                    dd4 = __builtin_fmaf(co[n], z1 - z2, dd4);
                    da4 = __builtin_fmaf(co[n], z3 - z4, da4);
                });
                const float der1 = hx * (dd1 + da1), der4 = hz * (dd4 - da4);
                const float div = der1 + der4;
                float rsum = Ld.r[0];
#pragma unroll
                for (int l = 1; l < L; ++l) rsum += Ld.r[l];
                float sn = __builtin_fmaf(Ld.ku, div, __builtin_fmaf(dt, rsum, s_old));
                float rn[L];
#pragma unroll
                for (int l = 0; l < L; ++l) rn[l] = __builtin_fmaf(-cl[l], __builtin_fmaf(Ld.kr, div, Ld.r[l]), Ld.r[l]);

                // Source (after the sponge of the x part, as addsrc follows
                // calpmlbclrtb).
                float add = 0.f;
                if (i >= a.sx0 && i < a.sx1 && src_col) add = amp * a.g[(i - a.sx0) * sg_w + (c - a.sz0)];

                // x strip row: the x parts of sigma and r_l.
                int xr = -1;
                if (i < a.xlo) xr = i;
                else if (i >= nx - a.xhi) xr = i - (nx - a.xhi) + a.xlo;
                if (xr >= 0) {
                    const float ax = ro(a.dsx, i + PAD), bx = 1.f - ax;
                    const uint32_t ob = o_xs + (uint32_t)xr * rowb;
                    const float t = __builtin_fmaf(Ld.ku, der1, AR.ld(vo_o, ob));
                    sn = __builtin_fmaf(-bx, t, sn);
                    AR.st(__builtin_fmaf(ax, t, add), vo_o, ob);
#pragma unroll
                    for (int l = 0; l < L; ++l) {
                        const uint32_t obl = ob + (uint32_t)(l + 1) * xs_fb;
                        const float tl = __builtin_fmaf(-cl[l] * Ld.kr, der1, AR.ld(vo_o, obl));
                        rn[l] = __builtin_fmaf(-bx, tl, rn[l]);
                        AR.st(ax * tl, vo_o, obl);
                    }
                }
                // z strip column: the z parts.
                if (ztile) {
                    const float az = dsz_c, bz = 1.f - az;
                    const uint32_t ob = o_zs + (uint32_t)i * zw * 4u;
                    const float t = __builtin_fmaf(Ld.ku, der4, AR.ld(vo_z, ob));
                    if (zc >= 0) sn = __builtin_fmaf(-bz, t, sn);
                    AR.st(az * t, vo_z, ob);
#pragma unroll
                    for (int l = 0; l < L; ++l) {
                        const uint32_t obl = ob + (uint32_t)(l + 1) * zs_fb;
                        const float tl = __builtin_fmaf(-cl[l] * Ld.kr, der4, AR.ld(vo_z, obl));
                        if (zc >= 0) rn[l] = __builtin_fmaf(-bz, tl, rn[l]);
                        AR.st(az * tl, vo_z, obl);
                    }
                }
                sn += add;
                AR.st(sn, vo_o, o_sout + so(i));
#pragma unroll
                for (int l = 0; l < L; ++l) AR.st(rn[l], vo_o, o_r + l * r_strb + so(i));

                if (a.rec_col >= 0) {
                    const int jlo = __builtin_amdgcn_readfirstlane(Ld.rjlo);
                    const int jhi = __builtin_amdgcn_readfirstlane(Ld.rjhi);
                    const int slot = __builtin_amdgcn_readfirstlane(Ld.rslot);
                    int more = __builtin_amdgcn_readfirstlane(Ld.rmore);
                    float* gk = a.rec_buf + (size_t)a.rec_col * a.nslot;
                    if (out && c >= jlo && c < jhi) gk[slot + (c - jlo)] = sn;
                    while (more >= 0) {
                        const int* sp = reinterpret_cast<const int*>(a.rec_more) + 4 * more;
                        const int mlo = ro(sp, 0), mhi = ro(sp, 1), mslot = ro(sp, 2);
                        if (out && c >= mlo && c < mhi) gk[mslot + (c - mlo)] = sn;
                        more = ro(sp, 3);
                    }
                }
            }
            __syncthreads();
        });
    }
}

using ViscoKernel = void (*)(ViscoStepArgs);

// Prefetch slots per radius: the slot count must divide 2R + 2.
template <int R>
constexpr int visco_slots() { return R == 5 ? 4 : 5; }

}  // namespace sfd
