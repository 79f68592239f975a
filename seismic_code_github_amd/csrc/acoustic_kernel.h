// Fused time-step kernel of the acoustic velocity-pressure staggered-grid
// propagator (reference: forward/Acoustic_FD_by_ss/src/Acoustic_1order_FD.cpp
// :247-377, one pass of its time loop).
//
// One launch advances every cell one step:
//   source injection      P[sx][sz] -= amp[k]                        (:247-251)
//   velocity half-step    Vx, Vz from P                              (:255-286)
//   pressure half-step    P from the new Vx, Vz                      (:289-314)
//   absorbing boundary    sponge factor on all fields (:329-340) or C-PML
//   receivers             absorbed new P at the receiver cells       (:376-377)
//
// Layout: fields are [nx][nz], z contiguous (the reference's layout).  A
// workgroup owns TZ = 64*NW consecutive z cells (one lane each) and a chunk of
// x rows, which every lane marches through in order:
//   * the x-direction stencils read register queues: the last 2R rows of P
//     (for the new Vx) and the last 2R rows of the new Vx (for the new P), so
//     x-neighbours cost no memory traffic;
//   * the z-direction stencils read LDS: a ring of 2R rows of P with a
//     2R-1-cell halo on each side, and one row of the new Vz with an R-cell
//     halo (recomputed by the halo lanes);
//   * every field is read once and written once from HBM per cell and step:
//     P, Vx, Vz in and out plus the two model fields = 32 B per cell-update.
// Chunks overlap by the stencil's reach: the first 4R-2 marched rows only
// fill the queues.  Outputs go to separate buffers (ping-pong), so a halo read
// never sees a value of the new time level.
//
// The reference updates cells in [R, N-R) on both axes and never writes the
// outer R cells, which therefore stay zero; this kernel does the same.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>

#include "device_util.h"

namespace sfd {

enum KernelMode : int {
    MODE_FAST_SPONGE = 0,
    MODE_FAST_CPML = 1,
    MODE_REF_SPONGE = 2,
    // Internal: fast arithmetic with no absorbing terms; the C-PML kernel runs
    // it on tiles that touch no PML strip.
    MODE_FAST_NONE = 3,
    // Tiled step only (acoustic_tiled.h).  Fast sponge step that also stores
    // div = D-Vx' + D-Vz' of every cell (the forward recomputation of the
    // gradient).
    MODE_FWD_DIV = 4,
    // Tiled step only.  Adjoint step (sponge): the same stencils on the
    // scaled adjoint fields (K lambda_P, -b lambda_V); no source injection;
    // accumulates grad += P_in * div_k and keeps the adjoint at the source
    // cell.
    MODE_ADJ = 5,
    // Tiled step only.  The fast sponge step run backwards on the cells of
    // the box [rev_x0, rev_x1) x [rev_z0, rev_z1), where the sponge factor is
    // 1: the pressure launch first (P_k = P_k+1 + kf div_k, storing div_k),
    // then the velocity launch (V_k = V_k+1 + bf D(P_k - source)).  Cells
    // outside the box are neither read as outputs nor written.
    MODE_REV = 6,
    // Tiled step only.  The C-PML step that also stores div = t_x + Wx' + t_z
    // + Wz' (the divergences with their memory terms) of every cell.
    MODE_FWD_DIV_CPML = 7,
    // Tiled step only.  Adjoint of the C-PML step (acoustic_adjoint_cpml.h).
    MODE_ADJ_CPML = 8,
};

struct AcousticStepArgs {
    // Every field pointer below (state, model, C-PML strips) lies in one
    // device allocation starting at `arena` (< 2 GiB), which the kernel
    // addresses through a single buffer descriptor.
    const float* arena;
    uint32_t arena_bytes;
    int nx, nz;
    int chunk;     // output rows per workgroup (x direction)
    int ntz, nchunk;  // z tiles and x chunks; the grid is 1-D with ntz*nchunk workgroups
    int k;         // step index: wavelet sample and gather column
    const float* __restrict__ p_in;
    const float* __restrict__ vx_in;
    const float* __restrict__ vz_in;
    float* __restrict__ p_out;
    float* __restrict__ vx_out;
    float* __restrict__ vz_out;
    // FAST: kf = rho vp^2 dt, bf = dt / rho.
    // REF:  kf = float(DEN*Vp*Vp*dt) evaluated as the reference does, bf = DEN.
    const float* __restrict__ kf;
    const float* __restrict__ bf;
    // Tiled step: nonzero = the bf of every updated cell (uniform density),
    // which the velocity launch then does not read from bf.
    float bf_const;
    // Sponge: ax[i] is the x-edge factor of row i, or 2.0 where row i is not
    // in an x-edge strip; az[j] is the z-edge factor of column j (1.0 inside).
    // Every per-row array (state, model, ax, rec_row) points at row 0 of an
    // allocation with kPadRows rows of padding before and after.
    const float* __restrict__ ax;
    const float* __restrict__ az;
    // C-PML: profiles [4][n] = b at nodes, a at nodes, b at half nodes, a at
    // half nodes; memory fields live in strips of sx_w rows / sz_w columns at
    // each end: psi_x* [2*sx_w][nz], psi_z* [nx][2*sz_w].
    const float* __restrict__ pml_x;
    const float* __restrict__ pml_z;
    int pml_xw, pml_zw;
    int sponge_w;  // sponge width (x-edge rows read ax[]); 0 without sponge
    // Tiled sponge step only: 1 when vx_in / vz_in hold velocities not yet
    // multiplied by their cell's sponge factor (the tiled step keeps them so
    // between steps and applies the factor when it reads them).
    int v_raw;
    const float* __restrict__ psi_px_in;
    float* __restrict__ psi_px_out;
    const float* __restrict__ psi_vx_in;
    float* __restrict__ psi_vx_out;
    const float* __restrict__ psi_pz_in;
    float* __restrict__ psi_pz_out;
    const float* __restrict__ psi_vz_in;
    float* __restrict__ psi_vz_out;
    // Source: P[sx][sz] -= amp[k] before the update.
    int sx, sz;
    const float* __restrict__ amp;
    // Receivers: rec_row[i] = (jlo, jhi, slot, more) is row i's first span;
    // the absorbed new P of cell (i, j), jlo <= j < jhi, goes to
    // rec_buf[k * nslot + slot + j - jlo].  more >= 0 indexes further spans
    // of the row in rec_more, chained by their 4th field (-1 ends).
    const int4* __restrict__ rec_row;
    const int4* __restrict__ rec_more;
    float* __restrict__ rec_buf;
    int nslot;
    // FAST: cx = C/dx and zx = dx/dz (z stencils use cx * zx).  REF: cx = C
    // (unscaled), zx unused.
    float cx[8];
    float zx;
    // REF only: the float quotients dt/dx, dt/dz and the double 1.0/dx, 1.0/dz.
    float dtdx, dtdz;
    double inv_dx, inv_dz;
    // Gradient modes (arena fields): div of step k at div + (k % div_period)
    // * div_stride floats; grad accumulates sum_k P_in * div_k; src_adj[k]
    // receives the adjoint P at the source cell.
    float* div;
    size_t div_stride;
    int div_period;
    float* grad;
    float* src_adj;
    // MODE_REV: the box it reconstructs.
    int rev_x0, rev_x1, rev_z0, rev_z1;
};

template <int R, int MODE>
struct Arith;

// FP32 with fused multiply-adds.
template <int R>
struct ArithFast {
    template <class F>
    static __device__ __forceinline__ float stencil(const float* c, F diff) {
        float s = c[0] * diff(1);
#pragma unroll
        for (int m = 2; m <= R; ++m) s = __builtin_fmaf(c[m - 1], diff(m), s);
        return s;
    }
    // New velocity from the old one, bf = dt/rho and the scaled derivative.
    static __device__ __forceinline__ float vel(float v, float bf, float d, float /*dtd*/) {
        return __builtin_fmaf(-bf, d, v);
    }
    static __device__ __forceinline__ float pres(float p, float kf, float dxs, float dzs,
                                                 double, double) {
        return __builtin_fmaf(-kf, dxs + dzs, p);
    }
};

// The reference's arithmetic: float products and sums in source order, no
// contraction, and the double-promoted sub-expressions of :259, :276, :293.
template <int R>
struct ArithRef {
    template <class F>
    static __device__ __forceinline__ float stencil(const float* c, F diff) {
#pragma clang fp contract(off)
        float s = c[0] * diff(1);
#pragma unroll
        for (int m = 2; m <= R; ++m) s = s + c[m - 1] * diff(m);
        return s;
    }
    // Vx2 - (1.0/DEN)*(dt/dx)*S in double, rounded to float.
    static __device__ __forceinline__ float vel(float v, float den, float s, float dtd) {
#pragma clang fp contract(off)
        const double q = (1.0 / (double)den) * (double)dtd;
        return (float)((double)v - q * (double)s);
    }
    // P2 - K*((1.0/dx)*Sx + (1.0/dz)*Sz) in double, rounded to float.
    static __device__ __forceinline__ float pres(float p, float kf, float sx, float sz,
                                                 double inv_dx, double inv_dz) {
#pragma clang fp contract(off)
        const double br = inv_dx * (double)sx + inv_dz * (double)sz;
        return (float)((double)p - (double)kf * br);
    }
};

template <int R, int MODE>
using ArithFor = typename std::conditional<MODE == MODE_REF_SPONGE, ArithRef<R>, ArithFast<R>>::type;

// Values one lane loads for marched iteration i (output row i).  The loads
// of iteration i + NS - 1 are issued while iteration i computes (software
// pipelining): every iteration issues the same loads, so NS - 1 rows of
// memory traffic stay in flight per lane.
struct RowLoads {
    float pf;           // P(i + 2R, j): the front row entering the queue
    float ph;           // P(i + 2R, halo column) on the P-halo lanes
    float vx, bx;       // Vx(g, j), bf(g, j), g = i + R
    float vz, bz;       // Vz(i + 1, j), bf(i + 1, j)
    float vzh, bzh;     // Vz(i + 1, jh), bf(i + 1, jh) on the Vz-halo lanes
    float kf;           // kf(i, j)
    int rjlo, rjhi, rslot, rmore;  // receiver span of row i (wave-uniform)
    float psvz;         // C-PML psi_vz(i, j)
    float pspz, pspzh;  // C-PML psi_pz(i + 1, j), psi_pz(i + 1, jh)
};

// Zero rows above and below every per-row allocation: the march's prefetch
// reads up to 4R + 2NS rows before a chunk and 2R + 2NS rows after it
// without clamping.
template <int R>
constexpr int kPadRows = 6 * R + 16;

// One launch = one time step.  Iteration i of a workgroup's march does three
// independent pieces of work, so their latencies overlap:
//   Vx(i + R)   from the P queue (rows i+1 .. i+2R)           -> qvx
//   Vz(i + 1)   from the LDS ring row i+1                     -> LDS Vz row
//   P(i)        from qvx rows i-R .. i+R-1 and LDS Vz row i
// then one barrier.  Vz rows rotate through VS LDS buffers (VS >= 3), so the
// write of row i+1 never meets a wave still reading row i-1.
template <int R, int NW>
struct StepLds {
    static constexpr int TZ = 64 * NW, Q = 2 * R, HR = 2 * R - 1;
    static constexpr int RW = TZ + 2 * HR, VW = TZ + 2 * R - 1;
    static constexpr int VS = (Q % 4 == 0) ? 4 : 5;
    float ring[Q][RW + 1];
    float vz[VS][VW + 1];
};

// Args is AcousticStepArgs, held in kernel arguments or (batched shots) in
// device memory read through the constant address space; k is the step.
template <int R, int NW, int MODE, int NS, class Args>
__device__ __forceinline__ void acoustic_march(const Args& a, int tile, int k, StepLds<R, NW>& lds) {
    constexpr int TZ = 64 * NW;         // z cells per workgroup
    constexpr int Q = 2 * R;            // register queue length
    constexpr int HR = 2 * R - 1;       // P halo on each side of the tile
    constexpr int RW = TZ + 2 * HR;     // LDS ring row width (+1 scratch slot)
    constexpr int VW = TZ + 2 * R - 1;  // LDS Vz row width: R left, R-1 right (+1 scratch)
    constexpr int VS = (Q % 4 == 0) ? 4 : 5;  // Vz row buffers; VS divides Q
    constexpr int PD = NS - 1;          // prefetch distance in rows
    constexpr int PADR = kPadRows<R>;
    constexpr bool CPML = (MODE == MODE_FAST_CPML);
    constexpr bool REF = (MODE == MODE_REF_SPONGE);
    constexpr bool NOABS = CPML || (MODE == MODE_FAST_NONE);
    static_assert(Q % NS == 0 && Q % VS == 0, "slot counts must divide the queue length");
    static_assert(TZ >= 2 * HR, "tile narrower than its halos");
    using A = ArithFor<R, MODE>;
    auto& s_ring = lds.ring;
    auto& s_vz = lds.vz;

    const int nx = a.nx, nz = a.nz;
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int j0 = R + (tile % a.ntz) * TZ;
    const int j = j0 + tid;
    const bool jupd = j < nz - R;  // lane's column is updated
    const int i0 = R + (tile / a.ntz) * a.chunk;
    const int i1 = min(i0 + a.chunk, nx - R);
    if (i0 >= i1) return;

    // Every per-row field starts kPadRows rows before row 0; all fields share
    // one buffer descriptor and differ by their byte offset in the arena.
    const uint32_t rowb = (uint32_t)nz * 4u;
    const Buf AR(a.arena, a.arena_bytes);
    auto fo = [&](const float* p) -> uint32_t {
        return (uint32_t)((const char*)p - (const char*)a.arena);
    };
    const uint32_t padb = (uint32_t)PADR * rowb;
    const uint32_t o_pin = fo(a.p_in) - padb, o_vxin = fo(a.vx_in) - padb, o_vzin = fo(a.vz_in) - padb;
    const uint32_t o_pout = fo(a.p_out) - padb, o_vxout = fo(a.vx_out) - padb, o_vzout = fo(a.vz_out) - padb;
    const uint32_t o_bf = fo(a.bf) - padb, o_kf = fo(a.kf) - padb;
    const Buf RR(reinterpret_cast<const int*>(a.rec_row) - 4 * PADR, (uint32_t)(nx + 2 * PADR) * 16u);
    auto so = [&](int r) -> uint32_t { return (uint32_t)(r + PADR) * rowb; };

    // Lane byte offsets.  P reads every stored column; the other fields only
    // updated columns (elsewhere they read 0 and stores are dropped).
    const int vo_p = (j < nz) ? 4 * j : kOOB;
    const int vo_u = jupd ? 4 * j : kOOB;
    // P-halo lanes: the first HR lanes read the columns left of the tile, the
    // last HR lanes those right of it.  Other lanes use the scratch slot.
    const bool has_ph = (tid < HR) || (tid >= TZ - HR);
    const int ph_col = (tid < HR) ? (j0 - HR + tid) : (j + HR);
    const int ph_pos = !has_ph ? RW : (tid < HR) ? tid : (tid + 2 * HR);
    const int vo_ph = (has_ph && ph_col >= 0 && ph_col < nz) ? 4 * ph_col : kOOB;
    const bool wave_halo = (wave == 0) || (wave == NW - 1);
    // Vz-halo lanes: R columns left of the tile, R-1 right of it.
    const bool has_vh = (tid < R) || (tid >= TZ - (R - 1));
    const int jh = (tid < R) ? (j0 - R + tid) : (j + R - 1);
    const bool vh_upd = has_vh && jh >= R && jh < nz - R;
    const int vh_pos = !has_vh ? VW : (tid < R) ? tid : (tid + 2 * R - 1);
    const int vh_rpos = !has_vh ? (tid + HR) : (tid < R) ? (tid + R - 1) : (tid + 3 * R - 2);
    const int vo_vh = vh_upd ? 4 * jh : kOOB;

    const float az_j = (j < nz) ? a.az[j] : 1.0f;
    const float amp = ro(a.amp, k);
    // One coefficient set in scalar registers: the z stencils use the x
    // coefficients and scale their sums by zx = dx/dz (fast arithmetic; the
    // reference arithmetic uses the unscaled C_m for both axes).
    float cx[R];
#pragma unroll
    for (int m = 0; m < R; ++m) cx[m] = in_vgpr(a.cx[m]);
    const float zx = REF ? 1.0f : in_vgpr(a.zx);

    // C-PML strips: psi_x* are [2*xw][nz] (x-strip rows), psi_z* [nx][2*zw].
    const int xw = a.pml_xw, zw = a.pml_zw, zw2 = 2 * zw;
    auto xstrip = [&](int r) -> int {
        if (r < xw) return r;
        if (r >= nx - xw) return r - (nx - xw) + xw;
        return -1;
    };
    auto zstrip = [&](int c) -> int {
        if (c < zw) return c;
        if (c >= nz - zw) return c - (nz - zw) + zw;
        return -1;
    };
    const int zcol = (CPML && jupd) ? zstrip(j) : -1;
    const int zcolh = (CPML && vh_upd) ? zstrip(jh) : -1;
    const int vo_zc = zcol >= 0 ? 4 * zcol : kOOB;
    const int vo_zch = zcolh >= 0 ? 4 * zcolh : kOOB;
    float pz_bi = 1.f, pz_ai = 0.f, pz_bh = 1.f, pz_ah = 0.f, ph_bh = 1.f, ph_ah = 0.f;
    if (CPML) {
        if (zcol >= 0) {
            pz_bi = a.pml_z[j];
            pz_ai = a.pml_z[nz + j];
            pz_bh = a.pml_z[2 * nz + j];
            pz_ah = a.pml_z[3 * nz + j];
        }
        if (zcolh >= 0) {
            ph_bh = a.pml_z[2 * nz + jh];
            ph_ah = a.pml_z[3 * nz + jh];
        }
    }
    uint32_t o_pxin = 0, o_pxout = 0, o_pvxin = 0, o_pvxout = 0, o_pzin = 0, o_pzout = 0, o_pvzin = 0,
             o_pvzout = 0;
    if (CPML) {
        o_pxin = fo(a.psi_px_in);
        o_pxout = fo(a.psi_px_out);
        o_pvxin = fo(a.psi_vx_in);
        o_pvxout = fo(a.psi_vx_out);
        o_pzin = fo(a.psi_pz_in);
        o_pzout = fo(a.psi_pz_out);
        o_pvzin = fo(a.psi_vz_in);
        o_pvzout = fo(a.psi_vz_out);
    }
    // Drain the lane-constant loads now: left pending into the march loop,
    // their first use in every row would wait for the whole prefetch queue.
    __builtin_amdgcn_s_waitcnt(0);

    // Sponge factor of row r at this lane: the x-edge strips (scalar table
    // read, edge rows only) override the z-edge strips in the corners, as the
    // reference's table does (:164-194).
    const int sw = a.sponge_w;
    auto absorb_row = [&](int r) -> float {
        if constexpr (NOABS) {
            return 1.0f;
        } else {
            if (r < sw || r >= nx - sw) return ro(a.ax, r);
            return az_j;
        }
    };

    const int i_first = i0 - 4 * R + 1;  // first marched iteration
    // Prefetches run past both ends of the chunk into the zero padding.
    // While the queues fill, fields the iteration does not use yet are read
    // from the first row the chunk will use (so they cost no extra traffic).
    auto load_row = [&](int i, RowLoads& L) {
        const uint32_t oi = so(max(i, i0));
        const uint32_t of = so(max(i + Q, i_first + Q));
        const uint32_t og = so(max(i + R, i0 - R));
        const uint32_t o1 = so(max(i + 1, i0));
        L.pf = AR.ld(vo_p, o_pin + of);
        L.ph = AR.ld(vo_ph, o_pin + of);
        L.vx = AR.ld(vo_u, o_vxin + og);
        L.bx = AR.ld(vo_u, o_bf + og);
        L.vz = AR.ld(vo_u, o_vzin + o1);
        L.bz = AR.ld(vo_u, o_bf + o1);
        L.vzh = AR.ld(vo_vh, o_vzin + o1);
        L.bzh = AR.ld(vo_vh, o_bf + o1);
        L.kf = AR.ld(vo_u, o_kf + oi);
        {
            const auto rv = __builtin_amdgcn_raw_buffer_load_b128(RR.r, 0, (int)((uint32_t)(max(i, i0) + PADR) * 16u), 0);
            L.rjlo = (int)rv[0];
            L.rjhi = (int)rv[1];
            L.rslot = (int)rv[2];
            L.rmore = (int)rv[3];
        }
        if constexpr (CPML) {
            // x-strip memory fields are read where used (strip rows only).
            const int ic = min(max(i, i0), nx - 1);
            const int i1c = min(max(i + 1, i0), nx - 1);
            L.psvz = AR.ld(vo_zc, o_pvzin + (uint32_t)ic * zw2 * 4u);
            L.pspz = AR.ld(vo_zc, o_pzin + (uint32_t)i1c * zw2 * 4u);
            L.pspzh = AR.ld(vo_zch, o_pzin + (uint32_t)i1c * zw2 * 4u);
        }
    };

    float qp[Q];   // qp[r % Q] = P(row r, j) with the source injected
    float qvx[Q];  // qvx[r % Q] = new Vx(row r, j), not yet absorbed
#pragma unroll
    for (int q = 0; q < Q; ++q) { qp[q] = 0.f; qvx[q] = 0.f; }
    RowLoads ld[NS];

    const int base0 = floor_div(i_first, Q) * Q;
    for (int base = base0 - Q; base < i1; base += Q) {
        static_for<Q>([&](auto U_) {
            constexpr int U = decltype(U_)::value;
            const int i = base + U;
            // Unconditional, so every iteration issues the same loads.
            load_row(i + PD, ld[(U + PD) % NS]);
            if (i < i_first || i >= i1) return;
            const RowLoads& L = ld[U % NS];

            // ---- P(i): x part from the Vx queue (before Vx(i+R) replaces
            //      row i-R in it) and the base value P(i) (before the front
            //      row replaces it in the P queue) ----
            const bool do_p = i >= i0;
            const float p_i = qp[U % Q];
            float sxp = 0.f;
            if (do_p) {
                sxp = A::stencil(cx, [&](int m) {
                    return qvx[(U + m - 1) % Q] - qvx[(U - m + Q) % Q];
                });
            }

            // ---- front row i + 2R into the P queue and the LDS ring ----
            const int f = i + Q;
            {
                float pf = L.pf, ph = L.ph;
                if (f == a.sx) {
                    if (j == a.sz) pf -= amp;
                    if (ph_col == a.sz && vo_ph != kOOB) ph -= amp;
                }
                qp[U % Q] = pf;
                // Rows before i0 are never read from the ring; writing them
                // is harmless (their slot is next used by row f + 2R).
                s_ring[U % Q][tid + HR] = pf;
                if (wave_halo && has_ph) s_ring[U % Q][ph_pos] = ph;
            }

            // ---- Vx(g), g = i + R, from P rows i+1 .. i+2R ----
            const int g = i + R;
            if (g >= i0 - R) {
                const float s = A::stencil(cx, [&](int m) {
                    return qp[(U + R + m) % Q] - qp[(U + R - m + 1) % Q];
                });
                float d = s;
                const bool own = g >= i0 && g < i1;  // rows below i0 belong to the previous chunk
                if constexpr (CPML) {
                    const int xr = xstrip(g);
                    if (xr >= 0) {
                        const float psx = AR.ld(vo_u, o_pxin + (uint32_t)xr * rowb);
                        const float psi = __builtin_fmaf(ro(a.pml_x, 3 * nx + g), s, ro(a.pml_x, 2 * nx + g) * psx);
                        d = s + psi;
                        if (own) AR.st(psi, vo_u, o_pxout + (uint32_t)xr * rowb);
                    }
                }
                float vxn = A::vel(L.vx, L.bx, d, a.dtdx);
                if constexpr (REF) vxn = (g >= R && g < nx - R && jupd) ? vxn : 0.f;
                qvx[(U + R) % Q] = vxn;
                if (own) AR.st(vxn * absorb_row(g), vo_u, o_vxout + so(g));
            }

            if (i < i0 - 1) return;
            // Ring rows written while the queues were filling become visible.
            if (i == i0 - 1) __syncthreads();

            // ---- Vz(i+1) (own lane and halo lanes) into LDS buffer (i+1) % VS ----
            const int i1r = i + 1;
            if (i1r < i1) {
                constexpr int sr = (U + 1) % Q;
                constexpr int vw = (U + 1) % VS;
                auto vz_new = [&](int rpos, float vold, float b, int zc, int vo, float bh, float ah,
                                  float psi_old, bool own) -> float {
                    const float s = zx * A::stencil(cx, [&](int m) {
                        return s_ring[sr][rpos + m] - s_ring[sr][rpos - m + 1];
                    });
                    float d = s;
                    if constexpr (CPML) {
                        if (zc >= 0) {
                            const float psi = __builtin_fmaf(ah, s, bh * psi_old);
                            d = s + psi;
                            if (own) AR.st(psi, vo, o_pzout + (uint32_t)i1r * zw2 * 4u);
                        }
                    }
                    return A::vel(vold, b, d, a.dtdz);
                };
                float vzn = vz_new(tid + HR, L.vz, L.bz, zcol, vo_zc, pz_bh, pz_ah, L.pspz, true);
                if constexpr (REF) vzn = jupd ? vzn : 0.f;
                s_vz[vw][tid + R] = vzn;
                AR.st(vzn * absorb_row(i1r), vo_u, o_vzout + so(i1r));
                if (wave_halo) {
                    float vh = vz_new(vh_rpos, L.vzh, L.bzh, zcolh, vo_zch, ph_bh, ph_ah, L.pspzh, false);
                    if constexpr (REF) vh = vh_upd ? vh : 0.f;
                    if (has_vh) s_vz[vw][vh_pos] = vh;
                }
            }

            // ---- P(i) ----
            if (do_p) {
                constexpr int vr = U % VS;
                float sx = sxp;
                float sz = zx * A::stencil(cx, [&](int m) {
                    return s_vz[vr][tid + R + m - 1] - s_vz[vr][tid + R - m];
                });
                if constexpr (CPML) {
                    const int xr = xstrip(i);
                    if (xr >= 0) {
                        const float psvx = AR.ld(vo_u, o_pvxin + (uint32_t)xr * rowb);
                        const float psi = __builtin_fmaf(ro(a.pml_x, nx + i), sx, ro(a.pml_x, i) * psvx);
                        AR.st(psi, vo_u, o_pvxout + (uint32_t)xr * rowb);
                        sx += psi;
                    }
                    if (zcol >= 0) {
                        const float psi = __builtin_fmaf(pz_ai, sz, pz_bi * L.psvz);
                        AR.st(psi, vo_zc, o_pvzout + (uint32_t)i * zw2 * 4u);
                        sz += psi;
                    }
                }
                const float pc = A::pres(p_i, L.kf, sx, sz, a.inv_dx, a.inv_dz);
                const float pn = pc * absorb_row(i);
                AR.st(pn, vo_u, o_pout + so(i));
                // Receivers of row i: the first span from the prefetch, further
                // spans (rare) from the chained list.
                {
                    const int jlo = __builtin_amdgcn_readfirstlane(L.rjlo);
                    const int jhi = __builtin_amdgcn_readfirstlane(L.rjhi);
                    const int slot = __builtin_amdgcn_readfirstlane(L.rslot);
                    int more = __builtin_amdgcn_readfirstlane(L.rmore);
                    float* gk = a.rec_buf + (size_t)k * a.nslot;
                    if (jupd && j >= jlo && j < jhi) gk[slot + (j - jlo)] = pn;
                    while (more >= 0) {
                        const int* sp = reinterpret_cast<const int*>(a.rec_more) + 4 * more;
                        const int mlo = ro(sp, 0), mhi = ro(sp, 1), mslot = ro(sp, 2);
                        if (jupd && j >= mlo && j < mhi) gk[mslot + (j - mlo)] = pn;
                        more = ro(sp, 3);
                    }
                }
            }
            __syncthreads();
        });
    }
}

template <int R, int NW, int MODE, int NS, class Args>
__device__ __forceinline__ void acoustic_tile(const Args& a, int tile, int k, StepLds<R, NW>& lds) {
    if constexpr (MODE == MODE_FAST_CPML) {
        // Tiles whose rows and columns (halos included) miss every PML strip
        // run without the memory-variable terms.
        constexpr int TZ = 64 * NW;
        const int ntz = a.ntz;
        const int j0 = R + (tile % ntz) * TZ;
        const int i0 = R + (tile / ntz) * a.chunk;
        const int nx = a.nx, nz = a.nz, xw = a.pml_xw, zw = a.pml_zw;
        const int i1 = min(i0 + a.chunk, nx - R);
        const bool xs = (i0 - 2 * R < xw) || (i1 + 2 * R > nx - xw);
        const bool zs = (j0 - 2 * R < zw) || (j0 + TZ + 2 * R > nz - zw);
        if (xs || zs) acoustic_march<R, NW, MODE_FAST_CPML, NS>(a, tile, k, lds);
        else acoustic_march<R, NW, MODE_FAST_NONE, NS>(a, tile, k, lds);
    } else {
        acoustic_march<R, NW, MODE, NS>(a, tile, k, lds);
    }
}

template <int R, int NW, int MODE, int NS>
__global__ __launch_bounds__(64 * NW) void acoustic_step_kernel(const AcousticStepArgs a) {
    __shared__ StepLds<R, NW> lds;
    acoustic_tile<R, NW, MODE, NS>(a, xcd_tile((int)blockIdx.x, a.ntz * a.nchunk), a.k, lds);
}

// Many shots in one launch.  shots[parity * nshot + s] is the argument block
// of shot s with its ping-pong buffers in the given order; the workgroups of
// one launch cover the tiles of every shot, tile_shot maps a global tile to
// its shot and tile_base[s] is the shot's first global tile.  The tiled
// kernels (in place, one block per shot) read tile_map[g] = (shot, the
// shot's tile) instead, so tiles of different shots may interleave.
struct BatchStepArgs {
    const AcousticStepArgs* shots;
    const int* tile_shot;
    const int* tile_base;
    const int2* tile_map;
    int nshot, ntile, k, parity;
};

template <int R, int NW, int MODE, int NS>
__global__ __launch_bounds__(64 * NW) void acoustic_batch_step_kernel(const BatchStepArgs b) {
    __shared__ StepLds<R, NW> lds;
    const int tile = xcd_tile((int)blockIdx.x, b.ntile);
    const int s = ro(b.tile_shot, tile);
    using CArgs = const __attribute__((address_space(4))) AcousticStepArgs;
    CArgs& a = *(CArgs*)(b.shots + (size_t)b.parity * b.nshot + s);
    acoustic_tile<R, NW, MODE, NS>(a, tile - ro(b.tile_base, s), b.k, lds);
}

// Prefetch slots per radius: the slot count must divide 2R.
template <int R>
constexpr int default_slots() { return R == 5 ? 5 : 4; }

}  // namespace sfd
