// Two-launch tiled time step of the acoustic velocity-pressure staggered-grid
// propagator in fast arithmetic: the same step as acoustic_step_kernel
// (reference loop forward/Acoustic_FD_by_ss/src/Acoustic_1order_FD.cpp:247-377)
// split at the reference's own seam:
//   acoustic_tiled_velocity  Vx, Vz from P (source injected)      (:247-286)
//   acoustic_tiled_pressure  P from the new Vx, Vz; receivers     (:289-314, :376-377)
// MODE is one of MODE_FAST_CPML, MODE_FAST_SPONGE (sponge of :329-340),
// MODE_FWD_DIV, MODE_FWD_DIV_CPML and MODE_ADJ (the gradient's recomputation
// and adjoint steps; the C-PML adjoint is acoustic_adjoint_cpml.h).
// A workgroup of NW waves owns a tile of TX rows x 64 columns of the updated
// region [R, nx - R) x [R, nz - R), one column per lane, TX / NW rows per
// wave.  The tile and the R-cell halo its stencils read are staged in LDS
// once (loads unrolled so all are in flight before the first store), tiles go
// to the XCDs in bands of whole rows.  Every field is streamed once per
// launch: 24 B per cell for the velocity launch (P, Vx, Vz, bf read; Vx, Vz
// written; 20 B with uniform density, whose bf is one kernel argument), 20 B
// for the pressure launch (Vx, Vz, P, kf read; P written);
// the fused march kernel needs 32 B per step but cannot keep as many bytes
// in flight.  Fields keep the plan's layout: rows of nz floats with zero rows
// above and below (columns past either edge wrap into the neighbouring row;
// no updated cell's stencil reaches them).  The plan runs these kernels in
// place (out == in): within a launch no workgroup reads a cell that another
// writes, and every output depends on its own cell's input.
//
// Sponge: the pressure launch's stencils need the new velocities before the
// sponge factor is applied, and in place they cannot be overwritten by their
// absorbed values while other workgroups still read them.  So the velocity
// fields keep the unabsorbed values between steps and the velocity launch
// multiplies its own cell's old velocity by the factor when it reads it (the
// same float product the fused kernel stores), while P is absorbed in place
// by the pressure launch.  Results are bitwise equal to the fused kernel's;
// state reads and snapshots apply the factor to the stored velocities.
#pragma once
#include <hip/hip_runtime.h>

#include "acoustic_kernel.h"
#include "device_util.h"

namespace sfd {

constexpr int kTiledTZ = kTileCols;

// Tile t of a shot: z tiles fastest (ntz of them), then row bands.
__device__ __forceinline__ void tiled_origin(int t, int ntz, int tx, int R, int* i0, int* j0) {
    *j0 = R + (t % ntz) * kTiledTZ;
    *i0 = R + (t / ntz) * tx;
}

// Sponge factor of cell (i, j): the x-edge rows' factor ax_i overrides the
// z-edge columns' az_j in the corners (the fused kernel's absorb_row).
__device__ __forceinline__ float tiled_sponge(int i, int nx, int sw, float ax_i, float az_j) {
    return (i < sw || i >= nx - sw) ? ax_i : az_j;
}

// Args is AcousticStepArgs, in kernel arguments or (batched shots) in device
// memory read through the constant address space; t is the shot's tile and
// k the step.  Fields are read and written through raw buffer views: a lane's
// offset is its column (lane * 4 bytes), a row's offset a wave-uniform scalar
// (each wave's rows are wave-uniform), so a cell costs no 64-bit address
// arithmetic.  Own-cell loads of rows past nx - R (the last row band) read
// the fields' padding rows and are not used.
template <int R, int TX, int NW, int MODE, class Args>
__device__ __forceinline__ void tiled_velocity(const Args& a, int t, int k, float* s_p) {
    constexpr bool CPML = MODE == MODE_FAST_CPML || MODE == MODE_FWD_DIV_CPML, ADJ = MODE == MODE_ADJ;
    constexpr bool REV = MODE == MODE_REV;
    constexpr int W = kTiledTZ + 2 * R, RPW = TX / NW;
    int i0, j0;
    tiled_origin(t, a.ntz, TX, R, &i0, &j0);
    const int nx = a.nx, nz = a.nz, lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int j = j0 + lane, rb = nz * 4, vl = lane * 4;
    const bool jok = j < nz - R;
    auto row = [&](int i) { return (uint32_t)i * (uint32_t)rb + (uint32_t)j0 * 4u; };  // byte offset of (i, j0)
    const Buf VXI = field_buf(a.vx_in), VZI = field_buf(a.vz_in);
    // This lane's own-cell inputs, loaded before the tile so that their
    // latency overlaps the staging.
    float vx0[RPW], vz0[RPW], bf[RPW], axr[RPW];
    const float bfc = a.bf_const;
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        const int i = i0 + wv + q * NW;
        vx0[q] = VXI.ld(vl, row(i));
        vz0[q] = VZI.ld(vl, row(i));
        if constexpr (!CPML && !REV) axr[q] = ro(a.ax, i);
    }
    // Uniform density: one bf for every cell, 4 B per cell less to read.
    if (bfc != 0.f) {
#pragma unroll
        for (int q = 0; q < RPW; ++q) bf[q] = bfc;
    } else {
        const Buf BF = field_buf(a.bf);
#pragma unroll
        for (int q = 0; q < RPW; ++q) bf[q] = BF.ld(vl, row(i0 + wv + q * NW));
    }
    float az_j = 1.f;
    if constexpr (!CPML && !REV) az_j = jok ? a.az[j] : 1.f;
    const float amp = ADJ ? 0.f : a.amp[k];
    tiled_stage<TX, NW, R, R, !ADJ>(s_p, field_buf(a.p_in), rb, i0, j0, wv, a.sx, a.sz, amp);
    __syncthreads();
    if (!jok) return;
    float cx[R];
#pragma unroll
    for (int m = 0; m < R; ++m) cx[m] = a.cx[m];
    const float zx = a.zx;
    const int xw = a.pml_xw, zw = a.pml_zw, sw = a.sponge_w;
    const bool raw = a.v_raw != 0;
    const Buf VXO = field_buf(a.vx_out), VZO = field_buf(a.vz_out);
    int zc = -1;
    float zb2 = 0.f, za2 = 0.f;
    if (CPML) {
        if (j < zw) zc = j;
        else if (j >= nz - zw) zc = j - (nz - zw) + zw;
        if (zc >= 0) {
            zb2 = a.pml_z[2 * nz + j];
            za2 = a.pml_z[3 * nz + j];
        }
    }
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        const int il = wv + q * NW, i = i0 + il;
        if (i >= nx - R) break;
        const float* p = s_p + (il + R) * W + lane + R;
        // Vx(i) sits between P(i) and P(i + 1); Vz(i, j) between P(i, j), P(i, j + 1).
        float sx = cx[0] * (p[W] - p[0]);
#pragma unroll
        for (int m = 2; m <= R; ++m) sx = __builtin_fmaf(cx[m - 1], p[m * W] - p[(1 - m) * W], sx);
        float sz = cx[0] * (p[1] - p[0]);
#pragma unroll
        for (int m = 2; m <= R; ++m) sz = __builtin_fmaf(cx[m - 1], p[m] - p[1 - m], sz);
        sz = zx * sz;
        float dx = sx, dz = sz;
        float v0x = vx0[q], v0z = vz0[q];
        if constexpr (CPML) {
            int xr = -1;
            if (i < xw) xr = i;
            else if (i >= nx - xw) xr = i - (nx - xw) + xw;
            if (xr >= 0) {
                const Buf PI = field_buf(a.psi_px_in);
                const float psi = __builtin_fmaf(ro(a.pml_x, 3 * nx + i), sx, ro(a.pml_x, 2 * nx + i) * PI.ld(vl, row(xr)));
                field_buf(a.psi_px_out).st(psi, vl, row(xr));
                dx = sx + psi;
            }
            if (zc >= 0) {
                const uint32_t zrow = (uint32_t)i * (uint32_t)(8 * zw);
                const float psi = __builtin_fmaf(za2, sz, zb2 * field_buf(a.psi_pz_in).ld(zc * 4, zrow));
                field_buf(a.psi_pz_out).st(psi, zc * 4, zrow);
                dz = sz + psi;
            }
        }
        if constexpr (REV) {
            // Inverse of the forward update (sponge factor 1 in the box).
            if (i < a.rev_x0 || i >= a.rev_x1 || j < a.rev_z0 || j >= a.rev_z1) continue;
            VXO.st(__builtin_fmaf(bf[q], dx, v0x), vl, row(i));
            VZO.st(__builtin_fmaf(bf[q], dz, v0z), vl, row(i));
        } else {
            if (!CPML && raw) {
                const float f = tiled_sponge(i, nx, sw, axr[q], az_j);
                v0x *= f;
                v0z *= f;
            }
            VXO.st(__builtin_fmaf(-bf[q], dx, v0x), vl, row(i));
            VZO.st(__builtin_fmaf(-bf[q], dz, v0z), vl, row(i));
        }
    }
}

template <int R, int TX, int NW, int MODE, class Args>
__device__ __forceinline__ void tiled_pressure(const Args& a, int t, int k, float* s_vx, float* s_vz) {
    constexpr bool CPML = MODE == MODE_FAST_CPML || MODE == MODE_FWD_DIV_CPML, ADJ = MODE == MODE_ADJ;
    constexpr bool DIVOUT = MODE == MODE_FWD_DIV || MODE == MODE_FWD_DIV_CPML;
    constexpr bool REV = MODE == MODE_REV;
    constexpr int WX = kTiledTZ, WZ = kTiledTZ + 2 * R, RPW = TX / NW;
    int i0, j0;
    tiled_origin(t, a.ntz, TX, R, &i0, &j0);
    const int nx = a.nx, nz = a.nz, lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int j = j0 + lane, rb = nz * 4, vl = lane * 4;
    const bool jok = j < nz - R;
    auto row = [&](int i) { return (uint32_t)i * (uint32_t)rb + (uint32_t)j0 * 4u; };  // byte offset of (i, j0)
    const float amp = ADJ ? 0.f : a.amp[k];
    // Divergences of step k (gradient modes).
    float* dv = nullptr;
    if constexpr (ADJ || DIVOUT || REV) dv = a.div + (size_t)(k % a.div_period) * a.div_stride;
    const Buf PI = field_buf(a.p_in), KF = field_buf(a.kf);
    // Own-cell inputs and the rows' receiver spans, before the tile.
    float p0[RPW], kf[RPW], axr[RPW], dk[RPW], gr[RPW];
    int4 span[RPW];
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        const int i = i0 + wv + q * NW;
        p0[q] = PI.ld(vl, row(i));
        kf[q] = KF.ld(vl, row(i));
        if constexpr (!CPML && !REV) axr[q] = ro(a.ax, i);
        if constexpr (ADJ) {
            dk[q] = field_buf(dv).ld(vl, row(i));
            gr[q] = field_buf(a.grad).ld(vl, row(i));
        } else if constexpr (!REV) {
            const int* rr = reinterpret_cast<const int*>(a.rec_row) + 4 * min(i, nx - 1);
            span[q] = make_int4(ro(rr, 0), ro(rr, 1), ro(rr, 2), ro(rr, 3));
        }
    }
    float az_j = 1.f;
    if constexpr (!CPML && !REV) az_j = jok ? a.az[j] : 1.f;
    tiled_stage<TX, NW, R, 0, false>(s_vx, field_buf(a.vx_out), rb, i0, j0, wv, 0, 0, 0.f);
    tiled_stage<TX, NW, 0, R, false>(s_vz, field_buf(a.vz_out), rb, i0, j0, wv, 0, 0, 0.f);
    __syncthreads();
    if (!jok) return;
    float cx[R];
#pragma unroll
    for (int m = 0; m < R; ++m) cx[m] = a.cx[m];
    const float zx = a.zx;
    const int xw = a.pml_xw, zw = a.pml_zw, sw = a.sponge_w;
    const Buf PO = field_buf(a.p_out);
    int zc = -1;
    float zb0 = 0.f, za0 = 0.f;
    if (CPML) {
        if (j < zw) zc = j;
        else if (j >= nz - zw) zc = j - (nz - zw) + zw;
        if (zc >= 0) {
            zb0 = a.pml_z[j];
            za0 = a.pml_z[nz + j];
        }
    }
    float* gk = a.rec_buf + (size_t)k * a.nslot;
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        const int il = wv + q * NW, i = i0 + il;
        if (i >= nx - R) break;
        const float* px = s_vx + (il + R) * WX + lane;
        const float* pz = s_vz + il * WZ + lane + R;
        float sx = cx[0] * (px[0] - px[-WX]);
#pragma unroll
        for (int m = 2; m <= R; ++m) sx = __builtin_fmaf(cx[m - 1], px[(m - 1) * WX] - px[-m * WX], sx);
        float sz = cx[0] * (pz[0] - pz[-1]);
#pragma unroll
        for (int m = 2; m <= R; ++m) sz = __builtin_fmaf(cx[m - 1], pz[m - 1] - pz[-m], sz);
        sz = zx * sz;
        if constexpr (CPML) {
            int xr = -1;
            if (i < xw) xr = i;
            else if (i >= nx - xw) xr = i - (nx - xw) + xw;
            if (xr >= 0) {
                const float psi = __builtin_fmaf(ro(a.pml_x, nx + i), sx, ro(a.pml_x, i) * field_buf(a.psi_vx_in).ld(vl, row(xr)));
                field_buf(a.psi_vx_out).st(psi, vl, row(xr));
                sx += psi;
            }
            if (zc >= 0) {
                const uint32_t zrow = (uint32_t)i * (uint32_t)(8 * zw);
                const float psi = __builtin_fmaf(za0, sz, zb0 * field_buf(a.psi_vz_in).ld(zc * 4, zrow));
                field_buf(a.psi_vz_out).st(psi, zc * 4, zrow);
                sz += psi;
            }
        }
        if constexpr (REV) {
            // Inverse of the forward update (sponge factor 1 in the box).
            if (i < a.rev_x0 || i >= a.rev_x1 || j < a.rev_z0 || j >= a.rev_z1) continue;
            const float s = sx + sz;
            float p = __builtin_fmaf(kf[q], s, p0[q]);
            if (i == a.sx && j == a.sz) p += amp;
            field_buf(dv).st(s, vl, row(i));
            PO.st(p, vl, row(i));
        } else {
            float p = p0[q];
            if (!ADJ && i == a.sx && j == a.sz) p -= amp;
            const float pc = __builtin_fmaf(-kf[q], sx + sz, p);
            if constexpr (DIVOUT) field_buf(dv).st(sx + sz, vl, row(i));
            if constexpr (ADJ) {
                field_buf(a.grad).st(__builtin_fmaf(p0[q], dk[q], gr[q]), vl, row(i));
                if (i == a.sx && j == a.sz) a.src_adj[k] = pc;
            }
            const float pn = CPML ? pc : pc * tiled_sponge(i, nx, sw, axr[q], az_j);
            PO.st(pn, vl, row(i));
            if constexpr (!ADJ) {
                // Receivers of row i: the first span, then further spans (rare).
                const int4 sp = span[q];
                if (j >= sp.x && j < sp.y) gk[sp.z + (j - sp.x)] = pn;
                for (int more = sp.w; more >= 0;) {
                    const int4 mp = a.rec_more[more];
                    if (j >= mp.x && j < mp.y) gk[mp.z + (j - mp.x)] = pn;
                    more = mp.w;
                }
            }
        }
    }
}

template <int R, int TX, int NW, int MODE>
__global__ __launch_bounds__(64 * NW) void acoustic_tiled_velocity(const AcousticStepArgs a) {
    __shared__ float s_p[(TX + 2 * R) * (kTiledTZ + 2 * R)];
    tiled_velocity<R, TX, NW, MODE>(a, xcd_tile((int)blockIdx.x, a.ntz * a.nchunk), a.k, s_p);
}

template <int R, int TX, int NW, int MODE>
__global__ __launch_bounds__(64 * NW) void acoustic_tiled_pressure(const AcousticStepArgs a) {
    __shared__ float s_vx[(TX + 2 * R) * kTiledTZ];
    __shared__ float s_vz[TX * (kTiledTZ + 2 * R)];
    tiled_pressure<R, TX, NW, MODE>(a, xcd_tile((int)blockIdx.x, a.ntz * a.nchunk), a.k, s_vx, s_vz);
}

// Many shots per launch: b.shots[s] is shot s's argument block (in place, so
// one block per shot) and b.tile_map[g] = (shot, the shot's tile) of global
// tile g.
template <int R, int TX, int NW, int MODE>
__global__ __launch_bounds__(64 * NW) void acoustic_tiled_velocity_batch(const BatchStepArgs b) {
    __shared__ float s_p[(TX + 2 * R) * (kTiledTZ + 2 * R)];
    const int* m = reinterpret_cast<const int*>(b.tile_map) + 2 * xcd_tile((int)blockIdx.x, b.ntile);
    using CArgs = const __attribute__((address_space(4))) AcousticStepArgs;
    CArgs& a = *(CArgs*)(b.shots + ro(m, 0));
    tiled_velocity<R, TX, NW, MODE>(a, ro(m, 1), b.k, s_p);
}

template <int R, int TX, int NW, int MODE>
__global__ __launch_bounds__(64 * NW) void acoustic_tiled_pressure_batch(const BatchStepArgs b) {
    __shared__ float s_vx[(TX + 2 * R) * kTiledTZ];
    __shared__ float s_vz[TX * (kTiledTZ + 2 * R)];
    const int* m = reinterpret_cast<const int*>(b.tile_map) + 2 * xcd_tile((int)blockIdx.x, b.ntile);
    using CArgs = const __attribute__((address_space(4))) AcousticStepArgs;
    CArgs& a = *(CArgs*)(b.shots + ro(m, 0));
    tiled_pressure<R, TX, NW, MODE>(a, ro(m, 1), b.k, s_vx, s_vz);
}

}  // namespace sfd
