// One-launch tiled time step of the acoustic velocity-pressure staggered-grid
// propagator in fast arithmetic (reference loop
// forward/Acoustic_FD_by_ss/src/Acoustic_1order_FD.cpp:247-377): the two
// half-steps of acoustic_tiled.h in one kernel, so each field crosses HBM
// once per step -- P, Vx, Vz read and written, kf read (28 B per cell with
// uniform density, 32 B with a per-cell bf) -- against the two-launch step's
// 40 B.  The price is redundant work on halos: a tile recomputes the new
// velocities its pressure stencils need beyond its own cells.
//
// A workgroup of NW waves owns TX rows x 64 columns (one column per lane):
//   1. P over rows [i0 - 2R + 1, i0 + TX + 2R - 1) is staged in LDS: the
//      tile's columns on every row, plus a 2R - 1 column halo on each side
//      of the tile's own rows;
//   2. the new Vz of the own rows over the 64 columns and an R / R - 1
//      column halo (z stencils on the staged rows), and the new Vx of rows
//      [i0 - R, i0 + TX + R - 1) (x stencils: each lane slides a register
//      window down its P column, so a P value is read from LDS once per
//      lane) go to LDS, the own cells' also to memory;
//   3. the new P of the own cells: x sums from a register window sliding
//      down the lane's Vx column, z sums from the Vz rows.
// Every new value is the same float expression, in the same order, as in
// acoustic_tiled.h, so results are bitwise equal to the two-launch step.
// Outputs go to separate buffers (ping-pong): other workgroups read this
// one's inputs as their halos.  Cells outside the updated region [R, n - R)
// keep their stored values, which both buffers hold alike.
#pragma once
#include <hip/hip_runtime.h>

#include "acoustic_kernel.h"
#include "acoustic_tiled.h"
#include "device_util.h"

namespace sfd {

template <int R, int TX, int NW>
struct FusedShape {
    static constexpr int T = TX / NW;          // own rows per wave (P pass)
    static constexpr int NPT = TX + 4 * R - 2; // staged P rows: i0 - 2R + 1 ..
    static constexpr int NVT = TX + 2 * R - 1; // new Vx rows: i0 - R ..
    static constexpr int VB = (NVT + NW - 1) / NW;  // Vx rows per wave
    static constexpr int WP = 64 + 4 * R - 2;  // staged P row: columns j0 - 2R + 1 ..
    static constexpr int WV = 64 + 2 * R - 1;  // Vz row: columns j0 - R ..
    static constexpr int HP = 4 * R - 2;       // z-halo P cells per own row
    static constexpr int HV = 2 * R - 1;       // z-halo Vz cells per own row
    static constexpr int LDS = NPT * WP + NVT * 64 + TX * WV;  // floats
};

// Tile t of a shot (TX rows x 64 columns, z tiles fastest).
template <int R, int TX, int NW, int MODE, class Args>
__device__ __forceinline__ void fused_step(const Args& a, int t, int k, float* lds) {
    using S = FusedShape<R, TX, NW>;
    constexpr int T = S::T;
    constexpr bool CPML = MODE == MODE_FAST_CPML;
    const int nx = a.nx, nz = a.nz, lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ntz = a.ntz;
    const int j0 = R + (t % ntz) * kTiledTZ;
    const int i0 = R + (t / ntz) * TX;
    const int j = j0 + lane;
    const bool jok = j < nz - R;
    float* pw = lds;                       // [NPT][WP], row r = global row i0 - 2R + 1 + r
    float* svx = pw + S::NPT * S::WP;      // [NVT][64], row r = global row i0 - R + r
    float* svz = svx + S::NVT * 64;        // [TX][WV], row r = global row i0 + r
    // State and model fields lie in one arena: raw buffer accesses, the
    // column as the vector offset and a wave-uniform row as the scalar one.
    // Offsets past the arena read 0.
    const Buf ar(a.arena, a.arena_bytes);
    const char* base = reinterpret_cast<const char*>(a.arena);
    auto foff = [&](const float* f) { return (int)(reinterpret_cast<const char*>(f) - base); };
    const int rb = nz * 4;
    const int o_p = foff(a.p_in), o_vx = foff(a.vx_in), o_vz = foff(a.vz_in), o_kf = foff(a.kf);
    const int o_po = foff(a.p_out), o_vxo = foff(a.vx_out), o_vzo = foff(a.vz_out);
    auto row = [&](int o, int i) { return (uint32_t)(o + i * rb); };
    const int vj = j * 4, vjm = jok ? vj : kOOB;
    const float bfc = a.bf_const;
    const float amp = a.amp[k];
    const int sxr = a.sx, szc = a.sz;
    const int pr0 = i0 - 2 * R + 1;
    // ---- 1. stage P.  Every lane loads its column on every staged row
    // (columns past nz - R are read by updated neighbours' z stencils;
    // columns past either edge wrap into the neighbouring row, which no
    // updated cell's stencil reaches), then the own rows' halo columns.
    constexpr int QP = (S::NPT + NW - 1) / NW;
    float pv[QP];
#pragma unroll
    for (int q = 0; q < QP; ++q) {
        const int r = wv + q * NW;
        pv[q] = r < S::NPT ? ar.ld(vj, row(o_p, pr0 + r)) : 0.f;
    }
    constexpr int NH = TX * S::HP, QH = (NH + 64 * NW - 1) / (64 * NW);
    float hv[QH];
    int hidx[QH];
#pragma unroll
    for (int q = 0; q < QH; ++q) {
        const int e = threadIdx.x + q * 64 * NW;
        const int r = e / S::HP, c0 = e - r * S::HP;
        const int c = c0 < 2 * R - 1 ? c0 : c0 + 64;  // column in the staged row
        const int gi = i0 + r, gj = j0 - 2 * R + 1 + c;
        float x = e < NH ? ar.ld(gj * 4, row(o_p, gi)) : 0.f;
        if (gi == sxr && gj == szc) x -= amp;
        hv[q] = x;
        hidx[q] = e < NH ? (r + 2 * R - 1) * S::WP + c : -1;
    }
    // Own-cell inputs of the later passes, loaded now so their latency
    // overlaps the staging.
    float kf[T], vz0[T];
    const int r0 = i0 + wv * T;  // this wave's own rows (P pass)
#pragma unroll
    for (int q = 0; q < T; ++q) {
        kf[q] = ar.ld(vjm, row(o_kf, r0 + q));
        vz0[q] = ar.ld(vj, row(o_vz, r0 + q));
    }
    const int vb0 = wv * S::VB;  // this wave's first Vx row (index into svx)
    float vx0[S::VB];
#pragma unroll
    for (int q = 0; q < S::VB; ++q) vx0[q] = ar.ld(vjm, row(o_vx, i0 - R + vb0 + q));
    // Halo Vz cells (the tile's rows x HV columns j0 - R + c', c' < R or
    // c' >= 64 + R), spread over all threads of the workgroup.
    constexpr int NZH = TX * S::HV, QZ = (NZH + 64 * NW - 1) / (64 * NW);
    float vzh[QZ];
    int zr[QZ], zcv[QZ];
#pragma unroll
    for (int q = 0; q < QZ; ++q) {
        const int e = threadIdx.x + q * 64 * NW;
        const int r = e / S::HV, c0 = e - r * S::HV;
        const int c = c0 < R ? c0 : c0 + 64;
        const int jj = j0 - R + c;
        const bool ok = e < NZH && i0 + r < nx - R && jj >= R && jj < nz - R;
        zr[q] = ok ? r : -1;  // -1: not an updated cell (keeps its stored value)
        zcv[q] = e < NZH ? r * S::WV + c : -1;
        vzh[q] = e < NZH ? ar.ld(jj * 4, row(o_vz, i0 + r)) : 0.f;
    }
    // Source on the staged column values.
    if (szc >= j0 && szc < j0 + 64 && sxr >= pr0 && sxr < pr0 + S::NPT) {
#pragma unroll
        for (int q = 0; q < QP; ++q)
            if (pr0 + wv + q * NW == sxr && j == szc) pv[q] -= amp;
    }
#pragma unroll
    for (int q = 0; q < QP; ++q) {
        const int r = wv + q * NW;
        if (r < S::NPT) pw[r * S::WP + 2 * R - 1 + lane] = pv[q];
    }
#pragma unroll
    for (int q = 0; q < QH; ++q)
        if (hidx[q] >= 0) pw[hidx[q]] = hv[q];
    __syncthreads();
    float cx[R];
#pragma unroll
    for (int m = 0; m < R; ++m) cx[m] = a.cx[m];
    const float zx = a.zx;
    const int xw = a.pml_xw, zw = a.pml_zw, sw = a.sponge_w;
    const bool raw = a.v_raw != 0;
    const float az_j = (!CPML && jok) ? a.az[j] : 1.f;
    auto bf_at = [&](int i, int jj) { return bfc != 0.f ? bfc : a.bf[(ptrdiff_t)i * nz + jj]; };
    // ---- 2a. new Vx of rows i0 - R + vb0 .. (this wave's block), sliding
    // down the lane's P column: Vx(i) reads P rows i - R + 1 .. i + R.
    {
        float w[S::VB + 2 * R - 1];
        const int ib = i0 - R + vb0;  // first row; P row i is pw row i - pr0
#pragma unroll
        for (int q = 0; q < S::VB + 2 * R - 1; ++q) {
            const int r = ib - R + 1 + q - pr0;  // = vb0 + q, always within the staged rows
            w[q] = r < S::NPT ? pw[r * S::WP + 2 * R - 1 + lane] : 0.f;
        }
#pragma unroll
        for (int q = 0; q < S::VB; ++q) {
            if (vb0 + q >= S::NVT) break;  // wave-uniform: past the last Vx row
            const int i = ib + q;
            float v = vx0[q];
            if (i >= R && i < nx - R && jok) {
                // P(i) = w[q + R - 1]
                float sx = cx[0] * (w[q + R] - w[q + R - 1]);
#pragma unroll
                for (int m = 2; m <= R; ++m) sx = __builtin_fmaf(cx[m - 1], w[q + R - 1 + m] - w[q + R - m], sx);
                float dx = sx;
                const bool own = i >= i0 && i < i0 + TX;
                if constexpr (CPML) {
                    int xr = -1;
                    if (i < xw) xr = i;
                    else if (i >= nx - xw) xr = i - (nx - xw) + xw;
                    if (xr >= 0) {
                        const size_t e = (size_t)xr * nz + j;
                        const float psi =
                            __builtin_fmaf(a.pml_x[3 * nx + i], sx, a.pml_x[2 * nx + i] * a.psi_px_in[e]);
                        if (own) a.psi_px_out[e] = psi;
                        dx = sx + psi;
                    }
                } else {
                    if (raw) v *= tiled_sponge(i, nx, sw, ro(a.ax, i), az_j);
                }
                v = __builtin_fmaf(-bf_at(i, j), dx, v);
                if (own) ar.st(v, vj, row(o_vxo, i));
            }
            svx[(vb0 + q) * 64 + lane] = v;
        }
    }
    // ---- 2b. new Vz of this wave's own rows over columns j0 - R ...
    int zc = -1;
    if constexpr (CPML) {
        if (j < zw) zc = j;
        else if (j >= nz - zw) zc = j - (nz - zw) + zw;
    }
#pragma unroll
    for (int q = 0; q < T; ++q) {
        const int i = r0 + q;
        if (i >= nx - R) break;  // wave-uniform: the rows below are not read
        const float fi = CPML ? 1.f : ro(a.ax, i);
        const float* prow = pw + (i - pr0) * S::WP;
        {
            const float* p = prow + lane + 2 * R - 1;  // P(j)
            float v = vz0[q];
            if (jok) {
                float sz = cx[0] * (p[1] - p[0]);
#pragma unroll
                for (int m = 2; m <= R; ++m) sz = __builtin_fmaf(cx[m - 1], p[m] - p[1 - m], sz);
                sz = zx * sz;
                float dz = sz;
                if constexpr (CPML) {
                    if (zc >= 0) {
                        const size_t e = (size_t)i * (2 * zw) + zc;
                        const float psi =
                            __builtin_fmaf(a.pml_z[3 * nz + j], sz, a.pml_z[2 * nz + j] * a.psi_pz_in[e]);
                        a.psi_pz_out[e] = psi;
                        dz = sz + psi;
                    }
                } else {
                    if (raw) v *= tiled_sponge(i, nx, sw, fi, az_j);
                }
                v = __builtin_fmaf(-bf_at(i, j), dz, v);
                ar.st(v, vj, row(o_vzo, i));
            }
            svz[(i - i0) * S::WV + lane + R] = v;
        }
    }
    // Halo Vz cells: P(jj) sits at staged column c + R - 1 of the row.
#pragma unroll
    for (int q = 0; q < QZ; ++q) {
        if (zcv[q] < 0) continue;
        float v = vzh[q];
        if (zr[q] >= 0) {
            const int r = zr[q], c = zcv[q] - r * S::WV, i = i0 + r, jj = j0 - R + c;
            const float* p = pw + (r + 2 * R - 1) * S::WP + c + R - 1;
            float sz = cx[0] * (p[1] - p[0]);
#pragma unroll
            for (int m = 2; m <= R; ++m) sz = __builtin_fmaf(cx[m - 1], p[m] - p[1 - m], sz);
            sz = zx * sz;
            float dz = sz;
            if constexpr (CPML) {
                int zch = -1;
                if (jj < zw) zch = jj;
                else if (jj >= nz - zw) zch = jj - (nz - zw) + zw;
                if (zch >= 0) {
                    const size_t e = (size_t)i * (2 * zw) + zch;
                    const float psi =
                        __builtin_fmaf(a.pml_z[3 * nz + jj], sz, a.pml_z[2 * nz + jj] * a.psi_pz_in[e]);
                    dz = sz + psi;
                }
            } else {
                if (raw) v *= tiled_sponge(i, nx, sw, a.ax[i], a.az[jj]);
            }
            v = __builtin_fmaf(-bf_at(i, jj), dz, v);
        }
        svz[zcv[q]] = v;
    }
    __syncthreads();
    // ---- 3. new P of this wave's own rows: x sums sliding down the lane's
    // Vx column (P(i) reads Vx rows i - R .. i + R - 1), z sums on Vz rows.
    float wx[T + 2 * R - 1];
#pragma unroll
    for (int q = 0; q < T + 2 * R - 1; ++q) wx[q] = svx[(r0 - i0 + q) * 64 + lane];  // row r0 - R + q
    float* gk = a.rec_buf + (size_t)k * a.nslot;
#pragma unroll
    for (int q = 0; q < T; ++q) {
        const int i = r0 + q;
        if (i >= nx - R) break;  // wave-uniform
        // Vx(i) = wx[q + R]
        float sx = cx[0] * (wx[q + R] - wx[q + R - 1]);
#pragma unroll
        for (int m = 2; m <= R; ++m) sx = __builtin_fmaf(cx[m - 1], wx[q + R + m - 1] - wx[q + R - m], sx);
        const float* pz = svz + (i - i0) * S::WV + lane + R;  // Vz(j)
        float sz = cx[0] * (pz[0] - pz[-1]);
#pragma unroll
        for (int m = 2; m <= R; ++m) sz = __builtin_fmaf(cx[m - 1], pz[m - 1] - pz[-m], sz);
        sz = zx * sz;
        const float p0 = pw[(i - pr0) * S::WP + 2 * R - 1 + lane];
        if (!jok) continue;
        if constexpr (CPML) {
            int xr = -1;
            if (i < xw) xr = i;
            else if (i >= nx - xw) xr = i - (nx - xw) + xw;
            if (xr >= 0) {
                const size_t e = (size_t)xr * nz + j;
                const float psi = __builtin_fmaf(a.pml_x[nx + i], sx, a.pml_x[i] * a.psi_vx_in[e]);
                a.psi_vx_out[e] = psi;
                sx += psi;
            }
            if (zc >= 0) {
                const size_t e = (size_t)i * (2 * zw) + zc;
                const float psi = __builtin_fmaf(a.pml_z[nz + j], sz, a.pml_z[j] * a.psi_vz_in[e]);
                a.psi_vz_out[e] = psi;
                sz += psi;
            }
        }
        const float pc = __builtin_fmaf(-kf[q], sx + sz, p0);
        const float pn = CPML ? pc : pc * tiled_sponge(i, nx, sw, ro(a.ax, i), az_j);
        ar.st(pn, vj, row(o_po, i));
        const int4 sp4 = a.rec_row[i];
        if (j >= sp4.x && j < sp4.y) gk[sp4.z + (j - sp4.x)] = pn;
        for (int more = sp4.w; more >= 0;) {
            const int4 mp = a.rec_more[more];
            if (j >= mp.x && j < mp.y) gk[mp.z + (j - mp.x)] = pn;
            more = mp.w;
        }
    }
}

template <int R, int TX, int NW, int MODE>
__global__ __launch_bounds__(64 * NW) void acoustic_fused_step(const AcousticStepArgs a) {
    __shared__ float lds[FusedShape<R, TX, NW>::LDS];
    fused_step<R, TX, NW, MODE>(a, xcd_tile((int)blockIdx.x, a.ntz * a.nchunk), a.k, lds);
}

// Many shots per launch: b.shots[parity * nshot + s] is shot s's argument
// block for the step's ping-pong order, b.tile_map[g] = (shot, the shot's
// tile) of global tile g.
template <int R, int TX, int NW, int MODE>
__global__ __launch_bounds__(64 * NW) void acoustic_fused_step_batch(const BatchStepArgs b) {
    __shared__ float lds[FusedShape<R, TX, NW>::LDS];
    const int* m = reinterpret_cast<const int*>(b.tile_map) + 2 * xcd_tile((int)blockIdx.x, b.ntile);
    using CArgs = const __attribute__((address_space(4))) AcousticStepArgs;
    CArgs& a = *(CArgs*)(b.shots + (size_t)b.parity * b.nshot + ro(m, 0));
    fused_step<R, TX, NW, MODE>(a, ro(m, 1), b.k, lds);
}

}  // namespace sfd
