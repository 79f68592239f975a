// Streaming one-launch time step of the acoustic velocity-pressure
// staggered-grid propagator in fast arithmetic (reference loop
// forward/Acoustic_FD_by_ss/src/Acoustic_1order_FD.cpp:247-377).
//
// One launch per step on ping-pong buffers, so each field crosses HBM about
// once per step (P, Vx, Vz read and written, kf read: 28 B per cell with
// uniform density) against the two-launch step's 40 B.  Unlike the tiled
// one-launch step (acoustic_fused.h), which recomputes the new velocities on
// an R-row halo of every tile, a workgroup here owns a strip of 64 columns
// over a chunk of rows and marches down it in blocks of B rows, keeping in LDS
//   * a ring of P rows: the WIN = B + 2R - 1 rows a block's new Vx read, and
//     the previous block's rows, which slower waves may still read;
//   * a ring of new Vx rows: the WIN rows a block's new P read;
//   * the new Vz of the block's rows over the strip and an R / R - 1 column
//     halo;
// so a P row is read from memory once per strip (plus the 4R - 2 rows that
// prime each chunk) and each new Vx row is computed once.  Every block
//   A  stores the P rows loaded during the previous block into the ring (and
//      the z-halo columns of the rows it updates);          barrier
//   B  issues the next block's loads (registers);
//   C  computes B new Vx rows (each lane slides a register window down its P
//      column) and the new Vz of its B rows;                 barrier
//   E  computes the new P of its B rows: x sums from a register window down
//      the lane's Vx column, z sums from the Vz rows.
// The fronts lag by the stencil's reach: a block updates P rows [u, u + B),
// computes Vx rows [u + R - 1, u + R - 1 + B) and loads P rows
// [u + 2R - 1, u + 2R - 1 + B); a chunk starts its march at u = c0 - 4R + 2.
// The rings' moduli are multiples of B and a block's windows start on a
// multiple of B, so a window never wraps: the first 2R - 1 slots are mirrored
// past each ring's end, and every LDS access of a block is a fixed offset from
// one per-block address.  Blocks wholly inside the chunk take a path without
// row masks; the chunk's first and last blocks mask the rows outside it.
// Every new value is the same float expression, in the same order, as in
// acoustic_tiled.h, so results are bitwise equal to the two-launch step.
// Cells outside the updated region [R, n - R) keep their stored values, which
// both buffers hold alike.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "acoustic_kernel.h"
#include "acoustic_tiled.h"
#include "device_util.h"

namespace sfd {

template <int R, int NW>
struct StreamShape {
    static constexpr int RPW = 4;                    // rows per wave per block
    static constexpr int B = RPW * NW;               // rows per block
    static constexpr int WIN = B + 2 * R - 1;        // rows of a block's P window (new Vx) and Vx window (new P)
    static constexpr int EXT = WIN - B;              // ring slots mirrored past the end
    static constexpr int MP = 3 * B;                 // P ring modulus (window + the previous block's rows)
    static constexpr int MV = 2 * B;                 // Vx ring modulus (window)
    static constexpr int WP = 64 + 4 * R - 2;        // P ring row: columns j0 - 2R + 1 ..
    static constexpr int WV = 64 + 2 * R - 1;        // Vz row: columns j0 - R ..
    static constexpr int HP = 4 * R - 2;             // z-halo P cells of an updated row
    static constexpr int HV = 2 * R - 1;             // z-halo Vz cells of an updated row
    static constexpr int NT = 64 * NW;               // threads
    static constexpr int QH = (B * HP + NT - 1) / NT;  // halo P cells per thread and block
    static constexpr int QZ = (B * HV + NT - 1) / NT;  // halo Vz cells per thread and block
    static constexpr int LDS = (MP + EXT) * WP + (MV + EXT) * 64 + B * WV;  // floats
    // A block's writes never reach the slots the previous block still reads.
    static_assert(B >= 2 * R, "rows per block must cover the stencil reach");
};

// Tile t of a shot: strip t % ntz (64 columns), chunk t / ntz (a.chunk rows).
// Loads a block consumes are issued before the next block's prefetch and
// every store is issued whatever the masks say (vector memory counters retire
// in order, so a wait on a load issued after the prefetch, or an unknown
// number of later stores, would also wait for the prefetch); per-row values
// are scalar loads.
// UBF: uniform density, every updated cell's bf is a.bf_const.
template <int R, int NW, int MODE, bool UBF, class Args>
__device__ __forceinline__ void stream_step(const Args& a, int t, int k, float* lds) {
    using S = StreamShape<R, NW>;
    constexpr int B = S::B, RPW = S::RPW, QH = S::QH, QZ = S::QZ, WP = S::WP, WV = S::WV;
    constexpr int MP = S::MP, MV = S::MV, EXT = S::EXT;
    constexpr bool CPML = MODE == MODE_FAST_CPML;
    const int nx = a.nx, nz = a.nz, lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wr = wv * RPW;  // this wave's first row within a block
    const int ntz = a.ntz;
    const int j0 = R + (t % ntz) * kTiledTZ;
    const int c0 = R + (t / ntz) * a.chunk;
    const int c1 = min(c0 + a.chunk, nx - R);  // own rows [c0, c1)
    const int j = j0 + lane;
    const bool jok = j < nz - R;
    float* pr = lds;                            // P ring [MP + EXT][WP]
    float* vr = pr + (MP + EXT) * WP;           // Vx ring [MV + EXT][64]
    float* zb = vr + (MV + EXT) * 64;           // Vz of the block's rows [B][WV]
    const Buf ar(a.arena, a.arena_bytes);
    const char* abase = reinterpret_cast<const char*>(a.arena);
    auto foff = [&](const float* f) { return (int)(reinterpret_cast<const char*>(f) - abase); };
    const int rb = nz * 4;
    const int o_p = foff(a.p_in), o_vx = foff(a.vx_in), o_vz = foff(a.vz_in), o_kf = foff(a.kf);
    const int o_po = foff(a.p_out), o_vxo = foff(a.vx_out), o_vzo = foff(a.vz_out), o_bf = foff(a.bf);
    const int vj = j * 4, vjm = jok ? vj : kOOB;
    const float bfc = a.bf_const;
    const float amp = a.amp[k];
    const int sxr = a.sx, szc = a.sz;
    const bool srcl = j == szc;  // this lane's column holds the source
    float cx[R];
#pragma unroll
    for (int m = 0; m < R; ++m) cx[m] = a.cx[m];
    const float zx = a.zx;
    const int xw = a.pml_xw, zw = a.pml_zw, sw = a.sponge_w;
    const bool raw = a.v_raw != 0;
    auto xstrip = [&](int i) { return i < xw ? i : i >= nx - xw ? i - (nx - xw) + xw : -1; };
    auto zstrip = [&](int jj) { return jj < zw ? jj : jj >= nz - zw ? jj - (nz - zw) + zw : -1; };
    // Own rows [c0, c1); P rows [c0 - 2R + 1, c1 + 2R - 1) and new Vx rows
    // [c0 - R, c1 + R - 1) are read by them.  Masked loads read 0 and cost no
    // traffic; masked stores are dropped.
    const int pfirst = c0 - 2 * R + 1, plast = c1 + 2 * R - 1;
    const int vfirst = c0 - R, vlast = c1 + R - 1;
    auto own = [&](int i) { return i >= c0 && i < c1; };
    // A block needs no masks when its rows, the Vx rows it computes and the
    // P rows it loads all lie inside those ranges (and so inside [R, nx - R)).
    auto inner = [&](int u) { return u >= c0 && u + B + R <= c1; };
    // ---- per-thread constants of the march: the own column's C-PML profile
    // or sponge factor, and those of this thread's halo Vz cells (a fixed
    // column per thread; its row moves with the block).
    const int zc = CPML ? zstrip(j) : -1;
    float zb0 = 0.f, za0 = 0.f, zb2 = 0.f, za2 = 0.f, az_j = 1.f;
    if constexpr (CPML) {
        if (zc >= 0) {
            zb0 = a.pml_z[j];
            za0 = a.pml_z[nz + j];
            zb2 = a.pml_z[2 * nz + j];
            za2 = a.pml_z[3 * nz + j];
        }
    } else {
        if (jok) az_j = a.az[j];
    }
    int hr[QZ], hc[QZ], hz[QZ];  // halo Vz cell: row in the block (-1: none), column in zb, C-PML strip column
    float hb2[QZ], ha2[QZ], haz[QZ];
#pragma unroll
    for (int q = 0; q < QZ; ++q) {
        const int e = threadIdx.x + q * S::NT;
        const int r = e / S::HV, ch = e - r * S::HV;
        const int c = ch < R ? ch : ch + 64;
        const int jj = j0 - R + c;
        const bool upd = e < B * S::HV && jj >= R && jj < nz - R;  // an updated column
        hr[q] = e < B * S::HV ? r : -1;
        hc[q] = c;
        hz[q] = upd && CPML ? zstrip(jj) : -1;
        hb2[q] = hz[q] >= 0 ? a.pml_z[2 * nz + jj] : 0.f;
        ha2[q] = hz[q] >= 0 ? a.pml_z[3 * nz + jj] : 0.f;
        haz[q] = upd && !CPML ? a.az[jj] : 1.f;
        if (!upd) hz[q] = -2;  // not updated: keeps its stored value
    }
    // Halo P cells a thread stages per block: row in the block, column in the
    // ring row (-1: none).
    int pr_r[QH], pr_c[QH];
#pragma unroll
    for (int q = 0; q < QH; ++q) {
        const int e = threadIdx.x + q * S::NT;
        const int r = e / S::HP, c0h = e - r * S::HP;
        pr_r[q] = e < B * S::HP ? r : -1;
        pr_c[q] = c0h < 2 * R - 1 ? c0h : c0h + 64;
    }
    // ---- the next block's stencil inputs (registers), issued one block ahead.
    float pp[RPW], ph[QH], vxo[RPW], kfv[RPW], vz0[RPW], vzh[QZ];
    auto prefetch = [&](int u) {
        const int bp = o_p + (u + 2 * R - 1 + wr) * rb;  // this wave's first P row
        const int bv = o_vx + (u + R - 1 + wr) * rb;     // first old Vx row
        const int bk = o_kf + (u + wr) * rb;             // first updated row
        const int bz = o_vz + (u + wr) * rb;
        const bool edge = !inner(u);  // rows outside the chunk's ranges are masked
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
            int op = vj, ov = vjm, ou = vjm, oz = vj;
            if (edge) {
                const int rp = u + 2 * R - 1 + wr + q, rv = u + R - 1 + wr + q, ru = u + wr + q;
                if (rp < pfirst || rp >= plast) op = kOOB;
                if (rv < vfirst || rv >= vlast) ov = kOOB;
                if (!own(ru)) ou = oz = kOOB;
            }
            pp[q] = ar.ld(op, (uint32_t)(bp + q * rb));
            vxo[q] = ar.ld(ov, (uint32_t)(bv + q * rb));
            kfv[q] = ar.ld(ou, (uint32_t)(bk + q * rb));
            vz0[q] = ar.ld(oz, (uint32_t)(bz + q * rb));
        }
        // Halo cells: the per-lane row goes into the vector offset (a per-lane
        // scalar offset would serialise the load over the lanes' rows).
        const int bh = o_p + u * rb, bzh = o_vz + u * rb;
#pragma unroll
        for (int q = 0; q < QH; ++q) {
            const int gj = j0 - 2 * R + 1 + pr_c[q];
            const bool ok = pr_r[q] >= 0 && own(u + pr_r[q]);
            ph[q] = ar.ld(ok ? pr_r[q] * rb + gj * 4 : kOOB, (uint32_t)bh);
        }
#pragma unroll
        for (int q = 0; q < QZ; ++q) {
            const bool ok = hr[q] >= 0 && own(u + hr[q]);
            vzh[q] = ar.ld(ok ? hr[q] * rb + (j0 - R + hc[q]) * 4 : kOOB, (uint32_t)bzh);
        }
    };
    // C-PML memory fields (in the arena): x strips [2 xw][nz], z strips [nx][2 zw].
    const int o_spx = CPML ? foff(a.psi_px_in) : 0, o_svx = CPML ? foff(a.psi_vx_in) : 0;
    const int o_spz = CPML ? foff(a.psi_pz_in) : 0, o_svz = CPML ? foff(a.psi_vz_in) : 0;
    const int o_spxo = CPML ? foff(a.psi_px_out) : 0, o_svxo = CPML ? foff(a.psi_vx_out) : 0;
    const int o_spzo = CPML ? foff(a.psi_pz_out) : 0, o_svzo = CPML ? foff(a.psi_vz_out) : 0;
    const int zrb = 2 * zw * 4;  // bytes per row of a z strip field
    // The workgroup's columns (with the Vz halo) reach a z strip.
    const bool zwg = CPML && (j0 - R < zw || j0 + 64 + R - 1 > nz - zw);
    // Registers of the block being computed.
    float vxc[RPW], kfc[RPW], vzc[RPW], vzhc[QZ];
    // Side inputs of the block (C-PML memory, per-cell bf, halo sponge rows).
    float spx[RPW], spz[RPW], svx[RPW], svz[RPW], bfv[RPW], bfz[RPW], spzh[QZ], bfh[QZ], axh[QZ];
    const int nblk = (c1 - c0 + 4 * R - 2 + B - 1) / B;
    int u = c0 - 4 * R + 2;
    int sp = 0, sv = 0;  // ring slots of the block's P window (row u) and Vx window (row u - R)
    bool xblk = false;   // the block's rows [u, u + B + R - 1) reach an x strip
    prefetch(u);
    // ---- A: P rows u + 2R - 1 + wr + q into the ring (and the mirror slot)
    // with the source term subtracted; z-halo columns of the updated rows;
    // the block's side inputs.  One code path for every block: the side
    // loads precede the prefetch, and a single path keeps the wait counts of
    // the vector memory counter exact.
    auto phase_a = [&]() {
        xblk = CPML && (u < xw || u + B + R - 1 > nx - xw);
        const int sq = sxr - (u + 2 * R - 1 + wr);  // the source row among this wave's P rows
        if (sq >= 0 && sq < RPW) {
            // x - (+0) is x bit for bit; every row is written, so no row
            // index is chosen at run time.
#pragma unroll
            for (int q = 0; q < RPW; ++q) pp[q] -= (q == sq && srcl) ? amp : 0.f;
        }
        // Row u + c goes to slot sp + c (below MP + EXT: contiguous with the
        // windows that start at sp) and, past MP, to its wrapped slot too.
        float* pa = pr + (sp + 2 * R - 1 + wr) * WP + 2 * R - 1 + lane;
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
            pa[q * WP] = pp[q];
            if (sp + 2 * R - 1 + wr + q >= MP) pa[(q - MP) * WP] = pp[q];
        }
        const bool srow = sxr >= u && sxr < u + B;  // the source row is among the block's rows
#pragma unroll
        for (int q = 0; q < QH; ++q) {
            if (pr_r[q] < 0) continue;
            float x = ph[q];
            if (srow && u + pr_r[q] == sxr && j0 - 2 * R + 1 + pr_c[q] == szc) x -= amp;
            pr[(sp + pr_r[q]) * WP + pr_c[q]] = x;
        }
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
            vxc[q] = vxo[q];
            kfc[q] = kfv[q];
            vzc[q] = vz0[q];
            const int iv = u + R - 1 + wr + q, iu = u + wr + q;
            const bool vok = jok && iv >= vfirst && iv < vlast && iv >= R && iv < nx - R;
            const bool uok = jok && own(iu);
            if constexpr (UBF) {
                bfv[q] = bfz[q] = bfc;
            } else {
                bfv[q] = ar.ld(vok ? o_bf + (iv * nz + j) * 4 : kOOB, 0);
                bfz[q] = ar.ld(uok ? o_bf + (iu * nz + j) * 4 : kOOB, 0);
            }
            if constexpr (CPML) {
                // (scalar offsets stay inside the field even for masked lanes)
                if (xblk) {
                    const int xv = xstrip(iv), xu = xstrip(iu);
                    spx[q] = ar.ld(vok && xv >= 0 ? vj : kOOB, (uint32_t)(o_spx + max(xv, 0) * rb));
                    svx[q] = ar.ld(uok && xu >= 0 ? vj : kOOB, (uint32_t)(o_svx + max(xu, 0) * rb));
                }
                if (zwg) {
                    spz[q] = ar.ld(uok && zc >= 0 ? zc * 4 : kOOB, (uint32_t)(o_spz + max(iu, 0) * zrb));
                    svz[q] = ar.ld(uok && zc >= 0 ? zc * 4 : kOOB, (uint32_t)(o_svz + max(iu, 0) * zrb));
                }
            }
        }
#pragma unroll
        for (int q = 0; q < QZ; ++q) {
            vzhc[q] = vzh[q];
            const int i = u + hr[q];
            const bool upd = hr[q] >= 0 && hz[q] != -2 && own(i);
            const int jj = j0 - R + hc[q];
            if constexpr (UBF) bfh[q] = bfc;
            else bfh[q] = ar.ld(upd ? o_bf + (i * nz + jj) * 4 : kOOB, 0);
            if constexpr (CPML) {
                if (zwg) spzh[q] = ar.ld(upd && hz[q] >= 0 ? i * zrb + hz[q] * 4 : kOOB, (uint32_t)o_spz);
            } else {
                axh[q] = upd && raw ? a.ax[i] : 1.f;
            }
        }
    };
    // ---- C (new velocities), barrier, E (new pressure) of one block; blocks
    // wholly inside the chunk skip the row masks.
    auto compute = [&](auto edge) {
        constexpr bool EDGE = decltype(edge)::value;
        const float* pw = pr + sp * WP + 2 * R - 1 + lane;  // P window: row u + r at pw[r * WP]
        const float* vw = vr + sv * 64 + lane;              // Vx window: row u - R + r at vw[r * 64]
        // ---- C1: new Vx of rows u + R - 1 + wr + q: P rows i - R + 1 .. i + R.
        {
            float w[RPW + 2 * R - 1];
#pragma unroll
            for (int q = 0; q < RPW + 2 * R - 1; ++q) w[q] = pw[(wr + q) * WP];  // P row u + wr + q
            // Vx row u - R + c goes to slot sv + c and, past MV, to its
            // wrapped slot too.
            float* va = vr + (sv + 2 * R - 1 + wr) * 64 + lane;
#pragma unroll
            for (int q = 0; q < RPW; ++q) {
                const int i = u + R - 1 + wr + q;
                float v = vxc[q];
                bool upd = jok, st = jok;
                if constexpr (EDGE) {
                    upd = upd && i >= vfirst && i < vlast && i >= R && i < nx - R;
                    st = upd && own(i);
                }
                if (upd) {
                    // P(i) = w[q + R - 1]
                    float sx = cx[0] * (w[q + R] - w[q + R - 1]);
#pragma unroll
                    for (int m = 2; m <= R; ++m)
                        sx = __builtin_fmaf(cx[m - 1], w[q + R - 1 + m] - w[q + R - m], sx);
                    float dx = sx;
                    if constexpr (CPML) {
                        const int xr = xblk ? xstrip(i) : -1;
                        if (xr >= 0) {
                            const float psi = __builtin_fmaf(ro(a.pml_x, 3 * nx + i), sx, ro(a.pml_x, 2 * nx + i) * spx[q]);
                            ar.st(psi, !EDGE || own(i) ? vj : kOOB, (uint32_t)(o_spxo + xr * rb));
                            dx = sx + psi;
                        }
                    } else {
                        if (raw) v *= tiled_sponge(i, nx, sw, ro(a.ax, i), az_j);
                    }
                    v = __builtin_fmaf(-bfv[q], dx, v);
                }
                ar.st(v, st ? vj : kOOB, (uint32_t)(o_vxo + i * rb));
                va[q * 64] = v;
                if (sv + 2 * R - 1 + wr + q >= MV) va[(q - MV) * 64] = v;
            }
        }
        // ---- C2: new Vz of the block's rows (own columns, then the halo).
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
            const int i = u + wr + q;
            bool upd = jok;
            if constexpr (EDGE) upd = upd && own(i);
            const float* p = pw + (wr + q) * WP;  // P(i, j)
            float v = vzc[q];
            if (upd) {
                float sz = cx[0] * (p[1] - p[0]);
#pragma unroll
                for (int m = 2; m <= R; ++m) sz = __builtin_fmaf(cx[m - 1], p[m] - p[1 - m], sz);
                sz = zx * sz;
                float dz = sz;
                if constexpr (CPML) {
                    if (zwg && zc >= 0) {
                        const float psi = __builtin_fmaf(za2, sz, zb2 * spz[q]);
                        ar.st(psi, zc * 4, (uint32_t)(o_spzo + i * zrb));
                        dz = sz + psi;
                    }
                } else {
                    if (raw) v *= tiled_sponge(i, nx, sw, ro(a.ax, i), az_j);
                }
                v = __builtin_fmaf(-bfz[q], dz, v);
            }
            ar.st(v, upd ? vj : kOOB, (uint32_t)(o_vzo + i * rb));
            zb[(wr + q) * WV + lane + R] = v;
        }
#pragma unroll
        for (int q = 0; q < QZ; ++q) {
            if (hr[q] < 0) continue;
            const int i = u + hr[q], c = hc[q];
            float v = vzhc[q];
            bool upd = hz[q] != -2;
            if constexpr (EDGE) upd = upd && own(i);
            if (upd) {
                const float* p = pr + (sp + hr[q]) * WP + c + R - 1;  // P(i, jj)
                float sz = cx[0] * (p[1] - p[0]);
#pragma unroll
                for (int m = 2; m <= R; ++m) sz = __builtin_fmaf(cx[m - 1], p[m] - p[1 - m], sz);
                sz = zx * sz;
                float dz = sz;
                if constexpr (CPML) {
                    if (zwg && hz[q] >= 0) dz = sz + __builtin_fmaf(ha2[q], sz, hb2[q] * spzh[q]);
                } else {
                    if (raw) v *= tiled_sponge(i, nx, sw, axh[q], haz[q]);
                }
                v = __builtin_fmaf(-bfh[q], dz, v);
            }
            zb[hr[q] * WV + c] = v;
        }
        __syncthreads();
        // ---- E: new P of the block's rows.
        float wx[RPW + 2 * R - 1];
#pragma unroll
        for (int q = 0; q < RPW + 2 * R - 1; ++q) wx[q] = vw[(wr + q) * 64];  // Vx row u - R + wr + q
        float* gk = a.rec_buf + (size_t)k * a.nslot;
        // The first receiver span of each of this wave's rows, (jlo, jhi,
        // slot, more) per row, in scalar loads ahead of any lane-dependent
        // branch (rows outside the chunk read the padding and go unused).
        const int* rsp = reinterpret_cast<const int*>(a.rec_row + u + wr);
        int span[4 * RPW];
#pragma unroll
        for (int f = 0; f < 4 * RPW; ++f) span[f] = ro(rsp, f);
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
            const int i = u + wr + q;
            const bool ow = EDGE ? own(i) : true;  // wave-uniform
            // Vx(i) = wx[q + R]
            float sx = cx[0] * (wx[q + R] - wx[q + R - 1]);
#pragma unroll
            for (int m = 2; m <= R; ++m) sx = __builtin_fmaf(cx[m - 1], wx[q + R + m - 1] - wx[q + R - m], sx);
            const float* pz = zb + (wr + q) * WV + lane + R;  // Vz(i, j)
            float sz = cx[0] * (pz[0] - pz[-1]);
#pragma unroll
            for (int m = 2; m <= R; ++m) sz = __builtin_fmaf(cx[m - 1], pz[m - 1] - pz[-m], sz);
            sz = zx * sz;
            const float p0 = pw[(wr + q) * WP];
            const bool upd = jok && ow;
            float pn = 0.f;
            if (upd) {
                if constexpr (CPML) {
                    const int xr = xblk ? xstrip(i) : -1;
                    if (xr >= 0) {
                        const float psi = __builtin_fmaf(ro(a.pml_x, nx + i), sx, ro(a.pml_x, i) * svx[q]);
                        ar.st(psi, vj, (uint32_t)(o_svxo + xr * rb));
                        sx += psi;
                    }
                    if (zwg && zc >= 0) {
                        const float psi = __builtin_fmaf(za0, sz, zb0 * svz[q]);
                        ar.st(psi, zc * 4, (uint32_t)(o_svzo + i * zrb));
                        sz += psi;
                    }
                }
                const float pc = __builtin_fmaf(-kfc[q], sx + sz, p0);
                pn = CPML ? pc : pc * tiled_sponge(i, nx, sw, ro(a.ax, i), az_j);
            }
            ar.st(pn, upd ? vj : kOOB, (uint32_t)(o_po + i * rb));
            // Receiver spans of row i, walked by the whole wave (the lanes of
            // a span store).
            int lo = span[4 * q], hi = span[4 * q + 1], slot = span[4 * q + 2], more = span[4 * q + 3];
            if (ow && (hi > lo || more >= 0)) {
                while (true) {
                    if (jok && j >= lo && j < hi) gk[slot + (j - lo)] = pn;
                    if (more < 0) break;
                    const int* spn = reinterpret_cast<const int*>(a.rec_more + more);
                    lo = __builtin_amdgcn_readfirstlane(ro(spn, 0));
                    hi = __builtin_amdgcn_readfirstlane(ro(spn, 1));
                    slot = __builtin_amdgcn_readfirstlane(ro(spn, 2));
                    more = __builtin_amdgcn_readfirstlane(ro(spn, 3));
                }
            }
        }
    };
    for (int blk = 0; blk < nblk; ++blk, u += B) {
        phase_a();
        __syncthreads();
        // ---- B: the next block's loads (after the last block every row
        // they name lies outside the chunk's ranges: masked, no traffic; an
        // unconditional issue keeps the wait counts exact).
        prefetch(u + B);
        if (inner(u)) compute(std::false_type{});
        else compute(std::true_type{});
        sp = sp + B == MP ? 0 : sp + B;
        sv = sv + B == MV ? 0 : sv + B;
    }
}

// Occupancy: 4 workgroups of 4 waves (2 of 8) per CU with uniform density
// (the LDS rings fit 4 x 40 KB), one fewer where the per-cell bf inputs would
// not fit the registers without spills.
template <int NW, bool UBF>
constexpr int kStreamBlocksPerCU = UBF ? 16 / NW : 12 / NW;

template <int R, int NW, int MODE, bool UBF>
__global__ __launch_bounds__(64 * NW, (kStreamBlocksPerCU<NW, UBF>)) void acoustic_stream_step(const AcousticStepArgs a) {
    __shared__ float lds[StreamShape<R, NW>::LDS];
    stream_step<R, NW, MODE, UBF>(a, xcd_tile((int)blockIdx.x, a.ntz * a.nchunk), a.k, lds);
}

// Many shots per launch: b.shots[parity * nshot + s] is shot s's argument
// block for the step's ping-pong order, b.tile_map[g] = (shot, the shot's
// tile) of global tile g.
template <int R, int NW, int MODE, bool UBF>
__global__ __launch_bounds__(64 * NW, (kStreamBlocksPerCU<NW, UBF>)) void acoustic_stream_step_batch(const BatchStepArgs b) {
    __shared__ float lds[StreamShape<R, NW>::LDS];
    const int* m = reinterpret_cast<const int*>(b.tile_map) + 2 * xcd_tile((int)blockIdx.x, b.ntile);
    using CArgs = const __attribute__((address_space(4))) AcousticStepArgs;
    CArgs& a = *(CArgs*)(b.shots + (size_t)b.parity * b.nshot + ro(m, 0));
    stream_step<R, NW, MODE, UBF>(a, ro(m, 1), b.k, lds);
}

}  // namespace sfd
