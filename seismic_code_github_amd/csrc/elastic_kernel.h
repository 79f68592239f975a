// Isotropic elastic standard-staggered-grid step with a free surface over
// topography: the two half-steps of isoewm2dfsbc
// (forward/yang_version_2016_2_2/Iso_multishot_fd_subroutine_fsbc_zy_m2_pri.f90:119-196).
//
// Grid positions (cell (i, k), z fastest): tau_xx, tau_zz at (i, k); vx at
// (i - 1/2, k); vz at (i, k + 1/2); tau_xz at (i - 1/2, k + 1/2).  x
// derivatives are 10th order everywhere (calux); z derivatives drop to orders
// 8, 6, 4, 2 on the rows next to the column's surface row t (caluz).
//
// One workgroup = NW waves over a tile of 64 rows x TX columns; each wave
// owns TX / NW columns, one row per lane.  The tile and its 5-cell halo of
// every field a stencil reads is staged in LDS once, so each field is read
// from HBM about once per launch; tiles are dealt to the XCDs in bands of
// whole x-columns, so the halo re-reads hit the XCD's own L2.
//
// Cells are updated unsplit, T += dt/h pa D, except in the sponge strips
// (columns of the left / right strips, rows of the bottom strip), where the
// reference damps the x- and z-parts of each field apart (calpmlbclrb,
// pub:895-934): there both parts are kept in `parts` and T = X + Z.
//
// Two modes, the reference's two loops: 'fsbc' (isoewm2dfsbc, ABC = false),
// a free surface on row t of every column, and 'abc' (isoewm2dabc, ABC =
// true), no free surface: rows from t - npzn down are updated with full-order
// stencils and the z parts of rows t - npzn .. t are damped (calpmlbclrtb).
//
// Free surface, as the reference's call order makes it:
//   tau_zz = 0 on and above the surface row t (the source still adds there);
//   tau_xz(t - 1) = -tau_xz(t) (written with the surface value);
//   vz(t - 1) = the previous step's vz(t) + lambda/(lambda + 2 mu)
//     (vx(i+1, t) - vx(i, t)): computed by the stress kernel into `ghost` and
//     stored by the next velocity kernel;
//   row -1 is never part of the fields (the reference sums rows 1..nz only).
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "device_util.h"

namespace sfd {

constexpr int kElTZ = 64;   // tile rows (one per lane)
constexpr int kElTXMax = 32;  // widest tile (columns); the field border covers it
constexpr int kElH = 5;     // stencil half-width
constexpr int kElPad = 8;   // zero rows / columns around the grid in every field

// Per-step arguments of both kernels.  Field pointers address cell (0, 0)
// inside a zero border of kElPad cells; `pitch` floats separate columns.
struct ElasticStepArgs {
    int nx, nz, pitch;
    float *vx, *vz, *sxx, *szz, *sxz;
    const float *b2e, *b1e, *lm, *ls, *mu;  // busg2e, busg1e, lambmusg, lambsg, musge
    const int *itopo, *kst, *srow;          // surface row; first updated row; first damped row of an x-strip column
    const float *fxn, *fxh, *ratio;         // per column: node / half x-sponge, lambsg/lambmusg at t
    const float *fzn, *fzh;                 // per row: node / half bottom sponge
    const float *tzn, *tzh;                 // 'abc' top sponge: node / half, by row - (t - npzn)
    int npzn;                               // 'abc': top sponge rows - 1
    float* ghost;                           // per column: vz(t - 1) for the next step
    float* parts;                           // [10][nstrip] x / z parts of vx, vz, sxx, szz, sxz
    int nstrip, nl, nr, kb0, nb, topbase;
    int ntz, ntiles;                        // z tiles, all tiles
    float cx, cz;                           // dt/dx, dt/dz
    float a10[5], a8[4], a6[3], a4[2];      // Taylor tables (module:9-26); a20 = 1
    int sx0, sx1, sz0, sz1, src_xx;         // source box, and whether tau_xx takes it
    const float *g, *amp;                   // box weights [sx1-sx0][sz1-sz0]; wavelet per step
    int step;
    float dt, dz2;
    int nrec, rec_col, nrt;                 // receivers; rec_col < 0: no sample this step
    const int *rx, *rz;
    float *rec_vx, *rec_vz;                 // [nrec][nrt]
};

// Strip cell index, or -1: left / right strip columns over all rows, then the
// bottom-strip rows of the columns between them, then ('abc') the npzn + 1
// rows t - npzn .. t of those columns.
template <bool ABC, class A>
__device__ __forceinline__ int el_strip(const A& a, int col, int k, int t) {
    if (col < a.nl) return col * a.nz + k;
    if (col >= a.nx - a.nr) return (a.nl + col - (a.nx - a.nr)) * a.nz + k;
    if (k >= a.kb0) return (a.nl + a.nr) * a.nz + (col - a.nl) * a.nb + (k - a.kb0);
    if (ABC && k <= t && k >= t - a.npzn) return a.topbase + (col - a.nl) * (a.npzn + 1) + (k - (t - a.npzn));
    return -1;
}

// Half-width of caluz's stencil on the row d cells below the surface: zflag 0
// (u(k+n) - u(k-n-1)) is 10th order from d = 5, zflag 1 (u(k+n+1) - u(k-n))
// from d = 4 (pub:1160-1337).
__device__ __forceinline__ int el_order(int d, int zflag) {
    return zflag == 0 ? (d >= 5 ? 5 : (d >= 1 ? d : 1)) : (d >= 4 ? 5 : d + 1);
}

// z stencil sum on an LDS column: p points at the output row's own sample;
// zflag 0 pairs p[n] - p[-n-1], zflag 1 pairs p[n+1] - p[-n].
template <int ZF, class A>
__device__ __forceinline__ float el_zsum(const A& a, const float* p, int d) {
    const int lim = ZF ? 4 : 5;
    if (d >= lim) {
        float s = a.a10[0] * (p[ZF] - p[ZF - 1]);
#pragma unroll
        for (int n = 1; n < 5; ++n) s += a.a10[n] * (p[n + ZF] - p[ZF - n - 1]);
        return s;
    }
    const int M = el_order(d, ZF);
    auto c = [&](int n) { return M == 4 ? a.a8[n] : M == 3 ? a.a6[n] : a.a4[n]; };
    if (M == 1) return p[ZF] - p[ZF - 1];
    float s = c(0) * (p[ZF] - p[ZF - 1]);
    for (int n = 1; n < M; ++n) s += c(n) * (p[n + ZF] - p[ZF - n - 1]);
    return s;
}

// x stencil sum over LDS columns `stride` floats apart: xflag 0 pairs
// q[n] - q[-n-1], xflag 1 pairs q[n+1] - q[-n].
template <int XF, class A>
__device__ __forceinline__ float el_xsum(const A& a, const float* q, int stride) {
    float s = a.a10[0] * (q[XF * stride] - q[(XF - 1) * stride]);
#pragma unroll
    for (int n = 1; n < 5; ++n) s += a.a10[n] * (q[(n + XF) * stride] - q[(XF - n - 1) * stride]);
    return s;
}

// Stages columns [x0 - HX, x0 + TX + HX) x rows [z0 - HZ, z0 + kElTZ + HZ)
// of a padded field into lds[c][r]: load() issues a fixed, unrolled count of
// loads per thread into registers and store() writes them to LDS, so the
// loads of every staged field are in flight before the first LDS store waits.
template <int TX, int NW, int HX, int HZ>
struct ElStage {
    static constexpr int W = TX + 2 * HX, Hh = kElTZ + 2 * HZ, N = W * Hh, T = 64 * NW;
    static constexpr int ITER = (N + T - 1) / T;
    float v[ITER];
    __device__ __forceinline__ void load(const float* f, int pitch, int x0, int z0) {
#pragma unroll
        for (int q = 0; q < ITER; ++q) {
            const int e = threadIdx.x + q * T;
            const int c = e / Hh, r = e - c * Hh;
            v[q] = (e < N) ? f[(ptrdiff_t)(x0 - HX + c) * pitch + (z0 - HZ + r)] : 0.f;
        }
    }
    __device__ __forceinline__ void store(float* lds) const {
#pragma unroll
        for (int q = 0; q < ITER; ++q) {
            const int e = threadIdx.x + q * T;
            if (e < N) lds[e] = v[q];
        }
    }
};

// Many shots per launch: shots[s] is shot s's argument block (read through
// the constant address space), tile_map[g] = (shot, the shot's tile) of
// global tile g; step and rec_col are shared by the shots (one time axis).
struct ElasticBatchArgs {
    const ElasticStepArgs* shots;
    const int2* tile_map;
    int ntile;
    int step, rec_col;
};

using ElasticShotArgs = const __attribute__((address_space(4))) ElasticStepArgs;

// Tile t of a shot: x band t / ntz, z tile t % ntz (tiles are dealt to the
// XCDs in order, z fastest, so an XCD holds whole x bands).
template <class A>
__device__ __forceinline__ void el_tile(const A& a, int tx, int t, int* z0, int* x0) {
    *z0 = (t % a.ntz) * kElTZ;
    *x0 = (t / a.ntz) * tx;
}

// Velocity half-step (pri:119-151; 'abc': pri:300-318): vx from tau_xx,
// tau_xz; vz from tau_xz, tau_zz; in 'fsbc' mode the pending vz(t - 1).
template <int TX>
struct ElVelLds {
    static constexpr int WX = TX + 2 * kElH, HZZ = kElTZ + 2 * kElH;
    float xx[WX * kElTZ], xz[WX * HZZ], zz[TX * HZZ];
};

template <int TX, int NW, bool ABC, class A>
__device__ __forceinline__ void el_velocity(const A& a, int tile, ElVelLds<TX>& lds) {
    constexpr int HX = kElH, HZ = kElH;
    constexpr int HZZ = kElTZ + 2 * HZ;
    constexpr int NC = TX / NW;  // columns per wave
    float* s_xx = lds.xx;
    float* s_xz = lds.xz;
    float* s_zz = lds.zz;
    int z0, x0;
    el_tile(a, TX, tile, &z0, &x0);
    const int tz = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int k = z0 + tz;
    const bool kin = k < a.nz;
    float* P = a.parts;
    const size_t n = (size_t)a.nstrip;
    // The bottom sponge's factors of this thread's row.
    const bool kbot = kin && k >= a.kb0;
    const float fzk = kbot ? a.fzn[k] : 1.f, fhk = kbot ? a.fzh[k] : 1.f;
    // Every column's surface row (scalar loads) and every updated cell's
    // inputs -- busg2e, busg1e and vx, vz, or the four split parts in a
    // sponge strip -- are loaded before the staging, so their latencies
    // overlap instead of adding up column after column.
    int tc[NC], sc[NC];
    bool uc[NC];
    float bxc[NC], bzc[NC], q0[NC], q1[NC], q2[NC], q3[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        const int col = x0 + wv + j * NW;
        const bool cin = col < a.nx;  // wave-uniform
        const int t = cin ? ro(a.itopo, col) : 0;
        const int k0 = ABC ? (cin ? ro(a.kst, col) : 0) : t;
        tc[j] = t;
        uc[j] = cin && kin && k >= k0;
        sc[j] = uc[j] ? el_strip<ABC>(a, col, k, t) : -1;
        q0[j] = q1[j] = q2[j] = q3[j] = bxc[j] = bzc[j] = 0.f;
        if (uc[j]) {
            const ptrdiff_t o = (ptrdiff_t)col * a.pitch + k;
            const int sj = sc[j];
            bxc[j] = a.b2e[o];
            bzc[j] = a.b1e[o];
            if (sj < 0) {
                q0[j] = a.vx[o];
                q1[j] = a.vz[o];
            } else {
                q0[j] = P[sj];
                q1[j] = P[n + sj];
                q2[j] = P[2 * n + sj];
                q3[j] = P[3 * n + sj];
            }
        }
    }
    {
        ElStage<TX, NW, HX, 0> gxx;
        ElStage<TX, NW, HX, HZ> gxz;
        ElStage<TX, NW, 0, HZ> gzz;
        gxx.load(a.sxx, a.pitch, x0, z0);
        gxz.load(a.sxz, a.pitch, x0, z0);
        gzz.load(a.szz, a.pitch, x0, z0);
        gxx.store(s_xx);
        gxz.store(s_xz);
        gzz.store(s_zz);
    }
    __syncthreads();
    if (!kin) return;
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        const int cl = wv + j * NW, col = x0 + cl;
        if (col >= a.nx) break;
        const int t = tc[j];
        const ptrdiff_t o = (ptrdiff_t)col * a.pitch + k;
        if (!uc[j]) {
            if (!ABC && k == t - 1) a.vz[o] = a.ghost[col];
            continue;
        }
        const int d = ABC ? kElTZ : k - t;  // 'abc': 10th order on every row
        const float* pxx = s_xx + (cl + HX) * kElTZ + tz;
        const float* pxz = s_xz + (cl + HX) * HZZ + tz + HZ;
        const float* pzz = s_zz + cl * HZZ + tz + HZ;
        const float bx = bxc[j], bz = bzc[j];
        const float ux = a.cx * bx * el_xsum<0>(a, pxx, kElTZ);   // calux xflag 0 on tau_xx
        const float uz = a.cz * bx * el_zsum<0>(a, pxz, d);       // caluz zflag 0 on tau_xz
        const float wx = a.cx * bz * el_xsum<1>(a, pxz, HZZ);     // calux xflag 1 on tau_xz
        const float wz = a.cz * bz * el_zsum<1>(a, pzz, d);       // caluz zflag 1 on tau_zz
        const int s = sc[j];
        if (s < 0) {
            a.vx[o] = q0[j] + ux + uz;
            a.vz[o] = q1[j] + wx + wz;
            continue;
        }
        const bool dx = k >= ro(a.srow, col);
        float fzv = 1.f, fzw = 1.f;  // z factors of vx (node table) and vz (half table)
        bool dz = false;
        if (kbot) {
            dz = true;
            fzv = fzk;
            fzw = fhk;
        } else if (ABC && k <= t && k >= t - a.npzn) {
            dz = true;
            fzv = a.tzn[k - (t - a.npzn)];
            fzw = a.tzh[k - (t - a.npzn)];
        }
        float X = q0[j] + ux, Z = q1[j] + uz;
        if (dx) X *= ro(a.fxn, col);
        if (dz) Z *= fzv;
        P[s] = X;
        P[n + s] = Z;
        a.vx[o] = X + Z;
        X = q2[j] + wx;
        Z = q3[j] + wz;
        if (dx) X *= ro(a.fxh, col);
        if (dz) Z *= fzw;
        P[2 * n + s] = X;
        P[3 * n + s] = Z;
        a.vz[o] = X + Z;
    }
}

// Stress half-step (pri:153-196; 'abc': pri:320-352) and the step's receiver
// samples.
template <int TX>
struct ElStrLds {
    static constexpr int WX = TX + 2 * kElH, HZZ = kElTZ + 2 * kElH;
    float vx[WX * HZZ], vz[WX * HZZ];
};

template <int TX, int NW, bool ABC, class A>
__device__ __forceinline__ void el_stress(const A& a, int tile, int step, int rec_col, ElStrLds<TX>& lds) {
    constexpr int HX = kElH, HZ = kElH;
    constexpr int HZZ = kElTZ + 2 * HZ;
    float* s_vx = lds.vx;
    float* s_vz = lds.vz;
    int z0, x0;
    el_tile(a, TX, tile, &z0, &x0);
    // The shot's first tile samples its receivers (the velocities are final).
    if (tile == 0 && rec_col >= 0) {
        for (int r = threadIdx.x; r < a.nrec; r += 64 * NW) {
            const ptrdiff_t o = (ptrdiff_t)a.rx[r] * a.pitch + a.rz[r];
            a.rec_vx[(size_t)r * a.nrt + rec_col] = a.vx[o];
            a.rec_vz[(size_t)r * a.nrt + rec_col] = a.vz[o];
        }
    }
    constexpr int NC = TX / NW;  // columns per wave
    const int tz = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int k = z0 + tz;
    const bool kin = k < a.nz;
    const bool kbot = kin && k >= a.kb0;
    const float fzk = kbot ? a.fzn[k] : 1.f, fhk = kbot ? a.fzh[k] : 1.f;
    // As in the velocity half-step: the column's surface row (scalar loads)
    // and every updated cell's lambmusg, lambsg, musge and tau_xx, tau_zz,
    // tau_xz, or the six split stress parts in a sponge strip, are loaded
    // before the staging.
    int tc[NC];
    bool uc[NC];
    float lmc[NC], lsc[NC], muc[NC], q[6][NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        const int col = x0 + wv + j * NW;
        const bool cin = col < a.nx;  // wave-uniform
        const int t = cin ? ro(a.itopo, col) : 0;
        const int k0 = ABC ? (cin ? ro(a.kst, col) : 0) : t;
        tc[j] = t;
        uc[j] = cin && kin && k >= k0;
        lmc[j] = lsc[j] = muc[j] = 0.f;
        for (int f = 0; f < 6; ++f) q[f][j] = 0.f;
        if (uc[j]) {
            const ptrdiff_t o = (ptrdiff_t)col * a.pitch + k;
            const int sj = el_strip<ABC>(a, col, k, t);
            lmc[j] = a.lm[o];
            lsc[j] = a.ls[o];
            muc[j] = a.mu[o];
            if (sj < 0) {
                q[0][j] = a.sxx[o];
                q[1][j] = a.szz[o];
                q[2][j] = a.sxz[o];
            } else {
                const float* P = a.parts;
                const size_t n = (size_t)a.nstrip;
#pragma unroll
                for (int f = 0; f < 6; ++f) q[f][j] = P[(4 + f) * n + sj];
            }
        }
    }
    {
        ElStage<TX, NW, HX, HZ> gvx, gvz;
        gvx.load(a.vx, a.pitch, x0, z0);
        gvz.load(a.vz, a.pitch, x0, z0);
        gvx.store(s_vx);
        gvz.store(s_vz);
    }
    __syncthreads();
    if (!kin) return;
    const bool in_box_z = k >= a.sz0 && k < a.sz1;
    const float w = a.amp[step];
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        const int cl = wv + j * NW, col = x0 + cl;
        if (col >= a.nx) break;
        const int t = tc[j];
        const ptrdiff_t o = (ptrdiff_t)col * a.pitch + k;
        // addsrc (pub:1429-1466): w g dt / dz^2, into tau_zz always and into
        // tau_xx in 'abc' mode or when the source lies below the surface.
        float src = 0.f;
        const bool in_box = in_box_z && col >= a.sx0 && col < a.sx1;
        if (in_box) src = ((w * a.g[(size_t)(col - a.sx0) * (a.sz1 - a.sz0) + (k - a.sz0)]) * a.dt) / a.dz2;
        const float srx = a.src_xx ? src : 0.f;
        const int s = el_strip<ABC>(a, col, k, t);
        float* P = a.parts;
        const size_t n = (size_t)a.nstrip;
        if (!uc[j]) {
            // Above the updated rows only the source acts: the x parts take it
            // (damped where a strip's rows reach up here); 'fsbc' resets tau_zz
            // on and above the surface to it, 'abc' lets it accumulate.
            if (s >= 0) {
                const float fx = k >= ro(a.srow, col) ? ro(a.fxh, col) : 1.f;
                float X = P[4 * n + s];
                if (k >= ro(a.srow, col)) X *= fx;
                X += srx;
                P[4 * n + s] = X;
                a.sxx[o] = X + P[5 * n + s];
                if (ABC) {
                    X = P[6 * n + s];
                    if (k >= ro(a.srow, col)) X *= fx;
                    X += src;
                    P[6 * n + s] = X;
                    a.szz[o] = X + P[7 * n + s];
                } else {
                    P[6 * n + s] = src;
                    P[7 * n + s] = 0.f;
                    a.szz[o] = src;
                }
            } else if (in_box) {
                a.sxx[o] = a.sxx[o] + srx;
                a.szz[o] = ABC ? a.szz[o] + src : src;
            }
            continue;
        }
        const int d = ABC ? kElTZ : k - t;
        const float* pvx = s_vx + (cl + HX) * HZZ + tz + HZ;
        const float* pvz = s_vz + (cl + HX) * HZZ + tz + HZ;
        const float dvx_x = el_xsum<1>(a, pvx, HZZ);  // calux xflag 1 on vx
        const float dvz_z = el_zsum<0>(a, pvz, d);    // caluz zflag 0 on vz
        const float dvz_x = el_xsum<0>(a, pvz, HZZ);  // calux xflag 0 on vz
        const float dvx_z = el_zsum<1>(a, pvx, d);    // caluz zflag 1 on vx
        const float lm = lmc[j], ls = lsc[j], mu = muc[j];
        const float ax = a.cx * lm * dvx_x, az = a.cz * ls * dvz_z;
        const float bx = a.cx * ls * dvx_x, bz = a.cz * lm * dvz_z;
        const float ex = a.cx * mu * dvz_x, ez = a.cz * mu * dvx_z;
        const bool top = !ABC && k == t;  // 'fsbc': tau_zz reset on the surface row
        float xz;
        if (s < 0) {
            a.sxx[o] = q[0][j] + ax + az + srx;
            a.szz[o] = top ? src : q[1][j] + bx + bz + src;
            xz = q[2][j] + ex + ez;
        } else {
            const bool dx = k >= ro(a.srow, col);
            float fzs = 1.f, fzx = 1.f;  // z factors of the normal stresses (node) and tau_xz (half)
            bool dz = false;
            if (kbot) {
                dz = true;
                fzs = fzk;
                fzx = fhk;
            } else if (ABC && k <= t && k >= t - a.npzn) {
                dz = true;
                fzs = a.tzn[k - (t - a.npzn)];
                fzx = a.tzh[k - (t - a.npzn)];
            }
            const float fh = ro(a.fxh, col);
            float X = q[0][j] + ax, Z = q[1][j] + az;
            if (dx) X *= fh;
            if (dz) Z *= fzs;
            X += srx;
            P[4 * n + s] = X;
            P[5 * n + s] = Z;
            a.sxx[o] = X + Z;
            X = q[2][j] + bx;
            Z = q[3][j] + bz;
            if (dx) X *= fh;
            if (dz) Z *= fzs;
            if (top) X = Z = 0.f;
            X += src;
            P[6 * n + s] = X;
            P[7 * n + s] = Z;
            a.szz[o] = X + Z;
            X = q[4][j] + ex;
            Z = q[5][j] + ez;
            if (dx) X *= ro(a.fxn, col);
            if (dz) Z *= fzx;
            P[8 * n + s] = X;
            P[9 * n + s] = Z;
            xz = X + Z;
        }
        a.sxz[o] = xz;
        if (ABC || k != t) continue;
        if (t >= 1) a.sxz[o - 1] = -xz;  // calfsbcstrxz (pub:1009-1023)
        // calfsbcv (pub:938-972) for the next step's vz(t - 1).  In a right
        // strip whose damped rows reach row t - 1 (they start at the mirrored
        // left column's surface, pub:917-922) the x-part is damped before the
        // sum, so the condition is applied to the parts there.
        float gh = 0.f;
        if (t >= 1 && col < a.nx - 1) {
            const float r = ro(a.ratio, col);
            if (col >= a.nx - a.nr && ro(a.srow, col) <= t - 1) {
                const int s1 = el_strip<ABC>(a, col + 1, t, a.itopo[col + 1]);
                const float gx = P[2 * n + s] + r * (P[s1] - P[s]);
                const float gz = P[3 * n + s] + r * (P[n + s1] - P[n + s]);
                gh = gx * ro(a.fxh, col) + gz;
            } else {
                gh = pvz[0] + r * (pvx[HZZ] - pvx[0]);
            }
        }
        a.ghost[col] = gh;
    }
}

template <int TX, int NW, bool ABC>
__global__ __launch_bounds__(64 * NW) void elastic_velocity_kernel(ElasticStepArgs a) {
    __shared__ ElVelLds<TX> lds;
    el_velocity<TX, NW, ABC>(a, xcd_tile(blockIdx.x, a.ntiles), lds);
}

template <int TX, int NW, bool ABC>
__global__ __launch_bounds__(64 * NW) void elastic_stress_kernel(ElasticStepArgs a) {
    __shared__ ElStrLds<TX> lds;
    el_stress<TX, NW, ABC>(a, xcd_tile(blockIdx.x, a.ntiles), a.step, a.rec_col, lds);
}

template <int TX, int NW, bool ABC>
__global__ __launch_bounds__(64 * NW) void elastic_velocity_batch(ElasticBatchArgs b) {
    __shared__ ElVelLds<TX> lds;
    const int* m = reinterpret_cast<const int*>(b.tile_map) + 2 * xcd_tile(blockIdx.x, b.ntile);
    ElasticShotArgs& a = *(ElasticShotArgs*)(b.shots + ro(m, 0));
    el_velocity<TX, NW, ABC>(a, ro(m, 1), lds);
}

template <int TX, int NW, bool ABC>
__global__ __launch_bounds__(64 * NW) void elastic_stress_batch(ElasticBatchArgs b) {
    __shared__ ElStrLds<TX> lds;
    const int* m = reinterpret_cast<const int*>(b.tile_map) + 2 * xcd_tile(blockIdx.x, b.ntile);
    ElasticShotArgs& a = *(ElasticShotArgs*)(b.shots + ro(m, 0));
    el_stress<TX, NW, ABC>(a, ro(m, 1), b.step, b.rec_col, lds);
}

}  // namespace sfd
