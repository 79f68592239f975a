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
template <bool ABC>
__device__ __forceinline__ int el_strip(const ElasticStepArgs& a, int col, int k, int t) {
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
template <int ZF>
__device__ __forceinline__ float el_zsum(const ElasticStepArgs& a, const float* p, int d) {
    const int lim = ZF ? 4 : 5;
    if (d >= lim) {
        float s = a.a10[0] * (p[ZF] - p[ZF - 1]);
#pragma unroll
        for (int n = 1; n < 5; ++n) s += a.a10[n] * (p[n + ZF] - p[ZF - n - 1]);
        return s;
    }
    const int M = el_order(d, ZF);
    const float* c = M == 4 ? a.a8 : M == 3 ? a.a6 : a.a4;
    if (M == 1) return p[ZF] - p[ZF - 1];
    float s = c[0] * (p[ZF] - p[ZF - 1]);
    for (int n = 1; n < M; ++n) s += c[n] * (p[n + ZF] - p[ZF - n - 1]);
    return s;
}

// x stencil sum over LDS columns `stride` floats apart: xflag 0 pairs
// q[n] - q[-n-1], xflag 1 pairs q[n+1] - q[-n].
template <int XF>
__device__ __forceinline__ float el_xsum(const ElasticStepArgs& a, const float* q, int stride) {
    float s = a.a10[0] * (q[XF * stride] - q[(XF - 1) * stride]);
#pragma unroll
    for (int n = 1; n < 5; ++n) s += a.a10[n] * (q[(n + XF) * stride] - q[(XF - n - 1) * stride]);
    return s;
}

// Stages columns [x0 - HX, x0 + TX + HX) x rows [z0 - HZ, z0 + kElTZ + HZ)
// of a padded field into lds[c][r] (a fixed, unrolled count per thread, so
// all loads are in flight before the first LDS store waits).
template <int TX, int NW, int HX, int HZ>
__device__ __forceinline__ void el_stage(float* lds, const float* f, int pitch, int x0, int z0) {
    constexpr int W = TX + 2 * HX, Hh = kElTZ + 2 * HZ, N = W * Hh, T = 64 * NW;
    constexpr int ITER = (N + T - 1) / T;
    float v[ITER];
#pragma unroll
    for (int q = 0; q < ITER; ++q) {
        const int e = threadIdx.x + q * T;
        const int c = e / Hh, r = e - c * Hh;
        v[q] = (e < N) ? f[(ptrdiff_t)(x0 - HX + c) * pitch + (z0 - HZ + r)] : 0.f;
    }
#pragma unroll
    for (int q = 0; q < ITER; ++q) {
        const int e = threadIdx.x + q * T;
        if (e < N) lds[e] = v[q];
    }
}

// Tile of this workgroup: XCD-aware order over (x band, z tile), z fastest.
__device__ __forceinline__ void el_tile(const ElasticStepArgs& a, int tx, int* z0, int* x0) {
    const int t = xcd_tile(blockIdx.x, a.ntiles);
    *z0 = (t % a.ntz) * kElTZ;
    *x0 = (t / a.ntz) * tx;
}

// Velocity half-step (pri:119-151; 'abc': pri:300-318): vx from tau_xx,
// tau_xz; vz from tau_xz, tau_zz; in 'fsbc' mode the pending vz(t - 1).
template <int TX, int NW, bool ABC>
__global__ __launch_bounds__(64 * NW) void elastic_velocity_kernel(ElasticStepArgs a) {
    constexpr int HX = kElH, HZ = kElH;
    constexpr int WX = TX + 2 * HX, HZZ = kElTZ + 2 * HZ;
    __shared__ float s_xx[WX * kElTZ];
    __shared__ float s_xz[WX * HZZ];
    __shared__ float s_zz[TX * HZZ];
    int z0, x0;
    el_tile(a, TX, &z0, &x0);
    el_stage<TX, NW, HX, 0>(s_xx, a.sxx, a.pitch, x0, z0);
    el_stage<TX, NW, HX, HZ>(s_xz, a.sxz, a.pitch, x0, z0);
    el_stage<TX, NW, 0, HZ>(s_zz, a.szz, a.pitch, x0, z0);
    __syncthreads();
    const int tz = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int k = z0 + tz;
    if (k >= a.nz) return;
    for (int j = 0; j < TX / NW; ++j) {
        const int cl = wv + j * NW, col = x0 + cl;
        if (col >= a.nx) break;
        const int t = a.itopo[col];
        const ptrdiff_t o = (ptrdiff_t)col * a.pitch + k;
        if (k < (ABC ? a.kst[col] : t)) {
            if (!ABC && k == t - 1) a.vz[o] = a.ghost[col];
            continue;
        }
        const int d = ABC ? kElTZ : k - t;  // 'abc': 10th order on every row
        const float* pxx = s_xx + (cl + HX) * kElTZ + tz;
        const float* pxz = s_xz + (cl + HX) * HZZ + tz + HZ;
        const float* pzz = s_zz + cl * HZZ + tz + HZ;
        const float bx = a.b2e[o], bz = a.b1e[o];
        const float ux = a.cx * bx * el_xsum<0>(a, pxx, kElTZ);   // calux xflag 0 on tau_xx
        const float uz = a.cz * bx * el_zsum<0>(a, pxz, d);       // caluz zflag 0 on tau_xz
        const float wx = a.cx * bz * el_xsum<1>(a, pxz, HZZ);     // calux xflag 1 on tau_xz
        const float wz = a.cz * bz * el_zsum<1>(a, pzz, d);       // caluz zflag 1 on tau_zz
        const int s = el_strip<ABC>(a, col, k, t);
        if (s < 0) {
            a.vx[o] = a.vx[o] + ux + uz;
            a.vz[o] = a.vz[o] + wx + wz;
            continue;
        }
        float* P = a.parts;
        const size_t n = (size_t)a.nstrip;
        const bool dx = k >= a.srow[col];
        float fzv = 1.f, fzw = 1.f;  // z factors of vx (node table) and vz (half table)
        bool dz = false;
        if (k >= a.kb0) {
            dz = true;
            fzv = a.fzn[k];
            fzw = a.fzh[k];
        } else if (ABC && k <= t && k >= t - a.npzn) {
            dz = true;
            fzv = a.tzn[k - (t - a.npzn)];
            fzw = a.tzh[k - (t - a.npzn)];
        }
        float X = P[s] + ux, Z = P[n + s] + uz;
        if (dx) X *= a.fxn[col];
        if (dz) Z *= fzv;
        P[s] = X;
        P[n + s] = Z;
        a.vx[o] = X + Z;
        X = P[2 * n + s] + wx;
        Z = P[3 * n + s] + wz;
        if (dx) X *= a.fxh[col];
        if (dz) Z *= fzw;
        P[2 * n + s] = X;
        P[3 * n + s] = Z;
        a.vz[o] = X + Z;
    }
}

// Stress half-step (pri:153-196; 'abc': pri:320-352) and the step's receiver
// samples.
template <int TX, int NW, bool ABC>
__global__ __launch_bounds__(64 * NW) void elastic_stress_kernel(ElasticStepArgs a) {
    constexpr int HX = kElH, HZ = kElH;
    constexpr int WX = TX + 2 * HX, HZZ = kElTZ + 2 * HZ;
    __shared__ float s_vx[WX * HZZ];
    __shared__ float s_vz[WX * HZZ];
    int z0, x0;
    el_tile(a, TX, &z0, &x0);
    if (blockIdx.x == 0 && a.rec_col >= 0) {
        for (int r = threadIdx.x; r < a.nrec; r += 64 * NW) {
            const ptrdiff_t o = (ptrdiff_t)a.rx[r] * a.pitch + a.rz[r];
            a.rec_vx[(size_t)r * a.nrt + a.rec_col] = a.vx[o];
            a.rec_vz[(size_t)r * a.nrt + a.rec_col] = a.vz[o];
        }
    }
    el_stage<TX, NW, HX, HZ>(s_vx, a.vx, a.pitch, x0, z0);
    el_stage<TX, NW, HX, HZ>(s_vz, a.vz, a.pitch, x0, z0);
    __syncthreads();
    const int tz = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int k = z0 + tz;
    if (k >= a.nz) return;
    const bool in_box_z = k >= a.sz0 && k < a.sz1;
    const float w = a.amp[a.step];
    for (int j = 0; j < TX / NW; ++j) {
        const int cl = wv + j * NW, col = x0 + cl;
        if (col >= a.nx) break;
        const int t = a.itopo[col];
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
        if (k < (ABC ? a.kst[col] : t)) {
            // Above the updated rows only the source acts: the x parts take it
            // (damped where a strip's rows reach up here); 'fsbc' resets tau_zz
            // on and above the surface to it, 'abc' lets it accumulate.
            if (s >= 0) {
                const float fx = k >= a.srow[col] ? a.fxh[col] : 1.f;
                float X = P[4 * n + s];
                if (k >= a.srow[col]) X *= fx;
                X += srx;
                P[4 * n + s] = X;
                a.sxx[o] = X + P[5 * n + s];
                if (ABC) {
                    X = P[6 * n + s];
                    if (k >= a.srow[col]) X *= fx;
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
        const float lm = a.lm[o], ls = a.ls[o], mu = a.mu[o];
        const float ax = a.cx * lm * dvx_x, az = a.cz * ls * dvz_z;
        const float bx = a.cx * ls * dvx_x, bz = a.cz * lm * dvz_z;
        const float ex = a.cx * mu * dvz_x, ez = a.cz * mu * dvx_z;
        const bool top = !ABC && k == t;  // 'fsbc': tau_zz reset on the surface row
        float xz;
        if (s < 0) {
            a.sxx[o] = a.sxx[o] + ax + az + srx;
            a.szz[o] = top ? src : a.szz[o] + bx + bz + src;
            xz = a.sxz[o] + ex + ez;
        } else {
            const bool dx = k >= a.srow[col];
            float fzs = 1.f, fzx = 1.f;  // z factors of the normal stresses (node) and tau_xz (half)
            bool dz = false;
            if (k >= a.kb0) {
                dz = true;
                fzs = a.fzn[k];
                fzx = a.fzh[k];
            } else if (ABC && k <= t && k >= t - a.npzn) {
                dz = true;
                fzs = a.tzn[k - (t - a.npzn)];
                fzx = a.tzh[k - (t - a.npzn)];
            }
            const float fh = a.fxh[col];
            float X = P[4 * n + s] + ax, Z = P[5 * n + s] + az;
            if (dx) X *= fh;
            if (dz) Z *= fzs;
            X += srx;
            P[4 * n + s] = X;
            P[5 * n + s] = Z;
            a.sxx[o] = X + Z;
            X = P[6 * n + s] + bx;
            Z = P[7 * n + s] + bz;
            if (dx) X *= fh;
            if (dz) Z *= fzs;
            if (top) X = Z = 0.f;
            X += src;
            P[6 * n + s] = X;
            P[7 * n + s] = Z;
            a.szz[o] = X + Z;
            X = P[8 * n + s] + ex;
            Z = P[9 * n + s] + ez;
            if (dx) X *= a.fxn[col];
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
            const float r = a.ratio[col];
            if (col >= a.nx - a.nr && a.srow[col] <= t - 1) {
                const int s1 = el_strip<ABC>(a, col + 1, t, a.itopo[col + 1]);
                const float gx = P[2 * n + s] + r * (P[s1] - P[s]);
                const float gz = P[3 * n + s] + r * (P[n + s1] - P[n + s]);
                gh = gx * a.fxh[col] + gz;
            } else {
                gh = pvz[0] + r * (pvx[HZZ] - pvx[0]);
            }
        }
        a.ghost[col] = gh;
    }
}

}  // namespace sfd
