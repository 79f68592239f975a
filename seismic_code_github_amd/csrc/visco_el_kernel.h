// Visco-elastic (vs > 0) and visco-acoustic (vs = 0) steps of the
// rotated-staggered-grid propagator:
// forward/Q_forward/Q_multishot_fd_subroutine_fsbc_zy_rotated.f90,
// extrapolation_one_step_10nd_absbc (:695-839) with L standard linear solids,
// as two launches per step (the reference's two half-steps):
//   velocity  vx += b (Dx sxx + Dz sxz),  vz += b (Dx sxz + Dz szz)   (:747-773)
//             with the RSG derivatives of calderxabs / calderzabs (xflag 0):
//             Dx = (0.5/dx)(dd + da), Dz = (0.5/dz)(dd - da), dd and da the
//             a_n-weighted sums along the two diagonals (:843-966)
//   stress    d1 = Dx vx, d2 = Dx vz, d3 = Dz vx, d4 = Dz vz (xflag 1, :779-782)
//             sxx += kup d1 + (kup - 2 kus) d4 + dt sum_l rxx_l      (calpstrxx)
//             szz += (kup - 2 kus) d1 + kup d4 + dt sum_l rzz_l      (calpstrzz)
//             sxz += kus (d2 + d3) + dt sum_l rxz_l                  (calpstrxz)
//   memory    rxx_l += -c_l (krp d1 + (krp - 2 krs) d4 + rxx_l)    (calprxx, :1135-1269)
//             rzz_l += -c_l ((krp - 2 krs) d1 + krp d4 + rzz_l), rxz_l += -c_l (krs (d2 + d3) + rxz_l)
//   with kup = dt M (1 + L taup), kus = dt mu (1 + L taus), krp = M taup,
//   krs = mu taus, c_l = dt / tau_sigma_l (pcalpas, :473-551).
//
// Sponge (calpmlbclrtb, :609-654): the x parts of every field are damped in
// the left/right strips, the z parts in the top/bottom strips; the memory
// terms of stresses and memory variables are not.  Fields are stored unsplit;
// a strip cell also keeps the damped part(s) only: new total = old total +
// increments - (1 - f) (old part + part increment), new part = f (...).
//
// Layout: fields [nx][nz] (x rows, z contiguous) with an 8-cell zero border
// (pitch floats between rows).  A workgroup of NW waves owns TX rows x 64
// columns, one column per lane; the fields a stencil reads are staged in LDS
// with an R-cell halo on every side.
#pragma once
#include <hip/hip_runtime.h>

#include "device_util.h"

namespace sfd {

constexpr int kVeTZ = 64;         // tile columns (z), one per lane
constexpr int kVePad = 8;         // zero border in every padded field
constexpr int kVeMaxMech = 3;

struct ViscoElArgs {
    int nx, nz, pitch, ntz, ntiles;
    float *vx, *vz, *sxx, *szz, *sxz;  // cell (0, 0) of each padded field
    float* r;                          // 3L padded fields: rxx_l, rzz_l, rxz_l for l = 0..L-1
    size_t rstride;                    // floats between two of them
    const float *b, *kup, *kus, *krp, *krs;
    float hx, hz, dt;
    float a[5];
    float c[kVeMaxMech];
    // Strips: x strips are rows [0, xlo) and [nx - xhi, nx) (all columns), z
    // strips columns [0, zlo) and [nz - zhi, nz) (all rows).  xs holds the x
    // parts [5 + 3L][xlo + xhi][nz] (vx, vz, sxx, szz, sxz, then the memory
    // variables in r's order); zs the z parts [5 + 3L][nx][zlo + zhi].
    int xlo, xhi, zlo, zhi;
    float *xs, *zs;
    // Damping per row (x parts) and per column (z parts), 1 outside: velocity
    // tables (spgx*, spgz*) and stress / memory tables (spg1*, spg2*).
    const float *dvx, *dsx, *dvz, *dsz;
    // Source: weights g over rows [sx0, sx1) x columns [sz0, sz1);
    // amp[k] = wavelet(k) dt / dz^2.
    int sx0, sx1, sz0, sz1;
    const float *g, *amp;
    int step;
    // Receivers: sxx, szz of the previous step gathered by workgroup 0 of the
    // velocity launch into column rec_col (< 0: none).
    int nrec, rec_col, nrt;
    const int *rx, *rz;
    float *rec_zz, *rec_xx;
};

// Several shots per launch: shots[s] is shot s's argument block (read
// through the constant address space), tile_map[g] = (shot, the shot's tile)
// of global tile g; step and rec_col are the shots' common step index and
// gather column.
struct ViscoBatchArgs {
    const ViscoElArgs* shots;
    const int2* tile_map;
    int ntile;
    int step, rec_col;
};

using ViscoShotArgs = const __attribute__((address_space(4))) ViscoElArgs;

// Tile t of a shot: row band t / ntz, z tile t % ntz.
template <class A>
__device__ __forceinline__ void ve_tile(const A& a, int tx, int t, int* i0, int* k0) {
    *k0 = (t % a.ntz) * kVeTZ;
    *i0 = (t / a.ntz) * tx;
}

// Rows [i0 - R, i0 + TX + R) x columns [k0 - R, k0 + 64 + R) of a padded
// field into lds[row][col] (unrolled so every load is in flight first).
template <int R, int TX, int NW>
__device__ __forceinline__ void ve_stage(float* lds, const float* f, int pitch, int i0, int k0) {
    constexpr int W = kVeTZ + 2 * R, Hh = TX + 2 * R, N = W * Hh, T = 64 * NW;
    constexpr int ITER = (N + T - 1) / T;
    float v[ITER];
#pragma unroll
    for (int q = 0; q < ITER; ++q) {
        const int e = threadIdx.x + q * T;
        const int rr = e / W, cc = e - rr * W;
        v[q] = (e < N) ? f[(ptrdiff_t)(i0 - R + rr) * pitch + (k0 - R + cc)] : 0.f;
    }
#pragma unroll
    for (int q = 0; q < ITER; ++q) {
        const int e = threadIdx.x + q * T;
        if (e < N) lds[e] = v[q];
    }
}

// Diagonal sums of the RSG derivative at (row, col) of an LDS tile `W` wide.
// SHIFT 0 (xflag 0): dd = sum a_n (u(i+n, k+n) - u(i-n-1, k-n-1)),
//                    da = sum a_n (u(i+n, k-n-1) - u(i-n-1, k+n));
// SHIFT 1 (xflag 1): the same one cell on: u(i+1+n, k+1+n) - u(i-n, k-n),
//                    u(i+1+n, k-n) - u(i-n, k+1+n).
template <int R, int SHIFT, int W>
__device__ __forceinline__ void ve_diag(const float* p, const float* co, float& dd, float& da) {
    dd = 0.f;
    da = 0.f;
#pragma unroll
    for (int n = 0; n < R; ++n) {
        const float u1 = p[(SHIFT + n) * W + (SHIFT + n)], u2 = p[(-n - 1 + SHIFT) * W + (-n - 1 + SHIFT)];
        const float u3 = p[(SHIFT + n) * W + (-n - 1 + SHIFT)], u4 = p[(-n - 1 + SHIFT) * W + (SHIFT + n)];
        dd = __builtin_fmaf(co[n], u1 - u2, dd);
        da = __builtin_fmaf(co[n], u3 - u4, da);
    }
}

// Bodies of the four launches for tile t of a shot (A = ViscoElArgs in
// kernel arguments, or a batch's block in the constant address space); the
// workgroup of tile 0 gathers the receivers (column rec_col, < 0: none).
template <int R, int TX, int NW, class A>
__device__ __forceinline__ void ve_el_velocity(const A& a, int t, int rec_col, float* s_xx, float* s_zz,
                                               float* s_xz) {
    constexpr int W = kVeTZ + 2 * R;
    int i0, k0;
    ve_tile(a, TX, t, &i0, &k0);
    if (t == 0 && rec_col >= 0) {
        for (int r = threadIdx.x; r < a.nrec; r += 64 * NW) {
            const ptrdiff_t o = (ptrdiff_t)a.rx[r] * a.pitch + a.rz[r];
            a.rec_zz[(size_t)r * a.nrt + rec_col] = a.szz[o];
            a.rec_xx[(size_t)r * a.nrt + rec_col] = a.sxx[o];
        }
    }
    ve_stage<R, TX, NW>(s_xx, a.sxx, a.pitch, i0, k0);
    ve_stage<R, TX, NW>(s_zz, a.szz, a.pitch, i0, k0);
    ve_stage<R, TX, NW>(s_xz, a.sxz, a.pitch, i0, k0);
    __syncthreads();
    float co[R];
#pragma unroll
    for (int m = 0; m < R; ++m) co[m] = a.a[m];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int k = k0 + lane;
    if (k >= a.nz) return;
    const int xw = a.xlo + a.xhi, zw = a.zlo + a.zhi;
    int zc = -1;
    if (k < a.zlo) zc = k;
    else if (k >= a.nz - a.zhi) zc = k - (a.nz - a.zhi) + a.zlo;
    for (int j = 0; j < TX / NW; ++j) {
        const int il = wv + j * NW, i = i0 + il;
        if (i >= a.nx) break;
        const int off = (il + R) * W + lane + R;
        float ddA, daA, ddB, daB, ddC, daC;
        ve_diag<R, 0, W>(s_xx + off, co, ddA, daA);
        ve_diag<R, 0, W>(s_xz + off, co, ddB, daB);
        ve_diag<R, 0, W>(s_zz + off, co, ddC, daC);
        const ptrdiff_t o = (ptrdiff_t)i * a.pitch + k;
        const float b = a.b[o];
        const float vxx = b * (a.hx * (ddA + daA)), vxz = b * (a.hz * (ddB - daB));  // calpvabs increments
        const float vzx = b * (a.hx * (ddB + daB)), vzz = b * (a.hz * (ddC - daC));
        float vx = a.vx[o] + vxx + vxz, vz = a.vz[o] + vzx + vzz;
        int xr = -1;
        if (i < a.xlo) xr = i;
        else if (i >= a.nx - a.xhi) xr = i - (a.nx - a.xhi) + a.xlo;
        if (xr >= 0) {
            const float f = a.dvx[i];
            const size_t q = (size_t)xw * a.nz, e = (size_t)xr * a.nz + k;
            const float tx = a.xs[e] + vxx, tz = a.xs[q + e] + vzx;
            vx -= (1.f - f) * tx;
            vz -= (1.f - f) * tz;
            a.xs[e] = f * tx;
            a.xs[q + e] = f * tz;
        }
        if (zc >= 0) {
            const float f = a.dvz[k];
            const size_t q = (size_t)a.nx * zw, e = (size_t)i * zw + zc;
            const float tx = a.zs[e] + vxz, tz = a.zs[q + e] + vzz;
            vx -= (1.f - f) * tx;
            vz -= (1.f - f) * tz;
            a.zs[e] = f * tx;
            a.zs[q + e] = f * tz;
        }
        a.vx[o] = vx;
        a.vz[o] = vz;
    }
}

template <int R, int L, int TX, int NW, class A>
__device__ __forceinline__ void ve_el_stress(const A& a, int t, int step, float* s_vx, float* s_vz) {
    constexpr int W = kVeTZ + 2 * R;
    int i0, k0;
    ve_tile(a, TX, t, &i0, &k0);
    ve_stage<R, TX, NW>(s_vx, a.vx, a.pitch, i0, k0);
    ve_stage<R, TX, NW>(s_vz, a.vz, a.pitch, i0, k0);
    __syncthreads();
    float co[R];
#pragma unroll
    for (int m = 0; m < R; ++m) co[m] = a.a[m];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int k = k0 + lane;
    if (k >= a.nz) return;
    const int xw = a.xlo + a.xhi, zw = a.zlo + a.zhi;
    int zc = -1;
    if (k < a.zlo) zc = k;
    else if (k >= a.nz - a.zhi) zc = k - (a.nz - a.zhi) + a.zlo;
    const float amp = a.amp[step];
    const bool src_col = k >= a.sz0 && k < a.sz1;
    for (int j = 0; j < TX / NW; ++j) {
        const int il = wv + j * NW, i = i0 + il;
        if (i >= a.nx) break;
        const int off = (il + R) * W + lane + R;
        float dd1, da1, dd4, da4;
        ve_diag<R, 1, W>(s_vx + off, co, dd1, da1);
        ve_diag<R, 1, W>(s_vz + off, co, dd4, da4);
        const float d1 = a.hx * (dd1 + da1), d3 = a.hz * (dd1 - da1);  // Dx vx, Dz vx
        const float d2 = a.hx * (dd4 + da4), d4 = a.hz * (dd4 - da4);  // Dx vz, Dz vz
        const ptrdiff_t o = (ptrdiff_t)i * a.pitch + k;
        const float kup = a.kup[o], kus = a.kus[o], krp = a.krp[o], krs = a.krs[o];
        const float kpm = kup - 2.f * kus, rpm = krp - 2.f * krs;
        float rold[3 * L], rsum[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < 3 * L; ++q) {
            rold[q] = a.r[(size_t)q * a.rstride + o];
            rsum[q % 3] += rold[q];
        }
        // Increments: x parts (d1, d2 terms), z parts (d4, d3 terms).
        float ix[3 + 3 * L], iz[3 + 3 * L], tot[3 + 3 * L];
        ix[0] = kup * d1;
        iz[0] = kpm * d4;
        ix[1] = kpm * d1;
        iz[1] = kup * d4;
        ix[2] = kus * d2;
        iz[2] = kus * d3;
        tot[0] = a.sxx[o] + ix[0] + iz[0] + a.dt * rsum[0];
        tot[1] = a.szz[o] + ix[1] + iz[1] + a.dt * rsum[1];
        tot[2] = a.sxz[o] + ix[2] + iz[2] + a.dt * rsum[2];
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const float cl = a.c[l];
            ix[3 + 3 * l] = -cl * (krp * d1);
            iz[3 + 3 * l] = -cl * (rpm * d4);
            ix[4 + 3 * l] = -cl * (rpm * d1);
            iz[4 + 3 * l] = -cl * (krp * d4);
            ix[5 + 3 * l] = -cl * (krs * d2);
            iz[5 + 3 * l] = -cl * (krs * d3);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const float rq = rold[3 * l + q];
                tot[3 + 3 * l + q] = rq + ix[3 + 3 * l + q] + iz[3 + 3 * l + q] - cl * rq;
            }
        }
        int xr = -1;
        if (i < a.xlo) xr = i;
        else if (i >= a.nx - a.xhi) xr = i - (a.nx - a.xhi) + a.xlo;
        // Source (after the x parts' sponge, as addsrc follows calpmlbclrtb):
        // into the x parts and the totals of sxx and szz.
        float add = 0.f;
        if (src_col && i >= a.sx0 && i < a.sx1) add = amp * a.g[(size_t)(i - a.sx0) * (a.sz1 - a.sz0) + (k - a.sz0)];
        if (xr >= 0) {
            const float f = a.dsx[i];
            const size_t q = (size_t)xw * a.nz, e = (size_t)xr * a.nz + k;
#pragma unroll
            for (int p = 0; p < 3 + 3 * L; ++p) {
                const float t = a.xs[(2 + p) * q + e] + ix[p];
                tot[p] -= (1.f - f) * t;
                a.xs[(2 + p) * q + e] = f * t + (p < 2 ? add : 0.f);
            }
        }
        if (zc >= 0) {
            const float f = a.dsz[k];
            const size_t q = (size_t)a.nx * zw, e = (size_t)i * zw + zc;
#pragma unroll
            for (int p = 0; p < 3 + 3 * L; ++p) {
                const float t = a.zs[(2 + p) * q + e] + iz[p];
                tot[p] -= (1.f - f) * t;
                a.zs[(2 + p) * q + e] = f * t;
            }
        }
        a.sxx[o] = tot[0] + add;
        a.szz[o] = tot[1] + add;
        a.sxz[o] = tot[2];
#pragma unroll
        for (int q = 0; q < 3 * L; ++q) a.r[(size_t)q * a.rstride + o] = tot[3 + q];
    }
}

// Visco-acoustic (vs = 0) step on the same padded layout: one stress sigma
// (= sxx = szz; a.sxx and a.szz both point at it), L memory variables, the
// model kup = ku, krp = kr; kus, krs and sxz are not used.  These are the
// elastic kernels' operations with mu = 0 (kup - 2 kus = kup, krp - 2 krs =
// krp) minus the terms that vanish: sigma_xz and its memory variables stay
// zero, so vx has an x part only and vz a z part only.  Strips: xs [2 + L]
// [xlo + xhi][nz] = vx, sigma, r_l; zs [2 + L][nx][zlo + zhi] = vz, sigma, r_l.
template <int R, int TX, int NW, class A>
__device__ __forceinline__ void ve_ac_velocity(const A& a, int t, int rec_col, float* s_s) {
    constexpr int W = kVeTZ + 2 * R;
    int i0, k0;
    ve_tile(a, TX, t, &i0, &k0);
    if (t == 0 && rec_col >= 0) {
        for (int r = threadIdx.x; r < a.nrec; r += 64 * NW)
            a.rec_zz[(size_t)r * a.nrt + rec_col] = a.szz[(ptrdiff_t)a.rx[r] * a.pitch + a.rz[r]];
    }
    ve_stage<R, TX, NW>(s_s, a.szz, a.pitch, i0, k0);
    __syncthreads();
    float co[R];
#pragma unroll
    for (int m = 0; m < R; ++m) co[m] = a.a[m];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int k = k0 + lane;
    if (k >= a.nz) return;
    const int zw = a.zlo + a.zhi;
    int zc = -1;
    if (k < a.zlo) zc = k;
    else if (k >= a.nz - a.zhi) zc = k - (a.nz - a.zhi) + a.zlo;
    for (int j = 0; j < TX / NW; ++j) {
        const int il = wv + j * NW, i = i0 + il;
        if (i >= a.nx) break;
        float dd, da;
        ve_diag<R, 0, W>(s_s + (il + R) * W + lane + R, co, dd, da);
        const ptrdiff_t o = (ptrdiff_t)i * a.pitch + k;
        const float b = a.b[o];
        const float vxx = b * (a.hx * (dd + da)), vzz = b * (a.hz * (dd - da));
        float vx = a.vx[o] + vxx, vz = a.vz[o] + vzz;
        int xr = -1;
        if (i < a.xlo) xr = i;
        else if (i >= a.nx - a.xhi) xr = i - (a.nx - a.xhi) + a.xlo;
        if (xr >= 0) {
            const float f = a.dvx[i];
            const size_t e = (size_t)xr * a.nz + k;
            const float t = a.xs[e] + vxx;
            vx -= (1.f - f) * t;
            a.xs[e] = f * t;
        }
        if (zc >= 0) {
            const float f = a.dvz[k];
            const size_t e = (size_t)i * zw + zc;
            const float t = a.zs[e] + vzz;
            vz -= (1.f - f) * t;
            a.zs[e] = f * t;
        }
        a.vx[o] = vx;
        a.vz[o] = vz;
    }
}

template <int R, int L, int TX, int NW, class A>
__device__ __forceinline__ void ve_ac_stress(const A& a, int t, int step, float* s_vx, float* s_vz) {
    constexpr int W = kVeTZ + 2 * R;
    int i0, k0;
    ve_tile(a, TX, t, &i0, &k0);
    ve_stage<R, TX, NW>(s_vx, a.vx, a.pitch, i0, k0);
    ve_stage<R, TX, NW>(s_vz, a.vz, a.pitch, i0, k0);
    __syncthreads();
    float co[R];
#pragma unroll
    for (int m = 0; m < R; ++m) co[m] = a.a[m];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int k = k0 + lane;
    if (k >= a.nz) return;
    const int xw = a.xlo + a.xhi, zw = a.zlo + a.zhi;
    int zc = -1;
    if (k < a.zlo) zc = k;
    else if (k >= a.nz - a.zhi) zc = k - (a.nz - a.zhi) + a.zlo;
    const float amp = a.amp[step];
    const bool src_col = k >= a.sz0 && k < a.sz1;
    for (int j = 0; j < TX / NW; ++j) {
        const int il = wv + j * NW, i = i0 + il;
        if (i >= a.nx) break;
        const int off = (il + R) * W + lane + R;
        float dd1, da1, dd4, da4;
        ve_diag<R, 1, W>(s_vx + off, co, dd1, da1);
        ve_diag<R, 1, W>(s_vz + off, co, dd4, da4);
        const float d1 = a.hx * (dd1 + da1), d4 = a.hz * (dd4 - da4);  // Dx vx, Dz vz
        const ptrdiff_t o = (ptrdiff_t)i * a.pitch + k;
        const float ku = a.kup[o], kr = a.krp[o];
        float rold[L], rsum = 0.f;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            rold[l] = a.r[(size_t)l * a.rstride + o];
            rsum += rold[l];
        }
        float ix[1 + L], iz[1 + L], tot[1 + L];
        ix[0] = ku * d1;
        iz[0] = ku * d4;
        tot[0] = a.szz[o] + ix[0] + iz[0] + a.dt * rsum;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const float cl = a.c[l];
            ix[1 + l] = -cl * (kr * d1);
            iz[1 + l] = -cl * (kr * d4);
            tot[1 + l] = rold[l] + ix[1 + l] + iz[1 + l] - cl * rold[l];
        }
        int xr = -1;
        if (i < a.xlo) xr = i;
        else if (i >= a.nx - a.xhi) xr = i - (a.nx - a.xhi) + a.xlo;
        // Source (after the x parts' sponge, as addsrc follows calpmlbclrtb):
        // into sigma's x part and total.
        float add = 0.f;
        if (src_col && i >= a.sx0 && i < a.sx1) add = amp * a.g[(size_t)(i - a.sx0) * (a.sz1 - a.sz0) + (k - a.sz0)];
        if (xr >= 0) {
            const float f = a.dsx[i];
            const size_t q = (size_t)xw * a.nz, e = (size_t)xr * a.nz + k;
#pragma unroll
            for (int p = 0; p < 1 + L; ++p) {
                const float t = a.xs[(1 + p) * q + e] + ix[p];
                tot[p] -= (1.f - f) * t;
                a.xs[(1 + p) * q + e] = f * t + (p == 0 ? add : 0.f);
            }
        }
        if (zc >= 0) {
            const float f = a.dsz[k];
            const size_t q = (size_t)a.nx * zw, e = (size_t)i * zw + zc;
#pragma unroll
            for (int p = 0; p < 1 + L; ++p) {
                const float t = a.zs[(1 + p) * q + e] + iz[p];
                tot[p] -= (1.f - f) * t;
                a.zs[(1 + p) * q + e] = f * t;
            }
        }
        a.szz[o] = tot[0] + add;
#pragma unroll
        for (int l = 0; l < L; ++l) a.r[(size_t)l * a.rstride + o] = tot[1 + l];
    }
}

template <int R, int TX, int NW>
__global__ __launch_bounds__(64 * NW) void visco_el_velocity_kernel(ViscoElArgs a) {
    constexpr int W = kVeTZ + 2 * R, HH = TX + 2 * R;
    __shared__ float s_xx[HH * W], s_zz[HH * W], s_xz[HH * W];
    ve_el_velocity<R, TX, NW>(a, xcd_tile(blockIdx.x, a.ntiles), a.rec_col, s_xx, s_zz, s_xz);
}

template <int R, int L, int TX, int NW>
__global__ __launch_bounds__(64 * NW) void visco_el_stress_kernel(ViscoElArgs a) {
    constexpr int W = kVeTZ + 2 * R, HH = TX + 2 * R;
    __shared__ float s_vx[HH * W], s_vz[HH * W];
    ve_el_stress<R, L, TX, NW>(a, xcd_tile(blockIdx.x, a.ntiles), a.step, s_vx, s_vz);
}

template <int R, int TX, int NW>
__global__ __launch_bounds__(64 * NW) void visco_ac_velocity_kernel(ViscoElArgs a) {
    constexpr int W = kVeTZ + 2 * R, HH = TX + 2 * R;
    __shared__ float s_s[HH * W];
    ve_ac_velocity<R, TX, NW>(a, xcd_tile(blockIdx.x, a.ntiles), a.rec_col, s_s);
}

template <int R, int L, int TX, int NW>
__global__ __launch_bounds__(64 * NW) void visco_ac_stress_kernel(ViscoElArgs a) {
    constexpr int W = kVeTZ + 2 * R, HH = TX + 2 * R;
    __shared__ float s_vx[HH * W], s_vz[HH * W];
    ve_ac_stress<R, L, TX, NW>(a, xcd_tile(blockIdx.x, a.ntiles), a.step, s_vx, s_vz);
}

// The same launches over the tiles of several shots.
template <int R, int TX, int NW>
__global__ __launch_bounds__(64 * NW) void visco_el_velocity_batch(ViscoBatchArgs b) {
    constexpr int W = kVeTZ + 2 * R, HH = TX + 2 * R;
    __shared__ float s_xx[HH * W], s_zz[HH * W], s_xz[HH * W];
    const int* m = reinterpret_cast<const int*>(b.tile_map) + 2 * xcd_tile(blockIdx.x, b.ntile);
    ViscoShotArgs& a = *(ViscoShotArgs*)(b.shots + ro(m, 0));
    ve_el_velocity<R, TX, NW>(a, ro(m, 1), b.rec_col, s_xx, s_zz, s_xz);
}

template <int R, int L, int TX, int NW>
__global__ __launch_bounds__(64 * NW) void visco_el_stress_batch(ViscoBatchArgs b) {
    constexpr int W = kVeTZ + 2 * R, HH = TX + 2 * R;
    __shared__ float s_vx[HH * W], s_vz[HH * W];
    const int* m = reinterpret_cast<const int*>(b.tile_map) + 2 * xcd_tile(blockIdx.x, b.ntile);
    ViscoShotArgs& a = *(ViscoShotArgs*)(b.shots + ro(m, 0));
    ve_el_stress<R, L, TX, NW>(a, ro(m, 1), b.step, s_vx, s_vz);
}

template <int R, int TX, int NW>
__global__ __launch_bounds__(64 * NW) void visco_ac_velocity_batch(ViscoBatchArgs b) {
    constexpr int W = kVeTZ + 2 * R, HH = TX + 2 * R;
    __shared__ float s_s[HH * W];
    const int* m = reinterpret_cast<const int*>(b.tile_map) + 2 * xcd_tile(blockIdx.x, b.ntile);
    ViscoShotArgs& a = *(ViscoShotArgs*)(b.shots + ro(m, 0));
    ve_ac_velocity<R, TX, NW>(a, ro(m, 1), b.rec_col, s_s);
}

template <int R, int L, int TX, int NW>
__global__ __launch_bounds__(64 * NW) void visco_ac_stress_batch(ViscoBatchArgs b) {
    constexpr int W = kVeTZ + 2 * R, HH = TX + 2 * R;
    __shared__ float s_vx[HH * W], s_vz[HH * W];
    const int* m = reinterpret_cast<const int*>(b.tile_map) + 2 * xcd_tile(blockIdx.x, b.ntile);
    ViscoShotArgs& a = *(ViscoShotArgs*)(b.shots + ro(m, 0));
    ve_ac_stress<R, L, TX, NW>(a, ro(m, 1), b.step, s_vx, s_vz);
}

// Receivers of the last step (the velocity launch of the next step gathers
// all others).
__global__ void visco_el_gather_kernel(ViscoElArgs a) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < a.nrec && a.rec_col >= 0) {
        const ptrdiff_t o = (ptrdiff_t)a.rx[r] * a.pitch + a.rz[r];
        a.rec_zz[(size_t)r * a.nrt + a.rec_col] = a.szz[o];
        if (a.rec_xx) a.rec_xx[(size_t)r * a.nrt + a.rec_col] = a.sxx[o];
    }
}

}  // namespace sfd
