// Adjoint of the fast-arithmetic C-PML time step (no counterpart in the
// reference, which has neither C-PML nor an adjoint).  The forward step of
// acoustic_tiled.h in C-PML mode is, per updated cell, in derivative units
// (cx = C/dx; z stencils scaled by zx = dx/dz):
//
//   P~  = P - amp_k e_s
//   d_x = D+x P~,  Qx' = bh_x Qx + ah_x d_x,  Vx' = Vx - b (d_x + Qx')   (z alike)
//   t_x = D-x Vx', Wx' = bi_x Wx + ai_x t_x                             (z alike)
//   P'  = P~ - K (t_x + Wx' + t_z + Wz')
//
// (Q = psi_p*, W = psi_v*, kept only in the strips; b = dt/rho, K = rho vp^2
// dt).  Its transpose runs as two in-place launches in the reverse order,
// on the unscaled adjoint fields (lP, lVx, lVz) and strip fields
// (lWx, lWz, qx, qz):
//
//   A1 (adjoint of the pressure half):  mu = K lP,
//        lt_x = -mu + ai_x (lWx - mu)          (staged with an x halo)
//        lVx <- lVx - D+x lt_x                 ((D-x)^T = -D+x)
//        qx  <- bh_x qx - b lVx                (z alike)
//   A2 (adjoint of the velocity half):
//        ls_x = -b lVx + ah_x qx               (staged with an x halo)
//        lP  <- lP - D-x ls_x - D-z ls_z       ((D+x)^T = -D-x)
//        lWx <- bi_x (lWx - mu), grad += mu (t_x + Wx' + t_z + Wz')_k,
//        src_adj[k] = K lP at the source cell
//
// qx holds the adjoint of Qx' before the factor bh_x: the adjoint of the
// memory field itself is bh_x qx, applied when A1 of the previous step reads
// it (qx is read by A2's neighbours within the step, so A1 cannot store the
// product).  Neither launch reads a cell another workgroup of the same launch
// writes: A1 reads lP, lW of the halo and writes lV, q; A2 reads lV, q of the
// halo and writes lP, lW.  The divergence of step k comes from the
// recomputation (MODE_FWD_DIV_CPML stores t_x + Wx' + t_z + Wz').  grad
// accumulates sum_k K lP' div_k, so dJ/dvp = -2 grad / vp as for the sponge
// adjoint.  The adjoint fields live in the arena like the state; lW*, q* in
// the second (otherwise unused) copies of the plan's strip fields.
#pragma once
#include <hip/hip_runtime.h>

#include "acoustic_kernel.h"
#include "acoustic_tiled.h"
#include "device_util.h"

namespace sfd {

// Strip row of row i (x strips: the first xw rows, then the last xw), -1
// outside; likewise the strip column of column j.
__device__ __forceinline__ int adj_xstrip(int i, int nx, int xw) {
    return i < xw ? i : i >= nx - xw ? i - (nx - xw) + xw : -1;
}

// Rows [i0 - HR, i0 + TX + HR) x columns [j0 - HC, j0 + 64 + HC) of the value
// f(i, j) (0 off the updated cells [R, nx - R) x [R, nz - R)) into
// lds[row][col]; all loads issued before the first store.
template <int TX, int NW, int HR, int HC, class F>
__device__ __forceinline__ void adj_stage(float* lds, int i0, int j0, int nx, int nz, int R, F f) {
    constexpr int W = kTiledTZ + 2 * HC, Hh = TX + 2 * HR, N = W * Hh, T = 64 * NW;
    constexpr int ITER = (N + T - 1) / T;
    float v[ITER];
#pragma unroll
    for (int q = 0; q < ITER; ++q) {
        const int e = threadIdx.x + q * T;
        const int r = e / W, c = e - r * W;
        const int gi = i0 - HR + r, gj = j0 - HC + c;
        const bool upd = e < N && gi >= R && gi < nx - R && gj >= R && gj < nz - R;
        v[q] = upd ? f(gi, gj) : 0.f;
    }
#pragma unroll
    for (int q = 0; q < ITER; ++q) {
        const int e = threadIdx.x + q * T;
        if (e < N) lds[e] = v[q];
    }
}

// A1.  Args: p_in = lP, vx_in/vx_out = lVx, vz = lVz, psi_vx_in = lWx,
// psi_vz_in = lWz, psi_px_in/out = qx, psi_pz_in/out = qz (in place).
template <int R, int TX, int NW, class Args>
__device__ __forceinline__ void adjc_a1(const Args& a, int t, float* s_x, float* s_z) {
    constexpr int WX = kTiledTZ, WZ = kTiledTZ + 2 * R, RPW = TX / NW;
    int i0, j0;
    tiled_origin(t, a.ntz, TX, R, &i0, &j0);
    const int nx = a.nx, nz = a.nz, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int j = j0 + lane;
    const bool jok = j < nz - R;
    const int xw = a.pml_xw, zw = a.pml_zw;
    const int zc = j < zw ? j : j >= nz - zw ? j - (nz - zw) + zw : -1;
    float vx0[RPW], vz0[RPW], bf[RPW], qx0[RPW], qz0[RPW];
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        const int i = i0 + wv + q * NW;
        const bool ok = jok && i < nx - R;
        const ptrdiff_t o = (ptrdiff_t)i * nz + j;
        vx0[q] = ok ? a.vx_in[o] : 0.f;
        vz0[q] = ok ? a.vz_in[o] : 0.f;
        bf[q] = a.bf_const != 0.f ? a.bf_const : ok ? a.bf[o] : 0.f;
        const int xr = ok ? adj_xstrip(i, nx, xw) : -1;
        qx0[q] = xr >= 0 ? a.psi_px_in[(size_t)xr * nz + j] : 0.f;
        qz0[q] = ok && zc >= 0 ? a.psi_pz_in[(size_t)i * (2 * zw) + zc] : 0.f;
    }
    // lt_x with an x halo, lt_z with a z halo.
    adj_stage<TX, NW, R, 0>(s_x, i0, j0, nx, nz, R, [&](int gi, int gj) {
        const ptrdiff_t o = (ptrdiff_t)gi * nz + gj;
        const float mu = a.kf[o] * a.p_in[o];
        const int xr = adj_xstrip(gi, nx, xw);
        if (xr < 0) return -mu;
        return __builtin_fmaf(a.pml_x[nx + gi], a.psi_vx_in[(size_t)xr * nz + gj] - mu, -mu);
    });
    adj_stage<TX, NW, 0, R>(s_z, i0, j0, nx, nz, R, [&](int gi, int gj) {
        const ptrdiff_t o = (ptrdiff_t)gi * nz + gj;
        const float mu = a.kf[o] * a.p_in[o];
        const int zz = gj < zw ? gj : gj >= nz - zw ? gj - (nz - zw) + zw : -1;
        if (zz < 0) return -mu;
        return __builtin_fmaf(a.pml_z[nz + gj], a.psi_vz_in[(size_t)gi * (2 * zw) + zz] - mu, -mu);
    });
    __syncthreads();
    if (!jok) return;
    float cx[R];
#pragma unroll
    for (int m = 0; m < R; ++m) cx[m] = a.cx[m];
    const float zx = a.zx;
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        const int il = wv + q * NW, i = i0 + il;
        if (i >= nx - R) break;
        // D+ of lt at the velocity node i (+1/2): rows i - R + 1 .. i + R.
        const float* px = s_x + (il + R) * WX + lane;
        const float* pz = s_z + il * WZ + lane + R;
        float sx = cx[0] * (px[WX] - px[0]);
#pragma unroll
        for (int m = 2; m <= R; ++m) sx = __builtin_fmaf(cx[m - 1], px[m * WX] - px[(1 - m) * WX], sx);
        float sz = cx[0] * (pz[1] - pz[0]);
#pragma unroll
        for (int m = 2; m <= R; ++m) sz = __builtin_fmaf(cx[m - 1], pz[m] - pz[1 - m], sz);
        sz = zx * sz;
        const ptrdiff_t o = (ptrdiff_t)i * nz + j;
        const float lvx = vx0[q] - sx, lvz = vz0[q] - sz;
        a.vx_out[o] = lvx;
        a.vz_out[o] = lvz;
        const int xr = adj_xstrip(i, nx, xw);
        if (xr >= 0)
            a.psi_px_out[(size_t)xr * nz + j] = __builtin_fmaf(a.pml_x[2 * nx + i], qx0[q], -bf[q] * lvx);
        if (zc >= 0)
            a.psi_pz_out[(size_t)i * (2 * zw) + zc] = __builtin_fmaf(a.pml_z[2 * nz + j], qz0[q], -bf[q] * lvz);
    }
}

// A2.  Args: p_in/p_out = lP, vx_in = lVx, vz_in = lVz, psi_px_in = qx,
// psi_pz_in = qz, psi_vx_in/out = lWx, psi_vz_in/out = lWz (in place), div,
// grad, src_adj.
template <int R, int TX, int NW, class Args>
__device__ __forceinline__ void adjc_a2(const Args& a, int t, int k, float* s_x, float* s_z) {
    constexpr int WX = kTiledTZ, WZ = kTiledTZ + 2 * R, RPW = TX / NW;
    int i0, j0;
    tiled_origin(t, a.ntz, TX, R, &i0, &j0);
    const int nx = a.nx, nz = a.nz, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int j = j0 + lane;
    const bool jok = j < nz - R;
    const int xw = a.pml_xw, zw = a.pml_zw;
    const int zc = j < zw ? j : j >= nz - zw ? j - (nz - zw) + zw : -1;
    const float* dv = a.div + (size_t)(k % a.div_period) * a.div_stride;
    const float bfc = a.bf_const;
    float p0[RPW], kf[RPW], dk[RPW], gr[RPW], wx0[RPW], wz0[RPW];
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        const int i = i0 + wv + q * NW;
        const bool ok = jok && i < nx - R;
        const ptrdiff_t o = (ptrdiff_t)i * nz + j;
        p0[q] = ok ? a.p_in[o] : 0.f;
        kf[q] = ok ? a.kf[o] : 0.f;
        dk[q] = ok ? dv[o] : 0.f;
        gr[q] = ok ? a.grad[o] : 0.f;
        const int xr = ok ? adj_xstrip(i, nx, xw) : -1;
        wx0[q] = xr >= 0 ? a.psi_vx_in[(size_t)xr * nz + j] : 0.f;
        wz0[q] = ok && zc >= 0 ? a.psi_vz_in[(size_t)i * (2 * zw) + zc] : 0.f;
    }
    // ls_x with an x halo, ls_z with a z halo.
    adj_stage<TX, NW, R, 0>(s_x, i0, j0, nx, nz, R, [&](int gi, int gj) {
        const ptrdiff_t o = (ptrdiff_t)gi * nz + gj;
        const float b = bfc != 0.f ? bfc : a.bf[o];
        const float nu = b * a.vx_in[o];
        const int xr = adj_xstrip(gi, nx, xw);
        if (xr < 0) return -nu;
        return __builtin_fmaf(a.pml_x[3 * nx + gi], a.psi_px_in[(size_t)xr * nz + gj], -nu);
    });
    adj_stage<TX, NW, 0, R>(s_z, i0, j0, nx, nz, R, [&](int gi, int gj) {
        const ptrdiff_t o = (ptrdiff_t)gi * nz + gj;
        const float b = bfc != 0.f ? bfc : a.bf[o];
        const float nu = b * a.vz_in[o];
        const int zz = gj < zw ? gj : gj >= nz - zw ? gj - (nz - zw) + zw : -1;
        if (zz < 0) return -nu;
        return __builtin_fmaf(a.pml_z[3 * nz + gj], a.psi_pz_in[(size_t)gi * (2 * zw) + zz], -nu);
    });
    __syncthreads();
    if (!jok) return;
    float cx[R];
#pragma unroll
    for (int m = 0; m < R; ++m) cx[m] = a.cx[m];
    const float zx = a.zx;
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        const int il = wv + q * NW, i = i0 + il;
        if (i >= nx - R) break;
        // D- of ls at the pressure node i: velocity nodes i - R .. i + R - 1.
        const float* px = s_x + (il + R) * WX + lane;
        const float* pz = s_z + il * WZ + lane + R;
        float sx = cx[0] * (px[0] - px[-WX]);
#pragma unroll
        for (int m = 2; m <= R; ++m) sx = __builtin_fmaf(cx[m - 1], px[(m - 1) * WX] - px[-m * WX], sx);
        float sz = cx[0] * (pz[0] - pz[-1]);
#pragma unroll
        for (int m = 2; m <= R; ++m) sz = __builtin_fmaf(cx[m - 1], pz[m - 1] - pz[-m], sz);
        sz = zx * sz;
        const ptrdiff_t o = (ptrdiff_t)i * nz + j;
        const float mu = kf[q] * p0[q];
        const float pn = p0[q] - (sx + sz);
        a.p_out[o] = pn;
        a.grad[o] = __builtin_fmaf(mu, dk[q], gr[q]);
        if (i == a.sx && j == a.sz) a.src_adj[k] = kf[q] * pn;
        const int xr = adj_xstrip(i, nx, xw);
        if (xr >= 0) a.psi_vx_out[(size_t)xr * nz + j] = a.pml_x[i] * (wx0[q] - mu);
        if (zc >= 0) a.psi_vz_out[(size_t)i * (2 * zw) + zc] = a.pml_z[j] * (wz0[q] - mu);
    }
}

template <int R, int TX, int NW>
__global__ __launch_bounds__(64 * NW) void acoustic_adjc_a1(const AcousticStepArgs a) {
    __shared__ float s_x[(TX + 2 * R) * kTiledTZ];
    __shared__ float s_z[TX * (kTiledTZ + 2 * R)];
    adjc_a1<R, TX, NW>(a, xcd_tile((int)blockIdx.x, a.ntz * a.nchunk), s_x, s_z);
}

template <int R, int TX, int NW>
__global__ __launch_bounds__(64 * NW) void acoustic_adjc_a2(const AcousticStepArgs a) {
    __shared__ float s_x[(TX + 2 * R) * kTiledTZ];
    __shared__ float s_z[TX * (kTiledTZ + 2 * R)];
    adjc_a2<R, TX, NW>(a, xcd_tile((int)blockIdx.x, a.ntz * a.nchunk), a.k, s_x, s_z);
}

template <int R, int TX, int NW>
__global__ __launch_bounds__(64 * NW) void acoustic_adjc_a1_batch(const BatchStepArgs b) {
    __shared__ float s_x[(TX + 2 * R) * kTiledTZ];
    __shared__ float s_z[TX * (kTiledTZ + 2 * R)];
    const int* m = reinterpret_cast<const int*>(b.tile_map) + 2 * xcd_tile((int)blockIdx.x, b.ntile);
    using CArgs = const __attribute__((address_space(4))) AcousticStepArgs;
    CArgs& a = *(CArgs*)(b.shots + ro(m, 0));
    adjc_a1<R, TX, NW>(a, ro(m, 1), s_x, s_z);
}

template <int R, int TX, int NW>
__global__ __launch_bounds__(64 * NW) void acoustic_adjc_a2_batch(const BatchStepArgs b) {
    __shared__ float s_x[(TX + 2 * R) * kTiledTZ];
    __shared__ float s_z[TX * (kTiledTZ + 2 * R)];
    const int* m = reinterpret_cast<const int*>(b.tile_map) + 2 * xcd_tile((int)blockIdx.x, b.ntile);
    using CArgs = const __attribute__((address_space(4))) AcousticStepArgs;
    CArgs& a = *(CArgs*)(b.shots + ro(m, 0));
    adjc_a2<R, TX, NW>(a, ro(m, 1), b.k, s_x, s_z);
}

}  // namespace sfd
