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

#include <type_traits>

namespace sfd {

// Strip row of row i (x strips: the first xw rows, then the last xw), -1
// outside; likewise the strip column of column j.
__device__ __forceinline__ int adj_xstrip(int i, int nx, int xw) {
    return i < xw ? i : i >= nx - xw ? i - (nx - xw) + xw : -1;
}

// Rows [i0 - HR, i0 + TX + HR) x columns [j0 - HC, j0 + 64 + HC) of the value
// f(i, j) (0 off the updated cells [R, nx - R) x [R, nz - R)) into
// lds[row][col]; all loads issued before the first store.  Each wave takes
// whole rows of the 64 tile columns (f called with a wave-uniform row and
// std::true_type), then the workgroup the 2 HC edge columns of every row (f
// called with std::false_type: row and column per lane).  wv: the wave's
// index, wave-uniform.
template <int TX, int NW, int HR, int HC, class F>
__device__ __forceinline__ void adj_stage(float* lds, int i0, int j0, int nx, int nz, int R, int wv, F f) {
    constexpr int W = kTiledTZ + 2 * HC, Hh = TX + 2 * HR, T = 64 * NW;
    constexpr int RI = (Hh + NW - 1) / NW;
    constexpr int EW = 2 * HC, E = EW * Hh, EWd = EW > 0 ? EW : 1;
    constexpr int EI = E > 0 ? (E + T - 1) / T : 0;
    const int lane = threadIdx.x & 63;
    const int gj = j0 + lane;
    const bool jin = gj >= R && gj < nz - R;
    float v[RI];
#pragma unroll
    for (int q = 0; q < RI; ++q) {
        const int r = wv + q * NW, gi = i0 - HR + r;
        v[q] = 0.f;
        if ((RI * NW == Hh || r < Hh) && gi >= R && gi < nx - R) {
            const float x = f(gi, jin ? gj : R, std::true_type{});  // every lane reads an updated cell
            v[q] = jin ? x : 0.f;
        }
    }
    float ve[EI > 0 ? EI : 1];
    int le[EI > 0 ? EI : 1];
#pragma unroll
    for (int q = 0; q < EI; ++q) {
        const int e = threadIdx.x + q * T;
        const int r = e / EWd, c = e - r * EWd;
        const int col = c < HC ? c : c + kTiledTZ;
        const int gi = i0 - HR + r, gjj = j0 - HC + col;
        le[q] = e < E ? r * W + col : -1;
        const bool upd = e < E && gi >= R && gi < nx - R && gjj >= R && gjj < nz - R;
        ve[q] = upd ? f(gi, gjj, std::false_type{}) : 0.f;
    }
#pragma unroll
    for (int q = 0; q < RI; ++q) {
        const int r = wv + q * NW;
        if (RI * NW == Hh || r < Hh) lds[r * W + HC + lane] = v[q];
    }
#pragma unroll
    for (int q = 0; q < EI; ++q)
        if (le[q] >= 0) lds[le[q]] = ve[q];
}

// A load of a field (row stride rs bytes) at (row, col) through a buffer
// view: the row's offset as a scalar where the row is wave-uniform (U), else
// both in the lane's offset.
template <class U>
__device__ __forceinline__ float adj_ld(const Buf& B, int row, int col, uint32_t rs, U) {
    if constexpr (U::value) return B.ld(col * 4, (uint32_t)row * rs);
    else return B.ld((int)((uint32_t)row * rs) + col * 4, 0u);
}
// A per-row or per-column profile value: through the constant address space
// where the index is wave-uniform.
template <class U>
__device__ __forceinline__ float adj_prof(const float* p, int idx, U) {
    if constexpr (U::value) return ro(p, idx);
    else return p[idx];
}

// A1.  Args: p_in = lP, vx_in/vx_out = lVx, vz = lVz, psi_vx_in = lWx,
// psi_vz_in = lWz, psi_px_in/out = qx, psi_pz_in/out = qz (in place).
// Fields through raw buffer views (a lane's offset its column, a row's a
// wave-uniform scalar) as in acoustic_tiled.h.
template <int R, int TX, int NW, class Args>
__device__ __forceinline__ void adjc_a1(const Args& a, int t, float* s_x, float* s_z) {
    constexpr int WX = kTiledTZ, WZ = kTiledTZ + 2 * R, RPW = TX / NW;
    int i0, j0;
    tiled_origin(t, a.ntz, TX, R, &i0, &j0);
    const int nx = a.nx, nz = a.nz, lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int j = j0 + lane, vl = j * 4;
    const bool jok = j < nz - R;
    const int xw = a.pml_xw, zw = a.pml_zw;
    const uint32_t rb = (uint32_t)nz * 4u, zb = (uint32_t)zw * 8u;  // row bytes of the fields, of the z strips
    const int zc = j < zw ? j : j >= nz - zw ? j - (nz - zw) + zw : -1;
    const Buf KF = field_buf(a.kf), LP = field_buf(a.p_in), LWX = field_buf(a.psi_vx_in), LWZ = field_buf(a.psi_vz_in);
    const Buf VXI = field_buf(a.vx_in), VZI = field_buf(a.vz_in), QX = field_buf(a.psi_px_in), QZ = field_buf(a.psi_pz_in);
    float vx0[RPW], vz0[RPW], bf[RPW], qx0[RPW], qz0[RPW];
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        const int i = i0 + wv + q * NW;
        const bool iok = i < nx - R;
        vx0[q] = VXI.ld(vl, (uint32_t)i * rb);
        vz0[q] = VZI.ld(vl, (uint32_t)i * rb);
        bf[q] = a.bf_const != 0.f ? a.bf_const : field_buf(a.bf).ld(vl, (uint32_t)i * rb);
        const int xr = iok ? adj_xstrip(i, nx, xw) : -1;
        qx0[q] = xr >= 0 ? QX.ld(jok ? vl : kOOB, (uint32_t)xr * rb) : 0.f;
        qz0[q] = iok && jok && zc >= 0 ? QZ.ld(zc * 4, (uint32_t)i * zb) : 0.f;
    }
    // lt_x with an x halo, lt_z with a z halo.
    adj_stage<TX, NW, R, 0>(s_x, i0, j0, nx, nz, R, wv, [&](int gi, int gj, auto u) {
        const float mu = adj_ld(KF, gi, gj, rb, u) * adj_ld(LP, gi, gj, rb, u);
        const int xr = adj_xstrip(gi, nx, xw);
        if (xr < 0) return -mu;
        return __builtin_fmaf(adj_prof(a.pml_x, nx + gi, u), adj_ld(LWX, xr, gj, rb, u) - mu, -mu);
    });
    adj_stage<TX, NW, 0, R>(s_z, i0, j0, nx, nz, R, wv, [&](int gi, int gj, auto u) {
        const float mu = adj_ld(KF, gi, gj, rb, u) * adj_ld(LP, gi, gj, rb, u);
        const int zz = gj < zw ? gj : gj >= nz - zw ? gj - (nz - zw) + zw : -1;
        if (zz < 0) return -mu;
        return __builtin_fmaf(a.pml_z[nz + gj], adj_ld(LWZ, gi, zz, zb, std::false_type{}) - mu, -mu);
    });
    __syncthreads();
    if (!jok) return;
    float cx[R];
#pragma unroll
    for (int m = 0; m < R; ++m) cx[m] = a.cx[m];
    const float zx = a.zx;
    const Buf VXO = field_buf(a.vx_out), VZO = field_buf(a.vz_out);
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
        const float lvx = vx0[q] - sx, lvz = vz0[q] - sz;
        VXO.st(lvx, vl, (uint32_t)i * rb);
        VZO.st(lvz, vl, (uint32_t)i * rb);
        const int xr = adj_xstrip(i, nx, xw);
        if (xr >= 0)
            field_buf(a.psi_px_out).st(__builtin_fmaf(ro(a.pml_x, 2 * nx + i), qx0[q], -bf[q] * lvx), vl,
                                       (uint32_t)xr * rb);
        if (zc >= 0)
            field_buf(a.psi_pz_out).st(__builtin_fmaf(a.pml_z[2 * nz + j], qz0[q], -bf[q] * lvz), zc * 4,
                                       (uint32_t)i * zb);
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
    const int nx = a.nx, nz = a.nz, lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int j = j0 + lane, vl = j * 4;
    const bool jok = j < nz - R;
    const int xw = a.pml_xw, zw = a.pml_zw;
    const uint32_t rb = (uint32_t)nz * 4u, zb = (uint32_t)zw * 8u;
    const int zc = j < zw ? j : j >= nz - zw ? j - (nz - zw) + zw : -1;
    const float* dv = a.div + (size_t)(k % a.div_period) * a.div_stride;
    const float bfc = a.bf_const;
    const Buf LP = field_buf(a.p_in), KF = field_buf(a.kf), DV = field_buf(dv), GR = field_buf(a.grad);
    const Buf LWX = field_buf(a.psi_vx_in), LWZ = field_buf(a.psi_vz_in);
    float p0[RPW], kf[RPW], dk[RPW], gr[RPW], wx0[RPW], wz0[RPW];
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        const int i = i0 + wv + q * NW;
        const bool iok = i < nx - R;
        const uint32_t o = (uint32_t)i * rb;
        p0[q] = LP.ld(vl, o);
        kf[q] = KF.ld(vl, o);
        dk[q] = DV.ld(vl, o);
        gr[q] = GR.ld(vl, o);
        const int xr = iok ? adj_xstrip(i, nx, xw) : -1;
        wx0[q] = xr >= 0 ? LWX.ld(jok ? vl : kOOB, (uint32_t)xr * rb) : 0.f;
        wz0[q] = iok && jok && zc >= 0 ? LWZ.ld(zc * 4, (uint32_t)i * zb) : 0.f;
    }
    // ls_x with an x halo, ls_z with a z halo.
    const Buf BF = field_buf(a.bf), LVX = field_buf(a.vx_in), LVZ = field_buf(a.vz_in);
    const Buf QX = field_buf(a.psi_px_in), QZ = field_buf(a.psi_pz_in);
    adj_stage<TX, NW, R, 0>(s_x, i0, j0, nx, nz, R, wv, [&](int gi, int gj, auto u) {
        const float b = bfc != 0.f ? bfc : adj_ld(BF, gi, gj, rb, u);
        const float nu = b * adj_ld(LVX, gi, gj, rb, u);
        const int xr = adj_xstrip(gi, nx, xw);
        if (xr < 0) return -nu;
        return __builtin_fmaf(adj_prof(a.pml_x, 3 * nx + gi, u), adj_ld(QX, xr, gj, rb, u), -nu);
    });
    adj_stage<TX, NW, 0, R>(s_z, i0, j0, nx, nz, R, wv, [&](int gi, int gj, auto u) {
        const float b = bfc != 0.f ? bfc : adj_ld(BF, gi, gj, rb, u);
        const float nu = b * adj_ld(LVZ, gi, gj, rb, u);
        const int zz = gj < zw ? gj : gj >= nz - zw ? gj - (nz - zw) + zw : -1;
        if (zz < 0) return -nu;
        return __builtin_fmaf(a.pml_z[3 * nz + gj], adj_ld(QZ, gi, zz, zb, std::false_type{}), -nu);
    });
    __syncthreads();
    if (!jok) return;
    float cx[R];
#pragma unroll
    for (int m = 0; m < R; ++m) cx[m] = a.cx[m];
    const float zx = a.zx;
    const Buf LPO = field_buf(a.p_out);
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
        const uint32_t o = (uint32_t)i * rb;
        const float mu = kf[q] * p0[q];
        const float pn = p0[q] - (sx + sz);
        LPO.st(pn, vl, o);
        GR.st(__builtin_fmaf(mu, dk[q], gr[q]), vl, o);
        if (i == a.sx && j == a.sz) a.src_adj[k] = kf[q] * pn;
        const int xr = adj_xstrip(i, nx, xw);
        if (xr >= 0) field_buf(a.psi_vx_out).st(ro(a.pml_x, i) * (wx0[q] - mu), vl, (uint32_t)xr * rb);
        if (zc >= 0) field_buf(a.psi_vz_out).st(a.pml_z[j] * (wz0[q] - mu), zc * 4, (uint32_t)i * zb);
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
