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

namespace sfd {

enum KernelMode : int {
    MODE_FAST_SPONGE = 0,
    MODE_FAST_CPML = 1,
    MODE_REF_SPONGE = 2,
};

struct AcousticStepArgs {
    int nx, nz;
    int chunk;     // output rows per workgroup (x direction)
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
    // Sponge: ax[i] is the x-edge factor of row i, or 2.0 where row i is not
    // in an x-edge strip; az[j] is the z-edge factor of column j (1.0 inside).
    const float* __restrict__ ax;
    const float* __restrict__ az;
    // C-PML: profiles [4][n] = b at nodes, a at nodes, b at half nodes, a at
    // half nodes; memory fields live in strips of sx_w rows / sz_w columns at
    // each end: psi_x* [2*sx_w][nz], psi_z* [nx][2*sz_w].
    const float* __restrict__ pml_x;
    const float* __restrict__ pml_z;
    int pml_xw, pml_zw;
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
    // Receivers: row i owns spans rec_first[i] .. rec_first[i+1]-1 of
    // rec_span = (jlo, jhi, slot) ; the absorbed new P of cell (i, j),
    // jlo <= j < jhi, goes to rec_buf[k * nslot + slot + j - jlo].
    const int* __restrict__ rec_first;
    const int4* __restrict__ rec_span;
    float* __restrict__ rec_buf;
    int nslot;
    // FAST: cx = C/dx, cz = C/dz.  REF: cx = C (unscaled), cz unused.
    float cx[8];
    float cz[8];
    // REF only: the float quotients dt/dx, dt/dz and the double 1.0/dx, 1.0/dz.
    float dtdx, dtdz;
    double inv_dx, inv_dz;
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

template <int N, class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl<N>(f, std::make_integer_sequence<int, N>{});
}

__device__ __forceinline__ int floor_div(int a, int b) {
    return (a >= 0) ? a / b : -((-a + b - 1) / b);
}

template <int R, int NW, int MODE>
__global__ __launch_bounds__(64 * NW) void acoustic_step_kernel(const AcousticStepArgs a) {
    constexpr int TZ = 64 * NW;      // z cells per workgroup
    constexpr int Q = 2 * R;         // register queue length
    constexpr int HR = 2 * R - 1;    // P halo on each side of the tile
    constexpr int RW = TZ + 2 * HR;  // LDS ring row width
    constexpr int VW = TZ + 2 * R - 1;  // LDS Vz row width: R left, R-1 right
    constexpr bool CPML = (MODE == MODE_FAST_CPML);
    using A = ArithFor<R, MODE>;

    __shared__ float s_ring[Q][RW];
    __shared__ float s_vz[2][VW];

    const int nx = a.nx, nz = a.nz;
    const int tid = threadIdx.x;
    const int j0 = R + (int)blockIdx.x * TZ;
    const int j = j0 + tid;
    const bool jload = j < nz;                // lane maps to a stored column
    const bool jupd = j < nz - R;             // lane's column is updated
    const int i0 = R + (int)blockIdx.y * a.chunk;
    const int i1 = min(i0 + a.chunk, nx - R);
    if (i0 >= i1) return;

    // Lane constants.
    const float az_j = jload ? a.az[j] : 1.0f;
    const float amp = a.amp[a.k];
    float cx[R], cz[R];
#pragma unroll
    for (int m = 0; m < R; ++m) {
        cx[m] = a.cx[m];
        cz[m] = (MODE == MODE_REF_SPONGE) ? a.cx[m] : a.cz[m];
    }
    // C-PML z profile of this lane and its strip column (-1 outside the strips).
    float pz_bi = 1.f, pz_ai = 0.f, pz_bh = 1.f, pz_ah = 0.f;
    int zcol = -1;
    if (CPML && jload) {
        if (j < a.pml_zw) zcol = j;
        else if (j >= nz - a.pml_zw) zcol = j - (nz - a.pml_zw) + a.pml_zw;
        if (zcol >= 0) {
            pz_bi = a.pml_z[j];
            pz_ai = a.pml_z[nz + j];
            pz_bh = a.pml_z[2 * nz + j];
            pz_ah = a.pml_z[3 * nz + j];
        }
    }
    const int zw2 = 2 * a.pml_zw;

    float qp[Q];   // qp[r % Q] = P(row r, j) with the source injected
    float qvx[Q];  // qvx[r % Q] = new Vx(row r, j), not yet absorbed
#pragma unroll
    for (int q = 0; q < Q; ++q) { qp[q] = 0.f; qvx[q] = 0.f; }

    const int i_first = i0 - 4 * R + 2;       // first marched row
    const int base0 = floor_div(i_first, Q) * Q;

    // Sponge factor of cell (r, j): the x-edge strips override the z-edge
    // strips in the corners, as the reference's table does (:164-194).
    auto absorb_row = [&](int r) -> float {
        if constexpr (CPML) {
            return 1.0f;
        } else {
            const float f = a.ax[r];
            return (f <= 1.0f) ? f : az_j;
        }
    };

    for (int base = base0; base < i1; base += Q) {
        static_for<Q>([&](auto U_) {
            constexpr int U = decltype(U_)::value;
            const int i = base + U;
            if (i < i_first || i >= i1) return;

            // ---- 1. front row of P: queue (and LDS ring for output rows) ----
            const int f = i + Q - 1;
            constexpr int sf = (U + Q - 1) % Q;
            {
                const bool frow = (f >= 0) && (f < nx);
                float pf = 0.f;
                if (frow && jload) pf = a.p_in[(size_t)f * nz + j];
                if (f == a.sx && j == a.sz) pf -= amp;
                qp[sf] = pf;
                if (f >= i0 && f < i1) {
                    s_ring[sf][tid + HR] = pf;
                    if (tid < HR) {
                        const int jl = j0 - HR + tid;
                        float v = 0.f;
                        if (jl >= 0) v = a.p_in[(size_t)f * nz + jl];
                        if (f == a.sx && jl == a.sz) v -= amp;
                        s_ring[sf][tid] = v;
                    }
                    if (tid >= TZ - HR) {
                        const int jr = j + HR;
                        float v = 0.f;
                        if (jr < nz) v = a.p_in[(size_t)f * nz + jr];
                        if (f == a.sx && jr == a.sz) v -= amp;
                        s_ring[sf][tid + 2 * HR] = v;
                    }
                }
            }

            // ---- 2. new Vx at row g = i + R - 1 from the P queue ----
            const int g = i + R - 1;
            constexpr int sg = (U + R - 1) % Q;
            if (g >= i0 - R && g < i1 + R - 1) {
                float vxn = 0.f;
                if (g >= R && g < nx - R && jupd) {
                    const size_t c = (size_t)g * nz + j;
                    const float s = A::stencil(cx, [&](int m) {
                        return qp[(U + R - 1 + m) % Q] - qp[(U + R - m) % Q];
                    });
                    float d = s;
                    if constexpr (CPML) {
                        const int xw = a.pml_xw;
                        int xrow = -1;
                        if (g < xw) xrow = g;
                        else if (g >= nx - xw) xrow = g - (nx - xw) + xw;
                        if (xrow >= 0) {
                            const size_t cs = (size_t)xrow * nz + j;
                            const float psi = __builtin_fmaf(a.pml_x[3 * nx + g], s,
                                                             a.pml_x[2 * nx + g] * a.psi_px_in[cs]);
                            d = s + psi;
                            if (g >= i0 && g < i1) a.psi_px_out[cs] = psi;
                        }
                    }
                    vxn = A::vel(a.vx_in[c], a.bf[c], d, a.dtdx);
                    if (g >= i0 && g < i1) a.vx_out[c] = vxn * absorb_row(g);
                }
                qvx[sg] = vxn;
            }

            if (i < i0) return;
            // Ring rows written while the queues were filling become visible.
            if (i == i0) __syncthreads();

            // ---- 3. new Vz at row i (own lane and halo lanes) into LDS ----
            constexpr int si = U % Q;
            const int vb = i & 1;
            const size_t rowoff = (size_t)i * nz;
            auto vz_at = [&](int jj, int rpos, bool own) -> float {
                // rpos: ring position of column jj
                if (jj < R || jj >= nz - R) return 0.f;
                const float s = A::stencil(cz, [&](int m) {
                    return s_ring[si][rpos + m] - s_ring[si][rpos - m + 1];
                });
                float d = s;
                if constexpr (CPML) {
                    int zc = own ? zcol : -1;
                    float bh = pz_bh, ah = pz_ah;
                    if (!own) {
                        if (jj < a.pml_zw) zc = jj;
                        else if (jj >= nz - a.pml_zw) zc = jj - (nz - a.pml_zw) + a.pml_zw;
                        if (zc >= 0) { bh = a.pml_z[2 * nz + jj]; ah = a.pml_z[3 * nz + jj]; }
                    }
                    if (zc >= 0) {
                        const size_t cs = (size_t)i * zw2 + zc;
                        const float psi = __builtin_fmaf(ah, s, bh * a.psi_pz_in[cs]);
                        d = s + psi;
                        if (own) a.psi_pz_out[cs] = psi;
                    }
                }
                return A::vel(a.vz_in[rowoff + jj], a.bf[rowoff + jj], d, a.dtdz);
            };
            {
                const float vzn = vz_at(j, tid + HR, true);
                s_vz[vb][tid + R] = vzn;
                if (jupd) a.vz_out[rowoff + j] = vzn * absorb_row(i);
                if (tid < R) {
                    const int jl = j0 - R + tid;
                    s_vz[vb][tid] = vz_at(jl, tid + R - 1, false);
                }
                if (tid >= TZ - (R - 1)) {
                    const int jr = j + R - 1;
                    s_vz[vb][tid + 2 * R - 1] = vz_at(jr, tid + 3 * R - 2, false);
                }
            }
            __syncthreads();

            // ---- 4. new P at row i ----
            if (jupd) {
                const size_t c = rowoff + j;
                float sx = A::stencil(cx, [&](int m) {
                    return qvx[(U + m - 1) % Q] - qvx[(U - m + Q) % Q];
                });
                float sz = A::stencil(cz, [&](int m) {
                    return s_vz[vb][tid + R + m - 1] - s_vz[vb][tid + R - m];
                });
                if constexpr (CPML) {
                    const int xw = a.pml_xw;
                    int xrow = -1;
                    if (i < xw) xrow = i;
                    else if (i >= nx - xw) xrow = i - (nx - xw) + xw;
                    if (xrow >= 0) {
                        const size_t cs = (size_t)xrow * nz + j;
                        const float psi = __builtin_fmaf(a.pml_x[nx + i], sx,
                                                         a.pml_x[i] * a.psi_vx_in[cs]);
                        a.psi_vx_out[cs] = psi;
                        sx += psi;
                    }
                    if (zcol >= 0) {
                        const size_t cs = (size_t)i * zw2 + zcol;
                        const float psi = __builtin_fmaf(pz_ai, sz, pz_bi * a.psi_vz_in[cs]);
                        a.psi_vz_out[cs] = psi;
                        sz += psi;
                    }
                }
                const float pn = A::pres(qp[si], a.kf[c], sx, sz, a.inv_dx, a.inv_dz) * absorb_row(i);
                a.p_out[c] = pn;
                const int r0 = a.rec_first[i], r1 = a.rec_first[i + 1];
                for (int r = r0; r < r1; ++r) {
                    const int4 sp = a.rec_span[r];
                    if (j >= sp.x && j < sp.y)
                        a.rec_buf[(size_t)a.k * a.nslot + sp.z + (j - sp.x)] = pn;
                }
            }
        });
    }
}

}  // namespace sfd
