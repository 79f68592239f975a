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

// Read-only tables indexed by wave-uniform values are read through the
// constant address space, so the compiler may use scalar loads (counted apart
// from the vector loads whose partial waits keep the prefetch queue full).
template <class T>
__device__ __forceinline__ T ro(const T* p, int i) {
    return ((const __attribute__((address_space(4))) T*)p)[i];
}

__device__ __forceinline__ int floor_div(int a, int b) {
    return (a >= 0) ? a / b : -((-a + b - 1) / b);
}

// Values one lane loads from global memory for marched row i.  The loads of
// row i + NS - 1 are issued while row i is computed (software pipelining), so
// NS - 1 rows of memory traffic are in flight per lane.
struct RowLoads {
    float pf;           // P(f, j), f = i + 2R - 1 (front row)
    float ph;           // P(f, halo column) on the P-halo lanes
    float vx, bx;       // Vx(g, j), bf(g, j), g = i + R - 1
    float vz, bz, kf;   // Vz(i, j), bf(i, j), kf(i, j)
    float vzh, bzh;     // Vz(i, jh), bf(i, jh) on the Vz-halo lanes
    float psx;          // C-PML psi_px(g, j)
    float psvx;         // C-PML psi_vx(i, j)
    float pspz, psvz;   // C-PML psi_pz(i, j), psi_vz(i, j)
    float pspzh;        // C-PML psi_pz(i, jh) on the Vz-halo lanes
};

template <int R, int NW, int MODE, int NS>
__global__ __launch_bounds__(64 * NW) void acoustic_step_kernel(const AcousticStepArgs a) {
    constexpr int TZ = 64 * NW;         // z cells per workgroup
    constexpr int Q = 2 * R;            // register queue length
    constexpr int HR = 2 * R - 1;       // P halo on each side of the tile
    constexpr int RW = TZ + 2 * HR;     // LDS ring row width
    constexpr int VW = TZ + 2 * R - 1;  // LDS Vz row width: R left, R-1 right
    constexpr int PD = NS - 1;          // prefetch distance in rows
    constexpr bool CPML = (MODE == MODE_FAST_CPML);
    static_assert(Q % NS == 0, "prefetch slots must divide the queue length");
    static_assert(TZ >= 2 * HR, "tile narrower than its halos");
    using A = ArithFor<R, MODE>;

    __shared__ float s_ring[Q][RW];
    __shared__ float s_vz[2][VW];

    const int nx = a.nx, nz = a.nz;
    const int tid = threadIdx.x;
    const int j0 = R + (int)blockIdx.x * TZ;
    const int j = j0 + tid;
    const bool jload = j < nz;       // lane maps to a stored column
    const bool jupd = j < nz - R;    // lane's column is updated
    const int i0 = R + (int)blockIdx.y * a.chunk;
    const int i1 = min(i0 + a.chunk, nx - R);
    if (i0 >= i1) return;

    // P-halo lanes: the first HR lanes load the columns left of the tile, the
    // last HR lanes those right of it (into ring positions ph_pos).
    const bool has_ph = (tid < HR) || (tid >= TZ - HR);
    const int ph_col = (tid < HR) ? (j0 - HR + tid) : (j + HR);
    const int ph_pos = (tid < HR) ? tid : (tid + 2 * HR);
    const bool ph_ok = has_ph && ph_col >= 0 && ph_col < nz;
    // Vz-halo lanes: R columns left of the tile, R-1 right of it.
    const bool has_vh = (tid < R) || (tid >= TZ - (R - 1));
    const int jh = (tid < R) ? (j0 - R + tid) : (j + R - 1);
    const int vh_pos = (tid < R) ? tid : (tid + 2 * R - 1);
    const int vh_rpos = (tid < R) ? (tid + R - 1) : (tid + 3 * R - 2);
    const bool vh_upd = has_vh && jh >= R && jh < nz - R;

    const float az_j = jload ? a.az[j] : 1.0f;
    const float amp = ro(a.amp, a.k);
    float cx[R], cz[R];
#pragma unroll
    for (int m = 0; m < R; ++m) {
        cx[m] = a.cx[m];
        cz[m] = (MODE == MODE_REF_SPONGE) ? a.cx[m] : a.cz[m];
    }
    // C-PML strip columns (-1 outside the strips) and z profiles.
    auto zstrip = [&](int c) -> int {
        if (!CPML) return -1;
        if (c < a.pml_zw) return c;
        if (c >= nz - a.pml_zw) return c - (nz - a.pml_zw) + a.pml_zw;
        return -1;
    };
    auto xstrip = [&](int r) -> int {
        if (!CPML) return -1;
        if (r < a.pml_xw) return r;
        if (r >= nx - a.pml_xw) return r - (nx - a.pml_xw) + a.pml_xw;
        return -1;
    };
    const int zcol = jload ? zstrip(j) : -1;
    const int zcolh = vh_upd ? zstrip(jh) : -1;
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
    const int zw2 = 2 * a.pml_zw;
    // Drain the lane-constant loads now: left pending into the march loop,
    // their first use in every row would wait for the whole prefetch queue.
    __builtin_amdgcn_s_waitcnt(0);

    const int i_first = i0 - 4 * R + 2;  // first marched row
    const int g_lo = max(R, i0 - R), g_hi = min(nx - R, i1 + R - 1);

    // Sponge factor of cell (r, j): the x-edge strips override the z-edge
    // strips in the corners, as the reference's table does (:164-194).
    // ax_r: the row's x-edge factor (2.0 when the row is interior).
    auto absorb_with = [&](float ax_r) -> float {
        if constexpr (CPML) {
            return 1.0f;
        } else {
            return (ax_r <= 1.0f) ? ax_r : az_j;
        }
    };

    // Every lane issues the same loads for every row (clamped, in-bounds
    // addresses; values are masked where they are used), so the count of
    // memory operations in flight is static and the waits stay partial.
    // Clamped rows and columns fall on lines the workgroup reads anyway.
    const int jc = min(j, nz - 1);
    const int phc = ph_ok ? ph_col : jc;
    const int jhc = vh_upd ? jh : jc;
    const int zc0 = max(zcol, 0), zch0 = max(zcolh, 0);
    auto load_row = [&](int i, RowLoads& L) {
        const int f = min(max(i + Q - 1, 0), nx - 1);
        L.pf = a.p_in[(size_t)f * nz + jc];
        L.ph = a.p_in[(size_t)f * nz + phc];
        const int g = min(max(i + R - 1, g_lo), g_hi - 1);
        L.vx = a.vx_in[(size_t)g * nz + jc];
        L.bx = a.bf[(size_t)g * nz + jc];
        const int ic = min(max(i, i0), i1 - 1);
        const size_t row = (size_t)ic * nz;
        L.vz = a.vz_in[row + jc];
        L.bz = a.bf[row + jc];
        L.kf = a.kf[row + jc];
        L.vzh = a.vz_in[row + jhc];
        L.bzh = a.bf[row + jhc];
        if constexpr (CPML) {
            L.psx = a.psi_px_in[(size_t)max(xstrip(g), 0) * nz + jc];
            L.psvx = a.psi_vx_in[(size_t)max(xstrip(ic), 0) * nz + jc];
            L.pspz = a.psi_pz_in[(size_t)ic * zw2 + zc0];
            L.psvz = a.psi_vz_in[(size_t)ic * zw2 + zc0];
            L.pspzh = a.psi_pz_in[(size_t)ic * zw2 + zch0];
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
            // Unconditional, so every marched row issues the same loads.
            load_row(i + PD, ld[(U + PD) % NS]);
            if (i < i_first || i >= i1) return;
            const RowLoads& L = ld[U % NS];

            // ---- 1. front row of P: queue (and LDS ring for output rows) ----
            const int f = i + Q - 1;
            constexpr int sf = (U + Q - 1) % Q;
            {
                float pf = (f >= 0 && f < nx && jload) ? L.pf : 0.f;
                if (f == a.sx && j == a.sz) pf -= amp;
                qp[sf] = pf;
                if (f >= i0 && f < i1) {
                    s_ring[sf][tid + HR] = pf;
                    if (has_ph) {
                        float v = ph_ok ? L.ph : 0.f;
                        if (f == a.sx && ph_col == a.sz) v -= amp;
                        s_ring[sf][ph_pos] = v;
                    }
                }
            }

            // ---- 2. new Vx at row g = i + R - 1 from the P queue ----
            const int g = i + R - 1;
            constexpr int sg = (U + R - 1) % Q;
            // Row-uniform tables, read in uniform control flow (scalar loads).
            const int gt = min(max(g, 0), nx - 1);
            const float ax_g = CPML ? 1.f : ro(a.ax, gt);
            const float pbh_g = CPML ? ro(a.pml_x, 2 * nx + gt) : 0.f;
            const float pah_g = CPML ? ro(a.pml_x, 3 * nx + gt) : 0.f;
            if (g >= i0 - R && g < i1 + R - 1) {
                float vxn = 0.f;
                if (g >= g_lo && g < g_hi && jupd) {
                    const float s = A::stencil(cx, [&](int m) {
                        return qp[(U + R - 1 + m) % Q] - qp[(U + R - m) % Q];
                    });
                    float d = s;
                    if constexpr (CPML) {
                        const int xr = xstrip(g);
                        if (xr >= 0) {
                            const float psi = __builtin_fmaf(pah_g, s, pbh_g * L.psx);
                            d = s + psi;
                            if (g >= i0 && g < i1) a.psi_px_out[(size_t)xr * nz + j] = psi;
                        }
                    }
                    vxn = A::vel(L.vx, L.bx, d, a.dtdx);
                    if (g >= i0 && g < i1) a.vx_out[(size_t)g * nz + j] = vxn * absorb_with(ax_g);
                }
                qvx[sg] = vxn;
            }

            if (i < i0) return;
            // Ring rows written while the queues were filling become visible.
            if (i == i0) __syncthreads();
            const float ax_i = CPML ? 1.f : ro(a.ax, i);
            const float pbi_i = CPML ? ro(a.pml_x, i) : 0.f;
            const float pai_i = CPML ? ro(a.pml_x, nx + i) : 0.f;
            const int rec0 = ro(a.rec_first, i), rec1 = ro(a.rec_first, i + 1);

            // ---- 3. new Vz at row i (own lane and halo lanes) into LDS ----
            constexpr int si = U % Q;
            const int vb = i & 1;
            const size_t rowoff = (size_t)i * nz;
            auto vz_new = [&](int rpos, float vold, float b, int zc, float bh, float ah, float psi_old,
                              bool own) -> float {
                const float s = A::stencil(cz, [&](int m) {
                    return s_ring[si][rpos + m] - s_ring[si][rpos - m + 1];
                });
                float d = s;
                if constexpr (CPML) {
                    if (zc >= 0) {
                        const float psi = __builtin_fmaf(ah, s, bh * psi_old);
                        d = s + psi;
                        if (own) a.psi_pz_out[(size_t)i * zw2 + zc] = psi;
                    }
                }
                return A::vel(vold, b, d, a.dtdz);
            };
            {
                float vzn = 0.f;
                if (jupd) {
                    vzn = vz_new(tid + HR, L.vz, L.bz, zcol, pz_bh, pz_ah, L.pspz, true);
                    a.vz_out[rowoff + j] = vzn * absorb_with(ax_i);
                }
                s_vz[vb][tid + R] = vzn;
                if (has_vh) {
                    float vh = 0.f;
                    if (vh_upd) vh = vz_new(vh_rpos, L.vzh, L.bzh, zcolh, ph_bh, ph_ah, L.pspzh, false);
                    s_vz[vb][vh_pos] = vh;
                }
            }
            __syncthreads();

            // ---- 4. new P at row i ----
            float pn = 0.f;
            if (jupd) {
                const size_t c = rowoff + j;
                float sx = A::stencil(cx, [&](int m) {
                    return qvx[(U + m - 1) % Q] - qvx[(U - m + Q) % Q];
                });
                float sz = A::stencil(cz, [&](int m) {
                    return s_vz[vb][tid + R + m - 1] - s_vz[vb][tid + R - m];
                });
                if constexpr (CPML) {
                    const int xr = xstrip(i);
                    if (xr >= 0) {
                        const float psi = __builtin_fmaf(pai_i, sx, pbi_i * L.psvx);
                        a.psi_vx_out[(size_t)xr * nz + j] = psi;
                        sx += psi;
                    }
                    if (zcol >= 0) {
                        const float psi = __builtin_fmaf(pz_ai, sz, pz_bi * L.psvz);
                        a.psi_vz_out[(size_t)i * zw2 + zcol] = psi;
                        sz += psi;
                    }
                }
                pn = A::pres(qp[si], L.kf, sx, sz, a.inv_dx, a.inv_dz) * absorb_with(ax_i);
                a.p_out[c] = pn;
            }
            // Receivers of row i (uniform loop over its spans).
            for (int r = rec0; r < rec1; ++r) {
                const int* spn = reinterpret_cast<const int*>(a.rec_span) + 4 * r;
                const int jlo = ro(spn, 0), jhi = ro(spn, 1), slot = ro(spn, 2);
                if (jupd && j >= jlo && j < jhi)
                    a.rec_buf[(size_t)a.k * a.nslot + slot + (j - jlo)] = pn;
            }
        });
    }
}

// Prefetch slots per radius: the slot count must divide 2R.
template <int R>
constexpr int default_slots() { return R == 5 ? 5 : 4; }

}  // namespace sfd
