// Device helpers shared by the step kernels: compile-time loops, scalar
// (constant address space) loads, raw buffer loads/stores with per-lane and
// per-row offsets, and the XCD-aware workgroup order.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

namespace sfd {

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

// A wave-uniform float moved to a scalar register (bit pattern preserved).
__device__ __forceinline__ float uniform_f(float x) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

// A wave-uniform float the compiler is told lives in a vector register: used
// for stencil coefficients, which would otherwise take scalar registers the
// march needs for its row offsets (and be reloaded from the kernel arguments
// with scalar loads whose waits also drain the LDS queue).
__device__ __forceinline__ float in_vgpr(float x) {
    asm volatile("" : "+v"(x));
    return x;
}

__device__ __forceinline__ int floor_div(int a, int b) {
    return (a >= 0) ? a / b : -((-a + b - 1) / b);
}

// Raw buffer view of a device array: loads and stores take a per-lane byte
// offset (VGPR) and a per-row byte offset (SGPR), so one marched row costs a
// scalar add per field instead of 64-bit address arithmetic per lane.  Lane
// offsets at or beyond the array size (kOOB) read 0 and drop stores, which
// masks lanes without branches.
constexpr int kOOB = (int)0x80000000u;

struct Buf {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ Buf(const void* p, uint32_t bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000)) {}
    __device__ __forceinline__ float ld(int voff, uint32_t soff) const {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, (int)soff, 0));
    }
    __device__ __forceinline__ void st(float v, int voff, uint32_t soff) const {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, voff, (int)soff, 0);
    }
};

// Tiles of the two-launch steps are 64 columns wide, one per lane.
constexpr int kTileCols = 64;

// Rows [i0 - HR, i0 + TX + HR) x columns [j0 - HC, j0 + 64 + HC) of a field
// (rows and columns counted from f's origin, rows rb bytes apart; the first
// staged row and column are >= 0) into lds[row][col], W = 64 + 2 HC
// floats a row.  Each wave copies whole rows of the 64 tile columns (lane =
// column, the row's byte offset a scalar, so no per-lane address arithmetic),
// then the workgroup copies the 2 HC edge columns of every row.  All loads
// are issued before the first LDS store.  With SRC, the source amplitude is
// taken off cell (sx, sz) on the way (same float subtraction as the fused
// kernel's).  wv is the wave's index, wave-uniform.
template <int TX, int NW, int HR, int HC, bool SRC>
__device__ __forceinline__ void tiled_stage(float* lds, const Buf& f, int rb, int i0, int j0, int wv, int sx, int sz,
                                            float amp) {
    constexpr int W = kTileCols + 2 * HC, Hh = TX + 2 * HR, T = 64 * NW;
    constexpr int RI = (Hh + NW - 1) / NW;           // rows per wave
    constexpr int EW = 2 * HC, E = EW * Hh;          // edge cells
    constexpr int EWd = EW > 0 ? EW : 1;
    constexpr int EI = E > 0 ? (E + T - 1) / T : 0;  // edge cells per thread
    const int lane = threadIdx.x & 63;
    const uint32_t srow = (uint32_t)(i0 - HR) * (uint32_t)rb;  // byte offset of the first staged row
    float v[RI];
#pragma unroll
    for (int q = 0; q < RI; ++q) {
        const int r = wv + q * NW;
        v[q] = (RI * NW == Hh || r < Hh) ? f.ld(lane * 4, srow + (uint32_t)r * rb + (uint32_t)j0 * 4u) : 0.f;
    }
    float ve[EI > 0 ? EI : 1];
    int le[EI > 0 ? EI : 1];
#pragma unroll
    for (int q = 0; q < EI; ++q) {
        const int e = threadIdx.x + q * T;
        const int r = e / EWd, c = e - r * EWd;
        const int col = c < HC ? c : c + kTileCols;  // column in the staged row
        le[q] = (e < E) ? r * W + col : -1;
        ve[q] = (e < E) ? f.ld(r * rb + col * 4, srow + (uint32_t)(j0 - HC) * 4u) : 0.f;
    }
    if constexpr (SRC) {
        // Wave-uniform tests first: only the wave (row) or edge holding the source does vector work.
        const int rs = sx - (i0 - HR), cs = sz - (j0 - HC);
        if (rs >= 0 && rs < Hh && cs >= 0 && cs < W) {
            if (cs >= HC && cs < HC + kTileCols) {
#pragma unroll
                for (int q = 0; q < RI; ++q)
                    if (wv + q * NW == rs && lane == cs - HC) v[q] -= amp;
            } else {
#pragma unroll
                for (int q = 0; q < EI; ++q)
                    if (le[q] == rs * W + cs) ve[q] -= amp;
            }
        }
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

// Raw buffer view of a field from its row-0 pointer (every offset the tiled
// kernels form from it is non-negative and below 2 GiB).
__device__ __forceinline__ Buf field_buf(const float* p) { return Buf(p, 0x7fffffffu); }

// Workgroups are dealt round-robin to the 8 XCDs (each with its own L2).
// Remap so each XCD gets a contiguous run of tiles in (x chunk, z tile)
// order, z fastest: an XCD holds a band of x chunks across all z tiles, so
// both halo directions are re-read from its own L2, and the boundary tiles
// spread over all XCDs.
__device__ __forceinline__ int xcd_tile(int L, int n) {
    const int xcd = L % 8, idx = L / 8, q = n / 8, rem = n % 8;
    return xcd * q + min(xcd, rem) + idx;
}

}  // namespace sfd
