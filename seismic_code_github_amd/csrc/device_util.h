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
