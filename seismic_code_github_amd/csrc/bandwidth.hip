// Streaming-bandwidth probe: the achievable HBM rate that the stencil steps'
// roofline fractions are read against.  Three hand-written kernels stream a
// device buffer in 16-byte vectors, each lane touching U vectors one block
// width apart (all loads issued before any use), workgroups dealt to the XCDs
// in contiguous runs so each XCD streams its own part of the buffer:
//   read   sum every vector (the sum is stored only if it equals a value no
//          finite input sums to, so the loads cannot be dropped)
//   write  store a constant
//   copy   dst = src
// Each kernel exists with plain and with nontemporal loads / stores; the
// probe reports the faster.  No counterpart in the reference (measurement
// only).
#include <algorithm>

#include "common.h"
#include "device_util.h"

namespace sfd {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kBwThreads = 256;
constexpr int kBwUnroll = 8;
constexpr int kBwPerBlock = kBwThreads * kBwUnroll;  // 16-byte vectors per workgroup

template <bool NT>
__device__ __forceinline__ v4f bw_load(const v4f* p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

template <bool NT>
__device__ __forceinline__ void bw_store(v4f v, v4f* p) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <bool NT>
__global__ __launch_bounds__(kBwThreads) void bw_read_kernel(const v4f* __restrict__ src, size_t n,
                                                             float* __restrict__ sink) {
    const size_t base = (size_t)xcd_tile((int)blockIdx.x, (int)gridDim.x) * kBwPerBlock + threadIdx.x;
    v4f v[kBwUnroll];
#pragma unroll
    for (int u = 0; u < kBwUnroll; ++u) {
        const size_t e = base + (size_t)u * kBwThreads;
        v[u] = e < n ? bw_load<NT>(src + e) : v4f{0.f, 0.f, 0.f, 0.f};
    }
    float s = 0.f;
#pragma unroll
    for (int u = 0; u < kBwUnroll; ++u) s += (v[u].x + v[u].y) + (v[u].z + v[u].w);
    if (s == -1.2345e-38f) sink[threadIdx.x] = s;
}

template <bool NT>
__global__ __launch_bounds__(kBwThreads) void bw_write_kernel(v4f* __restrict__ dst, size_t n, float value) {
    const size_t base = (size_t)xcd_tile((int)blockIdx.x, (int)gridDim.x) * kBwPerBlock + threadIdx.x;
    const v4f v = {value, value, value, value};
#pragma unroll
    for (int u = 0; u < kBwUnroll; ++u) {
        const size_t e = base + (size_t)u * kBwThreads;
        if (e < n) bw_store<NT>(v, dst + e);
    }
}

template <bool NT>
__global__ __launch_bounds__(kBwThreads) void bw_copy_kernel(const v4f* __restrict__ src, v4f* __restrict__ dst,
                                                             size_t n) {
    const size_t base = (size_t)xcd_tile((int)blockIdx.x, (int)gridDim.x) * kBwPerBlock + threadIdx.x;
    v4f v[kBwUnroll];
#pragma unroll
    for (int u = 0; u < kBwUnroll; ++u) {
        const size_t e = base + (size_t)u * kBwThreads;
        v[u] = e < n ? bw_load<NT>(src + e) : v4f{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < kBwUnroll; ++u) {
        const size_t e = base + (size_t)u * kBwThreads;
        if (e < n) bw_store<NT>(v[u], dst + e);
    }
}

}  // namespace
}  // namespace sfd

extern "C" {

int sfd_bandwidth_probe(sfd_context* ctx, size_t bytes, int kind, int reps, double* gbps) {
    using namespace sfd;
    if (!ctx || !gbps || kind < SFD_BW_READ || kind > SFD_BW_COPY || reps < 1 || bytes < 16 * (size_t)kBwPerBlock)
        return fail(ctx, SFD_EINVAL, "sfd_bandwidth_probe: bad arguments");
    const size_t n = bytes / 16;
    const size_t blocks = (n + kBwPerBlock - 1) / kBwPerBlock;
    if (blocks > 0x7fffffff) return fail(ctx, SFD_EINVAL, "sfd_bandwidth_probe: buffer too large");
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    DeviceBuffer a, b, sink;
    SFD_HIP(ctx, a.alloc(n * 16));
    if (kind == SFD_BW_COPY) SFD_HIP(ctx, b.alloc(n * 16));
    SFD_HIP(ctx, sink.alloc(kBwThreads * sizeof(float)));
    hipStream_t s = ctx->stream;
    // Initialise the source (also pages every buffer in).
    hipLaunchKernelGGL(bw_write_kernel<false>, dim3((unsigned)blocks), dim3(kBwThreads), 0, s, a.as<v4f>(), n, 1.0f);
    if (kind == SFD_BW_COPY)
        hipLaunchKernelGGL(bw_write_kernel<false>, dim3((unsigned)blocks), dim3(kBwThreads), 0, s, b.as<v4f>(), n, 0.f);
    SFD_HIP(ctx, hipGetLastError());
    double best = 0.0;
    for (int nt = 0; nt < 2; ++nt) {
        for (int r = 0; r <= reps; ++r) {  // the first pass warms up
            SFD_HIP(ctx, hipEventRecord(ctx->ev0, s));
            const dim3 g((unsigned)blocks), blk(kBwThreads);
            if (kind == SFD_BW_READ) {
                if (nt) hipLaunchKernelGGL(bw_read_kernel<true>, g, blk, 0, s, a.as<const v4f>(), n, sink.as<float>());
                else hipLaunchKernelGGL(bw_read_kernel<false>, g, blk, 0, s, a.as<const v4f>(), n, sink.as<float>());
            } else if (kind == SFD_BW_WRITE) {
                if (nt) hipLaunchKernelGGL(bw_write_kernel<true>, g, blk, 0, s, a.as<v4f>(), n, (float)r);
                else hipLaunchKernelGGL(bw_write_kernel<false>, g, blk, 0, s, a.as<v4f>(), n, (float)r);
            } else {
                if (nt) hipLaunchKernelGGL(bw_copy_kernel<true>, g, blk, 0, s, a.as<const v4f>(), b.as<v4f>(), n);
                else hipLaunchKernelGGL(bw_copy_kernel<false>, g, blk, 0, s, a.as<const v4f>(), b.as<v4f>(), n);
            }
            SFD_HIP(ctx, hipGetLastError());
            SFD_HIP(ctx, hipEventRecord(ctx->ev1, s));
            SFD_HIP(ctx, hipEventSynchronize(ctx->ev1));
            float ms = 0.f;
            SFD_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            const double moved = (double)n * 16.0 * (kind == SFD_BW_COPY ? 2.0 : 1.0);
            if (r > 0 && ms > 0.f) best = std::max(best, moved / (ms * 1e-3) / 1e9);
        }
    }
    *gbps = best;
    return SFD_OK;
}

}  // extern "C"
