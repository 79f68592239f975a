// Shared host-side helpers of libseisfd: context, error reporting, device
// buffers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "seisfd.h"

namespace sfd {
// A launch shape chosen by the autotuner for one grid class: kind 0 = the
// tiled step (a = tile rows, b = waves), kind 1 = the fused march kernel (a =
// waves, b = rows per chunk); us = device time per step when it was measured.
struct TunedShape {
    int kind = 0, a = 0, b = 0;
    float us = 0.f;
};
}  // namespace sfd

struct sfd_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int* d_flag = nullptr;  // device word of the non-finite watchdog
    std::string err;
    // Autotuner cache: grid class -> fastest measured shape on this device
    // (saved to and loaded from JSON by sfd_context_save/load_tuning).
    std::string arch;
    std::map<std::string, sfd::TunedShape> tuning;
};

namespace sfd {

// Records a formatted error on the context and returns `code`.
int fail(sfd_context* ctx, int code, const char* fmt, ...);

// Returns SFD_EDEVICE with the HIP error text when `e` is an error.
int hip_check(sfd_context* ctx, hipError_t e, const char* what);

#define SFD_HIP(ctx, call)                                         \
    do {                                                           \
        hipError_t e_ = (call);                                    \
        if (e_ != hipSuccess) return ::sfd::hip_check(ctx, e_, #call); \
    } while (0)

// Owning device buffer (hipMalloc / hipFree).
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t n) {
        release();
        if (n == 0) return hipSuccess;
        hipError_t e = hipMalloc(&ptr, n);
        if (e == hipSuccess) bytes = n;
        else ptr = nullptr;
        return e;
    }
    template <class T>
    T* as() const { return static_cast<T*>(ptr); }
};

// Non-owning view of a region inside a larger device allocation.
struct Field {
    void* ptr = nullptr;
    size_t bytes = 0;
    template <class T>
    T* as() const { return static_cast<T*>(ptr); }
};

// Carves consecutive 256-byte aligned fields out of one allocation: first a
// sizing pass (base == nullptr), then the real one.
struct Carver {
    char* base = nullptr;
    size_t used = 0;
    Field take(size_t bytes) {
        used = (used + 255) / 256 * 256;
        Field f;
        f.ptr = base ? base + used : nullptr;
        f.bytes = bytes;
        used += bytes;
        return f;
    }
};

// Allocates `b` for n elements and copies them from the host (synchronous).
template <class T>
hipError_t upload(DeviceBuffer& b, const T* h, size_t n, hipStream_t s) {
    hipError_t e = b.alloc(n * sizeof(T));
    if (e != hipSuccess || n == 0) return e;
    e = hipMemcpyAsync(b.ptr, h, n * sizeof(T), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

// Autotuner timing: device time per step of n steps from a reset state, the
// better of two timed runs after a two-step warm-up.  run(-1) resets the
// state (outside the timed region), run(m) enqueues m steps.
template <class Run>
int time_steps(sfd_context* ctx, int n, Run run, float* us) {
    float best = 3.4e38f;
    for (int rep = 0; rep < 3; ++rep) {
        const int m = rep == 0 ? (n < 2 ? n : 2) : n;
        int rc = run(-1);
        if (rc) return rc;
        SFD_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        rc = run(m);
        if (rc) return rc;
        SFD_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        SFD_HIP(ctx, hipEventSynchronize(ctx->ev1));
        float ms = 0.f;
        SFD_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        if (rep > 0 && ms * 1000.f / m < best) best = ms * 1000.f / m;
    }
    *us = best;
    return SFD_OK;
}

// Non-finite watchdog: SFD_ENAN (message naming `what`) when any of the n
// device floats at `a` is Inf or NaN, SFD_OK otherwise.  Synchronises the
// context stream; allocates nothing.
int check_finite(sfd_context* ctx, const float* a, size_t n, const char* what);

// Looks up the device address of the allocation-free entry point's scratch
// array on ctx's device (once per device; the device must be current).
int noalloc_init(sfd_context* ctx);

// Nearest power of 2 of n, as an exponent (autotuner grid classes).
int log2_round(double n);

// Staggered Taylor coefficients used when the caller passes none: the
// reference's own tables (Acoustic_1order_FD.cpp:38-45 for R = 8;
// Iso_multishot_fd_module_zy_m2.f90:17-26 for R = 4 and 5).
bool default_coefs(int radius, float* c);

// Receiver table of the step kernels.  One slot per distinct receiver cell
// (slot_of[r] for receiver r); runs of consecutive z cells in a row become
// spans (jlo, jhi, slot).  row[x + pad] holds row x's first span, further
// spans of the row go to `more`, chained by their 4th field (-1 ends).
void receiver_spans(int nx, int pad, int nrec, const int* rx, const int* rz, std::vector<int>& slot_of,
                    std::vector<int4>& row, std::vector<int4>& more, int* nslot);

}  // namespace sfd
