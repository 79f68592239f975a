// Context and shared helpers of libseisfd.
#include <algorithm>
#include <cmath>
#include <new>
#include <tuple>

#include "common.h"

namespace sfd {

int fail(sfd_context* ctx, int code, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

int hip_check(sfd_context* ctx, hipError_t e, const char* what) {
    const int code = (e == hipErrorOutOfMemory) ? SFD_ENOMEM : SFD_EDEVICE;
    return fail(ctx, code, "%s: %s", what, hipGetErrorString(e));
}

__global__ void nonfinite_kernel(const float* __restrict__ a, size_t n, int* __restrict__ flag) {
    int bad = 0;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x)
        bad |= !isfinite(a[c]);
    if (bad) atomicOr(flag, 1);
}

__global__ void clear_flag_kernel(int* flag) { *flag = 0; }

int check_finite(sfd_context* ctx, const float* a, size_t n, const char* what) {
    if (!a || n == 0) return SFD_OK;
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(clear_flag_kernel, dim3(1), dim3(1), 0, s, ctx->d_flag);
    const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(nonfinite_kernel, dim3(nb), dim3(256), 0, s, a, n, ctx->d_flag);
    SFD_HIP(ctx, hipGetLastError());
    int h = 0;
    SFD_HIP(ctx, hipMemcpyAsync(&h, ctx->d_flag, sizeof h, hipMemcpyDeviceToHost, s));
    SFD_HIP(ctx, hipStreamSynchronize(s));
    if (h) return fail(ctx, SFD_ENAN, "%s holds non-finite values: the run diverged or an input is not finite", what);
    return SFD_OK;
}

int log2_round(double n) { return (int)std::lround(std::log2(std::max(n, 1.0))); }

bool default_coefs(int radius, float* c) {
    switch (radius) {
    case 8: {
        // C1..C8, Acoustic_1order_FD.cpp:38-45.
        const float t[8] = {1.2340911f, -1.0664985e-1f, 2.3036367e-2f, -5.3423856e-3f,
                            1.0772712e-3f, -1.6641888e-4f, 1.7021711e-5f, -8.5234642e-7f};
        for (int m = 0; m < 8; ++m) c[m] = t[m];
        return true;
    }
    case 4: {
        // a80..a83, Iso_multishot_fd_module_zy_m2.f90:17-20.
        const float t[4] = {1225.0f / 1024.0f, -245.0f / 3072.0f, 49.0f / 5120.0f,
                            -5.0f / 7168.0f};
        for (int m = 0; m < 4; ++m) c[m] = t[m];
        return true;
    }
    case 5: {
        // a100..a104, Iso_multishot_fd_module_zy_m2.f90:22-26.
        const float t[5] = {19845.0f / 16384.0f, -735.0f / 8192.0f, 567.0f / 40960.0f,
                            -405.0f / 229376.0f, 35.0f / 294912.0f};
        for (int m = 0; m < 5; ++m) c[m] = t[m];
        return true;
    }
    default:
        return false;
    }
}


void receiver_spans(int nx, int pad, int nrec, const int* rx, const int* rz, std::vector<int>& slot_of,
                    std::vector<int4>& row, std::vector<int4>& more, int* nslot) {
    slot_of.assign(nrec, 0);
    row.assign((size_t)nx + 2 * pad, make_int4(0, 0, 0, -1));
    more.clear();
    std::vector<std::tuple<int, int, int>> cells(nrec);
    for (int r = 0; r < nrec; ++r) cells[r] = {rx[r], rz[r], r};
    std::sort(cells.begin(), cells.end());
    std::vector<int4> spans;  // (jlo, jhi, slot, row)
    int slot = -1, prev_x = -1, prev_z = -1;
    for (auto& [x, z, r] : cells) {
        if (x != prev_x || z != prev_z) {
            ++slot;
            if (!spans.empty() && x == prev_x && z == prev_z + 1) {
                spans.back().y = z + 1;
            } else {
                spans.push_back(make_int4(z, z + 1, slot, x));
            }
            prev_x = x;
            prev_z = z;
        }
        slot_of[r] = slot;
    }
    *nslot = slot + 1;
    // tail[row]: -2 = no span yet, -1 = head only, else last index in `more`.
    std::vector<int> tail(nx, -2);
    for (const int4& s4 : spans) {
        const int x = s4.w;
        const int4 sp = make_int4(s4.x, s4.y, s4.z, -1);
        if (tail[x] == -2) {
            row[x + pad] = sp;
            tail[x] = -1;
        } else {
            more.push_back(sp);
            const int idx = (int)more.size() - 1;
            if (tail[x] == -1) row[x + pad].w = idx;
            else more[tail[x]].w = idx;
            tail[x] = idx;
        }
    }
}

}  // namespace sfd

extern "C" {

int sfd_abi_version(void) { return SEISFD_ABI_VERSION; }

int sfd_context_create(int device, sfd_context** out) {
    if (!out) return SFD_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return SFD_EDEVICE;
    if (device < 0 || device >= n) return SFD_EINVAL;
    sfd_context* ctx = new (std::nothrow) sfd_context;
    if (!ctx) return SFD_ENOMEM;
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        hipMalloc(&ctx->d_flag, 256) != hipSuccess ||
        hipGetDeviceProperties(&prop, device) != hipSuccess) {
        sfd_context_destroy(ctx);
        return SFD_EDEVICE;
    }
    ctx->arch = prop.gcnArchName;
    if (sfd::noalloc_init(ctx) != SFD_OK) {
        sfd_context_destroy(ctx);
        return SFD_EDEVICE;
    }
    *out = ctx;
    return SFD_OK;
}

int sfd_context_destroy(sfd_context* ctx) {
    if (!ctx) return SFD_EINVAL;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->d_flag) (void)hipFree(ctx->d_flag);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return SFD_OK;
}

int sfd_device_alloc(sfd_context* ctx, size_t bytes, void** ptr) {
    if (!ctx || !ptr) return SFD_EINVAL;
    *ptr = nullptr;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    if (bytes == 0) return SFD_OK;
    const hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return e == hipErrorOutOfMemory ? sfd::fail(ctx, SFD_ENOMEM, "device allocation of %zu bytes failed", bytes)
                                        : sfd::hip_check(ctx, e, "hipMalloc");
    }
    return SFD_OK;
}

int sfd_device_copy(sfd_context* ctx, void* dst, const void* src, size_t bytes, int kind) {
    if (!ctx || (bytes && (!dst || !src))) return SFD_EINVAL;
    if (kind != 0 && kind != 1) return sfd::fail(ctx, SFD_EINVAL, "copy kind must be 0 (to device) or 1 (to host)");
    if (bytes == 0) return SFD_OK;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    SFD_HIP(ctx, hipMemcpyAsync(dst, src, bytes, kind == 0 ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost,
                                ctx->stream));
    SFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SFD_OK;
}

int sfd_device_free(sfd_context* ctx, void* ptr) {
    if (!ctx) return SFD_EINVAL;
    if (!ptr) return SFD_OK;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    SFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    SFD_HIP(ctx, hipFree(ptr));
    return SFD_OK;
}

const char* sfd_last_error(const sfd_context* ctx) {
    return ctx ? ctx->err.c_str() : "null context";
}

int sfd_window_model(const float* model, int mnx, int mnz, int x0, int nxe, int pad_xlo, int pad_xhi, int pad_zlo,
                     int pad_zhi, float* out) {
    if (!model || !out || mnx <= 0 || mnz <= 0 || nxe <= 0 || pad_xlo < 0 || pad_xhi < 0 || pad_zlo < 0 ||
        pad_zhi < 0)
        return SFD_EINVAL;
    const int nx = nxe + pad_xlo + pad_xhi, nz = mnz + pad_zlo + pad_zhi;
    for (int x = 0; x < nx; ++x) {
        // Window column (edge-replicated into the x padding), then model
        // column clamped to the model.
        const int w = std::min(std::max(x - pad_xlo, 0), nxe - 1);
        const int mx = std::min(std::max(x0 + w, 0), mnx - 1);
        const float* src = model + (size_t)mx * mnz;
        float* dst = out + (size_t)x * nz;
        for (int z = 0; z < pad_zlo; ++z) dst[z] = src[0];
        for (int z = 0; z < mnz; ++z) dst[pad_zlo + z] = src[z];
        for (int z = 0; z < pad_zhi; ++z) dst[pad_zlo + mnz + z] = src[mnz - 1];
    }
    return SFD_OK;
}

double sfd_acoustic_cfl_limit(int radius, const float* coef) {
    float c[8];
    if (radius < 1 || radius > 8) return 0.0;
    if (coef) {
        for (int m = 0; m < radius; ++m) c[m] = coef[m];
    } else if (!sfd::default_coefs(radius, c)) {
        return 0.0;
    }
    double s = 0.0;
    for (int m = 0; m < radius; ++m) s += std::fabs((double)c[m]);
    return s > 0.0 ? 1.0 / s : 0.0;
}

}  // extern "C"
