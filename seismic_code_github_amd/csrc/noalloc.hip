// Allocation-free acoustic shot: sfd_acoustic_forward_noalloc, the variant of
// sfd_acoustic_forward for callers that cannot allocate device memory (code
// under stream capture, serving loops, harnesses that forbid hipMalloc).  The
// caller passes inputs and outputs as device arrays; everything else the
// step kernels address -- wavefield, folded model, boundary tables, source
// amplitudes, checks -- lives in `scratch`, a file-scope device array of
// kScratchFloats floats that each GPU holds until the process ends.  Calls on
// one GPU share that array: the set-up kernels clear what must start at
// zero, and calls that overlap on one GPU race.  The time loop is the
// two-launch tiled step of acoustic_tiled.h, so gathers are bitwise those of
// sfd_acoustic_forward (same kernels, same folded coefficients).
#include <algorithm>
#include <cmath>
#include <mutex>

#include "acoustic_plan.h"
#include "acoustic_tiled.h"

namespace sfd {

constexpr size_t kScratchFloats = (size_t)8 << 20;  // 32 MiB per GPU

static __device__ __align__(256) float scratch[kScratchFloats];

namespace {

constexpr int kMaxDevices = 64;

// Device address of `scratch` on each GPU (the host-side name is only a
// stand-in): looked up once per device, with that device current.
std::mutex g_mu;
float* g_scratch[kMaxDevices] = {};

// Flags of the input check (scratch words).
enum : int { kFlagBadVp = 1, kFlagBadDen = 2, kFlagBadRec = 4 };

__global__ void na_fill_kernel(float* __restrict__ f, size_t n, float v) {
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x)
        f[c] = v;
}

// Receiver span rows with no span (jlo = jhi = 0, no further spans): this
// entry point gathers with its own kernel.
__global__ void na_no_spans_kernel(int4* __restrict__ row, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) row[t] = make_int4(0, 0, 0, -1);
}

// chk[0] = max vp (as int bits, vp > 0), chk[1] = flags.
__global__ void na_check_kernel(const float* __restrict__ vp, const float* __restrict__ den, size_t n,
                                const int* __restrict__ rx, const int* __restrict__ rz, int nrec, int nx, int nz,
                                int* __restrict__ chk) {
    int vmax = 0, flags = 0;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        const float v = vp[c];
        if (!(v > 0.f) || !isfinite(v)) flags |= kFlagBadVp;
        else vmax = max(vmax, __float_as_int(v));
        if (den) {
            const float d = den[c];
            if (!(d > 0.f) || !isfinite(d)) flags |= kFlagBadDen;
        }
    }
    for (int r = (int)(blockIdx.x * blockDim.x + threadIdx.x); r < nrec; r += gridDim.x * blockDim.x)
        if (rx[r] < 0 || rx[r] >= nx || rz[r] < 0 || rz[r] >= nz) flags |= kFlagBadRec;
    if (vmax) atomicMax(&chk[0], vmax);
    if (flags) atomicOr(&chk[1], flags);
}

// Folded model of the fast arithmetic on the padded layout: kf = rho vp^2 dt
// and bf = dt / rho on the updated cells [R, n - R)^2, zero elsewhere (the
// plan's own set-up, evaluated the same way).
__global__ void na_model_kernel(const float* __restrict__ vp, const float* __restrict__ den, int nx, int nz, int R,
                                float dt, float* __restrict__ kf, float* __restrict__ bf) {
#pragma clang fp contract(off)
    const size_t n = (size_t)nx * nz;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(c / nz), j = (int)(c - (size_t)i * nz);
        const bool upd = i >= R && i < nx - R && j >= R && j < nz - R;
        const float d = den ? den[c] : 2400.0f;
        kf[c] = upd ? d * vp[c] * vp[c] * dt : 0.f;
        if (bf) bf[c] = upd ? (float)((double)dt / (double)d) : 0.f;
    }
}

__device__ float na_sponge_factor(int d, int L) {
#pragma clang fp contract(off)
    return (float)(-0.38 * d * d / (1.0 * L * L) + 1);
}

// Sponge rows / columns (the plan's sponge_profiles), padded row table.
__global__ void na_sponge_kernel(float* __restrict__ ax, int nx, int pad, float* __restrict__ az, int nz, int L) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nx + 2 * pad) {
        const int i = t - pad;
        float f = 2.0f;
        if (L > 0 && i >= 0 && i < L) f = na_sponge_factor(L - i - 1, L);
        else if (L > 0 && i >= nx - L && i < nx) f = na_sponge_factor(i - nx + L, L);
        ax[t] = f;
    }
    if (t < nz) {
        float f = 1.0f;
        if (L > 0 && t < L) f = na_sponge_factor(L - t - 1, L);
        else if (L > 0 && t >= nz - L) f = na_sponge_factor(t - nz + L, L);
        az[t] = f;
    }
}

// C-PML profile of one axis (the plan's cpml_profile formula in double).
__global__ void na_cpml_kernel(float* __restrict__ prof, int n, int w, double h, double dt, double R, double vmax,
                               double f0) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * n) return;
    const int half = t / n, i = t - half * n;
    const double x = i + 0.5 * half;
    const double lo = w, hi = n - 1 - w;
    double depth = 0.0;
    if (x < lo) depth = lo - x;
    else if (x > hi) depth = x - hi;
    double b = 1.0, a = 0.0;
    if (w > 0 && depth > 0.0) {
        const double xi = fmin(depth / w, 1.0);
        const double d = -3.0 * vmax * log(R) / (2.0 * w * h) * xi * xi;
        const double alpha = M_PI * f0 * (1.0 - xi);
        b = exp(-(d + alpha) * dt);
        a = d * (b - 1.0) / (d + alpha);
    }
    prof[(size_t)(2 * half) * n + i] = (float)b;
    prof[(size_t)(2 * half + 1) * n + i] = (float)a;
}

// Source amplitude per step, dt * rho * vp * wavelet at the source cell.
__global__ void na_amp_kernel(const float* __restrict__ vp, const float* __restrict__ den, size_t s,
                              const float* __restrict__ wavelet, int nt, float dt, float* __restrict__ amp) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nt) return;
    const float d = den ? den[s] : 2400.0f;
    amp[k] = dt * d * vp[s] * wavelet[k];
}

// Receivers of step k: the absorbed new P at each receiver cell, written
// into the caller's trace-major gather.
__global__ void na_gather_kernel(const float* __restrict__ p, int nz, const int* __restrict__ rx,
                                 const int* __restrict__ rz, int nrec, float* __restrict__ record, int nt, int k) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrec) record[(size_t)r * nt + k] = p[(size_t)rx[r] * nz + rz[r]];
}

int launch_fill(hipStream_t s, float* f, size_t n, float v) {
    if (n == 0) return SFD_OK;
    const int nb = (int)std::min<size_t>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(na_fill_kernel, dim3(nb), dim3(256), 0, s, f, n, v);
    return hipGetLastError() == hipSuccess ? SFD_OK : SFD_EDEVICE;
}

}  // namespace

// Looks up this device's `scratch` once (the device must be current).
int noalloc_init(sfd_context* ctx) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return fail(ctx, SFD_EINVAL, "device ordinal too large");
    if (g_scratch[ctx->device]) return SFD_OK;
    void* ptr = nullptr;
    SFD_HIP(ctx, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(ctx, SFD_EDEVICE, "no device address for the scratch array");
    g_scratch[ctx->device] = static_cast<float*>(ptr);
    return SFD_OK;
}

}  // namespace sfd

extern "C" {

size_t sfd_noalloc_scratch_bytes(void) { return sfd::kScratchFloats * sizeof(float); }

int sfd_acoustic_forward_noalloc(sfd_context* ctx, const sfd_acoustic_params* p, const float* d_vp,
                                 const float* d_den, const float* d_wavelet, const int* d_rx, const int* d_rz,
                                 float* d_record) {
    using namespace sfd;
    if (!ctx) return SFD_EINVAL;
    if (!p || !d_vp || !d_wavelet) return fail(ctx, SFD_EINVAL, "null params, vp or wavelet");
    // Host-side checks of sfd_acoustic_forward, minus the host arrays this
    // entry point does not take (receivers come as device arrays).
    sfd_acoustic_params q = *p;
    q.nrec = 0;
    int rc = validate(ctx, &q, d_vp, d_wavelet);
    if (rc) return rc;
    if (p->arith != SFD_ARITH_FAST) return fail(ctx, SFD_ENOTSUP, "the allocation-free shot runs fast arithmetic only");
    if (p->nsnap != 0) return fail(ctx, SFD_ENOTSUP, "the allocation-free shot records gathers only (nsnap = 0)");
    if (p->nrec < 0 || (p->nrec > 0 && (!d_rx || !d_rz || !d_record)))
        return fail(ctx, SFD_EINVAL, "bad receiver list or record");
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    rc = noalloc_init(ctx);
    if (rc) return rc;
    float* const base = g_scratch[ctx->device];
    const int nx = p->nx, nz = p->nz, R = p->radius, nt = p->nt;
    const int mode = p->boundary == SFD_BOUNDARY_CPML ? MODE_FAST_CPML : MODE_FAST_SPONGE;
    const int pad = pad_rows(R);
    const bool cpml = mode == MODE_FAST_CPML;
    const int pw = cpml ? p->width + 1 : 0;
    // Layout in floats, 64-float (256 B) aligned pieces.
    const size_t npad = (size_t)(nx + 2 * pad) * nz;
    size_t used = 0;
    auto take = [&](size_t n) {
        used = (used + 63) / 64 * 64;
        const size_t o = used;
        used += n;
        return o;
    };
    const size_t o_p = take(npad), o_vx = take(npad), o_vz = take(npad), o_kf = take(npad);
    const size_t o_bf = d_den ? take(npad) : 0;
    const size_t o_ppx = take((size_t)2 * pw * nz), o_pvx = take((size_t)2 * pw * nz);
    const size_t o_ppz = take((size_t)2 * pw * nx), o_pvz = take((size_t)2 * pw * nx);
    const size_t o_ax = take((size_t)nx + 2 * pad), o_az = take(nz);
    const size_t o_pmx = take(cpml ? 4 * (size_t)nx : 0), o_pmz = take(cpml ? 4 * (size_t)nz : 0);
    const size_t o_row = take(4 * ((size_t)nx + 2 * pad)), o_amp = take(std::max(nt, 1)), o_chk = take(2);
    if (used > kScratchFloats)
        return fail(ctx, SFD_EINVAL,
                    "grid %dx%d needs %zu bytes of scratch; the allocation-free shot has %zu (use sfd_acoustic_forward)",
                    nx, nz, used * 4, kScratchFloats * 4);
    hipStream_t s = ctx->stream;
    const size_t n = (size_t)nx * nz;
    // Inputs: vp > 0 and finite, den likewise, receivers on the grid; vmax
    // for the stability bound.
    int* chk = reinterpret_cast<int*>(base + o_chk);
    rc = launch_fill(s, base + o_chk, 2, 0.f);
    if (rc) return fail(ctx, rc, "scratch fill failed");
    hipLaunchKernelGGL(na_check_kernel, dim3(256), dim3(256), 0, s, d_vp, d_den, n, d_rx, d_rz, p->nrec, nx, nz, chk);
    SFD_HIP(ctx, hipGetLastError());
    int hchk[2] = {0, 0};
    SFD_HIP(ctx, hipMemcpyAsync(hchk, chk, sizeof hchk, hipMemcpyDeviceToHost, s));
    SFD_HIP(ctx, hipStreamSynchronize(s));
    if (hchk[1] & kFlagBadVp) return fail(ctx, SFD_EINVAL, "vp must be finite and > 0");
    if (hchk[1] & kFlagBadDen) return fail(ctx, SFD_EINVAL, "den must be finite and > 0");
    if (hchk[1] & kFlagBadRec) return fail(ctx, SFD_EINVAL, "a receiver lies outside the grid");
    float coef[8] = {};
    if (p->coef) {
        for (int m = 0; m < R; ++m) coef[m] = p->coef[m];
    } else {
        default_coefs(R, coef);
    }
    const float vmax = __builtin_bit_cast(float, hchk[0]);
    const double courant = vmax * (double)p->dt * std::sqrt(1.0 / ((double)p->dx * p->dx) + 1.0 / ((double)p->dz * p->dz));
    const double limit = sfd_acoustic_cfl_limit(R, coef);
    if (courant > limit)
        return fail(ctx, SFD_ECFL, "unstable: vmax*dt*sqrt(1/dx^2+1/dz^2) = %.4f exceeds %.4f", courant, limit);
    // Set-up: zero wavefield and memory fields, folded model, tables.
    for (size_t o : {o_p, o_vx, o_vz, o_kf}) {
        rc = launch_fill(s, base + o, npad, 0.f);
        if (rc) return fail(ctx, rc, "scratch fill failed");
    }
    if (d_den) {
        rc = launch_fill(s, base + o_bf, npad, 0.f);
        if (rc) return fail(ctx, rc, "scratch fill failed");
    }
    if (cpml) {
        for (size_t o : {o_ppx, o_pvx}) rc |= launch_fill(s, base + o, (size_t)2 * pw * nz, 0.f);
        for (size_t o : {o_ppz, o_pvz}) rc |= launch_fill(s, base + o, (size_t)2 * pw * nx, 0.f);
    }
    if (rc) return fail(ctx, SFD_EDEVICE, "scratch fill failed");
    hipLaunchKernelGGL(na_no_spans_kernel, dim3((nx + 2 * pad + 255) / 256), dim3(256), 0, s,
                       reinterpret_cast<int4*>(base + o_row), nx + 2 * pad);
    float* kf = base + o_kf + (size_t)pad * nz;
    float* bf = d_den ? base + o_bf + (size_t)pad * nz : nullptr;
    hipLaunchKernelGGL(na_model_kernel, dim3(1024), dim3(256), 0, s, d_vp, d_den, nx, nz, R, p->dt, kf, bf);
    const int L = cpml ? 0 : p->width;
    const int nmax = std::max(nx + 2 * pad, nz);
    hipLaunchKernelGGL(na_sponge_kernel, dim3((nmax + 255) / 256), dim3(256), 0, s, base + o_ax, nx, pad, base + o_az,
                       nz, L);
    if (cpml) {
        hipLaunchKernelGGL(na_cpml_kernel, dim3((2 * nx + 255) / 256), dim3(256), 0, s, base + o_pmx, nx, p->width,
                           (double)p->dx, (double)p->dt, (double)p->pml_r, (double)vmax, (double)p->pml_f0);
        hipLaunchKernelGGL(na_cpml_kernel, dim3((2 * nz + 255) / 256), dim3(256), 0, s, base + o_pmz, nz, p->width,
                           (double)p->dz, (double)p->dt, (double)p->pml_r, (double)vmax, (double)p->pml_f0);
    }
    if (nt > 0)
        hipLaunchKernelGGL(na_amp_kernel, dim3((nt + 255) / 256), dim3(256), 0, s, d_vp, d_den,
                           (size_t)p->sx * nz + p->sz, d_wavelet, nt, p->dt, base + o_amp);
    SFD_HIP(ctx, hipGetLastError());
    // Step arguments: the plan's, pointing into scratch; in place.
    AcousticStepArgs a{};
    a.arena = base;
    a.arena_bytes = (uint32_t)(used * 4);
    a.nx = nx;
    a.nz = nz;
    a.p_in = a.p_out = base + o_p + (size_t)pad * nz;
    a.vx_in = a.vx_out = base + o_vx + (size_t)pad * nz;
    a.vz_in = a.vz_out = base + o_vz + (size_t)pad * nz;
    a.kf = kf;
    a.bf = bf;
    a.bf_const = d_den ? 0.f : (float)((double)p->dt / 2400.0);
    a.ax = base + o_ax + pad;
    a.az = base + o_az;
    a.pml_x = cpml ? base + o_pmx : nullptr;
    a.pml_z = cpml ? base + o_pmz : nullptr;
    a.pml_xw = a.pml_zw = pw;
    a.sponge_w = cpml ? 0 : p->width;
    a.psi_px_in = a.psi_px_out = base + o_ppx;
    a.psi_vx_in = a.psi_vx_out = base + o_pvx;
    a.psi_pz_in = a.psi_pz_out = base + o_ppz;
    a.psi_vz_in = a.psi_vz_out = base + o_pvz;
    a.sx = p->sx;
    a.sz = p->sz;
    a.amp = base + o_amp;
    a.rec_row = reinterpret_cast<const int4*>(base + o_row) + pad;
    a.rec_more = nullptr;
    a.rec_buf = base + o_chk;  // never written: no receiver spans
    a.nslot = 0;
    for (int m = 0; m < 8; ++m) a.cx[m] = coef[m] / p->dx;
    a.zx = p->dx / p->dz;
    a.ntz = (nz - 2 * R + kTiledTZ - 1) / kTiledTZ;
    a.nchunk = (nx - 2 * R + kTiledTX - 1) / kTiledTX;
    StepKernel kv, kp;
    if (!pick_tiled(R, kTiledTX, kTiledNW, mode, &kv, &kp)) return fail(ctx, SFD_ENOTSUP, "no step kernel");
    const dim3 grid(a.ntz * a.nchunk), block(64 * kTiledNW);
    for (int k = 0; k < nt; ++k) {
        a.k = k;
        // The sponge step reads its velocities as stored unabsorbed from the
        // first step on (the wavefield starts at zero).
        a.v_raw = mode == MODE_FAST_SPONGE ? 1 : 0;
        hipLaunchKernelGGL(kv, grid, block, 0, s, a);
        hipLaunchKernelGGL(kp, grid, block, 0, s, a);
        if (p->nrec > 0)
            hipLaunchKernelGGL(na_gather_kernel, dim3((p->nrec + 255) / 256), dim3(256), 0, s, a.p_out, nz, d_rx, d_rz,
                               p->nrec, d_record, nt, k);
    }
    SFD_HIP(ctx, hipGetLastError());
    SFD_HIP(ctx, hipStreamSynchronize(s));
    rc = check_finite(ctx, d_record, p->nrec > 0 ? (size_t)p->nrec * nt : 0, "the gather");
    if (rc) return rc;
    return check_finite(ctx, a.p_out, (size_t)nx * nz, "the pressure field");
}

}  // extern "C"
