// Convolution surface-related multiple prediction: the consumer of the
// modelled gathers in de-multiple/convolution_srm/convolution_srm.cpp:7-276.
//
// For shot is and output trace it the reference FFTs the shot gather's
// ntrace traces and nshot traces of the common-receiver file starting at
// trace is*ntrace + it*nshot (each zero-padded to ltt = 2 lt), sums the
// products of the spectra of trace k of the shot and trace is + k of that
// block (k < ntrace, is + k < nshot), divides by ntrace + nshot, inverse-FFTs
// and keeps the real part / ltt of the first lt samples.
//
// MI355X mapping.  Every trace is transformed once (batched hipFFT R2C of
// all shot and receiver traces, where the reference transforms nshot
// receiver traces per output trace), the spectra are transposed to
// frequency-major order, and one workgroup per (frequency, shot) forms the
// shot's ntrace output spectra: each wave takes one output trace at a time,
// lanes split the sum over k (contiguous windows of the frequency's row),
// and a wave reduction finishes it.  A frequency's rows (about 4 MB at the
// reference's 401 x 401 x 4000) stay in one XCD's L2 because the workgroups
// of a frequency are dealt to one XCD.  The outputs are transposed back and
// inverse-transformed by a batched C2R.
//
// When ntrace = nshot = N (the reference's card) the window of output
// (is, it) starts at (is + it) N + is, so per frequency the sum is a matrix
// product: M'[is][r] = sum_j A'[is][j] C[r][j] with A'[is][j] = A[is][j - is]
// (j >= is, else 0) and C[r][j] the receiver spectra, and output (is, it) =
// M'[is][is + it].  That path builds A' in its transpose, runs one rocBLAS
// strided-batched CGEMM over all frequencies (MFMA) and picks the outputs
// out of M'.  The product is split into blocks of shot rows that skip most
// of the zeros of A' and of the unused entries of M': 1.56x the
// multiply-adds of the direct sum at the card (4 blocks) instead of 4x, at
// matrix-core rate.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cstring>
#include <new>

#include "common.h"
#include "device_util.h"

namespace sfd {
namespace {

constexpr int kSrmMaxTrace = 8192;  // shot traces held in LDS (8 B each)
constexpr int kTr = 64;             // transpose tile edge

// out[c][r] = in[src(r)][c] for complex [rows][cols]; kTr x kTr tiles through
// LDS (one wave reads or writes a whole 512-byte tile row).  src(r) = r, or with shift_n = N > 0 the banded shift of the GEMM
// path: row r = is N + j reads row is N + j - is when j >= is, else 0.
__global__ __launch_bounds__(256) void srm_transpose_kernel(const float2* __restrict__ in, float2* __restrict__ out,
                                                            int rows, int cols, int shift_n) {
    __shared__ float2 tile[kTr][kTr + 1];
    const int c0 = blockIdx.x * kTr;
    for (int r0 = blockIdx.y * kTr; r0 < rows; r0 += gridDim.y * kTr) {
        for (int j = threadIdx.y; j < kTr; j += 256 / kTr) {
            const int r = r0 + j, c = c0 + threadIdx.x;
            if (r < rows && c < cols) {
                int src = r;
                if (shift_n) {
                    const int is = r / shift_n, jj = r - is * shift_n;
                    src = jj >= is ? r - is : -1;
                }
                tile[j][threadIdx.x] = src >= 0 ? in[(size_t)src * cols + c] : make_float2(0.f, 0.f);
            }
        }
        __syncthreads();
        for (int j = threadIdx.y; j < kTr; j += 256 / kTr) {
            const int c = c0 + j, r = r0 + threadIdx.x;
            if (r < rows && c < cols) out[(size_t)c * rows + r] = tile[threadIdx.x][j];
        }
        __syncthreads();
    }
}

// One workgroup per (frequency f, shot is), dealt to the XCDs in runs of
// whole frequencies:
//   m[f][is*nt + it] = sum_k a[f][is*nt + k] * b[f][is*(nt+1) + it*nshot + k] / (nt + nshot)
// over k < min(nt, nshot - is).
__global__ __launch_bounds__(256) void srm_contract_kernel(const float2* __restrict__ a, const float2* __restrict__ b,
                                                           float2* __restrict__ m, int nshot, int nt, int ncrg,
                                                           int nf) {
    extern __shared__ float2 s_a[];
    const int L = xcd_tile(blockIdx.x, nshot * nf);
    const int f = L / nshot, is = L - f * nshot;
    const int K = min(nt, nshot - is);
    const size_t ntr = (size_t)nshot * nt;
    const float2* af = a + (size_t)f * ntr + (size_t)is * nt;
    const float2* bf = b + (size_t)f * ncrg + (size_t)is * (nt + 1);
    for (int k = threadIdx.x; k < K; k += blockDim.x) s_a[k] = af[k];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    const float inv = (float)(nt + nshot);
    for (int it = wave; it < nt; it += nwave) {
        const float2* bw = bf + (size_t)it * nshot;
        float re = 0.f, im = 0.f;
        for (int k = lane; k < K; k += 64) {
            const float2 x = s_a[k], y = bw[k];
            re += x.x * y.x - x.y * y.y;
            im += x.x * y.y + x.y * y.x;
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            re += __shfl_xor(re, o, 64);
            im += __shfl_xor(im, o, 64);
        }
        if (lane == 0) m[(size_t)f * ntr + (size_t)is * nt + it] = make_float2(re / inv, im / inv);
    }
}

// GEMM path: out[f][is N + it] = mp[f][(is + it) N + is] / (2 N), mp the
// column-major N x (2N - 1) product of each frequency.
__global__ void srm_extract_kernel(const float2* __restrict__ mp, float2* __restrict__ out, int n, int nf) {
    const long long nn = (long long)n * n, total = nn * nf, ld = (long long)n * (2 * n - 1);
    const float div = (float)(2 * n);
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long f = e / nn, rem = e - f * nn;
        const int is = (int)(rem / n), it = (int)(rem - (long long)is * n);
        const float2 v = mp[f * ld + (long long)(is + it) * n + is];
        out[e] = make_float2(v.x / div, v.y / div);
    }
}

// The same through LDS, one workgroup per (64 r, 64 is) block of one
// frequency (blockIdx.z): the block of mp is read along is (contiguous), and
// each is writes its outputs it = r - is as one contiguous run.
__global__ __launch_bounds__(256) void srm_extract_tiled_kernel(const float2* __restrict__ mp,
                                                                float2* __restrict__ out, int n) {
    __shared__ float2 tile[64][65];
    const int r0 = blockIdx.x * 64, i0 = blockIdx.y * 64, f = blockIdx.z;
    // outputs need 0 <= r - is < n: skip blocks with none
    if (r0 + 63 < i0 || r0 > i0 + 63 + n - 1) return;
    const long long ld = (long long)n * (2 * n - 1);
    const float2* m = mp + (long long)f * ld;
    const int x = threadIdx.x & 63, y0 = threadIdx.x >> 6;
    for (int y = y0; y < 64; y += 4) {
        const int r = r0 + y, is = i0 + x;
        if (r < 2 * n - 1 && is < n) tile[y][x] = m[(long long)r * n + is];
    }
    __syncthreads();
    const float div = (float)(2 * n);
    float2* o = out + (long long)f * n * n;
    for (int y = y0; y < 64; y += 4) {
        const int is = i0 + y, r = r0 + x, it = r - is;
        if (is < n && it >= 0 && it < n) {
            const float2 v = tile[x][y];
            o[(long long)is * n + it] = make_float2(v.x / div, v.y / div);
        }
    }
}

// out[tr][t] = full[tr][t] / ltt, t < lt.
__global__ void srm_crop_kernel(const float* __restrict__ full, float* __restrict__ out, long long ntr, int lt,
                                int ltt) {
    const long long n = ntr * lt;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const long long tr = i / lt;
        const int t = (int)(i - tr * lt);
        out[i] = full[tr * ltt + t] / (float)ltt;
    }
}

const char* fft_err(hipfftResult r) {
    switch (r) {
    case HIPFFT_SUCCESS: return "success";
    case HIPFFT_ALLOC_FAILED: return "allocation failed";
    case HIPFFT_INVALID_VALUE: return "invalid value";
    case HIPFFT_INVALID_SIZE: return "invalid size";
    case HIPFFT_EXEC_FAILED: return "execution failed";
    case HIPFFT_SETUP_FAILED: return "setup failed";
    default: return "error";
    }
}

}  // namespace
}  // namespace sfd

using namespace sfd;

struct sfd_srm_plan {
    sfd_context* ctx = nullptr;
    int nshot = 0, nt = 0, lt = 0, ltt = 0, nf = 0, ncrg = 0;
    int method = SFD_SRM_AUTO;
    size_t ntr = 0;
    DeviceBuffer csg, crg, spec_s, spec_c, tr_s, tr_c, out;
    hipfftHandle fwd_s = 0, fwd_c = 0, inv = 0;
    bool has_fwd_s = false, has_fwd_c = false, has_inv = false;
    rocblas_handle blas = nullptr;
    ~sfd_srm_plan() {
        if (has_fwd_s) hipfftDestroy(fwd_s);
        if (has_fwd_c) hipfftDestroy(fwd_c);
        if (has_inv) hipfftDestroy(inv);
        if (blas) rocblas_destroy_handle(blas);
    }
    bool gemm() const { return nt == nshot && method != SFD_SRM_DIRECT; }
};

namespace sfd {
namespace {

#define SFD_FFT(ctx, call)                                                                        \
    do {                                                                                          \
        hipfftResult r_ = (call);                                                                 \
        if (r_ != HIPFFT_SUCCESS) return fail(ctx, SFD_EDEVICE, "%s: hipFFT %s", #call, fft_err(r_)); \
    } while (0)

int srm_transpose(sfd_context* ctx, const void* in, void* out, size_t rows, int cols, int shift_n = 0) {
    const dim3 block(kTr, 256 / kTr);
    const unsigned gy = (unsigned)std::min<size_t>((rows + kTr - 1) / kTr, 65535);
    const dim3 grid((cols + kTr - 1) / kTr, gy);
    hipLaunchKernelGGL(srm_transpose_kernel, grid, block, 0, ctx->stream, (const float2*)in, (float2*)out, (int)rows,
                       cols, shift_n);
    SFD_HIP(ctx, hipGetLastError());
    return SFD_OK;
}

// ntrace = nshot = N: spectra into A' and C (frequency-major), the product
// M' of every frequency into spec_c, outputs picked into tr_s.
int srm_contract_gemm(sfd_srm_plan* pl) {
    sfd_context* ctx = pl->ctx;
    const int n = pl->nshot;
    int rc = srm_transpose(ctx, pl->spec_s.ptr, pl->tr_s.ptr, pl->ntr, pl->nf, n);
    if (!rc) rc = srm_transpose(ctx, pl->spec_c.ptr, pl->tr_c.ptr, (size_t)pl->ncrg, pl->nf);
    if (rc) return rc;
    const rocblas_float_complex one(1.f, 0.f), zero(0.f, 0.f);
    const rocblas_stride sa = (rocblas_stride)n * n, sb = (rocblas_stride)pl->ncrg,
                         sc = (rocblas_stride)n * (2 * n - 1);
    // Shot rows in nb blocks: rows is in [i0, i1) need only j >= i0 (A' is
    // zero left of its band) and r = is + it < i1 + n - 1, which skips most of
    // the zeros and unused outputs of the full product (2 n^3 multiply-adds
    // -> (nb + 1)^2 / (2 nb^2) n^3).  Entries of M' outside these blocks are
    // never read by the extraction.
    const int nb = n >= 256 ? 4 : (n >= 64 ? 2 : 1);
    const auto* A = pl->tr_s.as<rocblas_float_complex>();
    const auto* B = pl->tr_c.as<rocblas_float_complex>();
    auto* C = pl->spec_c.as<rocblas_float_complex>();
    for (int b = 0; b < nb; ++b) {
        const int i0 = (int)((long long)n * b / nb), i1 = (int)((long long)n * (b + 1) / nb);
        const int m = i1 - i0, k = n - i0, cols = i1 + n - 1 - i0;
        const size_t off = (size_t)i0 * n + i0;  // (j, is) = (i0, i0) of A'; (j, r) = (i0, i0) of C; (is, r) of M'
        const rocblas_status st = rocblas_cgemm_strided_batched(
            pl->blas, rocblas_operation_transpose, rocblas_operation_none, m, cols, k, &one, A + off, n, sa, B + off,
            n, sb, &zero, C + off, n, sc, pl->nf);
        if (st != rocblas_status_success)
            return fail(ctx, SFD_EDEVICE, "rocblas_cgemm_strided_batched: %s", rocblas_status_to_string(st));
    }
    if (pl->nf <= 65535) {
        const dim3 grid((2 * n - 1 + 63) / 64, (n + 63) / 64, pl->nf);
        hipLaunchKernelGGL(srm_extract_tiled_kernel, grid, dim3(256), 0, ctx->stream, pl->spec_c.as<float2>(),
                           pl->tr_s.as<float2>(), n);
    } else {
        const long long total = (long long)n * n * pl->nf;
        const int blocks = (int)std::min<long long>((total + 255) / 256, 1 << 20);
        hipLaunchKernelGGL(srm_extract_kernel, dim3(blocks), dim3(256), 0, ctx->stream, pl->spec_c.as<float2>(),
                           pl->tr_s.as<float2>(), n, pl->nf);
    }
    SFD_HIP(ctx, hipGetLastError());
    return SFD_OK;
}

int srm_enqueue(sfd_srm_plan* pl) {
    sfd_context* ctx = pl->ctx;
    SFD_FFT(ctx, hipfftExecR2C(pl->fwd_s, pl->csg.as<hipfftReal>(), pl->spec_s.as<hipfftComplex>()));
    SFD_FFT(ctx, hipfftExecR2C(pl->fwd_c, pl->crg.as<hipfftReal>(), pl->spec_c.as<hipfftComplex>()));
    int rc;
    const void* spec_out;  // output spectra, frequency-major
    if (pl->gemm()) {
        rc = srm_contract_gemm(pl);
        spec_out = pl->tr_s.ptr;
    } else {
        rc = srm_transpose(ctx, pl->spec_s.ptr, pl->tr_s.ptr, pl->ntr, pl->nf);
        if (!rc) rc = srm_transpose(ctx, pl->spec_c.ptr, pl->tr_c.ptr, (size_t)pl->ncrg, pl->nf);
        if (rc) return rc;
        hipLaunchKernelGGL(srm_contract_kernel, dim3(pl->nshot * pl->nf), dim3(256),
                           (size_t)std::min(pl->nt, pl->nshot) * sizeof(float2), ctx->stream, pl->tr_s.as<float2>(),
                           pl->tr_c.as<float2>(), pl->spec_c.as<float2>(), pl->nshot, pl->nt, pl->ncrg, pl->nf);
        SFD_HIP(ctx, hipGetLastError());
        spec_out = pl->spec_c.ptr;
    }
    if (rc) return rc;
    // back to trace-major in spec_s for the inverse transform
    rc = srm_transpose(ctx, spec_out, pl->spec_s.ptr, (size_t)pl->nf, (int)pl->ntr);
    if (rc) return rc;
    SFD_FFT(ctx, hipfftExecC2R(pl->inv, pl->spec_s.as<hipfftComplex>(), pl->tr_c.as<hipfftReal>()));
    const long long n = (long long)pl->ntr * pl->lt;
    const int blocks = (int)std::min<long long>((n + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(srm_crop_kernel, dim3(blocks), dim3(256), 0, ctx->stream, pl->tr_c.as<float>(),
                       pl->out.as<float>(), (long long)pl->ntr, pl->lt, pl->ltt);
    SFD_HIP(ctx, hipGetLastError());
    return SFD_OK;
}

}  // namespace
}  // namespace sfd

extern "C" {

int sfd_srm_crg_traces(int nshot, int ntrace) {
    if (nshot < 1 || ntrace < 1) return SFD_EINVAL;
    return ntrace * (2 * nshot - 1);
}

int sfd_srm_plan_create(sfd_context* ctx, const sfd_srm_params* p, const float* csg, const float* crg,
                        sfd_srm_plan** out) {
    if (!ctx || !out) return SFD_EINVAL;
    *out = nullptr;
    if (!p || !csg || !crg) return fail(ctx, SFD_EINVAL, "params, csg and crg are required");
    if (p->nshot < 1 || p->ntrace < 1 || p->lt < 1) return fail(ctx, SFD_EINVAL, "nshot, ntrace and lt must be >= 1");
    if (std::min(p->ntrace, p->nshot) > kSrmMaxTrace)
        return fail(ctx, SFD_EINVAL, "min(ntrace, nshot) must be <= %d", kSrmMaxTrace);
    const long long need = (long long)p->ntrace * (2LL * p->nshot - 1);
    if (p->ncrg < need)
        return fail(ctx, SFD_EINVAL, "the common-receiver array needs ntrace (2 nshot - 1) = %lld traces, got %d", need,
                    p->ncrg);
    if ((long long)p->nshot * (p->lt + 1) > 0x7fffffffLL || (long long)p->ncrg * 2 * p->lt > (1LL << 40))
        return fail(ctx, SFD_EINVAL, "problem too large");
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    sfd_srm_plan* pl = new (std::nothrow) sfd_srm_plan;
    if (!pl) return fail(ctx, SFD_ENOMEM, "host allocation failed");
    auto bail = [&](int code) {
        delete pl;
        return code;
    };
    pl->ctx = ctx;
    pl->nshot = p->nshot;
    pl->nt = p->ntrace;
    pl->lt = p->lt;
    pl->ltt = 2 * p->lt;
    pl->nf = p->lt + 1;
    pl->ncrg = p->ncrg;
    pl->ntr = (size_t)p->nshot * p->ntrace;
    const size_t nc = (size_t)p->ncrg, ltt = (size_t)pl->ltt, nf = (size_t)pl->nf;
    hipError_t e = pl->csg.alloc(pl->ntr * ltt * sizeof(float));
    if (e == hipSuccess) e = pl->crg.alloc(nc * ltt * sizeof(float));
    if (e == hipSuccess) e = pl->spec_s.alloc(pl->ntr * nf * sizeof(float2));
    if (e == hipSuccess) e = pl->spec_c.alloc(nc * nf * sizeof(float2));
    if (e == hipSuccess) e = pl->tr_s.alloc(pl->ntr * nf * sizeof(float2));
    if (e == hipSuccess) e = pl->tr_c.alloc(nc * nf * sizeof(float2));
    if (e == hipSuccess) e = pl->out.alloc(pl->ntr * (size_t)p->lt * sizeof(float));
    if (e != hipSuccess) return bail(hip_check(ctx, e, "SRM buffers"));
    hipStream_t sm = ctx->stream;
    // traces zero-padded to ltt (the reference's us / ur rows, :99-100, :170-172)
    e = hipMemsetAsync(pl->csg.ptr, 0, pl->csg.bytes, sm);
    if (e == hipSuccess) e = hipMemsetAsync(pl->crg.ptr, 0, pl->crg.bytes, sm);
    if (e == hipSuccess)
        e = hipMemcpy2DAsync(pl->csg.ptr, ltt * sizeof(float), csg, (size_t)p->lt * sizeof(float),
                             (size_t)p->lt * sizeof(float), pl->ntr, hipMemcpyHostToDevice, sm);
    if (e == hipSuccess)
        e = hipMemcpy2DAsync(pl->crg.ptr, ltt * sizeof(float), crg, (size_t)p->lt * sizeof(float),
                             (size_t)p->lt * sizeof(float), nc, hipMemcpyHostToDevice, sm);
    if (e == hipSuccess) e = hipStreamSynchronize(sm);
    if (e != hipSuccess) return bail(hip_check(ctx, e, "SRM upload"));
    int n = pl->ltt;
    hipfftResult r = hipfftPlanMany(&pl->fwd_s, 1, &n, nullptr, 1, pl->ltt, nullptr, 1, pl->nf, HIPFFT_R2C, (int)pl->ntr);
    pl->has_fwd_s = r == HIPFFT_SUCCESS;
    if (r == HIPFFT_SUCCESS) {
        r = hipfftPlanMany(&pl->fwd_c, 1, &n, nullptr, 1, pl->ltt, nullptr, 1, pl->nf, HIPFFT_R2C, p->ncrg);
        pl->has_fwd_c = r == HIPFFT_SUCCESS;
    }
    if (r == HIPFFT_SUCCESS) {
        r = hipfftPlanMany(&pl->inv, 1, &n, nullptr, 1, pl->nf, nullptr, 1, pl->ltt, HIPFFT_C2R, (int)pl->ntr);
        pl->has_inv = r == HIPFFT_SUCCESS;
    }
    if (r == HIPFFT_SUCCESS) r = hipfftSetStream(pl->fwd_s, sm);
    if (r == HIPFFT_SUCCESS) r = hipfftSetStream(pl->fwd_c, sm);
    if (r == HIPFFT_SUCCESS) r = hipfftSetStream(pl->inv, sm);
    if (r != HIPFFT_SUCCESS) return bail(fail(ctx, SFD_EDEVICE, "hipFFT plans: %s", fft_err(r)));
    if (pl->nt == pl->nshot) {
        rocblas_status st = rocblas_create_handle(&pl->blas);
        if (st == rocblas_status_success) st = rocblas_set_stream(pl->blas, sm);
        if (st != rocblas_status_success)
            return bail(fail(ctx, SFD_EDEVICE, "rocBLAS handle: %s", rocblas_status_to_string(st)));
    }
    *out = pl;
    return SFD_OK;
}

int sfd_srm_plan_set_method(sfd_srm_plan* pl, int method) {
    if (!pl) return SFD_EINVAL;
    if (method != SFD_SRM_AUTO && method != SFD_SRM_DIRECT && method != SFD_SRM_GEMM)
        return fail(pl->ctx, SFD_EINVAL, "unknown SRM method %d", method);
    if (method == SFD_SRM_GEMM && pl->nt != pl->nshot)
        return fail(pl->ctx, SFD_EINVAL, "the GEMM path needs ntrace = nshot");
    pl->method = method;
    return SFD_OK;
}

int sfd_srm_plan_get_method(const sfd_srm_plan* pl) {
    if (!pl) return SFD_EINVAL;
    return pl->gemm() ? SFD_SRM_GEMM : SFD_SRM_DIRECT;
}

int sfd_srm_plan_run(sfd_srm_plan* pl, float* elapsed_ms) {
    if (!pl) return SFD_EINVAL;
    sfd_context* ctx = pl->ctx;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    SFD_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    int rc = srm_enqueue(pl);
    if (rc) return rc;
    SFD_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    SFD_HIP(ctx, hipEventSynchronize(ctx->ev1));
    if (elapsed_ms) SFD_HIP(ctx, hipEventElapsedTime(elapsed_ms, ctx->ev0, ctx->ev1));
    return check_finite(ctx, pl->out.as<float>(), pl->out.bytes / 4, "the predicted multiples");
}

int sfd_srm_plan_fetch(sfd_srm_plan* pl, float* out) {
    if (!pl || !out) return SFD_EINVAL;
    sfd_context* ctx = pl->ctx;
    SFD_HIP(ctx, hipSetDevice(ctx->device));
    SFD_HIP(ctx, hipMemcpyAsync(out, pl->out.ptr, pl->out.bytes, hipMemcpyDeviceToHost, ctx->stream));
    SFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SFD_OK;
}

int sfd_srm_plan_destroy(sfd_srm_plan* pl) {
    if (!pl) return SFD_EINVAL;
    (void)hipSetDevice(pl->ctx->device);
    (void)hipStreamSynchronize(pl->ctx->stream);
    delete pl;
    return SFD_OK;
}

int sfd_srm_predict(sfd_context* ctx, const sfd_srm_params* p, const float* csg, const float* crg, float* out) {
    sfd_srm_plan* pl = nullptr;
    int rc = sfd_srm_plan_create(ctx, p, csg, crg, &pl);
    if (rc) return rc;
    rc = sfd_srm_plan_run(pl, nullptr);
    if (!rc) rc = sfd_srm_plan_fetch(pl, out);
    sfd_srm_plan_destroy(pl);
    return rc;
}

}  // extern "C"
