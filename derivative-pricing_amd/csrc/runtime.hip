// Engine runtime: per-thread streams and scratch, the fixed-order tile merge, RNG parity
// kernels and the C-ABI engine entry points.
#include <math.h>
#include <stdio.h>

#include <atomic>
#include <mutex>

#include "runtime.hpp"

namespace mcdx {

namespace {
std::atomic<int> g_precision{MCDX_PREC_FP32};
thread_local ThreadCtx t_ctx;
thread_local bool t_ready = false;

// One context per thread; released when the thread exits.
struct CtxReaper {
    ~CtxReaper() {
        if (!t_ready) return;
        if (t_ctx.dev) (void)hipFree(t_ctx.dev);
        if (t_ctx.pinned) (void)hipHostFree(t_ctx.pinned);
        if (t_ctx.stream) (void)hipStreamDestroy(t_ctx.stream);
    }
};
thread_local CtxReaper t_reaper;
} // namespace

int status_of(hipError_t e) {
    if (e == hipSuccess) return MCDX_OK;
    if (e == hipErrorOutOfMemory) return MCDX_ENOMEM;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return MCDX_ENODEV;
    return MCDX_EDEVICE;
}

int thread_ctx(ThreadCtx** ctx) {
    if (!t_ready) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return MCDX_ENODEV;
        (void)&t_reaper;
        hipError_t e = hipStreamCreateWithFlags(&t_ctx.stream, hipStreamNonBlocking);
        if (e != hipSuccess) return status_of(e);
        t_ready = true;
    }
    *ctx = &t_ctx;
    return MCDX_OK;
}

int ensure_dev(ThreadCtx* ctx, size_t bytes) {
    if (bytes <= ctx->dev_bytes) return MCDX_OK;
    if (ctx->dev) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return status_of(e);
        (void)hipFree(ctx->dev);
        ctx->dev = nullptr;
        ctx->dev_bytes = 0;
    }
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes + bytes / 4;
    hipError_t e = hipMalloc(&ctx->dev, want);
    if (e != hipSuccess) return status_of(e);
    ctx->dev_bytes = want;
    return MCDX_OK;
}

int ensure_pinned(ThreadCtx* ctx, size_t bytes) {
    if (bytes <= ctx->pinned_bytes) return MCDX_OK;
    if (ctx->pinned) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return status_of(e);
        (void)hipHostFree(ctx->pinned);
        ctx->pinned = nullptr;
        ctx->pinned_bytes = 0;
    }
    size_t want = bytes < 4096 ? 4096 : bytes;
    hipError_t e = hipHostMalloc(&ctx->pinned, want, hipHostMallocDefault);
    if (e != hipSuccess) return status_of(e);
    ctx->pinned_bytes = want;
    return MCDX_OK;
}

mcdx_result finish(const Welford& w) {
    mcdx_result r;
    r.price = w.mean;
    const double var = w.n < 2 ? 0.0 : w.m2 / (double)(w.n - 1);
    r.standard_error = w.n == 0 ? 0.0 : sqrt(var / (double)w.n);
    r.path_count = w.n;
    return r;
}

// Fixed-order merge of one contract's tile partials: 1024 groups, group g folds the tiles
// chunk_bounds(g, n_tiles, 1024) left to right (the reference's static partition,
// thread_pool.cpp:9-18), then a pairwise tree over the groups in index order.
__global__ __launch_bounds__(kMergeGroups) void final_merge_kernel(
    const Welford* __restrict__ partials, const uint64_t* __restrict__ tile_offsets,
    uint64_t n_tiles_single, Welford* __restrict__ out) {
    __shared__ Welford lds[kMergeGroups];
    const uint32_t c = blockIdx.x;
    uint64_t base = (uint64_t)c * n_tiles_single, n_tiles = n_tiles_single;
    if (tile_offsets) {
        base = tile_offsets[c];
        n_tiles = tile_offsets[c + 1] - base;
    }
    const uint32_t g = threadIdx.x;
    const uint64_t q = n_tiles / kMergeGroups, rem = n_tiles % kMergeGroups;
    const uint64_t b = g * q + (g < rem ? g : rem);
    const uint64_t e = b + q + (g < rem ? 1 : 0);
    Welford acc = {0, 0.0, 0.0};
    for (uint64_t t = b; t < e; ++t) acc = welford_merge(acc, partials[base + t]);
    lds[g] = acc;
    __syncthreads();
    for (uint32_t s = 1; s < kMergeGroups; s <<= 1) {
        if ((g & (2 * s - 1)) == 0) lds[g] = welford_merge(lds[g], lds[g + s]);
        __syncthreads();
    }
    if (g == 0) out[c] = lds[0];
}

hipError_t launch_final_merge(const Welford* partials, const uint64_t* tile_offsets,
                              uint64_t n_tiles, uint32_t n_contracts, Welford* out,
                              hipStream_t stream) {
    hipLaunchKernelGGL(final_merge_kernel, dim3(n_contracts), dim3(kMergeGroups), 0, stream,
                       partials, tile_offsets, n_tiles, out);
    return hipGetLastError();
}

// ---- RNG parity kernels ----
__global__ __launch_bounds__(256) void philox_bits_kernel(uint64_t seed, uint64_t path0,
                                                          uint64_t n, uint32_t draw,
                                                          uint32_t stream, uint32_t* out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = philox_bits(seed, path0 + i, draw, stream);
}

__global__ __launch_bounds__(256) void normal_draws_kernel(uint64_t seed, uint64_t path0,
                                                           uint64_t n, uint32_t draw,
                                                           int fp64, double* out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t bits = philox_bits(seed, path0 + i, draw, 0);
    out[i] = fp64 ? inv_normal_f64(uniform_open01_f64(bits)) : (double)inv_normal_f32(bits);
}

} // namespace mcdx

using namespace mcdx;

extern "C" {

int mcdx_device_count(int* count) {
    if (!count) return MCDX_EINVAL;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = e == hipSuccess ? n : 0;
    return e == hipSuccess ? MCDX_OK : status_of(e);
}

int mcdx_set_precision(int precision) {
    if (precision != MCDX_PREC_FP32 && precision != MCDX_PREC_FP64) return MCDX_EINVAL;
    g_precision.store(precision);
    return MCDX_OK;
}

int mcdx_get_precision(void) { return g_precision.load(); }

const char* mcdx_version(void) { return "0.1.0"; }

int mcdx_philox_bits(uint64_t seed, uint64_t path0, uint64_t n, uint32_t draw, uint32_t stream,
                     uint32_t* out) {
    if (!out && n) return MCDX_EINVAL;
    if (n == 0) return MCDX_OK;
    if (n > (1ull << 31)) return MCDX_EINVAL;
    ThreadCtx* ctx;
    int st = thread_ctx(&ctx);
    if (st) return st;
    if ((st = ensure_dev(ctx, n * 4))) return st;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(philox_bits_kernel, dim3(blocks), dim3(256), 0, ctx->stream, seed, path0,
                       n, draw, stream, (uint32_t*)ctx->dev);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, ctx->dev, n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return status_of(e);
}

int mcdx_normal_draws(uint64_t seed, uint64_t path0, uint64_t n, uint32_t draw, int precision,
                      double* out) {
    if (!out && n) return MCDX_EINVAL;
    if (precision != MCDX_PREC_FP32 && precision != MCDX_PREC_FP64) return MCDX_EINVAL;
    if (n == 0) return MCDX_OK;
    if (n > (1ull << 31)) return MCDX_EINVAL;
    ThreadCtx* ctx;
    int st = thread_ctx(&ctx);
    if (st) return st;
    if ((st = ensure_dev(ctx, n * 8))) return st;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(normal_draws_kernel, dim3(blocks), dim3(256), 0, ctx->stream, seed, path0,
                       n, draw, precision == MCDX_PREC_FP64 ? 1 : 0, (double*)ctx->dev);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, ctx->dev, n * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return status_of(e);
}

} // extern "C"
