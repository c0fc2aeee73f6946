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
        if (t_ctx.enq_done) (void)hipEventSynchronize(t_ctx.enq_done);
        if (t_ctx.dev) (void)hipFree(t_ctx.dev);
        if (t_ctx.enq_dev) (void)hipFree(t_ctx.enq_dev);
        if (t_ctx.enq_done) (void)hipEventDestroy(t_ctx.enq_done);
        if (t_ctx.pinned) (void)hipHostFree(t_ctx.pinned);
        if (t_ctx.stream) (void)hipStreamDestroy(t_ctx.stream);
    }
};
thread_local CtxReaper t_reaper;
} // namespace

namespace {
thread_local int32_t t_fail_alloc = 0;  // countdown to an injected allocation failure (0: off)
thread_local int32_t t_fail_launch = 0; // countdown to an injected launch failure (0: off)
// Set when a call on this thread returned a HIP error: work it queued before the failure may
// still be in flight on the thread's stream (and its pinned staging), so the next call
// drains that stream before reusing anything.
thread_local bool t_unsettled = false;
} // namespace

bool inject_alloc_failure() { return t_fail_alloc > 0 && --t_fail_alloc == 0; }

hipError_t launch_check() {
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess && t_fail_launch > 0 && --t_fail_launch == 0) return hipErrorLaunchFailure;
    return e;
}

int status_of(hipError_t e) {
    if (e == hipSuccess) return MCDX_OK;
    t_unsettled = true;
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
    if (t_unsettled) {
        t_unsettled = false;
        (void)hipStreamSynchronize(t_ctx.stream);
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
    hipError_t e = dev_malloc(&ctx->dev, want);
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
    hipError_t e = inject_alloc_failure() ? hipErrorOutOfMemory
                                          : hipHostMalloc(&ctx->pinned, want, hipHostMallocDefault);
    if (e != hipSuccess) return status_of(e);
    ctx->pinned_bytes = want;
    return MCDX_OK;
}

int enqueue_scratch(ThreadCtx* ctx, hipStream_t s, size_t bytes, void** out) {
    hipError_t e;
    if (!ctx->enq_done) {
        e = hipEventCreateWithFlags(&ctx->enq_done, hipEventDisableTiming);
        if (e != hipSuccess) return status_of(e);
    }
    if (bytes > ctx->enq_bytes) {
        if (ctx->enq_dev) {
            // the last use, on whichever stream, must end before the buffer is freed
            if (ctx->enq_used && (e = hipEventSynchronize(ctx->enq_done)) != hipSuccess)
                return status_of(e);
            (void)hipFree(ctx->enq_dev);
            ctx->enq_dev = nullptr;
            ctx->enq_bytes = 0;
        }
        const size_t want = bytes < (1u << 20) ? (1u << 20) : bytes + bytes / 4;
        if ((e = dev_malloc(&ctx->enq_dev, want)) != hipSuccess) return status_of(e);
        ctx->enq_bytes = want;
        ctx->enq_used = false;
    } else if (ctx->enq_used && ctx->enq_last != s) {
        // work already queued on `s` is ordered after the previous use by `s` itself
        if ((e = hipStreamWaitEvent(s, ctx->enq_done, 0)) != hipSuccess) return status_of(e);
    }
    *out = ctx->enq_dev;
    return MCDX_OK;
}

int enqueue_scratch_done(ThreadCtx* ctx, hipStream_t s) {
    hipError_t e = hipEventRecord(ctx->enq_done, s);
    if (e != hipSuccess) return status_of(e);
    ctx->enq_last = s;
    ctx->enq_used = true;
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
    const Welford total = merge_tiles_fixed_order(partials + base, n_tiles, lds);
    if (threadIdx.x == 0) out[c] = total;
}

hipError_t launch_final_merge(const Welford* partials, const uint64_t* tile_offsets,
                              uint64_t n_tiles, uint32_t n_contracts, Welford* out,
                              hipStream_t stream) {
    hipLaunchKernelGGL(final_merge_kernel, dim3(n_contracts), dim3(kMergeGroups), 0, stream,
                       partials, tile_offsets, n_tiles, out);
    return launch_check();
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

__global__ __launch_bounds__(256) void inv_normal_words_kernel(const uint32_t* __restrict__ words,
                                                               uint64_t n, int fp64,
                                                               double* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t bits = words[i];
    out[i] = fp64 ? inv_normal_f64(uniform_open01_f64(bits)) : (double)inv_normal_f32(bits);
}

// Error of the FP32 inverse normal against the FP64 one (the reference's Acklam + Halley) over
// bit patterns begin .. begin + count - 1: e = |z32 - z64| / (1 + |z64|). Each block keeps
// its largest e, separately for central and tail patterns (is_tail), with the smallest bit
// pattern on ties; slot b of the outputs belongs to block b.
constexpr int kScanBlocks = 8192;
__global__ __launch_bounds__(256) void inv_normal_scan_kernel(uint64_t begin, uint64_t count,
                                                              double* __restrict__ err_out,
                                                              uint32_t* __restrict__ bits_out) {
    __shared__ double se[2][256];
    __shared__ uint32_t sb[2][256];
    double best[2] = {-1.0, -1.0};
    uint32_t arg[2] = {0u, 0u};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const uint32_t bits = (uint32_t)(begin + i);
        const double z32 = (double)inv_normal_f32(bits);
        const double z64 = inv_normal_f64(uniform_open01_f64(bits));
        const double e = fabs(z32 - z64) / (1.0 + fabs(z64));
        const int t = is_tail(bits) ? 1 : 0;
        if (e > best[t]) { // bits rise within a thread, so the first maximum is the smallest
            best[t] = e;
            arg[t] = bits;
        }
    }
    for (int t = 0; t < 2; ++t) {
        se[t][threadIdx.x] = best[t];
        sb[t][threadIdx.x] = arg[t];
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int t = 0; t < 2; ++t) {
                const double o = se[t][threadIdx.x + s];
                const uint32_t ob = sb[t][threadIdx.x + s];
                if (o > se[t][threadIdx.x] || (o == se[t][threadIdx.x] && ob < sb[t][threadIdx.x])) {
                    se[t][threadIdx.x] = o;
                    sb[t][threadIdx.x] = ob;
                }
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int t = 0; t < 2; ++t) {
            err_out[2 * blockIdx.x + t] = se[t][0];
            bits_out[2 * blockIdx.x + t] = sb[t][0];
        }
    }
}

} // namespace mcdx

using namespace mcdx;

extern "C" {

int mcdx_fault_inject(int site, int32_t countdown) {
    if (countdown < 0) return MCDX_EINVAL;
    if (site == MCDX_FAULT_ALLOC) t_fail_alloc = countdown;
    else if (site == MCDX_FAULT_LAUNCH) t_fail_launch = countdown;
    else return MCDX_EINVAL;
    return MCDX_OK;
}

int mcdx_fault_pending(int site) {
    if (site == MCDX_FAULT_ALLOC) return t_fail_alloc;
    if (site == MCDX_FAULT_LAUNCH) return t_fail_launch;
    return -MCDX_EINVAL;
}

int mcdx_inv_normal_words(const uint32_t* words, uint64_t n, int precision, double* out) {
    if ((!words || !out) && n) return MCDX_EINVAL;
    if (precision != MCDX_PREC_FP32 && precision != MCDX_PREC_FP64) return MCDX_EINVAL;
    if (n == 0) return MCDX_OK;
    if (n > (1ull << 28)) return MCDX_EINVAL;
    ThreadCtx* ctx;
    int st = thread_ctx(&ctx);
    if (st) return st;
    if ((st = ensure_dev(ctx, n * 12 + 256))) return st;
    double* dout = (double*)ctx->dev;
    uint32_t* dw = (uint32_t*)(dout + n);
    hipError_t e = hipMemcpyAsync(dw, words, n * 4, hipMemcpyHostToDevice, ctx->stream);
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(inv_normal_words_kernel, dim3(blocks), dim3(256), 0, ctx->stream, dw,
                           n, precision == MCDX_PREC_FP64 ? 1 : 0, dout);
        e = launch_check();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, n * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return status_of(e);
}

int mcdx_inv_normal_f32_scan(uint64_t begin, uint64_t count, double* max_err,
                             uint32_t* argmax_bits) {
    if (!max_err || !argmax_bits) return MCDX_EINVAL;
    if (begin > (1ull << 32) || count > (1ull << 32) - begin) return MCDX_EINVAL;
    max_err[0] = max_err[1] = -1.0;
    argmax_bits[0] = argmax_bits[1] = 0u;
    if (count == 0) return MCDX_OK;
    ThreadCtx* ctx;
    int st = thread_ctx(&ctx);
    if (st) return st;
    const size_t slots = 2 * (size_t)kScanBlocks;
    const size_t bytes = slots * (sizeof(double) + sizeof(uint32_t));
    if ((st = ensure_dev(ctx, bytes))) return st;
    if ((st = ensure_pinned(ctx, bytes))) return st;
    double* derr = (double*)ctx->dev;
    uint32_t* dbits = (uint32_t*)(derr + slots);
    hipLaunchKernelGGL(inv_normal_scan_kernel, dim3(kScanBlocks), dim3(256), 0, ctx->stream,
                       begin, count, derr, dbits);
    hipError_t e = launch_check();
    if (e == hipSuccess)
        e = hipMemcpyAsync(ctx->pinned, ctx->dev, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return status_of(e);
    const double* herr = (const double*)ctx->pinned;
    const uint32_t* hbits = (const uint32_t*)(herr + slots);
    for (size_t b = 0; b < (size_t)kScanBlocks; ++b) {
        for (int t = 0; t < 2; ++t) {
            const double v = herr[2 * b + t];
            const uint32_t vb = hbits[2 * b + t];
            if (v > max_err[t] || (v == max_err[t] && v >= 0.0 && vb < argmax_bits[t])) {
                max_err[t] = v;
                argmax_bits[t] = vb;
            }
        }
    }
    return MCDX_OK;
}

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
    hipError_t e = launch_check();
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
    hipError_t e = launch_check();
    if (e == hipSuccess) e = hipMemcpyAsync(out, ctx->dev, n * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return status_of(e);
}

} // extern "C"
