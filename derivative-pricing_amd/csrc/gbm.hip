// Per-contract GBM launches: one tile of 1024 paths per block, then the fixed-order merge.
// The path arithmetic lives in gbm_paths.hpp (shared with the batched book).
#include <atomic>
#include <mutex>

#include "gbm_paths.hpp"

namespace mcdx {

// Tile partials of the allocation-free entry point: one array per device, held until the
// process ends. Host code never uses the array's name as an address: each device's copy is
// reached through the pointer hipGetSymbolAddress returned for it (noalloc_scratch).
constexpr uint64_t kNoAllocMaxTiles = 65536; // 2^26 paths per call, 1.5 MiB
static __device__ __align__(256) Welford scratch[kNoAllocMaxTiles];

// One tile of 1024 paths per block: lane folds its 4 paths (Welford, stats.hpp:23-29),
// then the block's pairwise tree writes the tile partial.
template <int KIND, bool FP64, bool ANTI, int SUMS>
__device__ __forceinline__ void tile_body(const GbmArgs& a) {
    __shared__ Welford lds[4];
    const uint64_t tile = blockIdx.x;
    Welford w = {0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < kPathsPerLane; ++k) {
        // base is a multiple of 256 and threadIdx.x < 256: every lane's path >> 32 is base >> 32.
        const uint64_t base = tile * kTilePaths + (uint64_t)(k * kTileLanes);
        const uint64_t path = base + threadIdx.x;
        if (path < a.n_paths) {
            const double x = path_sample<KIND, FP64, ANTI, SUMS>(a, path, (uint32_t)(base >> 32));
            w.n += 1;
            const double d = x - w.mean;
            w.mean += d * (1.0 / (double)(k + 1));
            w.m2 += d * (x - w.mean);
        }
    }
    const Welford t = block_merge_256(w, lds);
    if (threadIdx.x == 0) a.partials[tile] = t;
}

template <int KIND, bool FP64, bool ANTI>
__global__ __launch_bounds__(kTileLanes) void gbm_tile_kernel(GbmArgs a) {
    tile_body<KIND, FP64, ANTI, kSumsAny>(a);
}

// FP32 Asian tiles with the sums fixed at compile time: one step loop and fewer registers
// than the per-path dispatch (65 instead of 86 VGPRs without antithetics: 7 waves per SIMD
// instead of 5). Holding it to 64 VGPRs (8 waves) spilled a register to scratch and measured
// no faster.
template <bool ANTI, int SUMS>
__global__ __launch_bounds__(kTileLanes) void gbm_asian_tile_kernel(GbmArgs a) {
    tile_body<kAsian, false, ANTI, SUMS>(a);
}

template <int KIND, bool FP64, bool ANTI>
__global__ __launch_bounds__(256) void gbm_samples_kernel(GbmArgs a, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a.samples[i] = path_sample<KIND, FP64, ANTI>(a, a.path0 + i, (uint32_t)((a.path0 + i) >> 32));
}

namespace {

template <bool FP64, bool ANTI>
void launch_tiles_k(int kind, const GbmArgs& a, unsigned grid, hipStream_t s) {
    switch (kind) {
    case kTerminal: hipLaunchKernelGGL((gbm_tile_kernel<kTerminal, FP64, ANTI>), dim3(grid), dim3(kTileLanes), 0, s, a); break;
    case kAsian:
        if constexpr (!FP64) {
            const int sums = asian_sums(a);
            if (sums == kSumsArith) hipLaunchKernelGGL((gbm_asian_tile_kernel<ANTI, kSumsArith>), dim3(grid), dim3(kTileLanes), 0, s, a);
            else if (sums == kSumsGeo) hipLaunchKernelGGL((gbm_asian_tile_kernel<ANTI, kSumsGeo>), dim3(grid), dim3(kTileLanes), 0, s, a);
            else hipLaunchKernelGGL((gbm_asian_tile_kernel<ANTI, kSumsBoth>), dim3(grid), dim3(kTileLanes), 0, s, a);
        } else {
            hipLaunchKernelGGL((gbm_tile_kernel<kAsian, FP64, ANTI>), dim3(grid), dim3(kTileLanes), 0, s, a);
        }
        break;
    case kBarrier: hipLaunchKernelGGL((gbm_tile_kernel<kBarrier, FP64, ANTI>), dim3(grid), dim3(kTileLanes), 0, s, a); break;
    case kBarrierBB: hipLaunchKernelGGL((gbm_tile_kernel<kBarrierBB, FP64, ANTI>), dim3(grid), dim3(kTileLanes), 0, s, a); break;
    default: hipLaunchKernelGGL((gbm_tile_kernel<kLookback, FP64, ANTI>), dim3(grid), dim3(kTileLanes), 0, s, a); break;
    }
}

template <bool FP64, bool ANTI>
void launch_samples_k(int kind, const GbmArgs& a, uint64_t n, hipStream_t s) {
    const unsigned grid = (unsigned)((n + 255) / 256);
    switch (kind) {
    case kTerminal: hipLaunchKernelGGL((gbm_samples_kernel<kTerminal, FP64, ANTI>), dim3(grid), dim3(256), 0, s, a, n); break;
    case kAsian: hipLaunchKernelGGL((gbm_samples_kernel<kAsian, FP64, ANTI>), dim3(grid), dim3(256), 0, s, a, n); break;
    case kBarrier: hipLaunchKernelGGL((gbm_samples_kernel<kBarrier, FP64, ANTI>), dim3(grid), dim3(256), 0, s, a, n); break;
    case kBarrierBB: hipLaunchKernelGGL((gbm_samples_kernel<kBarrierBB, FP64, ANTI>), dim3(grid), dim3(256), 0, s, a, n); break;
    default: hipLaunchKernelGGL((gbm_samples_kernel<kLookback, FP64, ANTI>), dim3(grid), dim3(256), 0, s, a, n); break;
    }
}

// Enqueue tiles + merge for one contract; scratch must hold tiles_for(N) + 1 Welfords.
hipError_t enqueue_contract(const mcdx_gbm_contract& c, int precision, Welford* partials,
                            Welford* result, hipStream_t s) {
    GbmArgs a = make_args(c);
    a.partials = partials;
    const uint64_t tiles = tiles_for(c.path_count);
    const int kind = kind_of(c);
    const bool fp64 = precision == MCDX_PREC_FP64;
    const bool anti = c.antithetic != 0;
    if (fp64) {
        if (anti) launch_tiles_k<true, true>(kind, a, (unsigned)tiles, s);
        else launch_tiles_k<true, false>(kind, a, (unsigned)tiles, s);
    } else {
        if (anti) launch_tiles_k<false, true>(kind, a, (unsigned)tiles, s);
        else launch_tiles_k<false, false>(kind, a, (unsigned)tiles, s);
    }
    hipError_t e = launch_check();
    if (e != hipSuccess) return e;
    return launch_final_merge(partials, nullptr, tiles, 1, result, s);
}

constexpr int kNoAllocMaxDevices = 64;
std::atomic<Welford*> g_scratch_ptr[kNoAllocMaxDevices];
std::mutex g_scratch_mu;

// The current device's scratch address, looked up once per device with that device current.
int noalloc_scratch(Welford** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return status_of(e);
    if (dev < 0 || dev >= kNoAllocMaxDevices) return MCDX_EDEVICE;
    Welford* p = g_scratch_ptr[dev].load(std::memory_order_acquire);
    if (!p) {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        p = g_scratch_ptr[dev].load(std::memory_order_relaxed);
        if (!p) {
            void* addr = nullptr;
            e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
            if (e != hipSuccess) return status_of(e);
            if (!addr) return MCDX_EDEVICE;
            p = (Welford*)addr;
            g_scratch_ptr[dev].store(p, std::memory_order_release);
        }
    }
    *out = p;
    return MCDX_OK;
}

} // namespace

// Shared with the batched book: price one contract synchronously on the thread's stream.
int price_gbm_sync(const mcdx_gbm_contract* c, int precision, mcdx_result* out) {
    int st = validate(c, precision);
    if (st) return st;
    if (!out) return MCDX_EINVAL;
    if (c->path_count == 0) {
        *out = mcdx_result{0.0, 0.0, 0};
        return MCDX_OK;
    }
    ThreadCtx* ctx;
    if ((st = thread_ctx(&ctx))) return st;
    const uint64_t tiles = tiles_for(c->path_count);
    if ((st = ensure_dev(ctx, (tiles + 1) * sizeof(Welford)))) return st;
    if ((st = ensure_pinned(ctx, sizeof(Welford)))) return st;
    Welford* partials = (Welford*)ctx->dev;
    hipError_t e = enqueue_contract(*c, precision, partials, partials + tiles, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ctx->pinned, partials + tiles, sizeof(Welford), hipMemcpyDeviceToHost,
                           ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return status_of(e);
    *out = finish(*(const Welford*)ctx->pinned);
    return MCDX_OK;
}

} // namespace mcdx

using namespace mcdx;

namespace {
mcdx_gbm_contract base_contract(double spot, double strike, double rate, double carry,
                                double vol, double time, uint8_t type, uint64_t n,
                                uint64_t seed, const mcdx_options* o) {
    mcdx_gbm_contract c{};
    c.spot = spot;
    c.strike = strike;
    c.rate = rate;
    c.carry_yield = carry;
    c.vol = vol;
    c.time = time;
    c.cash_amount = 1.0;
    c.type = type;
    c.path_count = n;
    c.seed = seed;
    c.knock = 1;
    if (o) {
        c.antithetic = o->antithetic;
        c.control_variate = o->control_variate;
        c.brownian_bridge = o->brownian_bridge;
    }
    return c;
}
} // namespace

extern "C" {

int mcdx_price_gbm(const mcdx_gbm_contract* contract, int precision, mcdx_result* out) {
    return price_gbm_sync(contract, precision, out);
}

int mcdx_price_gbm_enqueue(const mcdx_gbm_contract* c, int precision, void* stream,
                           void* dev_welford) {
    int st = validate(c, precision);
    if (st) return st;
    if (!dev_welford) return MCDX_EINVAL;
    ThreadCtx* ctx;
    if ((st = thread_ctx(&ctx))) return st;
    hipStream_t s = pick_stream(ctx, stream);
    if (c->path_count == 0) {
        const Welford zero = {0, 0.0, 0.0};
        return status_of(hipMemcpyAsync(dev_welford, &zero, sizeof(zero), hipMemcpyHostToDevice, s));
    }
    const uint64_t tiles = tiles_for(c->path_count);
    void* scratch;
    if ((st = enqueue_scratch(ctx, s, tiles * sizeof(Welford), &scratch))) return st;
    hipError_t e = enqueue_contract(*c, precision, (Welford*)scratch, (Welford*)dev_welford, s);
    // recorded whatever was enqueued, so a later user of the scratch waits for it even when a
    // launch reported an error after an earlier kernel of this call was queued
    const int done = enqueue_scratch_done(ctx, s);
    if (e != hipSuccess) return status_of(e);
    return done;
}

int mcdx_price_gbm_enqueue_noalloc(const mcdx_gbm_contract* c, int precision, void* stream,
                                   void* dev_welford) {
    int st = validate(c, precision);
    if (st) return st;
    if (!dev_welford) return MCDX_EINVAL;
    const uint64_t tiles = tiles_for(c->path_count);
    if (tiles > kNoAllocMaxTiles) return MCDX_EINVAL;
    Welford* partials;
    if ((st = noalloc_scratch(&partials))) return st;
    hipStream_t s = (hipStream_t)stream; // NULL: the device's null stream
    hipError_t e;
    if (tiles == 0) // a merge over no tiles writes the empty accumulator {0, 0, 0}
        e = launch_final_merge(partials, nullptr, 0, 1, (Welford*)dev_welford, s);
    else
        e = enqueue_contract(*c, precision, partials, (Welford*)dev_welford, s);
    return status_of(e);
}

int mcdx_gbm_path_samples(const mcdx_gbm_contract* c, int precision, uint64_t path0, uint64_t n,
                          double* out) {
    int st = validate(c, precision);
    if (st) return st;
    if (!out && n) return MCDX_EINVAL;
    if (n == 0) return MCDX_OK;
    if (n > (1ull << 28)) return MCDX_EINVAL;
    ThreadCtx* ctx;
    if ((st = thread_ctx(&ctx))) return st;
    if ((st = ensure_dev(ctx, n * sizeof(double)))) return st;
    GbmArgs a = make_args(*c);
    a.samples = (double*)ctx->dev;
    a.path0 = path0;
    const int kind = kind_of(*c);
    const bool anti = c->antithetic != 0;
    if (precision == MCDX_PREC_FP64) {
        if (anti) launch_samples_k<true, true>(kind, a, n, ctx->stream);
        else launch_samples_k<true, false>(kind, a, n, ctx->stream);
    } else {
        if (anti) launch_samples_k<false, true>(kind, a, n, ctx->stream);
        else launch_samples_k<false, false>(kind, a, n, ctx->stream);
    }
    hipError_t e = launch_check();
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, ctx->dev, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return status_of(e);
}

int mcdx_monte_carlo_european(double spot, double strike, double rate, double carry_yield,
                              double vol, double time, uint8_t type, uint64_t path_count,
                              uint64_t seed, const mcdx_options* options, mcdx_result* out) {
    mcdx_gbm_contract c = base_contract(spot, strike, rate, carry_yield, vol, time, type,
                                        path_count, seed, options);
    c.product = MCDX_EUROPEAN;
    return price_gbm_sync(&c, mcdx_get_precision(), out);
}

int mcdx_monte_carlo_digital(double spot, double strike, double rate, double carry_yield,
                             double vol, double time, uint8_t type, uint8_t style,
                             double cash_amount, uint64_t path_count, uint64_t seed,
                             const mcdx_options* options, mcdx_result* out) {
    mcdx_gbm_contract c = base_contract(spot, strike, rate, carry_yield, vol, time, type,
                                        path_count, seed, options);
    c.product = MCDX_DIGITAL;
    c.digital_style = style;
    c.cash_amount = cash_amount;
    return price_gbm_sync(&c, mcdx_get_precision(), out);
}

int mcdx_monte_carlo_asian(double spot, double strike, double rate, double carry_yield,
                           double vol, double time, uint8_t type, uint8_t strike_style,
                           uint8_t average_style, int32_t monitoring_points,
                           uint64_t path_count, uint64_t seed, const mcdx_options* options,
                           mcdx_result* out) {
    mcdx_gbm_contract c = base_contract(spot, strike, rate, carry_yield, vol, time, type,
                                        path_count, seed, options);
    c.product = MCDX_ASIAN;
    c.strike_style = strike_style;
    c.average_style = average_style;
    c.monitoring_points = monitoring_points;
    return price_gbm_sync(&c, mcdx_get_precision(), out);
}

int mcdx_monte_carlo_barrier(double spot, double strike, double barrier, double rate,
                             double carry_yield, double vol, double time, uint8_t type,
                             uint8_t direction, uint8_t knock, double rebate,
                             int32_t monitoring_points, uint64_t path_count, uint64_t seed,
                             const mcdx_options* options, mcdx_result* out) {
    mcdx_gbm_contract c = base_contract(spot, strike, rate, carry_yield, vol, time, type,
                                        path_count, seed, options);
    c.product = MCDX_BARRIER;
    c.barrier = barrier;
    c.direction = direction;
    c.knock = knock;
    c.rebate = rebate;
    c.monitoring_points = monitoring_points;
    return price_gbm_sync(&c, mcdx_get_precision(), out);
}

int mcdx_monte_carlo_lookback(double spot, double strike, double rate, double carry_yield,
                              double vol, double time, uint8_t type, uint8_t style,
                              int32_t monitoring_points, uint64_t path_count, uint64_t seed,
                              const mcdx_options* options, mcdx_result* out) {
    mcdx_gbm_contract c = base_contract(spot, strike, rate, carry_yield, vol, time, type,
                                        path_count, seed, options);
    c.product = MCDX_LOOKBACK;
    c.strike_style = style;
    c.monitoring_points = monitoring_points;
    return price_gbm_sync(&c, mcdx_get_precision(), out);
}

} // extern "C"
