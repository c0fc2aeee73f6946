// Batched book: thousands of GBM contracts with uneven path counts, maturities and payoff
// types priced by one persistent launch, replayed from a captured HIP graph.
//
// Work is split into (contract, tile) items, a tile being the same 1024 consecutive path
// indices every per-contract launch uses (device_common.hpp). Persistent blocks walk the
// item list round robin and write each tile's Welford partial to its contract's slot; the per-contract fixed-order merge then runs exactly as for a single contract. A
// contract's result therefore does not depend on the schedule and equals a per-contract
// launch bit for bit; only the item order and grid size change, and those are what the
// autotuner picks.
//
// The reference prices contracts one call at a time (src/pricers/monte_carlo.cpp:234-338);
// this entry point has no reference counterpart and exists for batched books.
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "gbm_paths.hpp"

struct mcdx_book {
    uint32_t n = 0;
    int precision = MCDX_PREC_FP32;
    uint64_t total_tiles = 0;
    std::vector<mcdx_gbm_contract> contracts;
    std::vector<uint64_t> tile_offsets;         // [n+1]
    std::vector<double> item_cost;              // per contract, cost of one tile
    // device
    mcdx::GbmArgs* d_args = nullptr;
    uint8_t* d_kind = nullptr;                  // kind | (antithetic << 4)
    uint64_t* d_toff = nullptr;
    uint64_t* d_items64 = nullptr;              // (contract << 40) | tile
    mcdx::Welford* d_partials = nullptr;
    mcdx::Welford* d_results = nullptr;
    // schedule
    uint32_t blocks_per_cu = 0;                 // 0 = occupancy maximum
    uint32_t order = 1;                         // 0 natural, 1 longest tiles first
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_valid = false;
    int cus = 256;
    int max_blocks_per_cu = 8;
    hipStream_t last_stream = nullptr;
};

namespace mcdx {

template <bool FP64>
__device__ __forceinline__ double book_sample(const GbmArgs& a, int kind, bool anti, uint64_t path) {
    switch (kind) {
    case kTerminal: return anti ? path_sample<kTerminal, FP64, true>(a, path) : path_sample<kTerminal, FP64, false>(a, path);
    case kAsian: return anti ? path_sample<kAsian, FP64, true>(a, path) : path_sample<kAsian, FP64, false>(a, path);
    case kBarrier: return anti ? path_sample<kBarrier, FP64, true>(a, path) : path_sample<kBarrier, FP64, false>(a, path);
    case kBarrierBB: return anti ? path_sample<kBarrierBB, FP64, true>(a, path) : path_sample<kBarrierBB, FP64, false>(a, path);
    default: return anti ? path_sample<kLookback, FP64, true>(a, path) : path_sample<kLookback, FP64, false>(a, path);
    }
}

// Persistent blocks over the item list, round robin: block b takes items b, b+G, b+2G, ...
// in the schedule's order (longest tiles first spreads the expensive tiles over all blocks).
// The loop trip count is uniform per block and the body has no thread-0-only region, so
// every lane runs every iteration in step with the block's barriers.
template <bool FP64>
__global__ __launch_bounds__(kTileLanes) void book_kernel(const GbmArgs* __restrict__ args,
                                                          const uint8_t* __restrict__ kinds,
                                                          const uint64_t* __restrict__ toff,
                                                          const uint64_t* __restrict__ items,
                                                          uint64_t n_items,
                                                          Welford* __restrict__ partials) {
    __shared__ Welford lds4[4];
    for (uint64_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const uint64_t packed = items[it];
        const uint32_t c = (uint32_t)(packed >> 40);
        const uint64_t tile = packed & ((1ull << 40) - 1);
        const GbmArgs a = args[c];
        const uint8_t kb = kinds[c];
        const int kind = kb & 15;
        const bool anti = (kb >> 4) != 0;
        Welford w = {0, 0.0, 0.0};
#pragma unroll 1
        for (int k = 0; k < kPathsPerLane; ++k) {
            const uint64_t path = tile * kTilePaths + (uint64_t)(k * kTileLanes) + threadIdx.x;
            if (path < a.n_paths) {
                const double x = book_sample<FP64>(a, kind, anti, path);
                w.n += 1;
                const double d = x - w.mean;
                w.mean += d * (1.0 / (double)(k + 1));
                w.m2 += d * (x - w.mean);
            }
        }
        const Welford t = block_merge_256_all(w, lds4);
        partials[toff[c] + tile] = t; // every lane stores the same value
    }
}

namespace {

std::mutex g_tune_mu;
std::map<uint64_t, std::pair<uint32_t, uint32_t>> g_tuned; // book shape -> (blocks/CU, order)

uint64_t shape_key(const mcdx_book& b) {
    // product mix, log2 path-count histogram and step histogram (FNV-1a over the counts)
    uint64_t hist[5 + 41 + 64] = {0};
    for (const auto& c : b.contracts) {
        hist[kind_of(c)]++;
        int lg = 0;
        while ((1ull << (lg + 1)) <= c.path_count && lg < 40) ++lg;
        hist[5 + lg]++;
        const int m = c.product == MCDX_EUROPEAN || c.product == MCDX_DIGITAL ? 1 : c.monitoring_points;
        hist[46 + std::min(63, m / 16)]++;
    }
    uint64_t h = 1469598103934665603ull ^ (uint64_t)b.precision;
    for (uint64_t v : hist) {
        h ^= v;
        h *= 1099511628211ull;
    }
    return h;
}

int upload_items(mcdx_book& b) {
    std::vector<uint64_t> items;
    items.reserve(b.total_tiles);
    for (uint32_t c = 0; c < b.n; ++c)
        for (uint64_t t = 0; t < b.tile_offsets[c + 1] - b.tile_offsets[c]; ++t)
            items.push_back(((uint64_t)c << 40) | t);
    if (b.order == 1)
        std::stable_sort(items.begin(), items.end(), [&](uint64_t x, uint64_t y) {
            return b.item_cost[x >> 40] > b.item_cost[y >> 40];
        });
    hipError_t e = hipMemcpy(b.d_items64, items.data(), items.size() * sizeof(uint64_t),
                             hipMemcpyHostToDevice);
    return status_of(e);
}

int grid_blocks(const mcdx_book& b) {
    const uint32_t per_cu = b.blocks_per_cu ? std::min<uint32_t>(b.blocks_per_cu, b.max_blocks_per_cu)
                                            : (uint32_t)b.max_blocks_per_cu;
    const uint64_t g = (uint64_t)per_cu * (uint64_t)b.cus;
    return (int)std::max<uint64_t>(1, std::min<uint64_t>(g, b.total_tiles));
}

hipError_t enqueue_raw(mcdx_book& b, hipStream_t s) {
    const int grid = grid_blocks(b);
    if (b.precision == MCDX_PREC_FP64)
        hipLaunchKernelGGL(book_kernel<true>, dim3(grid), dim3(kTileLanes), 0, s, b.d_args, b.d_kind,
                           b.d_toff, b.d_items64, b.total_tiles, b.d_partials);
    else
        hipLaunchKernelGGL(book_kernel<false>, dim3(grid), dim3(kTileLanes), 0, s, b.d_args, b.d_kind,
                           b.d_toff, b.d_items64, b.total_tiles, b.d_partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_final_merge(b.d_partials, b.d_toff, 0, b.n, b.d_results, s);
}

int rebuild_graph(mcdx_book& b) {
    if (b.exec) (void)hipGraphExecDestroy(b.exec);
    if (b.graph) (void)hipGraphDestroy(b.graph);
    b.exec = nullptr;
    b.graph = nullptr;
    b.graph_valid = false;
    hipError_t e = hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return status_of(e);
    hipError_t le = enqueue_raw(b, b.stream);
    e = hipStreamEndCapture(b.stream, &b.graph);
    if (le != hipSuccess) return status_of(le);
    if (e != hipSuccess) return status_of(e);
    e = hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return status_of(e);
    b.graph_valid = true;
    return MCDX_OK;
}

void free_book(mcdx_book* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    void* ptrs[] = {b->d_args, b->d_kind, b->d_toff, b->d_items64, b->d_partials, b->d_results};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

} // namespace
} // namespace mcdx

using namespace mcdx;

extern "C" {

int mcdx_book_create(const mcdx_gbm_contract* contracts, uint32_t n, int precision,
                     mcdx_book** out) {
    if (!out || (!contracts && n) || n == 0 || n > (1u << 24)) return MCDX_EINVAL;
    *out = nullptr;
    for (uint32_t i = 0; i < n; ++i) {
        const int st = validate(&contracts[i], precision);
        if (st) return st;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return MCDX_ENODEV;
    mcdx_book* b = new mcdx_book();
    b->n = n;
    b->precision = precision;
    b->contracts.assign(contracts, contracts + n);
    b->tile_offsets.resize(n + 1);
    b->item_cost.resize(n);
    std::vector<GbmArgs> args(n);
    std::vector<uint8_t> kinds(n);
    uint64_t acc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const mcdx_gbm_contract& c = contracts[i];
        b->tile_offsets[i] = acc;
        acc += tiles_for(c.path_count);
        args[i] = make_args(c);
        const int kind = kind_of(c);
        kinds[i] = (uint8_t)(kind | (c.antithetic ? 16 : 0));
        const double steps = kind == kTerminal ? 1.0 : (double)c.monitoring_points;
        b->item_cost[i] = steps * (c.antithetic ? 1.3 : 1.0) * (kind == kBarrierBB ? 1.5 : 1.0);
    }
    b->tile_offsets[n] = acc;
    b->total_tiles = acc;
    if (acc == 0 || acc >= (1ull << 40)) {
        delete b;
        return MCDX_EINVAL;
    }
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev);
    int occ = 0;
    if (e == hipSuccess)
        e = precision == MCDX_PREC_FP64
                ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, book_kernel<true>, kTileLanes, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, book_kernel<false>, kTileLanes, 0);
    b->max_blocks_per_cu = occ > 0 ? occ : 1;
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&b->d_args, n * sizeof(GbmArgs));
    if (e == hipSuccess) e = hipMalloc(&b->d_kind, n);
    if (e == hipSuccess) e = hipMalloc(&b->d_toff, (n + 1) * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(&b->d_items64, acc * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(&b->d_partials, acc * sizeof(Welford));
    if (e == hipSuccess) e = hipMalloc(&b->d_results, n * sizeof(Welford));
    if (e == hipSuccess) e = hipMemcpy(b->d_args, args.data(), n * sizeof(GbmArgs), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d_kind, kinds.data(), n, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_toff, b->tile_offsets.data(), (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice);
    int st = status_of(e);
    if (!st) {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = g_tuned.find(shape_key(*b));
        if (it != g_tuned.end()) {
            b->blocks_per_cu = it->second.first;
            b->order = it->second.second;
        }
    }
    if (!st) st = upload_items(*b);
    if (!st) st = rebuild_graph(*b);
    if (st) {
        free_book(b);
        return st;
    }
    *out = b;
    return MCDX_OK;
}

int mcdx_book_set_schedule(mcdx_book* b, uint32_t blocks_per_cu, uint32_t order) {
    if (!b || order > 1) return MCDX_EINVAL;
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e == hipSuccess && b->last_stream) e = hipStreamSynchronize(b->last_stream);
    if (e != hipSuccess) return status_of(e);
    b->blocks_per_cu = blocks_per_cu;
    if (order != b->order) {
        b->order = order;
        const int st = upload_items(*b);
        if (st) return st;
    }
    return rebuild_graph(*b);
}

int mcdx_book_get_schedule(const mcdx_book* b, uint32_t* blocks_per_cu, uint32_t* order,
                           uint32_t* grid) {
    if (!b) return MCDX_EINVAL;
    if (blocks_per_cu) *blocks_per_cu = b->blocks_per_cu ? b->blocks_per_cu : (uint32_t)b->max_blocks_per_cu;
    if (order) *order = b->order;
    if (grid) *grid = (uint32_t)grid_blocks(*b);
    return MCDX_OK;
}

int mcdx_book_enqueue(mcdx_book* b, void* stream) {
    if (!b || !b->graph_valid) return MCDX_EINVAL;
    hipStream_t s = stream ? (hipStream_t)stream : b->stream;
    b->last_stream = s;
    return status_of(hipGraphLaunch(b->exec, s));
}

int mcdx_book_results(mcdx_book* b, mcdx_result* out) {
    if (!b || !out) return MCDX_EINVAL;
    hipStream_t s = b->last_stream ? b->last_stream : b->stream;
    std::vector<Welford> h(b->n);
    hipError_t e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipMemcpy(h.data(), b->d_results, b->n * sizeof(Welford), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return status_of(e);
    for (uint32_t i = 0; i < b->n; ++i) out[i] = finish(h[i]);
    return MCDX_OK;
}

int mcdx_book_price(mcdx_book* b, mcdx_result* out) {
    int st = mcdx_book_enqueue(b, nullptr);
    if (st) return st;
    return mcdx_book_results(b, out);
}

// Times every candidate schedule (item order x blocks per CU) with HIP events on the book's
// stream, keeps the fastest, and remembers it for books of the same shape in this process.
int mcdx_book_autotune(mcdx_book* b, float* best_ms) {
    if (!b) return MCDX_EINVAL;
    hipEvent_t t0, t1;
    if (hipEventCreate(&t0) != hipSuccess) return MCDX_EDEVICE;
    if (hipEventCreate(&t1) != hipSuccess) {
        (void)hipEventDestroy(t0);
        return MCDX_EDEVICE;
    }
    float best = 1e30f;
    uint32_t best_bpc = b->blocks_per_cu, best_order = b->order;
    int st = MCDX_OK;
    for (uint32_t order = 0; order <= 1 && !st; ++order) {
        for (uint32_t bpc = 1; bpc <= (uint32_t)b->max_blocks_per_cu && !st; bpc *= 2) {
            if ((st = mcdx_book_set_schedule(b, bpc, order))) break;
            float ms = 0.f;
            hipError_t e = hipGraphLaunch(b->exec, b->stream); // warm-up
            if (e == hipSuccess) e = hipEventRecord(t0, b->stream);
            if (e == hipSuccess) e = hipGraphLaunch(b->exec, b->stream);
            if (e == hipSuccess) e = hipEventRecord(t1, b->stream);
            if (e == hipSuccess) e = hipEventSynchronize(t1);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
            if (e != hipSuccess) {
                st = status_of(e);
                break;
            }
            if (ms < best) {
                best = ms;
                best_bpc = bpc;
                best_order = order;
            }
        }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (st) return st;
    if ((st = mcdx_book_set_schedule(b, best_bpc, best_order))) return st;
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        g_tuned[shape_key(*b)] = {best_bpc, best_order};
    }
    if (best_ms) *best_ms = best;
    return MCDX_OK;
}

void mcdx_book_destroy(mcdx_book* b) { free_book(b); }

} // extern "C"
