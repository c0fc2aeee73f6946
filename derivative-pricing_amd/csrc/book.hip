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
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <string>
#include <map>
#include <tuple>
#include <mutex>
#include <vector>

#include "gbm_paths.hpp"

namespace mcdx {

// A shared-draw group: contracts of one kind, seed and step count (path counts descending),
// evaluated on the same paths from one normal per step. A price group holds up to
// kFusePrice contracts. A Greeks group holds up to kUnitCap units, a unit being a contract and
// its four bumped scenarios (spot up/down, vol up/down) as five consecutive members; each
// unit's per-path central differences go to three result slots and, for a European or Asian
// contract (no control variate), the pathwise delta and vega of the unbumped member
// (pathwise.cpp:28-147) to two more. How many units share one group's draws (the fusion
// width) is a schedule parameter: results do not depend on it.
constexpr int kFusePrice = 8;
constexpr int kUnitCap = 3;
constexpr int kUnitMembers = 5;
constexpr int kFuseCap = kUnitCap * kUnitMembers; // 15 >= kFusePrice
constexpr int kGreekSlots = 5; // FD delta, gamma, vega; pathwise delta, vega

enum PathwiseKind : int32_t { kPwNone = 0, kPwTerminal = 1, kPwAverage = 2 };

struct FusedGroup {
    uint32_t c[kFuseCap];
    uint32_t k, m;
    uint32_t units;              // Greek units (0: a price group)
    int32_t kind;
    uint32_t gslot[kUnitCap];    // each unit's first Greek slot
    double inv_2hs[kUnitCap], inv_hs2[kUnitCap], inv_2hv[kUnitCap];
    int32_t pw[kUnitCap];        // PathwiseKind of each unit's base contract
    double pw_sqrt_t[kUnitCap], pw_vol_t[kUnitCap]; // European: sqrt(T), vol T
    float pw_sdt[kUnitCap], pw_vdt[kUnitCap];       // Asian: sqrt(dt), vol dt
};

// Member 0's pathwise estimator: the reference has one for European calls and puts
// (pathwise.cpp:28-62) and for Asians (:64-147); none with a control variate.
inline int32_t pathwise_kind(const mcdx_gbm_contract& c) {
    if (c.product == MCDX_EUROPEAN) return kPwTerminal;
    if (c.product == MCDX_ASIAN && !c.control_variate) return kPwAverage;
    return kPwNone;
}

} // namespace mcdx

struct mcdx_book {
    uint32_t n = 0;                             // user contracts (result slots 0..n-1)
    uint32_t slots = 0;                         // + bumped scenarios + Greek slots
    uint32_t flags = 0;
    int precision = MCDX_PREC_FP32;
    uint64_t total_tiles = 0;
    std::vector<mcdx_gbm_contract> contracts;   // per slot (Greek slots copy their contract)
    std::vector<uint64_t> tile_offsets;         // [slots+1]
    std::vector<double> item_cost;              // per slot, cost of one tile
    std::vector<int32_t> group_of;              // per slot: fused group or -1
    std::vector<mcdx::FusedGroup> groups;
    std::vector<uint32_t> greek_slot;           // per user contract (greeks books)
    uint64_t n_items = 0;                       // plain items, then the group items of each
    uint64_t n_fused[4] = {0, 0, 0, 0};         // family (terminal, Asian, barrier, lookback)
    int max_fused_blocks_per_cu[4] = {1, 1, 1, 1}; // of the kernel the current width uses
    // device
    mcdx::GbmArgs* d_args = nullptr;
    uint8_t* d_kind = nullptr;                  // kind | (antithetic << 4)
    uint64_t* d_toff = nullptr;
    uint64_t* d_items64 = nullptr;              // (contract << 40) | tile, or fused (bit 63)
    mcdx::FusedGroup* d_groups = nullptr;
    mcdx::Welford* d_partials = nullptr;
    mcdx::Welford* d_results = nullptr;
    // schedule
    uint32_t blocks_per_cu = 0;                 // 0 = occupancy maximum
    uint32_t order = 1;                         // 0 natural, 1 longest tiles first
    uint32_t width = 2;                         // Greek units per shared-draw group (1..3)
    uint32_t tiles_per_item = 1;                // consecutive tiles one block takes per item
    std::vector<double> hs, hv;                 // Greeks books: bump sizes per contract
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
__device__ __forceinline__ double book_sample(const GbmArgs& a, int kind, bool anti, uint64_t path,
                                              uint32_t phh) {
    switch (kind) {
    case kTerminal: return anti ? path_sample<kTerminal, FP64, true>(a, path, phh) : path_sample<kTerminal, FP64, false>(a, path, phh);
    case kAsian: return anti ? path_sample<kAsian, FP64, true>(a, path, phh) : path_sample<kAsian, FP64, false>(a, path, phh);
    case kBarrier: return anti ? path_sample<kBarrier, FP64, true>(a, path, phh) : path_sample<kBarrier, FP64, false>(a, path, phh);
    case kBarrierBB: return anti ? path_sample<kBarrierBB, FP64, true>(a, path, phh) : path_sample<kBarrierBB, FP64, false>(a, path, phh);
    default: return anti ? path_sample<kLookback, FP64, true>(a, path, phh) : path_sample<kLookback, FP64, false>(a, path, phh);
    }
}

__device__ __forceinline__ void welford_add_k(Welford& w, double x, int k) {
    w.n += 1;
    const double d = x - w.mean;
    w.mean += d * (1.0 / (double)(k + 1));
    w.m2 += d * (x - w.mean);
}

// One tile of a shared-draw group with at most M members (U = 0: a price group) or U Greek
// units (M = 5U). Each member's state follows path_f32's arithmetic step for step (the same
// normals, the same log2-space update, observe and settle functions), so each member's tile
// partial equals the one its own launch writes. In a Greeks group only the unit's base
// contract keeps a price accumulator (the bumped scenarios enter through the differences).
template <int KIND, int M, int U>
__device__ __forceinline__ void fused_item(const GbmArgs* __restrict__ args, const FusedGroup& g,
                                           uint64_t tile, const uint64_t* __restrict__ toff,
                                           Welford* __restrict__ partials, Welford* lds4) {
#pragma clang fp contract(fast)
    static_assert(M <= kFuseCap && (U == 0 || M == kUnitMembers * U), "group shape");
    constexpr int NW = U == 0 ? M : U; // price accumulators
    const GbmArgs* A[M];
    int kact = 0;
#pragma unroll
    for (int i = 0; i < M; ++i) {
        A[i] = args + g.c[(uint32_t)i < g.k ? i : 0];
        if ((uint32_t)i < g.k && A[i]->n_paths > tile * kTilePaths) kact = i + 1;
    }
    const uint32_t k0 = (uint32_t)A[0]->seed, k1 = (uint32_t)(A[0]->seed >> 32);
    Welford w[NW], gw[U > 0 ? U : 1][kGreekSlots];
#pragma unroll
    for (int i = 0; i < NW; ++i) w[i] = Welford{0, 0.0, 0.0};
#pragma unroll
    for (int u = 0; u < (U > 0 ? U : 1); ++u)
#pragma unroll
        for (int j = 0; j < kGreekSlots; ++j) gw[u][j] = Welford{0, 0.0, 0.0};
#pragma unroll 1
    for (int k = 0; k < kPathsPerLane; ++k) {
        // base is a multiple of 256 and threadIdx.x < 256: every lane's path >> 32 is base >> 32.
        const uint64_t base = tile * kTilePaths + (uint64_t)(k * kTileLanes);
        const uint64_t path = base + threadIdx.x;
        const uint32_t plo = (uint32_t)path, phh = (uint32_t)(base >> 32);
        double v[M];
        double pwd[U > 0 ? U : 1], pwv[U > 0 ? U : 1]; // each unit's pathwise samples
#pragma unroll
        for (int u = 0; u < (U > 0 ? U : 1); ++u) pwd[u] = pwv[u] = 0.0;
        if constexpr (KIND == kTerminal) {
            const float z = inv_normal_f32(philox_w0(plo, phh, 0u, 0u, k0, k1));
#pragma unroll
            for (int i = 0; i < M; ++i)
                v[i] = i < kact ? A[i]->discount *
                                      terminal_payoff(*A[i], A[i]->spot * (double)fast_exp2(
                                                                 fmaf(A[i]->sig2, z, A[i]->mu2)))
                                : 0.0;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (g.pw[u] == kPwTerminal) { // pathwise.cpp:38-58 on this path's FP32 draw
                    const GbmArgs& a0 = *A[kUnitMembers * u];
                    const double st = a0.spot * (double)fast_exp2(fmaf(a0.sig2, z, a0.mu2));
                    const double phi = a0.type == 0 ? 1.0 : -1.0;
                    const double slope = phi * (st - a0.strike) > 0.0 ? phi : 0.0;
                    pwd[u] = a0.discount * slope * (st / a0.spot);
                    pwv[u] = a0.discount * slope * (st * (g.pw_sqrt_t[u] * (double)z - g.pw_vol_t[u]));
                }
            }
        } else {
            StateF32 s[M];
            // Step constants, read once per item for every member (A[i] is always a valid
            // contract), so the step loop runs on registers: observe_f32's arithmetic with
            // the member's own constants.
            bool na[M], ng[M];
            float mu2[M], sig2[M];
#pragma unroll
            for (int i = 0; i < M; ++i) {
                s[i] = StateF32{0.f, 0.f, 0.f, 1.f, 0.0, 0.0, false};
                if constexpr (KIND == kBarrier) { // extrema over the monitoring dates only
                    s[i].xmax = -__builtin_inff();
                    s[i].xmin = __builtin_inff();
                }
                na[i] = A[i]->average_style == 0 || A[i]->cv;
                ng[i] = A[i]->average_style == 1 || A[i]->cv;
                mu2[i] = A[i]->mu2;
                sig2[i] = A[i]->sig2;
            }
            // Each unit's pathwise state (Asian): the cumulative normal and step count are the
            // group's (same draws); d ln S / d vol at the last date and its sums over the
            // dates are the unit's (pathwise.cpp:84-97).
            float pw_cz = 0.f, pw_i = 0.f;
            float pw_dlv[U > 0 ? U : 1];
            double pw_ssv[U > 0 ? U : 1], pw_sdv[U > 0 ? U : 1];
#pragma unroll
            for (int u = 0; u < (U > 0 ? U : 1); ++u) {
                pw_dlv[u] = 0.f;
                pw_ssv[u] = pw_sdv[u] = 0.0;
            }
            auto apply = [&](float z) {
#pragma unroll
                for (int i = 0; i < M; ++i) {
                    if (i < kact) {
                        StateF32& t = s[i];
                        t.x = t.x + fmaf(sig2[i], z, mu2[i]);
                        if constexpr (KIND == kAsian) {
                            if (na[i]) {
                                const float uu = fast_exp2(t.x);
                                t.blk_u += uu;
                                t.last_u = uu;
                            }
                            if (ng[i]) t.sum_x += (double)t.x;
                        } else if constexpr (KIND == kBarrier || KIND == kLookback) {
                            t.xmax = fmaxf(t.xmax, t.x);
                            t.xmin = fminf(t.xmin, t.x);
                        }
                    }
                }
                if constexpr (KIND == kAsian && U > 0) {
                    pw_cz += z;
                    pw_i += 1.f;
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        if (g.pw[u] == kPwAverage && kUnitMembers * u < kact) {
                            // d ln S_i / d vol = -vol t_i + sqrt(dt) (z_1 + ... + z_i)
                            pw_dlv[u] = fmaf(g.pw_sdt[u], pw_cz, -g.pw_vdt[u] * pw_i);
                            if (na[kUnitMembers * u])
                                pw_ssv[u] += (double)(s[kUnitMembers * u].last_u * pw_dlv[u]);
                            else
                                pw_sdv[u] += (double)pw_dlv[u];
                        }
                    }
                }
            };
            const int m = (int)g.m;
            constexpr int B = 8;
            int step = 0;
            // flush_f32's block sums, at the same dates as path_f32's
            auto flush = [&]() {
                if constexpr (KIND == kAsian) {
#pragma unroll
                    for (int i = 0; i < M; ++i)
                        if (i < kact && na[i]) flush_f32<kAsian, true>(s[i]);
                }
            };
            for (; step + B <= m; step += B) {
                float z[B];
                normal_block_f32<B>(plo, phh, (uint32_t)step, k0, k1, z);
#pragma unroll
                for (int j = 0; j < B; ++j) apply(z[j]);
                flush();
            }
            for (; step < m; ++step) apply(inv_normal_f32(philox_w0(plo, phh, (uint32_t)step, 0u, k0, k1)));
            flush();
#pragma unroll
            for (int i = 0; i < M; ++i) v[i] = i < kact ? settle_f32<KIND>(*A[i], s[i]) : 0.0;
            if constexpr (KIND == kAsian) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (g.pw[u] == kPwAverage && kUnitMembers * u < kact) {
                        // pathwise.cpp:99-143, averages formed as settle_f32 forms them
                        const StateF32& s0 = s[kUnitMembers * u];
                        const GbmArgs& a0 = *A[kUnitMembers * u];
                        const bool arith = a0.average_style == 0;
                        const double mm = (double)a0.m;
                        const double last = a0.spot * (double)(arith ? s0.last_u : fast_exp2(s0.x));
                        const double avg = arith ? a0.spot * (s0.sum_u / mm)
                                                 : a0.spot * exp2(s0.sum_x / mm);
                        const double davg_s = avg / a0.spot;
                        const double davg_v = arith ? a0.spot * (pw_ssv[u] / mm) : avg * (pw_sdv[u] / mm);
                        const double ref = a0.strike_style == 0 ? avg - a0.strike : last - avg;
                        const double phi = a0.type == 0 ? 1.0 : -1.0;
                        const double slope = phi * ref > 0.0 ? phi : 0.0;
                        if (a0.strike_style == 0) {
                            pwd[u] = a0.discount * (slope * davg_s);
                            pwv[u] = a0.discount * (slope * davg_v);
                        } else {
                            pwd[u] = a0.discount * (slope * (last / a0.spot - davg_s));
                            pwv[u] = a0.discount * (slope * (last * (double)pw_dlv[u] - davg_v));
                        }
                    }
                }
            }
        }
        if constexpr (U == 0) {
#pragma unroll
            for (int i = 0; i < M; ++i)
                if (i < kact && path < A[i]->n_paths) welford_add_k(w[i], v[i], k);
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int b0 = kUnitMembers * u;
                if (b0 < kact && path < A[b0]->n_paths) { // finite_difference.cpp:58-62, path by path
                    welford_add_k(w[u], v[b0], k);
                    welford_add_k(gw[u][0], (v[b0 + 1] - v[b0 + 2]) * g.inv_2hs[u], k);
                    welford_add_k(gw[u][1], ((v[b0 + 1] - 2.0 * v[b0]) + v[b0 + 2]) * g.inv_hs2[u], k);
                    welford_add_k(gw[u][2], (v[b0 + 3] - v[b0 + 4]) * g.inv_2hv[u], k);
                    if (g.pw[u] != kPwNone) {
                        welford_add_k(gw[u][3], pwd[u], k);
                        welford_add_k(gw[u][4], pwv[u], k);
                    }
                }
            }
        }
    }
    if constexpr (U == 0) {
#pragma unroll
        for (int i = 0; i < M; ++i) {
            if (i < kact) { // uniform across the block
                const Welford t = block_merge_256_all(w[i], lds4);
                partials[toff[g.c[i]] + tile] = t;
            }
        }
    } else {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (kUnitMembers * u < kact) { // uniform across the block
                const Welford t = block_merge_256_all(w[u], lds4);
                partials[toff[g.c[kUnitMembers * u]] + tile] = t;
#pragma unroll
                for (int j = 0; j < kGreekSlots; ++j) {
                    const Welford x = block_merge_256_all(gw[u][j], lds4);
                    partials[toff[g.gslot[u] + j] + tile] = x;
                }
            }
        }
    }
}

// Persistent blocks over one payoff family's shared-draw items, round robin like book_kernel:
// an item is (group << 40 | first tile) and covers up to tiles_per_item consecutive tiles of
// the group. One instantiation per family and group shape keeps each at its own register
// budget. FP32 only.
template <int KIND, int M, int U>
__global__ __launch_bounds__(kTileLanes) void book_fused_kernel(const GbmArgs* __restrict__ args,
                                                                const uint64_t* __restrict__ toff,
                                                                const uint64_t* __restrict__ items,
                                                                uint64_t n_items,
                                                                const FusedGroup* __restrict__ groups,
                                                                Welford* __restrict__ partials,
                                                                uint32_t tiles_per_item) {
    __shared__ Welford lds4[4];
    for (uint64_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const uint64_t packed = items[it];
        const FusedGroup g = groups[packed >> 40];
        const uint64_t t0 = packed & ((1ull << 40) - 1);
        const uint64_t tiles = (args[g.c[0]].n_paths + kTilePaths - 1) / kTilePaths;
        const uint64_t t1 = t0 + tiles_per_item < tiles ? t0 + tiles_per_item : tiles; // uniform
        for (uint64_t t = t0; t < t1; ++t) fused_item<KIND, M, U>(args, g, t, toff, partials, lds4);
    }
}

constexpr int kFusedKinds[4] = {kTerminal, kAsian, kBarrier, kLookback};

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
            const uint64_t base = tile * kTilePaths + (uint64_t)(k * kTileLanes);
            const uint64_t path = base + threadIdx.x;
            if (path < a.n_paths) {
                const double x = book_sample<FP64>(a, kind, anti, path, (uint32_t)(base >> 32));
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

// A book schedule: blocks per CU (0 = occupancy maximum), item order, fusion width (Greek
// units per shared-draw group) and consecutive tiles per item. None changes any result.
struct Schedule {
    bool valid = false;
    uint32_t blocks_per_cu = 0, order = 1, width = 2, tiles_per_item = 1;
};

bool schedule_ok(const Schedule& s) {
    return s.blocks_per_cu <= 64 && s.order <= 1 && s.width >= 1 && s.width <= (uint32_t)kUnitCap &&
           (s.tiles_per_item == 1 || s.tiles_per_item == 2 || s.tiles_per_item == 4 ||
            s.tiles_per_item == 8);
}

// Tuned schedules by book shape: this process's, and those of earlier processes kept in a
// text file ($MCDX_BOOK_SCHEDULE_CACHE, else ~/.cache/mcdx/book_schedules.txt; one line per
// tuning, "mcdx-book-schedule 1 <shape key hex> <blocks/CU> <order> <width> <tiles/item>",
// the last line of a shape wins). Lines that do not parse or hold values out of range are
// ignored, so a damaged file falls back to the default schedule.
std::mutex g_tune_mu;
std::map<uint64_t, Schedule> g_tuned;
bool g_cache_read = false;
std::string g_cache_path; // the file g_tuned was read from

std::string cache_path() {
    if (const char* p = getenv("MCDX_BOOK_SCHEDULE_CACHE")) return p;
    if (const char* h = getenv("HOME")) return std::string(h) + "/.cache/mcdx/book_schedules.txt";
    return std::string();
}

// (Re)reads the cache file when it has not been read or its path changed since.
void read_cache_locked() {
    const std::string path = cache_path();
    if (g_cache_read && path == g_cache_path) return;
    g_cache_read = true;
    g_cache_path = path;
    g_tuned.clear();
    if (path.empty()) return;
    FILE* f = fopen(path.c_str(), "r");
    if (!f) return;
    char line[256];
    while (fgets(line, sizeof line, f)) {
        unsigned long long key = 0;
        unsigned ver = 0, bpc = 0, order = 0, width = 0, tpi = 0;
        char tail = 0;
        if (sscanf(line, "mcdx-book-schedule %u %llx %u %u %u %u %c", &ver, &key, &bpc, &order,
                   &width, &tpi, &tail) != 6 || ver != 1)
            continue;
        Schedule sc;
        sc.blocks_per_cu = bpc;
        sc.order = order;
        sc.width = width;
        sc.tiles_per_item = tpi;
        sc.valid = schedule_ok(sc);
        if (sc.valid) g_tuned[(uint64_t)key] = sc;
    }
    fclose(f);
}

Schedule tuned_schedule(uint64_t key) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    read_cache_locked();
    auto it = g_tuned.find(key);
    return it == g_tuned.end() ? Schedule{} : it->second;
}

void remember_schedule(uint64_t key, const Schedule& sc) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    read_cache_locked();
    g_tuned[key] = sc;
    const std::string path = cache_path();
    if (path.empty()) return;
    const size_t slash = path.rfind('/');
    if (slash != std::string::npos && slash > 0) {
        // create the directory chain (best effort: a missing cache only costs a retune)
        std::string dir = path.substr(0, slash);
        for (size_t i = 1; i <= dir.size(); ++i)
            if (i == dir.size() || dir[i] == '/') (void)mkdir(dir.substr(0, i).c_str(), 0755);
    }
    if (FILE* f = fopen(path.c_str(), "a")) {
        fprintf(f, "mcdx-book-schedule 1 %llx %u %u %u %u\n", (unsigned long long)key,
                sc.blocks_per_cu, sc.order, sc.width, sc.tiles_per_item);
        fclose(f);
    }
}

uint64_t shape_key(const mcdx_book& b) {
    // product mix, log2 path-count histogram and step histogram of the user contracts
    // (FNV-1a over the counts), plus the sharing / Greeks flags and the contract count
    uint64_t hist[5 + 41 + 64 + 2] = {0};
    hist[5 + 41 + 64] = b.flags;
    hist[5 + 41 + 64 + 1] = b.n;
    for (uint32_t i = 0; i < b.n; ++i) {
        const mcdx_gbm_contract& c = b.contracts[i];
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

// The shared-draw kernel of family F for the book's group shape: price groups of up to
// kFusePrice members, or Greeks groups of `width` units.
template <int F>
const void* fused_kernel_for(uint32_t units) {
    switch (units) {
    case 0: return (const void*)book_fused_kernel<kFusedKinds[F], kFusePrice, 0>;
    case 1: return (const void*)book_fused_kernel<kFusedKinds[F], kUnitMembers, 1>;
    case 2: return (const void*)book_fused_kernel<kFusedKinds[F], 2 * kUnitMembers, 2>;
    default: return (const void*)book_fused_kernel<kFusedKinds[F], 3 * kUnitMembers, 3>;
    }
}

const void* fused_kernel(int family, uint32_t units) {
    switch (family) {
    case 0: return fused_kernel_for<0>(units);
    case 1: return fused_kernel_for<1>(units);
    case 2: return fused_kernel_for<2>(units);
    default: return fused_kernel_for<3>(units);
    }
}

uint32_t group_units(const mcdx_book& b) {
    return (b.flags & MCDX_BOOK_GREEKS) ? b.width : 0u;
}

// Shared-draw groups for the book's fusion width, and the occupancy of the kernels they use.
hipError_t build_groups(mcdx_book& b) {
    const uint32_t n = b.n;
    const bool greeks = (b.flags & MCDX_BOOK_GREEKS) != 0;
    const bool share = (b.flags & MCDX_BOOK_NO_SHARING) == 0 && b.precision == MCDX_PREC_FP32;
    b.groups.clear();
    std::fill(b.group_of.begin(), b.group_of.end(), -1);
    // key (seed, kind, steps) -> user contracts by path count, descending
    std::map<std::tuple<uint64_t, int, int>, std::vector<uint32_t>> buckets;
    for (uint32_t i = 0; i < n; ++i) {
        const mcdx_gbm_contract& c = b.contracts[i];
        const int kind = kind_of(c);
        if (b.precision != MCDX_PREC_FP32 || c.antithetic || kind == kBarrierBB) continue;
        buckets[{c.seed, kind, kind == kTerminal ? 1 : c.monitoring_points}].push_back(i);
    }
    for (auto& kv : buckets) {
        auto& mem = kv.second;
        std::stable_sort(mem.begin(), mem.end(), [&](uint32_t x, uint32_t y) {
            return b.contracts[x].path_count > b.contracts[y].path_count;
        });
        const int kind = std::get<1>(kv.first);
        const uint32_t m = (uint32_t)std::get<2>(kv.first);
        if (greeks) {
            for (size_t o = 0; o < mem.size(); o += b.width) {
                FusedGroup g{};
                g.kind = kind;
                g.m = m;
                g.units = (uint32_t)std::min<size_t>(b.width, mem.size() - o);
                g.k = kUnitMembers * g.units;
                const int32_t gi = (int32_t)b.groups.size();
                for (uint32_t u = 0; u < g.units; ++u) {
                    const uint32_t i = mem[o + u];
                    const mcdx_gbm_contract& c = b.contracts[i];
                    g.c[kUnitMembers * u] = i;
                    for (int j = 0; j < 4; ++j) g.c[kUnitMembers * u + 1 + j] = n + 4 * i + j;
                    g.gslot[u] = b.greek_slot[i];
                    g.inv_2hs[u] = 1.0 / (2.0 * b.hs[i]);
                    g.inv_hs2[u] = 1.0 / (b.hs[i] * b.hs[i]);
                    g.inv_2hv[u] = 1.0 / (2.0 * b.hv[i]);
                    g.pw[u] = pathwise_kind(c);
                    const double dt = kind == kTerminal ? c.time : c.time / (double)c.monitoring_points;
                    g.pw_sqrt_t[u] = sqrt(c.time);
                    g.pw_vol_t[u] = c.vol * c.time;
                    g.pw_sdt[u] = (float)sqrt(dt);
                    g.pw_vdt[u] = (float)(c.vol * dt);
                    for (int j = 0; j < kUnitMembers; ++j) b.group_of[g.c[kUnitMembers * u + j]] = gi;
                    for (int j = 0; j < kGreekSlots; ++j) b.group_of[g.gslot[u] + j] = -2;
                }
                b.groups.push_back(g);
            }
        } else if (share && mem.size() >= 2) {
            for (size_t o = 0; o < mem.size(); o += kFusePrice) {
                FusedGroup g{};
                g.k = (uint32_t)std::min<size_t>(kFusePrice, mem.size() - o);
                if (g.k < 2) break; // a lone tail member stays a plain item
                for (uint32_t j = 0; j < g.k; ++j) g.c[j] = mem[o + j];
                g.kind = kind;
                g.m = m;
                for (uint32_t j = 0; j < g.k; ++j) b.group_of[g.c[j]] = (int32_t)b.groups.size();
                b.groups.push_back(g);
            }
        }
    }
    const uint32_t units = group_units(b);
    hipError_t e = hipSuccess;
    for (int f = 0; f < 4 && e == hipSuccess; ++f) {
        int occf = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occf, fused_kernel(f, units), kTileLanes, 0);
        b.max_fused_blocks_per_cu[f] = occf > 0 ? occf : 1;
    }
    return e;
}

int upload_groups(mcdx_book& b) {
    if (b.groups.empty()) return MCDX_OK;
    return status_of(hipMemcpy(b.d_groups, b.groups.data(), b.groups.size() * sizeof(FusedGroup),
                               hipMemcpyHostToDevice));
}

int upload_items(mcdx_book& b) {
    std::vector<uint64_t> items;
    std::vector<double> cost;
    items.reserve(b.total_tiles);
    for (uint32_t c = 0; c < b.slots; ++c) {
        if (b.group_of[c] != -1) continue; // in a shared-draw group (>= 0) or a Greek slot (-2)
        for (uint64_t t = 0; t < b.tile_offsets[c + 1] - b.tile_offsets[c]; ++t) {
            items.push_back(((uint64_t)c << 40) | t);
            cost.push_back(b.item_cost[c]);
        }
    }
    std::vector<uint64_t> fused[4];
    std::vector<double> fcost[4];
    const uint64_t tpi = b.tiles_per_item ? b.tiles_per_item : 1;
    for (uint32_t gi = 0; gi < b.groups.size(); ++gi) {
        const FusedGroup& g = b.groups[gi];
        int f = 0;
        while (kFusedKinds[f] != g.kind) ++f;
        const uint64_t tiles = tiles_for(b.contracts[g.c[0]].path_count);
        const double c1 = (double)g.m * (1.0 + 0.12 * (double)g.k);
        for (uint64_t t = 0; t < tiles; t += tpi) {
            fused[f].push_back(((uint64_t)gi << 40) | t);
            fcost[f].push_back(c1 * (double)std::min<uint64_t>(tpi, tiles - t));
        }
    }
    auto order_by_cost = [&](std::vector<uint64_t>& it, const std::vector<double>& c) {
        if (b.order != 1) return;
        std::vector<uint32_t> idx(it.size());
        for (uint32_t i = 0; i < idx.size(); ++i) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](uint32_t x, uint32_t y) { return c[x] > c[y]; });
        std::vector<uint64_t> sorted(it.size());
        for (size_t i = 0; i < idx.size(); ++i) sorted[i] = it[idx[i]];
        it.swap(sorted);
    };
    order_by_cost(items, cost);
    b.n_items = items.size();
    for (int f = 0; f < 4; ++f) {
        order_by_cost(fused[f], fcost[f]);
        b.n_fused[f] = fused[f].size();
        items.insert(items.end(), fused[f].begin(), fused[f].end());
    }
    hipError_t e = hipMemcpy(b.d_items64, items.data(), items.size() * sizeof(uint64_t),
                             hipMemcpyHostToDevice);
    return status_of(e);
}

// Grid of the plain kernel (family < 0) or of one family's shared-draw kernel.
int grid_blocks(const mcdx_book& b, int family = -1) {
    const uint32_t lim = family >= 0 ? (uint32_t)b.max_fused_blocks_per_cu[family]
                                     : (uint32_t)b.max_blocks_per_cu;
    const uint32_t per_cu = b.blocks_per_cu ? std::min<uint32_t>(b.blocks_per_cu, lim) : lim;
    const uint64_t g = (uint64_t)per_cu * (uint64_t)b.cus;
    const uint64_t n = family >= 0 ? b.n_fused[family] : b.n_items;
    return (int)std::max<uint64_t>(1, std::min<uint64_t>(g, n));
}

hipError_t launch_family(const mcdx_book& b, int f, uint64_t offset, hipStream_t s) {
    if (!b.n_fused[f]) return hipSuccess;
    const GbmArgs* args = b.d_args;
    const uint64_t* toff = b.d_toff;
    const uint64_t* items = b.d_items64 + offset;
    uint64_t n_items = b.n_fused[f];
    const FusedGroup* groups = b.d_groups;
    Welford* partials = b.d_partials;
    uint32_t tpi = b.tiles_per_item;
    void* params[] = {&args, &toff, &items, &n_items, &groups, &partials, &tpi};
    return hipLaunchKernel(fused_kernel(f, group_units(b)), dim3(grid_blocks(b, f)),
                           dim3(kTileLanes), params, 0, s);
}

hipError_t enqueue_raw(mcdx_book& b, hipStream_t s) {
    if (b.n_items) {
        const int grid = grid_blocks(b);
        if (b.precision == MCDX_PREC_FP64)
            hipLaunchKernelGGL(book_kernel<true>, dim3(grid), dim3(kTileLanes), 0, s, b.d_args,
                               b.d_kind, b.d_toff, b.d_items64, b.n_items, b.d_partials);
        else
            hipLaunchKernelGGL(book_kernel<false>, dim3(grid), dim3(kTileLanes), 0, s, b.d_args,
                               b.d_kind, b.d_toff, b.d_items64, b.n_items, b.d_partials);
    }
    uint64_t off = b.n_items;
    for (int f = 0; f < 4; ++f) {
        hipError_t le = launch_family(b, f, off, s);
        if (le != hipSuccess) return le;
        off += b.n_fused[f];
    }
    hipError_t e = launch_check();
    if (e != hipSuccess) return e;
    return launch_final_merge(b.d_partials, b.d_toff, 0, b.slots, b.d_results, s);
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
    void* ptrs[] = {b->d_args, b->d_kind, b->d_toff, b->d_items64, b->d_groups, b->d_partials,
                    b->d_results};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

} // namespace
} // namespace mcdx

using namespace mcdx;

namespace {
// The Greek slots of every contract (kGreekSlots each, consecutive from greek_slot[0]).
int book_greek_results(mcdx_book* b, std::vector<Welford>& h) {
    hipStream_t s = b->last_stream ? b->last_stream : b->stream;
    h.resize((size_t)kGreekSlots * b->n);
    hipError_t e = hipStreamSynchronize(s);
    if (e == hipSuccess)
        e = hipMemcpy(h.data(), b->d_results + b->greek_slot[0], h.size() * sizeof(Welford),
                      hipMemcpyDeviceToHost);
    return status_of(e);
}
} // namespace

extern "C" {

int mcdx_book_create_ex(const mcdx_gbm_contract* contracts, uint32_t n, int precision,
                        uint32_t flags, mcdx_book** out) {
    if (!out || (!contracts && n) || n == 0 || n > (1u << 22) || (flags & ~3u)) return MCDX_EINVAL;
    *out = nullptr;
    const bool greeks = (flags & MCDX_BOOK_GREEKS) != 0;
    const bool share = (flags & MCDX_BOOK_NO_SHARING) == 0 && precision == MCDX_PREC_FP32;
    for (uint32_t i = 0; i < n; ++i) {
        const int st = validate(&contracts[i], precision);
        if (st) return st;
        const int kind = kind_of(contracts[i]);
        if (greeks && (precision != MCDX_PREC_FP32 || contracts[i].antithetic ||
                       kind == kBarrierBB || contracts[i].spot <= 0.0 || contracts[i].vol <= 0.0))
            return MCDX_EINVAL;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return MCDX_ENODEV;
    mcdx_book* b = new mcdx_book();
    b->n = n;
    b->flags = flags;
    b->precision = precision;
    b->contracts.assign(contracts, contracts + n);
    // Greeks: four bumped scenarios per contract (default_bump_sizes,
    // finite_difference.cpp:31-38), then kGreekSlots result slots (FD delta, gamma, vega;
    // pathwise delta, vega).
    std::vector<double>& hs = b->hs;
    std::vector<double>& hv = b->hv;
    hs.assign(n, 0.0);
    hv.assign(n, 0.0);
    if (greeks) {
        for (uint32_t i = 0; i < n; ++i) {
            const mcdx_gbm_contract c = contracts[i];
            hs[i] = std::max(c.spot * 0.01, 1e-4);
            hv[i] = std::max(c.vol * 0.01, 1e-4);
            mcdx_gbm_contract v = c;
            v.spot = c.spot + hs[i];
            b->contracts.push_back(v);
            v.spot = c.spot - hs[i];
            b->contracts.push_back(v);
            v = c;
            v.vol = c.vol + hv[i];
            b->contracts.push_back(v);
            v.vol = c.vol - hv[i];
            b->contracts.push_back(v);
        }
        b->greek_slot.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            b->greek_slot[i] = (uint32_t)b->contracts.size();
            for (int j = 0; j < kGreekSlots; ++j) b->contracts.push_back(contracts[i]);
        }
    }
    const uint32_t slots = (uint32_t)b->contracts.size();
    b->slots = slots;
    b->tile_offsets.resize(slots + 1);
    b->item_cost.resize(slots);
    b->group_of.assign(slots, -1);
    std::vector<GbmArgs> args(slots);
    std::vector<uint8_t> kinds(slots);
    uint64_t acc = 0;
    for (uint32_t i = 0; i < slots; ++i) {
        const mcdx_gbm_contract& c = b->contracts[i];
        b->tile_offsets[i] = acc;
        // a bumped scenario of a Greeks book has no price of its own to merge: its paths
        // enter only through its unit's differences
        const bool bumped = greeks && i >= n && i < 5 * n;
        acc += bumped ? 0 : tiles_for(c.path_count);
        args[i] = make_args(c);
        const int kind = kind_of(c);
        kinds[i] = (uint8_t)(kind | (c.antithetic ? 16 : 0));
        const double steps = kind == kTerminal ? 1.0 : (double)c.monitoring_points;
        b->item_cost[i] = steps * (c.antithetic ? 1.3 : 1.0) * (kind == kBarrierBB ? 1.5 : 1.0);
    }
    b->tile_offsets[slots] = acc;
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
    (void)share;
    if (e == hipSuccess) {
        const Schedule sc = tuned_schedule(shape_key(*b));
        if (sc.valid) {
            b->blocks_per_cu = sc.blocks_per_cu;
            b->order = sc.order;
            b->width = sc.width;
            b->tiles_per_item = sc.tiles_per_item;
        }
    }
    if (e == hipSuccess) e = build_groups(*b);
    // the most groups any width gives: one unit (Greeks) or two contracts (prices) each
    const size_t ng = std::max<size_t>(1, greeks ? n : n / 2 + 1);

    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = dev_malloc(&b->d_args, slots * sizeof(GbmArgs));
    if (e == hipSuccess) e = dev_malloc(&b->d_kind, slots);
    if (e == hipSuccess) e = dev_malloc(&b->d_toff, (slots + 1) * sizeof(uint64_t));
    if (e == hipSuccess) e = dev_malloc(&b->d_items64, acc * sizeof(uint64_t));
    if (e == hipSuccess) e = dev_malloc(&b->d_groups, ng * sizeof(FusedGroup));
    if (e == hipSuccess) e = dev_malloc(&b->d_partials, acc * sizeof(Welford));
    if (e == hipSuccess) e = dev_malloc(&b->d_results, slots * sizeof(Welford));
    if (e == hipSuccess) e = hipMemcpy(b->d_args, args.data(), slots * sizeof(GbmArgs), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d_kind, kinds.data(), slots, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_toff, b->tile_offsets.data(), (slots + 1) * sizeof(uint64_t), hipMemcpyHostToDevice);
    int st = status_of(e);
    if (!st) st = upload_groups(*b);
    if (!st) st = upload_items(*b);
    if (!st) st = rebuild_graph(*b);
    if (st) {
        free_book(b);
        return st;
    }
    *out = b;
    return MCDX_OK;
}

int mcdx_book_create(const mcdx_gbm_contract* contracts, uint32_t n, int precision,
                     mcdx_book** out) {
    return mcdx_book_create_ex(contracts, n, precision, 0u, out);
}

int mcdx_book_set_schedule2(mcdx_book* b, uint32_t blocks_per_cu, uint32_t order,
                            uint32_t fusion_width, uint32_t tiles_per_item) {
    if (!b) return MCDX_EINVAL;
    Schedule sc;
    sc.blocks_per_cu = blocks_per_cu;
    sc.order = order;
    sc.width = fusion_width ? fusion_width : b->width;
    sc.tiles_per_item = tiles_per_item ? tiles_per_item : b->tiles_per_item;
    if (!schedule_ok(sc)) return MCDX_EINVAL;
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e == hipSuccess && b->last_stream) e = hipStreamSynchronize(b->last_stream);
    if (e != hipSuccess) return status_of(e);
    const bool regroup = sc.width != b->width && (b->flags & MCDX_BOOK_GREEKS);
    const bool reitem = regroup || sc.order != b->order || sc.tiles_per_item != b->tiles_per_item;
    b->blocks_per_cu = sc.blocks_per_cu;
    b->order = sc.order;
    b->width = sc.width;
    b->tiles_per_item = sc.tiles_per_item;
    int st = MCDX_OK;
    if (regroup) {
        st = status_of(build_groups(*b));
        if (!st) st = upload_groups(*b);
    }
    if (!st && reitem) st = upload_items(*b);
    if (!st) st = rebuild_graph(*b);
    return st;
}

int mcdx_book_set_schedule(mcdx_book* b, uint32_t blocks_per_cu, uint32_t order) {
    if (!b || order > 1) return MCDX_EINVAL;
    return mcdx_book_set_schedule2(b, blocks_per_cu, order, 0, 0);
}

int mcdx_book_get_schedule2(const mcdx_book* b, uint32_t* blocks_per_cu, uint32_t* order,
                            uint32_t* grid, uint32_t* fusion_width, uint32_t* tiles_per_item) {
    if (!b) return MCDX_EINVAL;
    if (blocks_per_cu) *blocks_per_cu = b->blocks_per_cu ? b->blocks_per_cu : (uint32_t)b->max_blocks_per_cu;
    if (order) *order = b->order;
    if (grid) *grid = (uint32_t)grid_blocks(*b);
    if (fusion_width) *fusion_width = b->width;
    if (tiles_per_item) *tiles_per_item = b->tiles_per_item;
    return MCDX_OK;
}

int mcdx_book_get_schedule(const mcdx_book* b, uint32_t* blocks_per_cu, uint32_t* order,
                           uint32_t* grid) {
    return mcdx_book_get_schedule2(b, blocks_per_cu, order, grid, nullptr, nullptr);
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

int mcdx_book_greeks(mcdx_book* b, mcdx_result* out) {
    if (!b || !out || !(b->flags & MCDX_BOOK_GREEKS)) return MCDX_EINVAL;
    std::vector<Welford> h;
    const int st = book_greek_results(b, h);
    if (st) return st;
    for (uint32_t i = 0; i < b->n; ++i)
        for (int j = 0; j < 3; ++j) out[3 * i + j] = finish(h[(size_t)kGreekSlots * i + j]);
    return MCDX_OK;
}

int mcdx_book_pathwise(mcdx_book* b, mcdx_result* out) {
    if (!b || !out || !(b->flags & MCDX_BOOK_GREEKS)) return MCDX_EINVAL;
    std::vector<Welford> h;
    const int st = book_greek_results(b, h);
    if (st) return st;
    for (uint32_t i = 0; i < b->n; ++i) {
        const bool has = pathwise_kind(b->contracts[i]) != kPwNone;
        for (int j = 0; j < 2; ++j)
            out[2 * i + j] = has ? finish(h[(size_t)kGreekSlots * i + 3 + j])
                                 : mcdx_result{NAN, NAN, 0};
    }
    return MCDX_OK;
}

int mcdx_book_price(mcdx_book* b, mcdx_result* out) {
    int st = mcdx_book_enqueue(b, nullptr);
    if (st) return st;
    return mcdx_book_results(b, out);
}

// Searches the schedule one dimension at a time, starting from the current schedule: the
// fusion width (Greeks books: 1, 2 or 3 units per group), the tiles per item (1, 2, 4), then
// blocks per CU (powers of two up to the occupancy limit) with each item order. Every
// candidate's replay is timed with HIP events on the book's stream after a warm-up replay;
// the fastest is kept, applied, and remembered for books of the same shape, in this process
// and in the schedule cache file.
int mcdx_book_autotune(mcdx_book* b, float* best_ms) {
    if (!b) return MCDX_EINVAL;
    hipEvent_t t0, t1;
    if (hipEventCreate(&t0) != hipSuccess) return MCDX_EDEVICE;
    if (hipEventCreate(&t1) != hipSuccess) {
        (void)hipEventDestroy(t0);
        return MCDX_EDEVICE;
    }
    Schedule best;
    best.blocks_per_cu = b->blocks_per_cu;
    best.order = b->order;
    best.width = b->width;
    best.tiles_per_item = b->tiles_per_item;
    float best_t = 1e30f;
    int st = MCDX_OK;
    auto trial = [&](const Schedule& sc) {
        if (st) return;
        if ((st = mcdx_book_set_schedule2(b, sc.blocks_per_cu, sc.order, sc.width, sc.tiles_per_item)))
            return;
        float ms = 0.f;
        hipError_t e = hipGraphLaunch(b->exec, b->stream); // warm-up
        if (e == hipSuccess) e = hipEventRecord(t0, b->stream);
        if (e == hipSuccess) e = hipGraphLaunch(b->exec, b->stream);
        if (e == hipSuccess) e = hipEventRecord(t1, b->stream);
        if (e == hipSuccess) e = hipEventSynchronize(t1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
        if (e != hipSuccess) {
            st = status_of(e);
            return;
        }
        if (ms < best_t) {
            best_t = ms;
            best = sc;
        }
    };
    trial(best);
    if (b->flags & MCDX_BOOK_GREEKS) {
        const Schedule base = best;
        for (uint32_t w = 1; w <= (uint32_t)kUnitCap; ++w) {
            Schedule sc = base;
            sc.width = w;
            if (w != base.width) trial(sc);
        }
    }
    {
        const Schedule base = best;
        for (uint32_t tpi : {1u, 2u, 4u}) {
            Schedule sc = base;
            sc.tiles_per_item = tpi;
            if (tpi != base.tiles_per_item) trial(sc);
        }
    }
    {
        const Schedule base = best;
        uint32_t top = (uint32_t)b->max_blocks_per_cu;
        for (int f = 0; f < 4; ++f) top = std::max(top, (uint32_t)b->max_fused_blocks_per_cu[f]);
        for (uint32_t order = 0; order <= 1; ++order)
            for (uint32_t bpc = 1; bpc <= top; bpc *= 2) {
                Schedule sc = base;
                sc.order = order;
                sc.blocks_per_cu = bpc;
                trial(sc);
            }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (st) return st;
    if ((st = mcdx_book_set_schedule2(b, best.blocks_per_cu, best.order, best.width,
                                      best.tiles_per_item)))
        return st;
    best.valid = true;
    remember_schedule(shape_key(*b), best);
    if (best_ms) *best_ms = best_t;
    return MCDX_OK;
}

void mcdx_book_destroy(mcdx_book* b) { free_book(b); }

} // extern "C"
