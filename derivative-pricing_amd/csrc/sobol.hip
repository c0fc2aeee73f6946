// Sobol sequences, Brownian-bridge path construction and the QMC pricers.
//
// Reference behaviour restated here:
//   sobol_point            src/core/sobol.cpp:83-97 (Gray code, x = XOR of v_{b+1} over set
//                          bits b of i^(i>>1), u = x / 2^32, index 0 -> 0)
//   direction numbers      src/core/sobol.cpp:22-68 (dims 0-6: the reference's polynomials
//                          with m_k = 1, Bratley-Fox recurrence)
//   bridge                 src/models/brownian_bridge_path.cpp:20-77 (BFS bisection: draw 0
//                          drives W(T), draw 1 the midpoint, ...)
//   qmc_sobol_european     src/pricers/qmc.cpp:15-30
//   qmc_sobol_asian        src/pricers/qmc.cpp:32-57 (capped at 7 dimensions there)
//
// Dimensions 7 and up have no reference definition. This engine extends the table to 1024
// dimensions: primitive polynomials (LFSR maximal-period test, the reference's own method)
// in order of degree and coefficient bits, skipping the six the reference uses, with
// initial direction integers m_1..m_s odd and < 2^k drawn from Philox key 0x50B01 at
// counter {dim, 0, k, 3 + attempt}. The attempt per dimension (sobol_attempts.inc) was chosen
// offline by sobol/search_initial_numbers.cpp among 256, keeping only attempts whose first
// eight direction integers differ from every earlier dimension's: the one with the smallest
// worst t-value of the two-dimensional projections onto (earlier dimension, this one) at
// 2^12 points, then the smallest sum of 4^t over earlier dimensions and 2^6, 2^9, 2^12 points.
//
// Kernel layout: one lane per Sobol index. A wave covers 64 consecutive indices, so the
// Gray code's bits >= 6 are wave-uniform: x = Xhi(dim) ^ Xlo(dim, lane), where Xhi folds
// up to 26 direction words on the scalar unit and Xlo at most six per lane. The bridge is
// replayed from a host-built depth-first program over a few LDS slots per lane (depth of
// the bisection tree + 2 values live at once), and each bridge point's spot feeds the
// average as soon as it exists, so no lane ever holds the 252-point path.
#include <math.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "qmc_store.hpp"
#include "runtime.hpp"

namespace mcdx {

constexpr unsigned kSobolRefDims = 7;
constexpr unsigned kSobolDims = 1024;
constexpr int kMaxBridgeSlots = 24;

// ---------------------------------------------------------------------------
// Direction tables (host)
// ---------------------------------------------------------------------------
namespace {

struct Poly {
    unsigned degree;
    uint32_t interior; // bit j-1 = a_j (sobol.cpp:14-17)
};
const Poly kRefPolys[kSobolRefDims] = {{0, 0}, {1, 0}, {2, 1}, {3, 1}, {4, 1}, {5, 2}, {6, 1}};

bool primitive(unsigned s, uint32_t interior) {
    const uint32_t poly = (1u << s) | (interior << 1) | 1u;
    const uint32_t period = (1u << s) - 1;
    uint32_t state = 1;
    for (uint32_t i = 1; i <= period; ++i) {
        state <<= 1;
        if (state & (1u << s)) state ^= poly;
        if (state == 1) return i == period;
    }
    return false;
}

void fill_directions(const Poly& p, const uint32_t* m_init, uint32_t* v /* [32], v[k-1] */) {
    uint32_t m[33] = {0};
    const unsigned s = p.degree;
    if (s == 0) {
        for (unsigned k = 1; k <= 32; ++k) m[k] = 1;
    } else {
        for (unsigned k = 1; k <= s; ++k) m[k] = m_init[k - 1];
        for (unsigned k = s + 1; k <= 32; ++k) { // sobol.cpp:52-59
            uint32_t mk = m[k - s] ^ (m[k - s] << s);
            for (unsigned j = 1; j < s; ++j)
                if ((p.interior >> (j - 1)) & 1u) mk ^= (m[k - j] << j);
            m[k] = mk;
        }
    }
    for (unsigned k = 1; k <= 32; ++k) v[k - 1] = m[k] << (32 - k);
}

// Per dimension d >= 7, the Philox attempt of its initial direction integers, chosen offline
// (sobol/search_initial_numbers.cpp) for the best two-dimensional projections with every
// earlier dimension.
const uint8_t kSobolAttempt[kSobolDims] = {
#include "sobol_attempts.inc"
};

std::vector<uint32_t> build_tables() {
    std::vector<uint32_t> v(kSobolDims * 32);
    uint32_t ones[32];
    for (auto& o : ones) o = 1;
    for (unsigned d = 0; d < kSobolRefDims; ++d) fill_directions(kRefPolys[d], ones, &v[d * 32]);
    unsigned d = kSobolRefDims;
    for (unsigned s = 1; s <= 16 && d < kSobolDims; ++s) {
        const uint32_t ncoef = s > 1 ? (1u << (s - 1)) : 1u;
        for (uint32_t c = 0; c < ncoef && d < kSobolDims; ++c) {
            if (!primitive(s, c)) continue;
            bool used = false;
            for (unsigned r = 1; r < kSobolRefDims; ++r)
                used = used || (kRefPolys[r].degree == s && kRefPolys[r].interior == c);
            if (used) continue;
            const uint32_t attempt = kSobolAttempt[d];
            uint32_t m_init[32];
            for (unsigned k = 1; k <= s; ++k) {
                const uint32_t bits = philox_bits(0x50B01ull, d, k, 3 + attempt);
                m_init[k - 1] = (k >= 32 ? bits : (bits & ((1u << k) - 1u))) | 1u;
            }
            fill_directions(Poly{s, c}, m_init, &v[d * 32]);
            ++d;
        }
    }
    return v;
}

std::mutex g_table_mu;
uint32_t* g_dev_tables[64] = {nullptr}; // per device, created once, kept for the process

int device_tables(const uint32_t** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return status_of(e);
    if (dev < 0 || dev >= 64) return MCDX_EDEVICE;
    std::lock_guard<std::mutex> lock(g_table_mu);
    if (!g_dev_tables[dev]) {
        const std::vector<uint32_t> host = build_tables();
        uint32_t* d = nullptr;
        e = dev_malloc(&d, host.size() * sizeof(uint32_t));
        if (e != hipSuccess) return status_of(e);
        e = hipMemcpy(d, host.data(), host.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(d);
            return status_of(e);
        }
        g_dev_tables[dev] = d;
    }
    *out = g_dev_tables[dev];
    return MCDX_OK;
}

} // namespace

// ---------------------------------------------------------------------------
// Bridge program (host): depth-first replay of the reference's BFS bisection
// ---------------------------------------------------------------------------
struct BridgeOp {
    double frac;     // (t_mid - t_left) / (t_right - t_left)
    double sd;       // sqrt((t_mid - t_left)(t_right - t_mid) / (t_right - t_left))
    double drift_t;  // (r - q - vol^2/2) * t_{point+1}
    float frac_f, sd_f, drift2_f;
    uint32_t point;  // the date index (0..n-1) whose bridge value this op produces
    // Whole 32-bit words, so the device reads them through the scalar unit (it has no
    // sub-word loads): the Sobol dimension (= the reference's draw index for this split) and
    // the LDS slots packed as left | right << 8 | output << 16.
    uint32_t dim;
    uint32_t slots;
};

namespace {

struct Interval {
    int l, r;
};

// Returns the number of slots used (slot 0 holds W(0) = 0), or -1 if more than allowed.
int build_bridge(int n, double spot_unused, double rate, double carry, double vol, double time,
                 std::vector<BridgeOp>& ops) {
    (void)spot_unused;
    std::vector<double> t(n + 1);
    for (int i = 0; i <= n; ++i) t[i] = (double)i * time / (double)n; // bridge.cpp:23-26
    // BFS draw assignment exactly as bridge.cpp:33-66.
    std::vector<int> draw_of_mid(n, -1);
    {
        std::vector<Interval> q;
        q.reserve(2 * n + 1);
        size_t head = 0;
        q.push_back({-1, n - 1});
        int draw = 1;
        while (head < q.size() && draw < n) {
            const Interval iv = q[head++];
            if (iv.r - iv.l <= 1) continue;
            const int mid = (iv.l + iv.r) / 2;
            draw_of_mid[mid] = draw++;
            q.push_back({iv.l, mid});
            q.push_back({mid, iv.r});
        }
    }
    const double drift_rate = rate - carry - 0.5 * vol * vol;
    const double log2e = 1.44269504088896340735992468100189214;
    auto t_of = [&](int idx) { return idx == -1 ? 0.0 : t[idx + 1]; };
    // Slot allocator: slot 0 = origin; a node's slot lives until both children are done.
    std::vector<int> free_slots;
    for (int s = kMaxBridgeSlots - 1; s >= 1; --s) free_slots.push_back(s);
    int used = 1;
    bool overflow = false;
    auto alloc = [&]() {
        if (free_slots.empty()) {
            overflow = true;
            return 1;
        }
        const int s = free_slots.back();
        free_slots.pop_back();
        used = s + 1 > used ? s + 1 : used;
        return s;
    };
    auto emit = [&](int dim, int sl, int sr, int so, double frac, double sd, int point) {
        BridgeOp op{};
        op.frac = frac;
        op.sd = sd;
        op.drift_t = drift_rate * t[point + 1];
        op.frac_f = (float)frac;
        op.sd_f = (float)sd;
        op.drift2_f = (float)(op.drift_t * log2e);
        op.dim = (uint32_t)dim;
        op.slots = (uint32_t)sl | (uint32_t)sr << 8 | (uint32_t)so << 16;
        op.point = (uint32_t)point;
        ops.push_back(op);
    };
    // Terminal point: w[n-1] = sqrt(t[n]) * z_0 (bridge.cpp:30).
    const int s_term = alloc();
    emit(0, 0, 0, s_term, 0.0, sqrt(t[n]), n - 1);
    // Depth-first over the same intervals; slot_of[idx+1] tracks live endpoint slots.
    std::vector<int> slot_of(n + 1, -1);
    slot_of[0] = 0;
    slot_of[n] = s_term;
    struct Frame {
        int l, r, stage;
    };
    std::vector<Frame> st;
    st.push_back({-1, n - 1, 0});
    while (!st.empty()) {
        Frame& f = st.back();
        if (f.r - f.l <= 1) {
            st.pop_back();
            continue;
        }
        const int mid = (f.l + f.r) / 2;
        if (f.stage == 0) {
            const double tl = t_of(f.l), tr = t_of(f.r), tm = t_of(mid);
            const double frac = (tm - tl) / (tr - tl);
            const double var = (tm - tl) * (tr - tm) / (tr - tl);
            const int so = alloc();
            emit(draw_of_mid[mid], slot_of[f.l + 1], slot_of[f.r + 1], so, frac, sqrt(var), mid);
            slot_of[mid + 1] = so;
            f.stage = 1;
            st.push_back({f.l, mid, 0});
        } else if (f.stage == 1) {
            f.stage = 2;
            st.push_back({mid, f.r, 0});
        } else {
            free_slots.push_back(slot_of[mid + 1]);
            slot_of[mid + 1] = -1;
            st.pop_back();
        }
    }
    return overflow ? -1 : used;
}

} // namespace

// ---------------------------------------------------------------------------
// Device side
// ---------------------------------------------------------------------------
// XOR of the direction words of `dim` over the set bits of `g`.
__device__ __forceinline__ uint32_t sobol_fold(const uint32_t* __restrict__ v, uint32_t g) {
    uint32_t x = 0;
    while (g) {
        const int b = __builtin_ctz(g);
        x ^= v[b];
        g &= g - 1;
    }
    return x;
}

// Sobol integer of (dim, i) split by who shares each Gray-code bit of i: bits >= 10 are the
// same for a whole 1024-point tile (their XOR per dimension is precomputed once per tile in
// LDS), bits 6..9 are wave-uniform (one wave of 64 consecutive indices) and bits 0..5 vary
// by lane. Per dimension that is one LDS broadcast read, ten direction words read through the
// scalar unit, and ten masked XORs, instead of a loop over the set bits.
struct SobolLane {
    uint32_t lo[6]; // all-ones where Gray bit b (b < 6) of this lane's index is set
    uint32_t mid[4]; // the same for bits 6..9 (wave-uniform values)
};

__device__ __forceinline__ SobolLane sobol_lane(uint32_t g_lo, uint32_t g_mid) {
    SobolLane L;
#pragma unroll
    for (int b = 0; b < 6; ++b) L.lo[b] = 0u - ((g_lo >> b) & 1u);
#pragma unroll
    for (int b = 0; b < 4; ++b) L.mid[b] = 0u - ((g_mid >> b) & 1u);
    return L;
}

__device__ __forceinline__ uint32_t sobol_x(const uint32_t* __restrict__ vd, const SobolLane& L,
                                            uint32_t tile_part) {
    uint32_t x = tile_part;
#pragma unroll
    for (int b = 0; b < 6; ++b) x ^= L.lo[b] & vd[b];
#pragma unroll
    for (int b = 0; b < 4; ++b) x ^= L.mid[b] & vd[6 + b];
    return x;
}

// tdir[d] = XOR of direction words 10.. of dimension d over the set bits of the tile's Gray
// bits (tile_gray = Gray code of the tile index, i.e. Gray bits >= 10 of every index in it).
__device__ __forceinline__ void sobol_tile_parts(const uint32_t* __restrict__ v, uint32_t dims,
                                                 uint32_t tile_gray, uint32_t* tdir) {
    for (uint32_t d = threadIdx.x; d < dims; d += blockDim.x) {
        uint32_t x = 0;
        for (uint32_t g = tile_gray; g; g &= g - 1) x ^= v[32 * d + 10 + __builtin_ctz(g)];
        tdir[d] = x;
    }
}

// u = x / 2^32 (sobol.cpp:96). FP32 mode uses the engine's FP32 inverse normal
// (device_common.hpp) on m = min(u, 1 - u) = min(x, 2^32 - x) / 2^32, formed exactly from the
// integer; Sobol points have no half-integer offset, so 4m = m_int * 2^-30.
__device__ __forceinline__ float sobol_normal_f32(uint32_t x) {
    const bool upper = x >= 0x80000000u;
    const uint32_t mi = upper ? 0u - x : x; // 2^32 - x on the upper half
    const float mf = (float)mi;
    const float m4 = mf * 0x1p-30f;
    const float y = fast_log2(__builtin_fmaf(-0.25f * m4, m4, m4));
    const float za = mi < kTailInt ? tail_poly(y)
                                   : __builtin_fmaf(mf, -0x1p-32f, 0.5f) * central_poly(y);
    return upper ? za : -za;
}

struct QmcArgs {
    double spot, strike, discount, drift, vsd, vol;
    float vol2, pad;       // FP32: vol * log2(e)
    float mu2, sig2;       // FP32 terminal: drift*log2e, vsd*log2e
    uint64_t n_points;     // Sobol indices 1..n_points
    const uint32_t* v;     // direction table [dims][32]
    const BridgeOp* ops;   // bridge program (Asian)
    int32_t n_ops, m, slots;
    uint8_t type, strike_style, european;
    uint8_t product;       // 0 arithmetic Asian, 1 barrier, 2 lookback (over the bridge)
    uint8_t direction, knock;
    double barrier, rebate;
    float barrier_rel;     // FP32: H / S0
    Welford* partials;     // [replicate][tile]
    const uint32_t* shifts; // randomized runs: digital shifts [replicate][dims]
    uint32_t dims;          // dimensions per point (1 for the European)
};

__device__ __forceinline__ double smax_d(double a, double b) { return (a < b) ? b : a; }

// One point of the QMC pricer (index i >= 1). FP64 mode mirrors qmc.cpp / bridge.cpp.
// Products of the QMC kernel: one kernel instance each, so the bridge loop carries only its
// own observation (no per-date product tests).
enum QmcProduct : int { kQmcAsian = 0, kQmcBarrier = 1, kQmcLookback = 2, kQmcEuropean = 3 };

// One point of the QMC pricer (index i >= 1). FP64 mode mirrors qmc.cpp / bridge.cpp.
// Normal of a Sobol integer. Plain runs: u = x / 2^32 (sobol.cpp:96). Randomized runs
// (digitally shifted x): u = (x + 0.5) / 2^32, the Philox uniform's mapping (rng.cpp:34-37),
// so a shifted point never reaches u = 0; the FP32 inverse normal is written for it.
template <bool FP64, bool SHIFTED>
__device__ __forceinline__ double qmc_normal_f64(uint32_t x) {
    return SHIFTED ? inv_normal_f64(uniform_open01_f64(x)) : inv_normal_f64((double)x / 4294967296.0);
}
template <bool SHIFTED>
__device__ __forceinline__ float qmc_normal_f32(uint32_t x) {
    return SHIFTED ? inv_normal_f32(x) : sobol_normal_f32(x);
}

template <bool FP64, int PROD, bool SHIFTED>
__device__ double qmc_point(const QmcArgs& a, const SobolLane& L, const uint32_t* tdir,
                            float* lds_f, double* lds_d, int tid, int nthreads) {
    const double phi = a.type == 0 ? 1.0 : -1.0;
    if constexpr (PROD == kQmcEuropean) { // qmc.cpp:24-28, dimension 0
        const uint32_t x = sobol_x(a.v, L, tdir[0]);
        double st;
        if constexpr (FP64) {
            const double z = qmc_normal_f64<true, SHIFTED>(x);
            st = a.spot * exp(a.drift + a.vsd * z);
        } else {
            const float z = qmc_normal_f32<SHIFTED>(x);
            st = a.spot * (double)fast_exp2(__builtin_fmaf(a.sig2, z, a.mu2));
        }
        return a.discount * smax_d(phi * (st - a.strike), 0.0);
    } else {
        // Asian over the bridge (qmc.cpp:44-54): replay the depth-first program. Barriers and
        // lookbacks observe the same dates (every date once; breach, maximum and minimum do
        // not depend on the visiting order; the program's first point is the last date).
        // FP32 mode observes s / S0 (log2-space state) and scales by S0 at the end.
        double sum = 0.0, last = 0.0;
        double mx = FP64 ? a.spot : 1.0, mn = mx; // lookback extrema seeded at S0
        bool breached = false;
        const bool up = a.direction == 0;
        if constexpr (FP64) lds_d[tid] = 0.0;
        else lds_f[tid] = 0.f;
        for (int k = 0; k < a.n_ops; ++k) {
            const BridgeOp& op = a.ops[k];
            const uint32_t x = sobol_x(a.v + 32 * op.dim, L, tdir[op.dim]);
            const uint32_t sl = op.slots & 0xffu, sr = (op.slots >> 8) & 0xffu, so = op.slots >> 16;
            double s;
            if constexpr (FP64) {
                const double z = qmc_normal_f64<true, SHIFTED>(x);
                const double wl = lds_d[sl * nthreads + tid];
                const double wr = lds_d[sr * nthreads + tid];
                const double mean = wl + op.frac * (wr - wl);
                const double w = mean + op.sd * z;
                lds_d[so * nthreads + tid] = w;
                s = a.spot * exp(op.drift_t + a.vol * w);
                if constexpr (PROD == kQmcBarrier)
                    breached = breached || (up ? s >= a.barrier : s <= a.barrier);
            } else {
                const float z = qmc_normal_f32<SHIFTED>(x);
                const float wl = lds_f[sl * nthreads + tid];
                const float wr = lds_f[sr * nthreads + tid];
                const float w = __builtin_fmaf(op.sd_f, z, __builtin_fmaf(op.frac_f, wr - wl, wl));
                lds_f[so * nthreads + tid] = w;
                const float su = fast_exp2(__builtin_fmaf(a.vol2, w, op.drift2_f));
                if constexpr (PROD == kQmcBarrier)
                    breached = breached || (up ? su >= a.barrier_rel : su <= a.barrier_rel);
                s = (double)su;
            }
            if (k == 0) last = s; // uniform: the first op is the last date
            if constexpr (PROD == kQmcAsian) sum += s;
            if constexpr (PROD == kQmcLookback) {
                mx = smax_d(mx, s);
                mn = (s < mn) ? s : mn;
            }
        }
        if constexpr (!FP64) {
            sum *= a.spot;
            last *= a.spot;
            mx *= a.spot;
            mn *= a.spot;
        }
        if constexpr (PROD == kQmcBarrier) { // barrier.hpp:32-37
            const double vanilla = smax_d(phi * (last - a.strike), 0.0);
            const bool pays = a.knock == 1 ? !breached : breached;
            return a.discount * (pays ? vanilla : a.rebate);
        } else if constexpr (PROD == kQmcLookback) { // lookback.hpp:32-35, 54-56
            double r;
            if (a.strike_style == 0)
                r = a.type == 0 ? smax_d(mx - a.strike, 0.0) : smax_d(a.strike - mn, 0.0);
            else
                r = a.type == 0 ? (last - mn) : (mx - last);
            return a.discount * r;
        } else {
            const double avg = sum / (double)a.m;
            const double ref = a.strike_style == 0 ? (avg - a.strike) : (last - avg);
            return a.discount * smax_d(phi * ref, 0.0);
        }
    }
}

// Tile of 1024 Sobol indices per block, one point per thread: thread t prices index
// tile*1024 + t, and the tile's Welford partial is then rebuilt in the canonical order (lane
// l < 256 folds points l, l+256, l+512, l+768, then the 256-lane tree). The points priced are
// indices 1..n_points (index 0 is skipped, qmc.cpp:24). A wave's 64 indices are 64-aligned,
// so Gray-code bits >= 6 are wave-uniform.
template <bool FP64, int PROD, bool SHIFTED>
__global__ __launch_bounds__(kTilePaths) void qmc_tile_kernel(QmcArgs a) {
    extern __shared__ unsigned char smem[];
    __shared__ double vals[kTilePaths];
    __shared__ unsigned char oks[kTilePaths];
    __shared__ Welford lds4[4];
    __shared__ uint32_t tdir[kSobolDims];
    float* lds_f = (float*)smem;
    double* lds_d = (double*)smem;
    const uint64_t tile = blockIdx.x;
    const uint32_t t = threadIdx.x;
    // The wave's first index, read as wave-uniform so that the branch below and the bridge
    // loop inside it stay scalar (the op table is then read through the scalar unit).
    const uint64_t wave_base = tile * kTilePaths + __builtin_amdgcn_readfirstlane(t & ~63u);
    sobol_tile_parts(a.v, a.dims, (uint32_t)(tile ^ (tile >> 1)), tdir);
    if (SHIFTED) { // replicate blockIdx.y: its digital shift joins the tile part
        __syncthreads();
        for (uint32_t d = t; d < a.dims; d += blockDim.x)
            tdir[d] ^= a.shifts[(uint64_t)blockIdx.y * a.dims + d];
    }
    __syncthreads();
    double x = 0.0;
    bool valid = false;
    if (wave_base <= a.n_points) { // wave-uniform
        const uint64_t i_raw = wave_base + (t & 63u);
        valid = i_raw >= 1 && i_raw <= a.n_points;
        const uint64_t i = valid ? i_raw : (wave_base | 1u); // finite stand-in, discarded
        const uint64_t gray = i ^ (i >> 1);
        const SobolLane L = sobol_lane((uint32_t)(gray & 63u),
                                       __builtin_amdgcn_readfirstlane((uint32_t)((gray >> 6) & 15u)));
        x = qmc_point<FP64, PROD, SHIFTED>(a, L, tdir, lds_f, lds_d, (int)t, kTilePaths);
    }
    vals[t] = x;
    oks[t] = valid ? 1 : 0;
    __syncthreads();
    Welford w = {0, 0.0, 0.0};
    if (t < (uint32_t)kTileLanes) {
#pragma unroll
        for (int k = 0; k < kPathsPerLane; ++k) {
            const uint32_t j = (uint32_t)(k * kTileLanes) + t;
            if (oks[j]) {
                const double xv = vals[j];
                w.n += 1;
                const double d = xv - w.mean;
                w.mean += d / (double)w.n;
                w.m2 += d * (xv - w.mean);
            }
        }
    }
    // block_merge_256's tree over lanes 0..255, in waves 0..3 (wave-uniform)
    const int lane = t & 63;
    if (t < (uint32_t)kTileLanes) {
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            Welford o;
            o.n = __shfl_down(w.n, off, 64);
            o.mean = __shfl_down(w.mean, off, 64);
            o.m2 = __shfl_down(w.m2, off, 64);
            if ((lane & (2 * off - 1)) == 0) w = welford_merge(w, o);
        }
        if (lane == 0) lds4[t >> 6] = w;
    }
    __syncthreads();
    if (t == 0)
        a.partials[(uint64_t)blockIdx.y * gridDim.x + tile] =
            welford_merge(welford_merge(lds4[0], lds4[1]), welford_merge(lds4[2], lds4[3]));
}

__global__ __launch_bounds__(256) void sobol_ints_kernel(const uint32_t* __restrict__ v,
                                                         uint64_t index0, uint64_t n,
                                                         uint32_t* out) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint64_t i = index0 + k;
    out[k] = i == 0 ? 0u : sobol_fold(v, (uint32_t)(i ^ (i >> 1)));
}

} // namespace mcdx

namespace mcdx {
namespace {

template <bool FP64, int PROD, bool SHIFTED>
hipError_t launch_qmc_as(unsigned tiles, unsigned reps, size_t shmem, hipStream_t s,
                         const QmcArgs& a) {
    hipError_t e = hipFuncSetAttribute((const void*)qmc_tile_kernel<FP64, PROD, SHIFTED>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((qmc_tile_kernel<FP64, PROD, SHIFTED>), dim3(tiles, reps), dim3(kTilePaths),
                       shmem, s, a);
    return launch_check();
}

template <bool FP64, bool SHIFTED>
hipError_t launch_qmc_prec(int prod, unsigned tiles, unsigned reps, size_t shmem, hipStream_t s,
                           const QmcArgs& a) {
    switch (prod) {
    case kQmcEuropean: return launch_qmc_as<FP64, kQmcEuropean, SHIFTED>(tiles, reps, shmem, s, a);
    case kQmcBarrier: return launch_qmc_as<FP64, kQmcBarrier, SHIFTED>(tiles, reps, shmem, s, a);
    case kQmcLookback: return launch_qmc_as<FP64, kQmcLookback, SHIFTED>(tiles, reps, shmem, s, a);
    default: return launch_qmc_as<FP64, kQmcAsian, SHIFTED>(tiles, reps, shmem, s, a);
    }
}

hipError_t launch_qmc(bool fp64, bool shifted, int prod, unsigned tiles, unsigned reps,
                      size_t shmem, hipStream_t s, const QmcArgs& a) {
    if (shifted)
        return fp64 ? launch_qmc_prec<true, true>(prod, tiles, reps, shmem, s, a)
                    : launch_qmc_prec<false, true>(prod, tiles, reps, shmem, s, a);
    return fp64 ? launch_qmc_prec<true, false>(prod, tiles, reps, shmem, s, a)
                : launch_qmc_prec<false, false>(prod, tiles, reps, shmem, s, a);
}

// Runs `reps` replicates of the point set (reps = 1 and no shifts: the plain QMC estimate;
// otherwise replicate r digitally shifted by shifts[r * dims ..]) and returns each
// replicate's Welford result in out[r].
int qmc_run(QmcArgs a, int precision, const std::vector<BridgeOp>& ops, int slots,
            uint32_t reps, const std::vector<uint32_t>& shifts, mcdx_result* out) {
    ThreadCtx* ctx;
    int st = thread_ctx(&ctx);
    if (st) return st;
    if ((st = device_tables(&a.v))) return st;
    a.dims = a.european ? 1u : (uint32_t)a.m;
    const bool shifted = !shifts.empty();
    if (shifted && shifts.size() != (size_t)reps * a.dims) return MCDX_EINVAL;
    if (reps == 0 || (!shifted && reps != 1)) return MCDX_EINVAL;
    // Sobol indices 0 .. n_points live in tiles of 1024 (index 0 skipped).
    const uint64_t tiles = (a.n_points + 1 + kTilePaths - 1) / kTilePaths;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t part_bytes = al(((size_t)reps * tiles + reps) * sizeof(Welford));
    const size_t off_bytes = al((size_t)(reps + 1) * sizeof(uint64_t));
    const size_t ops_bytes = al(ops.size() * sizeof(BridgeOp));
    const size_t shift_bytes = al(shifts.size() * sizeof(uint32_t));
    if ((st = ensure_dev(ctx, part_bytes + off_bytes + ops_bytes + shift_bytes))) return st;
    if ((st = ensure_pinned(ctx, off_bytes + shift_bytes + reps * sizeof(Welford)))) return st;
    char* dev = (char*)ctx->dev;
    Welford* partials = (Welford*)dev;
    Welford* merged = partials + (size_t)reps * tiles;
    uint64_t* doffs = (uint64_t*)(dev + part_bytes);
    BridgeOp* dops = (BridgeOp*)(dev + part_bytes + off_bytes);
    uint32_t* dshift = (uint32_t*)(dev + part_bytes + off_bytes + ops_bytes);
    hipError_t e = hipStreamSynchronize(ctx->stream); // the pinned staging is free
    char* h = (char*)ctx->pinned;
    uint64_t* hoffs = (uint64_t*)h;
    for (uint32_t r = 0; r <= reps; ++r) hoffs[r] = (uint64_t)r * tiles;
    if (shifted) memcpy(h + off_bytes, shifts.data(), shifts.size() * sizeof(uint32_t));
    if (e == hipSuccess && reps > 1)
        e = hipMemcpyAsync(doffs, hoffs, (size_t)(reps + 1) * sizeof(uint64_t),
                           hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && !ops.empty())
        e = hipMemcpyAsync(dops, ops.data(), ops.size() * sizeof(BridgeOp), hipMemcpyHostToDevice,
                           ctx->stream);
    if (e == hipSuccess && shifted)
        e = hipMemcpyAsync(dshift, h + off_bytes, shifts.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice, ctx->stream);
    a.ops = dops;
    a.n_ops = (int32_t)ops.size();
    a.slots = slots;
    a.partials = partials;
    a.shifts = shifted ? dshift : nullptr;
    const size_t shmem = (size_t)(slots > 0 ? slots : 1) * kTilePaths *
                         (precision == MCDX_PREC_FP64 ? sizeof(double) : sizeof(float));
    if (shmem + 16 * 1024 > 160 * 1024) return MCDX_EINVAL; // bridge slots beyond LDS (not reached for <= 1024 dates)
    const int prod = a.european ? kQmcEuropean : a.product == 1 ? kQmcBarrier
                                               : a.product == 2 ? kQmcLookback : kQmcAsian;
    const bool fp64 = precision == MCDX_PREC_FP64;
    if (e == hipSuccess)
        e = launch_qmc(fp64, shifted, prod, (unsigned)tiles, reps, shmem, ctx->stream, a);
    if (e == hipSuccess)
        e = launch_final_merge(partials, reps > 1 ? doffs : nullptr, tiles, reps, merged, ctx->stream);
    Welford* hres = (Welford*)(h + off_bytes + shift_bytes);
    if (e == hipSuccess)
        e = hipMemcpyAsync(hres, merged, reps * sizeof(Welford), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return status_of(e);
    for (uint32_t r = 0; r < reps; ++r) out[r] = finish(hres[r]);
    return MCDX_OK;
}

int qmc_run(QmcArgs a, int precision, const std::vector<BridgeOp>& ops, int slots,
            mcdx_result* out) {
    return qmc_run(a, precision, ops, slots, 1, {}, out);
}

QmcArgs qmc_args(double spot, double strike, double rate, double carry, double vol, double time,
                 uint8_t type, uint64_t n) {
    QmcArgs a{};
    const double log2e = 1.44269504088896340735992468100189214;
    a.spot = spot;
    a.strike = strike;
    a.discount = exp(-rate * time);
    a.drift = (rate - carry - 0.5 * vol * vol) * time; // gbm.cpp:8-9 with T
    a.vsd = vol * sqrt(time);
    a.vol = vol;
    a.vol2 = (float)(vol * log2e);
    a.mu2 = (float)(a.drift * log2e);
    a.sig2 = (float)(a.vsd * log2e);
    a.n_points = n;
    a.type = type;
    return a;
}

} // namespace

int qmc_asian(double spot, double strike, double rate, double carry, double vol, double time,
              uint8_t type, uint8_t strike_style, int32_t m, uint64_t n, int precision,
              mcdx_result* out) {
    if (!out || type > 1 || strike_style > 1 || m < 1 || m > (int32_t)kSobolDims) return MCDX_EINVAL;
    if (precision != MCDX_PREC_FP32 && precision != MCDX_PREC_FP64) return MCDX_EINVAL;
    if (n > (1ull << 32) - 2) return MCDX_EINVAL;
    if (n == 0) { // reference: sum / 0 (qmc.cpp:56)
        *out = mcdx_result{NAN, 0.0, 0};
        return MCDX_OK;
    }
    std::vector<BridgeOp> ops;
    const int slots = build_bridge(m, spot, rate, carry, vol, time, ops);
    if (slots < 0) return MCDX_EINVAL;
    QmcArgs a = qmc_args(spot, strike, rate, carry, vol, time, type, n);
    a.m = m;
    a.strike_style = strike_style;
    a.european = 0;
    return qmc_run(a, precision, ops, slots, out);
}

int qmc_path(const mcdx_gbm_contract* c, int precision, mcdx_result* out) {
    if (!out || !c) return MCDX_EINVAL;
    if (precision != MCDX_PREC_FP32 && precision != MCDX_PREC_FP64) return MCDX_EINVAL;
    const int32_t m = c->monitoring_points;
    if (c->type > 1 || c->strike_style > 1 || c->direction > 1 || c->knock > 1 || m < 1 ||
        m > (int32_t)kSobolDims || c->path_count > (1ull << 32) - 2)
        return MCDX_EINVAL;
    uint8_t product;
    if (c->product == MCDX_ASIAN && c->average_style == 0) product = 0;
    else if (c->product == MCDX_BARRIER && !c->brownian_bridge) product = 1;
    else if (c->product == MCDX_LOOKBACK) product = 2;
    else return MCDX_EINVAL;
    if (c->path_count == 0) {
        *out = mcdx_result{NAN, 0.0, 0};
        return MCDX_OK;
    }
    std::vector<BridgeOp> ops;
    const int slots = build_bridge(m, c->spot, c->rate, c->carry_yield, c->vol, c->time, ops);
    if (slots < 0) return MCDX_EINVAL;
    QmcArgs a = qmc_args(c->spot, c->strike, c->rate, c->carry_yield, c->vol, c->time, c->type,
                         c->path_count);
    a.m = m;
    a.strike_style = c->strike_style;
    a.european = 0;
    a.product = product;
    a.direction = c->direction;
    a.knock = c->knock;
    a.barrier = c->barrier;
    a.rebate = c->rebate;
    a.barrier_rel = (float)(c->barrier / c->spot);
    return qmc_run(a, precision, ops, slots, out);
}

// One Sobol index per thread, 256 threads, bridge values in LDS slots [slot][thread]; each
// op's spot goes to its date's store column (coalesced across the block). Block b holds the
// raw indices i = 256 b + t and path p = i - 1 (index 0 has no path), so the Gray-code bits
// of i split as in the pricing tiles: bits >= 8 are the Gray code of b (their XOR per
// dimension, with the run's shift, is formed once per block in LDS), bits 6..7 are
// wave-uniform and bits 0..5 vary by lane. One dimension then costs an LDS read and eight
// masked XORs instead of a loop over the index's set bits.
template <bool FP64, bool SHIFTED>
__global__ __launch_bounds__(256) void qmc_store_kernel(const uint32_t* __restrict__ v,
                                                        const BridgeOp* __restrict__ ops,
                                                        int32_t n_ops,
                                                        const uint32_t* __restrict__ shift,
                                                        double spot, double vol, float vol2,
                                                        uint64_t n, void* store) {
    extern __shared__ unsigned char smem[];
    __shared__ uint32_t tdir[kSobolDims];
    // grid y = run: its shift words and its store follow the previous run's
    shift += (size_t)blockIdx.y * (size_t)n_ops;
    store = (char*)store + (size_t)blockIdx.y * (size_t)n_ops * n * (FP64 ? 8 : 4);
    const uint32_t t = threadIdx.x;
    const uint32_t bg = blockIdx.x ^ (blockIdx.x >> 1); // Gray bits >= 8 of the block's indices
    for (int32_t d = (int32_t)t; d < n_ops; d += (int32_t)blockDim.x) { // dims 0 .. n_ops - 1
        uint32_t x = SHIFTED ? shift[d] : 0u;
        for (uint32_t g = bg; g; g &= g - 1) x ^= v[32 * d + 8 + __builtin_ctz(g)];
        tdir[d] = x;
    }
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * 256u + t;
    if (i == 0 || i > n) return; // after the only barrier
    const uint64_t p = i - 1;
    const uint32_t gl = (uint32_t)((i ^ (i >> 1)) & 255u); // Gray bits 0..7
    const SobolLane L = sobol_lane(gl & 63u, __builtin_amdgcn_readfirstlane(gl >> 6));
    if constexpr (FP64) {
        double* w_lds = (double*)smem;
        w_lds[t] = 0.0;
        for (int k = 0; k < n_ops; ++k) {
            const BridgeOp& op = ops[k];
            const uint32_t x = sobol_x(v + 32 * op.dim, L, tdir[op.dim]);
            const double z = qmc_normal_f64<true, SHIFTED>(x);
            const uint32_t sl = op.slots & 0xffu, sr = (op.slots >> 8) & 0xffu, so = op.slots >> 16;
            const double wl = w_lds[sl * 256 + t], wr = w_lds[sr * 256 + t];
            const double w = (wl + op.frac * (wr - wl)) + op.sd * z;
            w_lds[so * 256 + t] = w;
            ((double*)store)[(uint64_t)op.point * n + p] = spot * exp(op.drift_t + vol * w);
        }
    } else {
        float* w_lds = (float*)smem;
        w_lds[t] = 0.f;
        const float s0 = (float)spot;
        for (int k = 0; k < n_ops; ++k) {
            const BridgeOp& op = ops[k];
            const uint32_t x = sobol_x(v + 32 * op.dim, L, tdir[op.dim]);
            const float z = qmc_normal_f32<SHIFTED>(x);
            const uint32_t sl = op.slots & 0xffu, sr = (op.slots >> 8) & 0xffu, so = op.slots >> 16;
            const float wl = w_lds[sl * 256 + t], wr = w_lds[sr * 256 + t];
            const float w = __builtin_fmaf(op.sd_f, z, __builtin_fmaf(op.frac_f, wr - wl, wl));
            w_lds[so * 256 + t] = w;
            ((float*)store)[(uint64_t)op.point * n + p] =
                s0 * fast_exp2(__builtin_fmaf(vol2, w, op.drift2_f));
        }
    }
}

constexpr uint32_t kMaxReplicates = 4096;

// Randomized QMC: `replicates` independent digital shifts of the same point set; the price
// is the mean of the replicate means and its standard error their spread, the error
// estimate plain QMC does not give.
int rqmc_path(const mcdx_gbm_contract* c, int precision, uint32_t replicates,
              uint64_t shift_seed, mcdx_result* out, double* replicate_prices) {
    if (!out || !c || replicates < 2 || replicates > kMaxReplicates) return MCDX_EINVAL;
    if (precision != MCDX_PREC_FP32 && precision != MCDX_PREC_FP64) return MCDX_EINVAL;
    if (c->type > 1 || c->path_count > (1ull << 32) - 2) return MCDX_EINVAL;
    const bool european = c->product == MCDX_EUROPEAN;
    const int32_t m = european ? 1 : c->monitoring_points;
    uint8_t product = 0;
    if (!european) {
        if (c->strike_style > 1 || c->direction > 1 || c->knock > 1 || m < 1 ||
            m > (int32_t)kSobolDims)
            return MCDX_EINVAL;
        if (c->product == MCDX_ASIAN && c->average_style == 0) product = 0;
        else if (c->product == MCDX_BARRIER && !c->brownian_bridge) product = 1;
        else if (c->product == MCDX_LOOKBACK) product = 2;
        else return MCDX_EINVAL;
    }
    if (c->path_count == 0) {
        *out = mcdx_result{NAN, 0.0, 0};
        return MCDX_OK;
    }
    std::vector<BridgeOp> ops;
    int slots = 0;
    if (!european) {
        slots = build_bridge(m, c->spot, c->rate, c->carry_yield, c->vol, c->time, ops);
        if (slots < 0) return MCDX_EINVAL;
    }
    QmcArgs a = qmc_args(c->spot, c->strike, c->rate, c->carry_yield, c->vol, c->time, c->type,
                         c->path_count);
    a.m = m;
    a.european = european ? 1 : 0;
    a.strike_style = c->strike_style;
    a.product = product;
    a.direction = c->direction;
    a.knock = c->knock;
    a.barrier = c->barrier;
    a.rebate = c->rebate;
    a.barrier_rel = (float)(c->barrier / c->spot);
    // shift of replicate r, dimension d: Philox4x32-10 word 0 at key shift_seed, counter
    // {r, 0, d, 3} (stream 3: no reference stream uses it)
    std::vector<uint32_t> shifts((size_t)replicates * (size_t)m);
    for (uint32_t r = 0; r < replicates; ++r)
        for (int32_t d = 0; d < m; ++d)
            shifts[(size_t)r * m + d] = philox_bits(shift_seed, r, (uint32_t)d, 3u);
    std::vector<mcdx_result> reps(replicates);
    const int st = qmc_run(a, precision, ops, slots, replicates, shifts, reps.data());
    if (st) return st;
    double sum = 0.0;
    for (uint32_t r = 0; r < replicates; ++r) sum += reps[r].price;
    const double mean = sum / (double)replicates;
    double ss = 0.0;
    for (uint32_t r = 0; r < replicates; ++r) ss += (reps[r].price - mean) * (reps[r].price - mean);
    const double se = sqrt(ss / (double)(replicates - 1) / (double)replicates);
    if (replicate_prices)
        for (uint32_t r = 0; r < replicates; ++r) replicate_prices[r] = reps[r].price;
    *out = mcdx_result{mean, se, c->path_count * replicates};
    return MCDX_OK;
}

size_t qmc_store_scratch_bytes(int32_t m, uint32_t runs) {
    const size_t ops = ((size_t)(m > 0 ? m : 1) * sizeof(BridgeOp) + 255) & ~(size_t)255;
    const size_t sh = ((size_t)(m > 0 ? m : 1) * runs * sizeof(uint32_t) + 255) & ~(size_t)255;
    return ops + sh;
}

int qmc_fill_store(double spot, double rate, double carry, double vol, double time, int32_t m,
                   uint64_t n, bool fp64, bool shifted, uint64_t shift_seed, uint32_t replicate,
                   uint32_t runs, void* store, void* scratch, hipStream_t s) {
    if (m < 1 || m > (int32_t)kSobolDims || n == 0 || n > (1ull << 32) - 2) return MCDX_EINVAL;
    if (runs < 1 || runs > 65535 || (runs > 1 && !shifted)) return MCDX_EINVAL;
    const uint32_t* v;
    int st = device_tables(&v);
    if (st) return st;
    std::vector<BridgeOp> ops;
    const int slots = build_bridge(m, spot, rate, carry, vol, time, ops);
    // one op per date: the kernel's per-run strides are ops.size() shift words and dates
    if (slots < 0 || ops.size() != (size_t)m) return MCDX_EINVAL;
    const size_t shmem = (size_t)slots * 256 * (fp64 ? sizeof(double) : sizeof(float));
    if (shmem > 64 * 1024) return MCDX_EINVAL;
    std::vector<uint32_t> shift((size_t)m * runs, 0u);
    if (shifted)
        for (uint32_t k = 0; k < runs; ++k)
            for (int32_t d = 0; d < m; ++d)
                shift[(size_t)k * m + d] = philox_bits(shift_seed, replicate + k, (uint32_t)d, 3u);
    BridgeOp* dops = (BridgeOp*)scratch;
    uint32_t* dshift = (uint32_t*)((char*)scratch + (((size_t)m * sizeof(BridgeOp) + 255) & ~(size_t)255));
    hipError_t e = hipMemcpyAsync(dops, ops.data(), ops.size() * sizeof(BridgeOp),
                                  hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dshift, shift.data(), shift.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s); // the host vectors end with this call
    if (e != hipSuccess) return status_of(e);
    const double log2e = 1.44269504088896340735992468100189214;
    const unsigned grid = (unsigned)((n + 1 + 255) / 256); // raw indices 0 .. n
    const float vol2 = (float)(vol * log2e);
    if (fp64) {
        if (shifted)
            hipLaunchKernelGGL((qmc_store_kernel<true, true>), dim3(grid, runs), dim3(256), shmem, s, v, dops,
                               (int32_t)ops.size(), dshift, spot, vol, vol2, n, store);
        else
            hipLaunchKernelGGL((qmc_store_kernel<true, false>), dim3(grid), dim3(256), shmem, s, v, dops,
                               (int32_t)ops.size(), dshift, spot, vol, vol2, n, store);
    } else {
        if (shifted)
            hipLaunchKernelGGL((qmc_store_kernel<false, true>), dim3(grid, runs), dim3(256), shmem, s, v, dops,
                               (int32_t)ops.size(), dshift, spot, vol, vol2, n, store);
        else
            hipLaunchKernelGGL((qmc_store_kernel<false, false>), dim3(grid), dim3(256), shmem, s, v, dops,
                               (int32_t)ops.size(), dshift, spot, vol, vol2, n, store);
    }
    return status_of(launch_check());
}

} // namespace mcdx

using namespace mcdx;

extern "C" {

int mcdx_qmc_sobol_path(const mcdx_gbm_contract* contract, int precision, mcdx_result* out) {
    return qmc_path(contract, precision, out);
}

int mcdx_rqmc_sobol_path(const mcdx_gbm_contract* contract, int precision, uint32_t replicates,
                         uint64_t shift_seed, mcdx_result* out, double* replicate_prices) {
    return rqmc_path(contract, precision, replicates, shift_seed, out, replicate_prices);
}

uint32_t mcdx_sobol_max_dimensions(void) { return kSobolRefDims; }
uint32_t mcdx_sobol_extended_dimensions(void) { return kSobolDims; }

int mcdx_sobol_ints(uint32_t dim, uint64_t index0, uint64_t n, uint32_t* out) {
    if (dim >= kSobolDims || (!out && n)) return MCDX_EINVAL;
    if (n == 0) return MCDX_OK;
    if (n > (1ull << 31) || index0 + n > (1ull << 32)) return MCDX_EINVAL;
    ThreadCtx* ctx;
    int st = thread_ctx(&ctx);
    if (st) return st;
    const uint32_t* v;
    if ((st = device_tables(&v))) return st;
    if ((st = ensure_dev(ctx, n * 4))) return st;
    hipLaunchKernelGGL(sobol_ints_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       ctx->stream, v + 32 * dim, index0, n, (uint32_t*)ctx->dev);
    hipError_t e = launch_check();
    if (e == hipSuccess) e = hipMemcpyAsync(out, ctx->dev, n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return status_of(e);
}

int mcdx_qmc_sobol_european_p(double spot, double strike, double rate, double carry_yield,
                              double vol, double time, uint8_t type, uint64_t path_count,
                              int precision, mcdx_result* out) {
    if (!out || type > 1) return MCDX_EINVAL;
    if (precision != MCDX_PREC_FP32 && precision != MCDX_PREC_FP64) return MCDX_EINVAL;
    if (path_count > (1ull << 32) - 2) return MCDX_EINVAL;
    if (path_count == 0) {
        *out = mcdx_result{NAN, 0.0, 0};
        return MCDX_OK;
    }
    QmcArgs a = qmc_args(spot, strike, rate, carry_yield, vol, time, type, path_count);
    a.european = 1;
    a.m = 1;
    std::vector<BridgeOp> none;
    return qmc_run(a, precision, none, 0, out);
}

int mcdx_qmc_sobol_european(double spot, double strike, double rate, double carry_yield,
                            double vol, double time, uint8_t type, uint64_t path_count,
                            double* price) {
    if (!price) return MCDX_EINVAL;
    mcdx_result r;
    const int st = mcdx_qmc_sobol_european_p(spot, strike, rate, carry_yield, vol, time, type,
                                             path_count, mcdx_get_precision(), &r);
    if (st == MCDX_OK) *price = r.price;
    return st;
}

int mcdx_qmc_sobol_asian(double spot, double strike, double rate, double carry_yield, double vol,
                         double time, uint8_t type, uint8_t strike_style,
                         int32_t monitoring_points, uint64_t path_count, double* price) {
    if (!price || monitoring_points < 1 || monitoring_points > (int32_t)kSobolRefDims)
        return MCDX_EINVAL;
    mcdx_result r;
    const int st = qmc_asian(spot, strike, rate, carry_yield, vol, time, type, strike_style,
                             monitoring_points, path_count, mcdx_get_precision(), &r);
    if (st == MCDX_OK) *price = r.price;
    return st;
}

int mcdx_qmc_sobol_asian_extended(double spot, double strike, double rate, double carry_yield,
                                  double vol, double time, uint8_t type, uint8_t strike_style,
                                  int32_t monitoring_points, uint64_t path_count, int precision,
                                  mcdx_result* out) {
    return qmc_asian(spot, strike, rate, carry_yield, vol, time, type, strike_style,
                     monitoring_points, path_count, precision, out);
}

} // extern "C"
