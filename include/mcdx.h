/*
 * mcdx -- C ABI of the MI355X (gfx950) Monte Carlo path engine, libmcd_hip.so.
 *
 * Plain C types only: scalars, POD structs and caller-owned host pointers. Every entry
 * point returns an mcdx_status and never throws or aborts. Results are written to
 * caller-allocated out-structs. Each function names the reference `mcd` interface it
 * replaces (file:line in FredTheMay/derivative-pricing).
 *
 * Enum arguments use the reference's uint8_t values (include/mcd/core/types.hpp:12-17):
 *   type 0 Call / 1 Put; direction 0 Up / 1 Down; knock 0 In / 1 Out;
 *   strike_style 0 Fixed / 1 Floating; digital_style 0 CashOrNothing / 1 AssetOrNothing;
 *   average_style 0 Arithmetic / 1 Geometric.
 *
 * Precision: MCDX_PREC_FP32 generates paths in FP32 (Philox bits exact, FP32 inverse
 * normal and log-space GBM stepping) and accumulates every payoff in FP64;
 * MCDX_PREC_FP64 repeats the reference's FP64 arithmetic operation for operation. The
 * reference-signature entry points use the process default (mcdx_set_precision).
 *
 * Threading: every entry point is reentrant. Each calling thread gets its own HIP stream
 * and scratch; no lock is held while kernels run. The reference's `num_threads` option is
 * accepted and ignored (the GPU schedule does not depend on it).
 */
#ifndef MCDX_H
#define MCDX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum mcdx_status {
    MCDX_OK = 0,
    MCDX_EINVAL = 1,  /* an argument is outside what the entry point accepts */
    MCDX_EDEVICE = 2, /* a HIP call failed */
    MCDX_ENOMEM = 3,  /* device or pinned allocation failed */
    MCDX_ENODEV = 4   /* no HIP device is visible */
} mcdx_status;

enum { MCDX_PREC_FP32 = 0, MCDX_PREC_FP64 = 1 };

/* McResult, include/mcd/pricers/monte_carlo.hpp:9-13. */
typedef struct mcdx_result {
    double price;
    double standard_error;
    uint64_t path_count;
} mcdx_result;

/* McOptions, include/mcd/pricers/monte_carlo.hpp:18-33 (num_threads ignored on GPU). */
typedef struct mcdx_options {
    uint8_t antithetic;
    uint8_t control_variate;
    uint8_t brownian_bridge;
    uint8_t reserved;
    uint32_t num_threads;
} mcdx_options;

/* One GBM contract: the union of every GBM pricer's arguments
 * (include/mcd/pricers/monte_carlo.hpp:38-80) plus its path count and seed. */
enum {
    MCDX_EUROPEAN = 0,
    MCDX_DIGITAL = 1,
    MCDX_ASIAN = 2,
    MCDX_BARRIER = 3,
    MCDX_LOOKBACK = 4
};
typedef struct mcdx_gbm_contract {
    double spot, strike, barrier, rebate, rate, carry_yield, vol, time, cash_amount;
    uint64_t path_count;
    uint64_t seed;
    int32_t monitoring_points; /* ignored for European/digital */
    uint8_t product;           /* MCDX_EUROPEAN .. MCDX_LOOKBACK */
    uint8_t type, direction, knock, strike_style, average_style, digital_style;
    uint8_t antithetic, control_variate, brownian_bridge;
    uint8_t reserved[2];
} mcdx_gbm_contract;

/* ---- engine ---- */
int mcdx_device_count(int* count);
int mcdx_set_precision(int precision); /* process default for reference-signature calls */
int mcdx_get_precision(void);
const char* mcdx_version(void); /* mcd::version(), src/core/types.cpp:5 */

/* ---- RNG streams (src/core/rng.cpp:41-74) ---- */
/* out[i] = word 0 of Philox4x32-10(counter {path0+i, draw, stream}, key seed), i < n. */
int mcdx_philox_bits(uint64_t seed, uint64_t path0, uint64_t n, uint32_t draw,
                     uint32_t stream, uint32_t* out);
/* out[i] = standard_normal_variate(seed, path0+i, draw) at `precision` (rng.cpp:59-66). */
int mcdx_normal_draws(uint64_t seed, uint64_t path0, uint64_t n, uint32_t draw, int precision,
                      double* out);
/* out[i] = inverse normal of u = (words[i] + 0.5) / 2^32 at `precision` (the FP64 mode is
 * inverse_standard_normal_cdf, src/core/normal.cpp:70-82; FP32 is this engine's), i < n. */
int mcdx_inv_normal_words(const uint32_t* words, uint64_t n, int precision, double* out);
/* Accuracy scan of the FP32-mode inverse normal (this engine's approximation; no reference
 * counterpart; the reference's FP64 Acklam + Halley is src/core/normal.cpp:70-82): over the
 * Philox output words b in [begin, begin + count) (count <= 2^32 - begin), the largest
 * |z32(b) - z64(b)| / (1 + |z64(b)|) against the FP64 inverse of u = (b + 0.5) / 2^32, with
 * its word (smallest on ties); index 0 = central-branch words, 1 = tail-branch words
 * (-1 where the range holds none). */
int mcdx_inv_normal_f32_scan(uint64_t begin, uint64_t count, double max_err[2],
                             uint32_t argmax_bits[2]);

/* ---- Error-path test seam (no reference counterpart) ----
 * Arms the calling thread so that its `countdown`-th device allocation (site
 * MCDX_FAULT_ALLOC: scratch growth, book buffers, table uploads, pinned staging) reports
 * MCDX_ENOMEM without allocating, or its `countdown`-th launch check (MCDX_FAULT_LAUNCH, made
 * after every kernel launch) reports MCDX_EDEVICE after consuming the real launch status.
 * countdown 0 disarms; the armed state is per thread and fires once. Kernels queued before
 * an injected launch failure still run; the entry point reports the error and its state
 * stays usable. mcdx_fault_pending returns the remaining countdown of a site. */
#define MCDX_FAULT_ALLOC 0
#define MCDX_FAULT_LAUNCH 1
int mcdx_fault_inject(int site, int32_t countdown);
int mcdx_fault_pending(int site);

/* ---- VALU issue-rate probe (no reference counterpart) ----
 * Times `iterations` steps of eight independent chains of one instruction class in every lane
 * of 8 waves per SIMD (all SIMDs) and reports the seconds and the wave64 instructions issued:
 * the class's sustained issue rate, the peak of the path kernels' lane-issue roofline. */
#define MCDX_ISSUE_FMA_F32 0     /* v_fma_f32 */
#define MCDX_ISSUE_MAD_U64_U32 1 /* 32x32 -> 64-bit integer multiply-add (Philox products) */
#define MCDX_ISSUE_MUL_HI_U32 2  /* high word of a 32x32 product */
#define MCDX_ISSUE_XOR3 3        /* three-input XOR */
#define MCDX_ISSUE_EXP2_F32 4    /* transcendental exp2 */
#define MCDX_ISSUE_FMA_F64 5     /* v_fma_f64 */
#define MCDX_ISSUE_MIX_FMA_MAD 6 /* alternate chains of classes 0 and 1 */
#define MCDX_ISSUE_MIX_FMA_EXP2 7 /* alternate chains of classes 0 and 4 */
int mcdx_issue_probe(int instruction_class, uint32_t iterations, double* seconds,
                     uint64_t* wave_instructions);

/* ---- GBM Monte Carlo pricers (src/pricers/monte_carlo.cpp:234-338) ---- */
int mcdx_monte_carlo_european(double spot, double strike, double rate, double carry_yield,
                              double vol, double time, uint8_t type, uint64_t path_count,
                              uint64_t seed, const mcdx_options* options, mcdx_result* out);
int mcdx_monte_carlo_digital(double spot, double strike, double rate, double carry_yield,
                             double vol, double time, uint8_t type, uint8_t style,
                             double cash_amount, uint64_t path_count, uint64_t seed,
                             const mcdx_options* options, mcdx_result* out);
int mcdx_monte_carlo_asian(double spot, double strike, double rate, double carry_yield,
                           double vol, double time, uint8_t type, uint8_t strike_style,
                           uint8_t average_style, int32_t monitoring_points,
                           uint64_t path_count, uint64_t seed, const mcdx_options* options,
                           mcdx_result* out);
int mcdx_monte_carlo_barrier(double spot, double strike, double barrier, double rate,
                             double carry_yield, double vol, double time, uint8_t type,
                             uint8_t direction, uint8_t knock, double rebate,
                             int32_t monitoring_points, uint64_t path_count, uint64_t seed,
                             const mcdx_options* options, mcdx_result* out);
int mcdx_monte_carlo_lookback(double spot, double strike, double rate, double carry_yield,
                              double vol, double time, uint8_t type, uint8_t style,
                              int32_t monitoring_points, uint64_t path_count, uint64_t seed,
                              const mcdx_options* options, mcdx_result* out);

/* Any GBM contract at an explicit precision (the pricers above route through this). */
int mcdx_price_gbm(const mcdx_gbm_contract* contract, int precision, mcdx_result* out);
/* Stream-ordered form for callers that own a HIP stream (hipStream_t passed as void*, NULL
 * = the calling thread's stream): enqueues the pricing and writes {n, mean, m2} as
 * (uint64_t, double, double) to the 24-byte device buffer `dev_welford`; returns at once. */
int mcdx_price_gbm_enqueue(const mcdx_gbm_contract* contract, int precision, void* stream,
                           void* dev_welford);
/* mcdx_price_gbm_enqueue without any allocation, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc); it has
 * no counterpart in the reference. Same kernels and results; the tile partials live in a
 * fixed per-device array of the library (1.5 MiB, held until the process ends), so
 * path_count is at most 2^26 (MCDX_EINVAL beyond). NULL stream = the device's null stream.
 * Every call on one GPU shares that array: calls that overlap on one GPU (any streams, any
 * threads) race and must be ordered by the caller. */
#define MCDX_NOALLOC_MAX_PATHS (1ull << 26)
int mcdx_price_gbm_enqueue_noalloc(const mcdx_gbm_contract* contract, int precision,
                                   void* stream, void* dev_welford);
/* Discounted per-path samples exactly as the pricer accumulates them (antithetic pairs
 * averaged) for paths [path0, path0+n): the per-path parity hook. */
int mcdx_gbm_path_samples(const mcdx_gbm_contract* contract, int precision, uint64_t path0,
                          uint64_t n, double* out);

/* ---- Sobol + Brownian bridge QMC (src/core/sobol.cpp, src/pricers/qmc.cpp) ---- */
/* kSobolMaxDimensions (include/mcd/core/sobol.hpp:14): 7, the drop-in cap. */
uint32_t mcdx_sobol_max_dimensions(void);
/* Dimensions of this engine's extended table (dims 0-6 equal the reference's). */
uint32_t mcdx_sobol_extended_dimensions(void);
/* out[k] = Sobol integer x of (dim, index0 + k); point = x / 2^32 (sobol.cpp:83-97). */
int mcdx_sobol_ints(uint32_t dim, uint64_t index0, uint64_t n, uint32_t* out);
/* qmc_sobol_european (src/pricers/qmc.cpp:15-30): mean over indices 1..path_count. */
int mcdx_qmc_sobol_european(double spot, double strike, double rate, double carry_yield,
                            double vol, double time, uint8_t type, uint64_t path_count,
                            double* price);
int mcdx_qmc_sobol_european_p(double spot, double strike, double rate, double carry_yield,
                              double vol, double time, uint8_t type, uint64_t path_count,
                              int precision, mcdx_result* out);
/* qmc_sobol_asian (src/pricers/qmc.cpp:32-57); EINVAL unless 1 <= monitoring_points <= 7,
 * as the reference binding raises (bindings/python/src/module.cpp:268-272). */
int mcdx_qmc_sobol_asian(double spot, double strike, double rate, double carry_yield, double vol,
                         double time, uint8_t type, uint8_t strike_style,
                         int32_t monitoring_points, uint64_t path_count, double* price);
/* The same pricer over up to mcdx_sobol_extended_dimensions() monitoring points. */
int mcdx_qmc_sobol_asian_extended(double spot, double strike, double rate, double carry_yield,
                                  double vol, double time, uint8_t type, uint8_t strike_style,
                                  int32_t monitoring_points, uint64_t path_count, int precision,
                                  mcdx_result* out);
/* Sobol + Brownian-bridge QMC for a path-dependent contract: arithmetic Asian, discrete
 * barrier (no bridge correction) or lookback, monitoring_points 1..1024 (one Sobol
 * dimension per date), Sobol indices 1..path_count. Composes the reference's pieces
 * (sobol.cpp:83-97, brownian_bridge_path.cpp:20-77, the payoff headers, qmc.cpp:44-56);
 * the reference prices only the arithmetic Asian this way. The seed is ignored; the
 * standard_error field is the points' sample standard error, which does not measure the
 * error of a deterministic point set. */
int mcdx_qmc_sobol_path(const mcdx_gbm_contract* contract, int precision, mcdx_result* out);
/* Randomized QMC (no reference counterpart: the reference's QMC returns a plain mean with no
 * error estimate, qmc.hpp:9-16). `replicates` (2..4096) digital shifts of the Sobol indices
 * 1..path_count: replicate r XORs dimension d's integer with Philox4x32-10 word 0 at key
 * shift_seed, counter {r, 0, d, 3}, and maps it to u = (x + 0.5) / 2^32. out->price = mean of
 * the replicate means, out->standard_error = their sample deviation / sqrt(replicates),
 * out->path_count = path_count * replicates; replicate_prices (optional, [replicates])
 * receives each replicate's mean. Products: European (dimension 0), arithmetic Asian, discrete
 * barrier and lookback over the bridge (as mcdx_qmc_sobol_path). */
int mcdx_rqmc_sobol_path(const mcdx_gbm_contract* contract, int precision, uint32_t replicates,
                         uint64_t shift_seed, mcdx_result* out, double* replicate_prices);

/* ---- Heston QE (src/pricers/heston.cpp) ---- */
/* HestonParams, include/mcd/models/heston.hpp:8-18 (same field order). */
typedef struct mcdx_heston_params {
    double spot, rate, carry_yield, v0, kappa, theta, xi, rho, time;
} mcdx_heston_params;
/* heston_qe_european (src/pricers/heston.cpp:120-156); precision = process default. */
int mcdx_heston_qe_european(const mcdx_heston_params* params, double strike, uint8_t type,
                            int32_t monitoring_points, uint64_t path_count, uint64_t seed,
                            mcdx_result* out);
/* Up to 32 strikes priced from one pass over the same paths; out[s] equals the
 * single-strike result for strikes[s] bit for bit. */
int mcdx_heston_qe_strikes(const mcdx_heston_params* params, const double* strikes,
                           int32_t n_strikes, uint8_t type, int32_t monitoring_points,
                           uint64_t path_count, uint64_t seed, int precision, mcdx_result* out);
/* Discounted per-path payoffs at `strike` for paths [path0, path0+n): the parity hook. */
int mcdx_heston_path_samples(const mcdx_heston_params* params, double strike, uint8_t type,
                             int32_t monitoring_points, uint64_t seed, int precision,
                             uint64_t path0, uint64_t n, double* out);

/* ---- Longstaff-Schwartz (src/pricers/lsm.cpp) ---- */
/* QR-LSM on Sobol + Brownian-bridge paths (no reference counterpart: the reference's LSM
 * simulates Philox paths, lsm.cpp:35-58). Path p is Sobol index p + 1 through the bridge, the
 * regression and decisions are mcdx_lsm_american's. replicates == 0: one run (the standard
 * error is the spread of the path values, not an error of the estimate); replicates in
 * 2..4096: that many runs with digitally shifted points (shift words and cell-midpoint mapping
 * as mcdx_rqmc_sobol_path), out = the mean of the run prices with their standard error and
 * path_count * replicates; replicate_prices (optional, [replicates]) gets each run's price. */
int mcdx_lsm_american_qmc(double spot, double strike, double rate, double carry_yield, double vol,
                          double time, uint8_t type, int32_t monitoring_points,
                          uint64_t path_count, uint32_t replicates, uint64_t shift_seed,
                          int precision, mcdx_result* out, double* replicate_prices);
/* monte_carlo_lsm_american (src/pricers/lsm.cpp:62-149) at `precision`. The policy
 * (LsmPolicy, include/mcd/pricers/lsm.hpp:16-24) is returned as coefficients[(M+1)*4]
 * (row d = date d), has_coefficients[M+1] and exercise_at_inception; any may be NULL. */
int mcdx_lsm_american(double spot, double strike, double rate, double carry_yield, double vol,
                      double time, uint8_t type, int32_t monitoring_points, uint64_t path_count,
                      uint64_t seed, int precision, mcdx_result* out, double* coefficients,
                      uint8_t* has_coefficients, uint8_t* exercise_at_inception);
/* Reference signature (src/pricers/lsm.cpp:62-66); precision = process default. */
int mcdx_monte_carlo_lsm_american(double spot, double strike, double rate, double carry_yield,
                                  double vol, double time, uint8_t type,
                                  int32_t monitoring_points, uint64_t path_count, uint64_t seed,
                                  mcdx_result* out);
/* reprice_against_frozen_policy (src/pricers/lsm.cpp:151-213); out->price is the price. */
int mcdx_lsm_reprice_frozen(double spot, double strike, double rate, double carry_yield,
                            double vol, double time, uint8_t type, int32_t monitoring_points,
                            uint64_t path_count, uint64_t seed, const double* coefficients,
                            const uint8_t* has_coefficients, uint8_t exercise_at_inception,
                            int precision, mcdx_result* out);

/* Backward-induction schedule (no reference counterpart): MCDX_LSM_AUTO (the default) runs
 * every date of a regression in one launch whose blocks stay resident when the run fits (at
 * most 2^20 paths per run, at most 1024 dates, every block of the batch resident at once on
 * the device), else one launch per date; MCDX_LSM_PER_DATE always launches per date. Both give
 * the same bits. Process-wide; returns the previous schedule, or -MCDX_EINVAL. */
#define MCDX_LSM_AUTO 0
#define MCDX_LSM_PER_DATE 1
int mcdx_lsm_schedule(int schedule);

/* American Greeks by central differences of frozen-policy prices (reprice_against_frozen_policy,
 * src/pricers/lsm.cpp:151-213, as the reference's tests/lsm_test.cpp:138-170 forms delta):
 * one QR-LSM fit at the base parameters, then the base, spot +/- spot_bump and vol +/- vol_bump
 * scenarios of every path, from the same draws, exercised by the frozen policy in one
 * streaming pass. out4 = price, delta, gamma, vega, each with the standard error of its
 * per-path estimator. Bumps <= 0 select default_bump_sizes (1% of spot, 1% of vol). */
int mcdx_lsm_american_greeks(double spot, double strike, double rate, double carry_yield,
                            double vol, double time, uint8_t type, int32_t monitoring_points,
                            uint64_t path_count, uint64_t seed, int precision, double spot_bump,
                            double vol_bump, mcdx_result* out4);

/* ---- Greeks (src/greeks), FP64 arithmetic as the reference ---- */
/* pathwise_european (src/greeks/pathwise.cpp:28-62): out3 = delta, vega, rho, each as
 * {value, standard_error, path_count}. */
int mcdx_pathwise_european(double spot, double strike, double rate, double carry_yield,
                           double vol, double time, uint8_t type, uint64_t path_count,
                           uint64_t seed, mcdx_result* out3);
/* pathwise_asian (src/greeks/pathwise.cpp:64-147): out3 = delta, vega, rho. */
int mcdx_pathwise_asian(double spot, double strike, double rate, double carry_yield, double vol,
                        double time, uint8_t type, uint8_t strike_style, uint8_t average_style,
                        int32_t monitoring_points, uint64_t path_count, uint64_t seed,
                        mcdx_result* out3);
/* likelihood_ratio_european / _digital (src/greeks/likelihood_ratio.cpp:77-99):
 * out5 = delta, gamma, vega, rho, theta. */
int mcdx_likelihood_ratio_european(double spot, double strike, double rate, double carry_yield,
                                   double vol, double time, uint8_t type, uint64_t path_count,
                                   uint64_t seed, mcdx_result* out5);
int mcdx_likelihood_ratio_digital(double spot, double strike, double rate, double carry_yield,
                                  double vol, double time, uint8_t type, uint8_t style,
                                  double cash_amount, uint64_t path_count, uint64_t seed,
                                  mcdx_result* out5);
/* likelihood_ratio_barrier (src/greeks/likelihood_ratio.cpp:105-154): out4 = delta, gamma,
 * vega, rho (no theta, as the reference). */
int mcdx_likelihood_ratio_barrier(double spot, double strike, double barrier, double rate,
                                  double carry_yield, double vol, double time, uint8_t type,
                                  uint8_t direction, uint8_t knock, double rebate,
                                  int32_t monitoring_points, uint64_t path_count, uint64_t seed,
                                  mcdx_result* out4);
/* The nine common-random-number European prices finite_difference_european needs
 * (src/greeks/finite_difference.cpp:40-65), from one pass: base, spot+, spot-, vol+, vol-,
 * rate+, rate-, time+, time-. */
int mcdx_fd_european_prices(double spot, double strike, double rate, double carry_yield,
                            double vol, double time, uint8_t type, uint64_t path_count,
                            uint64_t seed, double bump_spot, double bump_vol, double bump_rate,
                            double bump_time, mcdx_result* out9);

/* ---- Learned continuation value (no reference counterpart) ----
 * A 2 -> 64 -> 64 -> 1 SiLU network over (S/K, t/T) with a flat FP32 parameter vector:
 * W1[64][2], b1[64], W2[64][64] (out, in), b2[64], W3[64], b3. */
uint32_t mcdx_mlp_param_count(void);
/* Gradient of the batch-mean squared error through the device's BF16-MFMA backward pass;
 * n <= 128 samples. loss2[0..1] = summed squared errors of the two 64-sample halves. */
int mcdx_mlp_gradients(const float* weights, const float* xs, const float* ts, const float* ys,
                       uint32_t n, float* grad, float* loss2);
/* out[i] = network(xs[i], ts[i]) through the device's BF16-MFMA forward pass. */
int mcdx_mlp_forward(const float* weights, const float* xs, const float* ts, uint64_t n,
                     float* out);
/* `steps` Adam steps (128 samples each, drawn by Philox(seed); NaN targets are skipped) from
 * `weights`, updated in place, with zero initial moments. losses[step] (optional) receives
 * the mean squared error of each step's first 64 samples. */
int mcdx_mlp_train(float* weights, const float* xs, const float* ts, const float* ys,
                   uint64_t n, int32_t steps, float learning_rate, float lr_decay,
                   uint64_t seed, float* losses);

/* ---- Learned-continuation LSM (no reference counterpart; extends src/pricers/lsm.cpp) ----
 * The QR-LSM backward pass captures every (path, date) regression sample; the network
 * C(4 (S/K - 1), 2 t/T - 1) ~ 10 cf/K is trained on them on the device, tabulated per
 * exercise date on a uniform S/K grid, and a streaming kernel exercises each path at the
 * first date where the intrinsic value beats the tabulated continuation (the QR
 * coefficients decide outside the table range). The same kernel replays the QR policy
 * forward on the same paths, which reproduces the backward-induction price. */
typedef struct mcdx_lsm_learn_options {
    uint64_t init_seed;        /* weight initialisation (Philox) */
    uint64_t sample_seed;      /* minibatch sampling (Philox) */
    int32_t steps;             /* Adam steps, 128 samples each */
    float learning_rate;       /* first step */
    float final_learning_rate; /* geometric decay to this by the last step */
    uint32_t grid_points;      /* continuation table points per exercise date, 2..4096 */
    uint32_t average_steps;    /* the network used is the mean of the last average_steps
                                  Adam iterates (Polyak tail averaging; all iterates when
                                  steps is smaller); 0 or 1: the last iterate */
    uint32_t reserved;
} mcdx_lsm_learn_options;

typedef struct mcdx_lsm_learned_result {
    mcdx_result qr_lsm;       /* QR-LSM backward induction (NaN from mcdx_lsm_learned_price) */
    mcdx_result learned;      /* first exercise by the network's continuation value */
    mcdx_result qr_forward;   /* first exercise by the QR coefficients, same paths */
    uint64_t itm_decisions;   /* in-the-money path-dates reached by both policies */
    uint64_t disagreements;   /* of those, where the two exercise decisions differ */
    uint64_t fallbacks;       /* path-dates decided by QR because S/K left the table range */
    double final_loss;        /* training MSE over the last min(100, steps) steps, (cf/K)^2 */
    double boundary_gap;      /* max over dates of |boundary_net - boundary_qr| in S/K */
    uint32_t boundary_dates;  /* dates where both boundaries lie on the grid */
    uint32_t reserved;
    double table_lo, table_hi; /* S/K range of the table */
} mcdx_lsm_learned_result;

void mcdx_lsm_learn_default_options(mcdx_lsm_learn_options* options);
/* Fit (QR-LSM on `path_count` paths of `seed`), train, tabulate and price on the same paths.
 * weights[4417], coefficients[(M+1)*4] and has_coefficients[M+1] receive the trained
 * network and the QR policy when non-null. */
int mcdx_lsm_american_learned(double spot, double strike, double rate, double carry_yield,
                              double vol, double time, uint8_t type, int32_t monitoring_points,
                              uint64_t path_count, uint64_t seed, int precision,
                              const mcdx_lsm_learn_options* options, float* weights,
                              double* coefficients, uint8_t* has_coefficients,
                              mcdx_lsm_learned_result* out);
/* Price with a trained network on fresh paths: streaming, no path store. With judge == 1 the
 * QR policy also runs on the same paths (qr_forward, decisions, disagreements) and the
 * boundary gap is measured; with judge == 0 it only decides outside the table range,
 * qr_forward and boundary_gap are NaN and boundary_dates is 0; judge == 2 serves the QR
 * policy alone by the same streaming kernel (qr_forward; weights may be NULL, learned NaN). */
int mcdx_lsm_learned_price(double spot, double strike, double rate, double carry_yield,
                           double vol, double time, uint8_t type, int32_t monitoring_points,
                           uint64_t path_count, uint64_t seed, int precision,
                           const float* weights, uint32_t grid_points, double table_lo,
                           double table_hi, const double* coefficients,
                           const uint8_t* has_coefficients, uint8_t judge,
                           mcdx_lsm_learned_result* out);

/* ---- Batched book: many GBM contracts, one persistent graph-captured launch ----
 * No reference counterpart (the reference prices one contract per call,
 * src/pricers/monte_carlo.cpp:234-338); each contract's result equals mcdx_price_gbm on
 * that contract bit for bit, whatever the schedule. */
typedef struct mcdx_book mcdx_book;
int mcdx_book_create(const mcdx_gbm_contract* contracts, uint32_t n, int precision,
                     mcdx_book** out);
/* Flags of mcdx_book_create_ex. By default, FP32 contracts that share a seed, a payoff
 * family (terminal, Asian, discrete barrier, lookback) and a step count are evaluated up to
 * eight at a time on the same paths from one normal per step (common random numbers); each
 * still equals its own mcdx_price_gbm bit for bit. */
#define MCDX_BOOK_NO_SHARING 1u /* one contract per work item */
#define MCDX_BOOK_GREEKS 2u     /* also delta, gamma, vega per contract (FP32 books; no
                                   antithetic or bridge-corrected barriers): central
                                   differences with the reference's default bumps
                                   (src/greeks/finite_difference.cpp:31-62), the bumped
                                   scenarios sharing each path's draws */
int mcdx_book_create_ex(const mcdx_gbm_contract* contracts, uint32_t n, int precision,
                        uint32_t flags, mcdx_book** out);
/* Search the schedule (fusion width for Greeks books, tiles per item, blocks per CU x item
 * order, one dimension at a time, each candidate timed by replay), keep the fastest, and
 * remember it for books of the same shape: in this process and in the schedule cache file
 * ($MCDX_BOOK_SCHEDULE_CACHE, else ~/.cache/mcdx/book_schedules.txt; unreadable or damaged
 * entries fall back to the default schedule). best_ms may be NULL. */
int mcdx_book_autotune(mcdx_book* book, float* best_ms);
/* order 0 = contract order, 1 = longest tiles first; blocks_per_cu 0 = occupancy limit. */
int mcdx_book_set_schedule(mcdx_book* book, uint32_t blocks_per_cu, uint32_t order);
int mcdx_book_get_schedule(const mcdx_book* book, uint32_t* blocks_per_cu, uint32_t* order,
                           uint32_t* grid);
/* The full schedule: fusion_width = Greek units (contract + its 4 bumped scenarios) sharing
 * each path's draws in one work item, 1..3 (Greeks books; 0 keeps the current one);
 * tiles_per_item = consecutive 1024-path tiles one block takes per item, 1, 2, 4 or 8 (0
 * keeps the current one). No schedule changes any result. */
int mcdx_book_set_schedule2(mcdx_book* book, uint32_t blocks_per_cu, uint32_t order,
                            uint32_t fusion_width, uint32_t tiles_per_item);
int mcdx_book_get_schedule2(const mcdx_book* book, uint32_t* blocks_per_cu, uint32_t* order,
                            uint32_t* grid, uint32_t* fusion_width, uint32_t* tiles_per_item);
/* Replay the captured graph on `stream` (NULL = the book's own stream); returns at once. */
int mcdx_book_enqueue(mcdx_book* book, void* stream);
/* Wait for the last enqueue and copy one result per contract into out[0..n). */
int mcdx_book_results(mcdx_book* book, mcdx_result* out);
int mcdx_book_price(mcdx_book* book, mcdx_result* out);
/* Books created with MCDX_BOOK_GREEKS: out[3 i + {0, 1, 2}] = delta, gamma, vega of contract
 * i, each the mean of its per-path central difference with that mean's standard error. */
int mcdx_book_greeks(mcdx_book* book, mcdx_result* out);
/* Books created with MCDX_BOOK_GREEKS: out[2 i + {0, 1}] = pathwise delta and vega of contract
 * i on the same paths (pathwise_european / pathwise_asian, src/greeks/pathwise.cpp:28-147,
 * with the book's FP32 draws); {NaN, NaN, 0} for contracts the reference has no pathwise
 * estimator for (digital, barrier, lookback, control-variate Asian). */
int mcdx_book_pathwise(mcdx_book* book, mcdx_result* out);
void mcdx_book_destroy(mcdx_book* book);

#ifdef __cplusplus
}
#endif
#endif /* MCDX_H */
