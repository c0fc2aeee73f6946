/*
 * CPU oracle for the Monte Carlo path engine -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64 restatement of the reference `mcd` algorithms, used as the
 * checker for the HIP kernels (tests, smoke() and the cpu_baseline leg of
 * bench.py). Nothing in the product library links or calls this code.
 *
 * Every function cites the reference file:line it restates. Enum values match
 * include/mcd/core/types.hpp:12-17 (declaration order 0, 1).
 */
#ifndef MCD_ORACLE_H
#define MCD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG (src/core/rng.cpp) ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint32_t orc_philox_bits(uint64_t seed, uint64_t path, uint32_t draw, uint32_t stream);
void orc_philox_bits_range(uint64_t seed, uint64_t path0, uint64_t n, uint32_t draw,
                           uint32_t stream, uint32_t* out);
double orc_uniform_open01(uint32_t bits);
double orc_standard_normal_variate(uint64_t seed, uint64_t path, uint32_t draw);
double orc_uniform_variate(uint64_t seed, uint64_t path, uint32_t draw);

/* ---- Normal distribution (src/core/normal.cpp) ---- */
double orc_normal_pdf(double x);
double orc_normal_cdf(double x);
double orc_inverse_normal_cdf(double u);
double orc_acklam(double u); /* Acklam's rational approximation without the Halley step */

/* ---- Welford (include/mcd/core/stats.hpp) ---- */
typedef struct orc_welford {
    uint64_t n;
    double mean;
    double m2;
} orc_welford;
void orc_welford_add(orc_welford* w, double x);
void orc_welford_merge(orc_welford* w, const orc_welford* o);
double orc_welford_variance(const orc_welford* w);
double orc_welford_se(const orc_welford* w);
void orc_chunk_bounds(unsigned chunk, uint64_t total, unsigned chunks, uint64_t* begin,
                      uint64_t* end);

/* ---- GBM products (src/pricers/monte_carlo.cpp, include/mcd/payoffs) ---- */
enum {
    ORC_EUROPEAN = 0,
    ORC_DIGITAL = 1,
    ORC_ASIAN = 2,
    ORC_BARRIER = 3,
    ORC_LOOKBACK = 4
};
typedef struct orc_product {
    double spot, strike, barrier, rebate, rate, carry_yield, vol, time, cash_amount;
    int32_t monitoring_points;
    uint8_t product;       /* ORC_* */
    uint8_t type;          /* 0 call, 1 put */
    uint8_t direction;     /* 0 up, 1 down */
    uint8_t knock;         /* 0 in, 1 out */
    uint8_t strike_style;  /* 0 fixed, 1 floating */
    uint8_t average_style; /* 0 arithmetic, 1 geometric */
    uint8_t digital_style; /* 0 cash-or-nothing, 1 asset-or-nothing */
    uint8_t antithetic, control_variate, brownian_bridge;
} orc_product;

typedef struct orc_result {
    double price;
    double standard_error;
    uint64_t path_count;
} orc_result;

double orc_gbm_terminal_spot(double spot, double rate, double carry, double vol, double time,
                             double z);
/* One per-path sample exactly as accumulate_paths feeds Welford (antithetic pairs averaged). */
double orc_path_sample(const orc_product* p, uint64_t seed, uint64_t path);
void orc_path_samples(const orc_product* p, uint64_t seed, uint64_t path0, uint64_t n,
                      double* out);
/* Chunked pricing: `chunks` = reference ThreadPool::logical_chunk_count(), `threads` workers. */
orc_result orc_price(const orc_product* p, uint64_t path_count, uint64_t seed, unsigned chunks,
                     unsigned threads);

/* ---- Sobol + Brownian bridge + QMC (src/core/sobol.cpp, src/models, src/pricers/qmc.cpp) ---- */
uint32_t orc_sobol_int(unsigned dim, uint64_t index);
double orc_sobol_point(unsigned dim, uint64_t index);
unsigned orc_sobol_max_dims(void);
int orc_sobol_direction_table(unsigned dim, uint32_t* v32);
void orc_brownian_bridge_path(double spot, double rate, double carry, double vol, double time,
                              int n, const double* normals, double* spots);
double orc_qmc_sobol_european(double spot, double strike, double rate, double carry, double vol,
                              double time, int type, uint64_t path_count);
/* QMC over Sobol + bridge for Asian (arithmetic), barrier and lookback products: the
 * reference's pieces composed (no reference counterpart for barriers and lookbacks). */
double orc_qmc_sobol_path(const orc_product* p, uint64_t path_count);
/* Randomized QMC (this engine's extension; no reference counterpart): replicate r uses the
 * digital shift d -> orc_rqmc_shift(shift_seed, r, d) and the cell-midpoint mapping
 * u = (x ^ shift + 0.5) / 2^32; returns that replicate's plain mean over indices 1..n.
 * ORC_EUROPEAN uses dimension 0 and the terminal spot; the others the bridge path. */
uint32_t orc_rqmc_shift(uint64_t shift_seed, uint32_t replicate, unsigned dim);
double orc_rqmc_replicate(const orc_product* p, uint64_t path_count, uint64_t shift_seed,
                          uint32_t replicate);
double orc_qmc_sobol_asian(double spot, double strike, double rate, double carry, double vol,
                           double time, int type, int strike_style, int monitoring_points,
                           uint64_t path_count);

/* ---- Heston (src/pricers/heston.cpp) ---- */
typedef struct orc_heston {
    double spot, rate, carry_yield, v0, kappa, theta, xi, rho, time;
} orc_heston;
double orc_qe_variance_step(double v, double kappa, double theta, double xi, double dt,
                            double u);
double orc_heston_path_sample(const orc_heston* h, double strike, int type, int steps,
                              uint64_t seed, uint64_t path);
orc_result orc_heston_qe_european(const orc_heston* h, double strike, int type, int steps,
                                  uint64_t path_count, uint64_t seed);
void orc_heston_qe_multi(const orc_heston* h, const double* strikes, int n_strikes, int type,
                         int steps, uint64_t path_count, uint64_t seed, unsigned threads,
                         orc_result* out);
double orc_heston_semi_analytic(const orc_heston* h, double strike, int type);
void orc_gauss_legendre_rule(int order, double* nodes, double* weights);

/* ---- LSM (src/pricers/lsm.cpp, src/core/linalg.cpp) ---- */
void orc_householder_ls(double* a, int rows, int cols, double* y, double* beta);
/* coeffs: (steps+1) x 4, has_coeffs: steps+1 flags; returns price, se. */
orc_result orc_lsm_american(double spot, double strike, double rate, double carry, double vol,
                            double time, int type, int steps, uint64_t path_count,
                            uint64_t seed, double* coeffs, uint8_t* has_coeffs,
                            uint8_t* exercise_at_inception);
/* LSM on Sobol + Brownian-bridge paths (this engine's extension): path i = Sobol index i + 1;
 * shifted != 0 applies replicate `replicate`'s digital shift (orc_rqmc_shift) at the cell
 * midpoint. The regression is orc_lsm_american's. */
orc_result orc_lsm_american_qmc(double spot, double strike, double rate, double carry,
                                double vol, double time, int type, int steps,
                                uint64_t path_count, int shifted, uint64_t shift_seed,
                                uint32_t replicate, double* coeffs, uint8_t* has_coeffs,
                                uint8_t* exercise_at_inception);
/* orc_lsm_american with the paths split over `threads` (a CPU baseline): per-chunk QR blocks
 * stacked in thread order; equal to the serial price up to rounding-flipped decisions. */
orc_result orc_lsm_american_mt(double spot, double strike, double rate, double carry,
                               double vol, double time, int type, int steps,
                               uint64_t path_count, uint64_t seed, unsigned threads,
                               double* coeffs, uint8_t* has_coeffs,
                               uint8_t* exercise_at_inception);
double orc_lsm_reprice_frozen(double spot, double strike, double rate, double carry, double vol,
                              double time, int type, int steps, uint64_t path_count,
                              uint64_t seed, const double* coeffs, const uint8_t* has_coeffs,
                              uint8_t exercise_at_inception);

/* ---- Analytic + CRR (src/pricers/analytic.cpp, src/pricers/binomial.cpp) ---- */
double orc_bsm(double s, double k, double r, double q, double v, double t, int type);
double orc_black76(double f, double k, double r, double v, double t, int type);
double orc_kemna_vorst(double s, double k, double r, double q, double v, double t, int type);
double orc_reiner_rubinstein(double s, double k, double h, double r, double q, double v,
                             double t, int type, int direction, int knock, double rebate);
double orc_lookback_fixed(double s, double k, double r, double q, double v, double t, int type);
double orc_lookback_floating(double s, double r, double q, double v, double t, int type);
double orc_digital(double s, double k, double r, double q, double v, double t, int type,
                   int style, double cash);
double orc_crr(double s, double k, double r, double q, double v, double t, int steps, int type,
               double* pi, double* up, double* down);
double orc_crr_american(double s, double k, double r, double q, double v, double t, int steps,
                        int type);
/* Bermudan on the CRR tree: exercise at levels that are multiples of exercise_every. */
double orc_crr_bermudan(double s, double k, double r, double q, double v, double t, int steps,
                        int type, int exercise_every);

/* ---- Greeks (src/greeks) ---- */
typedef struct orc_greek {
    double value, standard_error;
} orc_greek;
/* out: delta, vega, rho */
void orc_pathwise_european(double s, double k, double r, double q, double v, double t, int type,
                           uint64_t n, uint64_t seed, orc_greek out[3]);
void orc_pathwise_asian(double s, double k, double r, double q, double v, double t, int type,
                        int strike_style, int average_style, int m, uint64_t n, uint64_t seed,
                        orc_greek out[3]);
/* out: delta, gamma, vega, rho, theta (theta unused for barrier) */
void orc_lr_terminal(double s, double k, double r, double q, double v, double t, int type,
                     int digital, int digital_style, double cash, uint64_t n, uint64_t seed,
                     orc_greek out[5]);
void orc_lr_barrier(double s, double k, double h, double r, double q, double v, double t,
                    int type, int direction, int knock, double rebate, int m, uint64_t n,
                    uint64_t seed, orc_greek out[4]);

#ifdef __cplusplus
}
#endif
#endif
