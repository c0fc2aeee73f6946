/* Drives every function of the CPU oracle (oracle/mcd_oracle.c) at small sizes, including its
 * edge cases (zero and ragged path counts, one monitoring date, degenerate least squares,
 * deep in- and out-of-the-money, extreme uniforms, several threads), for a build under
 * AddressSanitizer and UndefinedBehaviorSanitizer. Exits 0 when every result is finite where
 * it must be; the sanitizers abort on the first memory or undefined-behaviour error. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "mcd_oracle.h"

static int failures = 0;

static void finite(const char* what, double x) {
    if (!isfinite(x)) {
        printf("not finite: %s = %g\n", what, x);
        ++failures;
    }
}

static orc_product product(int kind) {
    orc_product p = {100.0, 100.0, 0.0, 0.0, 0.05, 0.01, 0.25, 1.0, 1.0, 1, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0};
    p.product = (uint8_t)kind;
    return p;
}

static void rng_and_normal(void) {
    const uint32_t ctr[4] = {1, 2, 3, 4}, key[2] = {5, 6};
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    uint32_t bits[37];
    orc_philox_bits_range(7, 0xFFFFFFF0ull, 37, 3, 1, bits); /* crosses 2^32 */
    finite("uniform lo", orc_uniform_open01(0));
    finite("uniform hi", orc_uniform_open01(0xFFFFFFFFu));
    const double us[] = {1e-300, 1e-20, 0.5 / 4294967296.0, 0.02424, 0.02425, 0.5, 0.97575,
                         1 - 1e-16, 1 - 0.5 / 4294967296.0};
    for (unsigned i = 0; i < sizeof us / sizeof us[0]; ++i) {
        finite("inverse normal", orc_inverse_normal_cdf(us[i]));
        finite("acklam", orc_acklam(us[i]));
    }
    finite("pdf", orc_normal_pdf(40.0));
    finite("cdf", orc_normal_cdf(-40.0));
    finite("draw", orc_standard_normal_variate(~0ull, ~0ull, ~0u));
    finite("uniform draw", orc_uniform_variate(0, 0, 0));
}

static void welford_and_chunks(void) {
    orc_welford a = {0, 0.0, 0.0}, b = {0, 0.0, 0.0};
    orc_welford_merge(&a, &b); /* empty with empty */
    for (int i = 0; i < 5; ++i) orc_welford_add(&a, i);
    orc_welford_merge(&b, &a);
    finite("variance", orc_welford_variance(&b));
    finite("se", orc_welford_se(&b));
    uint64_t lo, hi;
    for (unsigned c = 0; c < 9; ++c) orc_chunk_bounds(c, 5, 9, &lo, &hi);
}

static void mc_products(void) {
    const uint64_t counts[] = {0, 1, 3, 1025};
    for (int kind = ORC_EUROPEAN; kind <= ORC_LOOKBACK; ++kind) {
        for (int variant = 0; variant < 4; ++variant) {
            orc_product p = product(kind);
            p.monitoring_points = variant == 0 ? 1 : 7 + variant;
            p.type = (uint8_t)(variant & 1);
            p.antithetic = (uint8_t)(variant >> 1);
            p.strike_style = (uint8_t)(variant & 1);
            p.average_style = (uint8_t)((variant >> 1) & 1);
            p.digital_style = (uint8_t)(variant & 1);
            p.control_variate = (uint8_t)(kind == ORC_ASIAN && variant == 3);
            p.barrier = variant & 1 ? 80.0 : 120.0;
            p.direction = (uint8_t)(variant & 1);
            p.knock = (uint8_t)(variant >> 1);
            p.brownian_bridge = (uint8_t)(kind == ORC_BARRIER && variant >= 2);
            for (unsigned c = 0; c < sizeof counts / sizeof counts[0]; ++c) {
                const orc_result r = orc_price(&p, counts[c], 42, 4, 3);
                if (counts[c] > 1) finite("price", r.price);
            }
            double s[5];
            orc_path_samples(&p, 9, 0xFFFFFFFEull, 5, s);
            finite("path sample", orc_path_sample(&p, 9, 3));
            if (kind != ORC_EUROPEAN && kind != ORC_DIGITAL) {
                finite("qmc path", orc_qmc_sobol_path(&p, 257));
                finite("rqmc", orc_rqmc_replicate(&p, 129, 77, 2));
            }
        }
    }
    orc_product e = product(ORC_EUROPEAN);
    finite("rqmc european", orc_rqmc_replicate(&e, 100, 1, 0));
}

static void sobol_and_bridge(void) {
    const unsigned dims = orc_sobol_max_dims();
    uint32_t v[32];
    for (unsigned d = 0; d < dims; d += 37) {
        (void)orc_sobol_direction_table(d, v);
        finite("sobol", orc_sobol_point(d, (1ull << 31) + d));
        (void)orc_sobol_int(d, 0);
    }
    (void)orc_sobol_direction_table(dims, v); /* out of range: an error status */
    const int ns[] = {1, 2, 3, 252};
    for (unsigned k = 0; k < sizeof ns / sizeof ns[0]; ++k) {
        const int n = ns[k];
        double* z = (double*)malloc(sizeof(double) * (size_t)n);
        double* sp = (double*)malloc(sizeof(double) * (size_t)n);
        for (int i = 0; i < n; ++i) z[i] = orc_standard_normal_variate(3, 5, (uint32_t)i);
        orc_brownian_bridge_path(100, 0.05, 0.0, 0.2, 1.0, n, z, sp);
        finite("bridge", sp[n - 1]);
        free(z);
        free(sp);
    }
    finite("qmc european", orc_qmc_sobol_european(100, 100, .05, 0, .2, 1, 1, 1000));
    finite("qmc asian", orc_qmc_sobol_asian(100, 100, .05, 0, .2, 1, 0, 1, 7, 1000));
    (void)orc_rqmc_shift(1, 2, 3);
}

static void heston(void) {
    const orc_heston h = {100.0, 0.03, 0.0, 0.04, 2.0, 0.04, 0.5, -0.7, 1.0};
    finite("qe step", orc_qe_variance_step(0.04, 2.0, 0.04, 0.5, 1.0 / 252, 0.3));
    finite("qe step zero", orc_qe_variance_step(0.0, 2.0, 0.04, 0.5, 1.0 / 252, 0.999));
    finite("heston sample", orc_heston_path_sample(&h, 100, 0, 12, 4, 5));
    const orc_result r = orc_heston_qe_european(&h, 90, 1, 12, 301, 7);
    finite("heston", r.price);
    const double ks[3] = {80, 100, 120};
    orc_result rs[3];
    orc_heston_qe_multi(&h, ks, 3, 0, 12, 301, 7, 5, rs);
    orc_heston_qe_multi(&h, ks, 3, 0, 12, 3, 7, 8, rs); /* more threads than paths */
    finite("semi-analytic", orc_heston_semi_analytic(&h, 100, 0));
    double nodes[64], weights[64];
    orc_gauss_legendre_rule(1, nodes, weights);
    orc_gauss_legendre_rule(64, nodes, weights);
}

static void lsm(void) {
    double co[21 * 4];
    uint8_t has[21], inc;
    double a[12] = {1, 2, 3, 0, 0, 0, 1, 2, 3, 4, 5, 6}, y[4] = {1, 2, 3, 4}, beta[3];
    orc_householder_ls(a, 4, 3, y, beta); /* a zero row and a repeated row */
    double z[6] = {0, 0, 0, 0, 0, 0}, yz[3] = {1, 1, 1}, bz[2];
    orc_householder_ls(z, 3, 2, yz, bz); /* all-zero columns */
    const uint64_t ns[] = {0, 1, 5, 2000};
    for (unsigned k = 0; k < sizeof ns / sizeof ns[0]; ++k) {
        orc_result r = orc_lsm_american(100, 100, .06, 0, .2, 1, 1, 20, ns[k], 3, co, has, &inc);
        if (ns[k] > 1) finite("lsm", r.price);
        r = orc_lsm_american_mt(100, 100, .06, 0, .2, 1, 1, 20, ns[k], 3, 7, co, has, &inc);
        if (ns[k] > 1) finite("lsm mt", r.price);
        finite("reprice", orc_lsm_reprice_frozen(101, 100, .06, 0, .21, 1, 1, 20, ns[k] + 1, 4,
                                                 co, has, inc));
    }
    (void)orc_lsm_american(30, 100, .06, 0, .2, 1, 1, 5, 100, 3, co, has, &inc); /* inception */
    (void)orc_lsm_american(300, 100, .06, 0, .2, 1, 1, 5, 100, 3, co, has, &inc); /* no ITM */
    (void)orc_lsm_american_qmc(100, 100, .06, 0, .2, 1, 1, 20, 500, 0, 0, 0, co, has, &inc);
    (void)orc_lsm_american_qmc(100, 100, .06, 0, .2, 1, 1, 20, 500, 1, 9, 3, co, has, &inc);
}

static void analytic_and_greeks(void) {
    for (int type = 0; type < 2; ++type) {
        finite("bsm", orc_bsm(100, 100, .05, .01, .2, 1, type));
        finite("black76", orc_black76(100, 90, .05, .2, 1, type));
        finite("kemna-vorst", orc_kemna_vorst(100, 100, .05, .01, .2, 1, type));
        for (int dir = 0; dir < 2; ++dir)
            for (int knock = 0; knock < 2; ++knock)
                finite("reiner-rubinstein",
                       orc_reiner_rubinstein(100, 100, dir ? 80 : 120, .05, .01, .2, 1, type, dir,
                                             knock, 1.0));
        finite("lookback fixed", orc_lookback_fixed(100, 100, .05, .01, .2, 1, type));
        finite("lookback floating", orc_lookback_floating(100, .05, .01, .2, 1, type));
        finite("digital", orc_digital(100, 100, .05, .01, .2, 1, type, type, 1.0));
        double pi, up, down;
        finite("crr", orc_crr(100, 100, .05, .01, .2, 1, 1, type, &pi, &up, &down));
        finite("crr null outs", orc_crr(100, 100, .05, .01, .2, 1, 50, type, NULL, NULL, NULL));
        finite("crr american", orc_crr_american(100, 100, .05, .01, .2, 1, 1, type));
        finite("crr bermudan", orc_crr_bermudan(100, 100, .05, .01, .2, 1, 60, type, 7));
        orc_greek g[5];
        orc_pathwise_european(100, 100, .05, .01, .2, 1, type, 999, 3, g);
        orc_pathwise_asian(100, 100, .05, .01, .2, 1, type, type, 1 - type, 5, 999, 3, g);
        orc_lr_terminal(100, 100, .05, .01, .2, 1, type, 0, 0, 1.0, 999, 3, g);
        orc_lr_terminal(100, 100, .05, .01, .2, 1, type, 1, type, 2.0, 999, 3, g);
        orc_lr_barrier(100, 100, 120, .05, .01, .2, 1, type, 0, 1, 0.5, 9, 999, 3, g);
        orc_lr_barrier(100, 100, 80, .05, .01, .2, 1, type, 1, 0, 0.0, 1, 1, 3, g);
    }
}

int main(void) {
    rng_and_normal();
    welford_and_chunks();
    mc_products();
    sobol_and_bridge();
    heston();
    lsm();
    analytic_and_greeks();
    printf("oracle sanitizer driver: %d failures\n", failures);
    return failures ? 1 : 0;
}
