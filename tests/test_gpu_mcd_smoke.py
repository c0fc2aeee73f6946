"""The reference's Python smoke suite (bindings/python/tests/test_smoke.py:15-268), assertion
for assertion, against this repository's `mcd` module on the GPU."""
import threading
import time

import pytest

import mcd

pytestmark = pytest.mark.gpu
C, P = mcd.OptionType.Call, mcd.OptionType.Put


def test_every_product_prices_with_a_confidence_interval():
    seed, n = 42, 100_000
    res = [
        mcd.monte_carlo_european(100, 100, 0.05, 0.0, 0.2, 1.0, C, n, seed),
        mcd.monte_carlo_digital(100, 100, 0.05, 0.0, 0.2, 1.0, C, mcd.DigitalStyle.CashOrNothing,
                                1.0, n, seed),
        mcd.monte_carlo_asian(100, 100, 0.05, 0.0, 0.2, 1.0, C, mcd.StrikeStyle.Fixed,
                              mcd.AverageStyle.Geometric, 12, n, seed),
        mcd.monte_carlo_barrier(100, 100, 120, 0.05, 0.0, 0.2, 1.0, C, mcd.BarrierDirection.Up,
                                mcd.BarrierKnock.Out, 0.0, 20, n, seed),
        mcd.monte_carlo_lookback(100, 100, 0.05, 0.0, 0.2, 1.0, C, mcd.StrikeStyle.Floating, 20,
                                 n, seed),
    ]
    for r in res:
        assert r.standard_error > 0.0 and r.path_count == n


def test_european_within_three_se_of_analytic():
    mc = mcd.monte_carlo_european(100, 100, 0.05, 0.0, 0.2, 1.0, C, 500_000, 42)
    analytic = mcd.black_scholes_merton(100, 100, 0.05, 0.0, 0.2, 1.0, C)
    assert abs(mc.price - analytic) < 3.0 * mc.standard_error


def test_american_lsm_prices():
    r = mcd.monte_carlo_lsm_american(100, 100, 0.05, 0.0, 0.25, 1.0, P, 20, 50_000, 7)
    assert r.price > 0.0 and r.standard_error >= 0.0


def test_finite_difference_european_greeks():
    bumps = mcd.default_bump_sizes(100, 0.2, 1.0)
    g = mcd.finite_difference_european(100, 100, 0.05, 0.0, 0.2, 1.0, C, 300_000, 42, bumps)
    assert 0.0 < g.delta < 1.0 and g.vega > 0.0


def test_likelihood_ratio_greeks():
    lr = mcd.likelihood_ratio_european(100, 100, 0.05, 0.0, 0.2, 1.0, C, 300_000, 42)
    assert 0.0 < lr.delta.value < 1.0 and lr.delta.standard_error > 0.0
    assert lr.theta is not None
    d = mcd.likelihood_ratio_digital(100, 100, 0.05, 0.0, 0.2, 1.0, C,
                                     mcd.DigitalStyle.CashOrNothing, 1.0, 200_000, 7)
    assert d.gamma.standard_error > 0.0
    b = mcd.likelihood_ratio_barrier(100, 100, 120, 0.05, 0.0, 0.2, 1.0, C,
                                     mcd.BarrierDirection.Up, mcd.BarrierKnock.Out, 0.0, 50,
                                     100_000, 11)
    assert b.gamma.standard_error > 0.0 and b.theta is None


def test_lr_gamma_se_below_fd_gamma_se_at_the_money_digital():
    s, k, r, q, sigma, t = 100.0, 100.0, 0.05, 0.0, 0.25, 1.0
    n, seed = 200_000, 7
    lr = mcd.likelihood_ratio_digital(s, k, r, q, sigma, t, C, mcd.DigitalStyle.CashOrNothing,
                                      1.0, n, seed)
    b = mcd.default_bump_sizes(s, sigma, t)
    up = mcd.monte_carlo_digital(s + b.spot, k, r, q, sigma, t, C, mcd.DigitalStyle.CashOrNothing, 1.0, n, seed)
    mid = mcd.monte_carlo_digital(s, k, r, q, sigma, t, C, mcd.DigitalStyle.CashOrNothing, 1.0, n, seed)
    dn = mcd.monte_carlo_digital(s - b.spot, k, r, q, sigma, t, C, mcd.DigitalStyle.CashOrNothing, 1.0, n, seed)
    fd_se = ((up.standard_error ** 2 + 4 * mid.standard_error ** 2 + dn.standard_error ** 2)
             ** 0.5) / (b.spot ** 2)
    assert lr.gamma.standard_error < fd_se


def test_pathwise_greeks():
    pw = mcd.pathwise_european(100, 100, 0.05, 0.0, 0.2, 1.0, C, 300_000, 42)
    assert 0.0 < pw.delta.value < 1.0 and pw.delta.standard_error > 0.0
    assert not hasattr(pw, "gamma") and not hasattr(pw, "theta")
    pa = mcd.pathwise_asian(100, 100, 0.05, 0.0, 0.2, 1.0, C, mcd.StrikeStyle.Fixed,
                            mcd.AverageStyle.Arithmetic, 12, 100_000, 42)
    assert pa.delta.standard_error > 0.0


def test_qmc():
    analytic = mcd.black_scholes_merton(100, 100, 0.05, 0.02, 0.2, 1.0, C)
    q = mcd.qmc_sobol_european(100, 100, 0.05, 0.02, 0.2, 1.0, C, 200_000)
    assert abs(q.price - analytic) < 0.02 and not hasattr(q, "standard_error")
    mc = mcd.monte_carlo_asian(100, 100, 0.05, 0.02, 0.25, 1.0, C, mcd.StrikeStyle.Fixed,
                               mcd.AverageStyle.Arithmetic, 7, 100_000, 42)
    qa = mcd.qmc_sobol_asian(100, 100, 0.05, 0.02, 0.25, 1.0, C, mcd.StrikeStyle.Fixed, 7, 100_000)
    assert abs(qa.price - mc.price) < 3.0 * mc.standard_error + 0.1
    with pytest.raises(ValueError):
        mcd.qmc_sobol_asian(100, 100, 0.05, 0.02, 0.25, 1.0, C, mcd.StrikeStyle.Fixed,
                            mcd.SOBOL_MAX_DIMENSIONS + 1, 10_000)


def _lewis():
    p = mcd.HestonParams()
    p.spot, p.rate, p.carry_yield = 100.0, 0.01, 0.02
    p.v0, p.kappa, p.theta, p.xi, p.rho = 0.04, 4.0, 0.25, 1.0, -0.5
    p.time = 1.0
    return p


def test_heston():
    p = _lewis()
    assert abs(mcd.heston_semi_analytic(p, 100.0, C) - 16.070154917028834278) < 1e-6
    mc = mcd.heston_qe_european(p, 100.0, C, 50, 100_000, 42)
    assert abs(mc.price - mcd.heston_semi_analytic(p, 100.0, C)) < 3.0 * mc.standard_error
    a = mcd.heston_qe_european(p, 100.0, C, 20, 10_000, 7)
    b = mcd.heston_qe_european(p, 100.0, C, 20, 10_000, 7)
    assert a.price == b.price


def test_benchmark_harness():
    r = mcd.benchmark_european(100, 100, 0.05, 0.0, 0.2, 1.0, C, 200_000, 42, mcd.McOptions())
    for key in ("price", "elapsed_seconds", "paths_per_second"):
        assert key in r
    assert r["elapsed_seconds"] > 0.0 and r["paths_per_second"] > 0.0
    r = mcd.benchmark_european(100, 100, 0.05, 0.0, 0.2, 1.0, C, 10_000, 42, mcd.McOptions())
    assert r["simd_enabled"] == mcd.HAS_NEON
    a = mcd.monte_carlo_european(100, 100, 0.05, 0.02, 0.2, 1.0, C, 50_000, 7)
    b = mcd.monte_carlo_european(100, 100, 0.05, 0.02, 0.2, 1.0, C, 50_000, 7)
    assert a.price == b.price


def test_gil_released_during_pricing():
    # The reference (test_smoke.py:230-264) checks that pricing releases the GIL by timing two
    # concurrent calls against one. On one GPU two calls share the device, so their wall time
    # mostly measures the device; this checks the property itself, with a margin that does not
    # depend on timing noise: while the main thread is inside a pricing call of at least 50 ms,
    # a pure-Python thread keeps counting at most of its solo rate. Were the GIL held for the
    # call, it could count only until the call took the GIL (one 5 ms switch interval at most).
    def call(n):
        mcd.monte_carlo_asian(100, 100, 0.05, 0.0, 0.2, 1.0, C, mcd.StrikeStyle.Fixed,
                              mcd.AverageStyle.Arithmetic, 252, n, 1)

    n = 1 << 22
    call(n)  # warm: per-thread engine state and the kernel's first load
    while True:
        t0 = time.perf_counter()
        call(n)
        busy = time.perf_counter() - t0
        if busy >= 0.05 or n >= 1 << 28:
            break
        n *= 2
    assert busy >= 0.05

    def counted(seconds_or_call):
        stop, started, out = threading.Event(), threading.Event(), []

        def spin():
            c = 0
            started.set()
            while not stop.is_set():
                c += 1
            out.append(c)

        t = threading.Thread(target=spin)
        t.start()
        started.wait()
        t0 = time.perf_counter()
        if callable(seconds_or_call):
            seconds_or_call()
        else:
            time.sleep(seconds_or_call)
        dt = time.perf_counter() - t0
        stop.set()
        t.join()
        return out[0] / dt

    solo = counted(busy)                 # the main thread sleeps: the spinner has the GIL
    during = counted(lambda: call(n))    # the main thread prices
    assert during > 0.5 * solo, (during, solo, busy)
