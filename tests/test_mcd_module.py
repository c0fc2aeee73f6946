"""CPU checks of the drop-in `mcd` Python module: it imports without a GPU, exposes the
reference's names (bindings/python/src/module.cpp:64-293), its host-side closed forms match the
oracle and the reference fixtures, and engine-backed calls fail loudly without a device."""
import math

import pytest

import mcd
import oracle_lib as O
import json
import os

L = O.lib()
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_small.json")))

REFERENCE_NAMES = [
    "OptionType", "BarrierDirection", "BarrierKnock", "StrikeStyle", "DigitalStyle",
    "AverageStyle", "McResult", "McOptions", "BinomialResult", "LsmResult", "EuropeanGreeks",
    "BumpSizes", "LrGreeksResult", "LrGreeks", "PathwiseGreeksResult", "PathwiseGreeks",
    "QmcResult", "HestonParams", "HestonMcResult", "SOBOL_MAX_DIMENSIONS", "HAS_NEON",
    "black_scholes_merton", "black76", "kemna_vorst", "reiner_rubinstein",
    "lookback_fixed_strike", "lookback_floating_strike", "digital", "forward_price",
    "forward_value", "futures_mark_to_market", "crr_binomial", "crr_binomial_american",
    "monte_carlo_european", "monte_carlo_digital", "monte_carlo_asian", "monte_carlo_barrier",
    "monte_carlo_lookback", "monte_carlo_lsm_american", "default_bump_sizes",
    "finite_difference_european", "likelihood_ratio_european", "likelihood_ratio_digital",
    "likelihood_ratio_barrier", "pathwise_european", "pathwise_asian", "qmc_sobol_european",
    "qmc_sobol_asian", "heston_qe_european", "heston_semi_analytic", "benchmark_european"]


def test_module_exposes_reference_api():
    missing = [n for n in REFERENCE_NAMES if not hasattr(mcd, n)]
    assert not missing, missing
    assert mcd.SOBOL_MAX_DIMENSIONS == 7 and mcd.HAS_NEON is False
    assert not hasattr(mcd, "pathwise_digital_delta_naive_and_broken")
    assert int(mcd.OptionType.Put) == 1 and int(mcd.BarrierKnock.Out) == 1


# Every attribute the reference's Python callers read from the module's results
# (infra/lambda/request.py:166-222, 253-310 and tools/generate_report.py): with these the
# callers run unchanged on this module.
CALLER_FIELDS = {
    "McResult": ["price", "standard_error", "path_count"],
    "BinomialResult": ["price", "up_factor", "down_factor", "risk_neutral_probability"],
    "EuropeanGreeks": ["delta", "gamma", "vega", "rho", "theta"],
    "LrGreeks": ["delta", "gamma", "vega", "rho", "theta"],
    "LrGreeksResult": ["value", "standard_error"],
    "PathwiseGreeks": ["delta", "vega", "rho"],
    "PathwiseGreeksResult": ["value", "standard_error"],
    "QmcResult": ["price"],
    "LsmResult": ["price", "standard_error"],
    "HestonMcResult": ["price", "standard_error", "path_count"],
}


def test_reference_callers_find_every_result_field():
    for cls, fields in CALLER_FIELDS.items():
        missing = [f for f in fields if not hasattr(getattr(mcd, cls), f)]
        assert not missing, (cls, missing)
    # PathwiseGreeks has no gamma or theta at all (pathwise.hpp:19-25), as the callers expect
    assert not hasattr(mcd.PathwiseGreeks, "gamma") and not hasattr(mcd.PathwiseGreeks, "theta")
    o = mcd.McOptions()  # request.py:215-222 sets the variance-reduction toggles
    o.antithetic, o.control_variate = True, True
    assert o.antithetic and o.control_variate and not o.brownian_bridge


def test_closed_forms_match_reference():
    C, P = mcd.OptionType.Call, mcd.OptionType.Put
    assert abs(mcd.black_scholes_merton(100, 100, 0.05, 0.0, 0.2, 1.0, C) - float.fromhex(GOLD["bsm_c1"])) < 1e-13
    assert abs(mcd.kemna_vorst(100, 100, .05, 0, .2, 1, C) - float.fromhex(GOLD["kemna_vorst"])) < 1e-12
    assert abs(mcd.lookback_fixed_strike(100, 90, .05, .01, .2, 1, C) - float.fromhex(GOLD["lookback_fixed"])) < 1e-12
    assert abs(mcd.lookback_floating_strike(100, .05, .01, .2, 1, P) - float.fromhex(GOLD["lookback_floating"])) < 1e-12
    dirs = [mcd.BarrierDirection.Up, mcd.BarrierDirection.Down]
    knocks = [mcd.BarrierKnock.In, mcd.BarrierKnock.Out]
    for h, t, d, k, v in GOLD["reiner_rubinstein"]:
        got = mcd.reiner_rubinstein(100, 100, h, .05, .01, .25, 1, [C, P][t], dirs[d], knocks[k], 1.5)
        assert abs(got - float.fromhex(v)) < 1e-11
    for n, v in GOLD["crr_american"].items():
        assert abs(mcd.crr_binomial_american(100, 100, .06, 0, .2, 1, int(n), P) - float.fromhex(v)) < 1e-11
    assert abs(mcd.digital(100, 100, .05, 0, .2, 1, C, mcd.DigitalStyle.CashOrNothing) -
               L.orc_digital(100, 100, .05, 0, .2, 1, 0, 0, 1.0)) < 1e-13
    assert abs(mcd.black76(100, 95, .05, .2, 1, P) - L.orc_black76(100, 95, .05, .2, 1, 1)) < 1e-13
    # reference bindings/python/tests/test_smoke.py:15-28
    assert abs(mcd.forward_price(100, 0.05, 0.02, 1.0) - 100 * math.exp(0.03)) < 1e-9
    b = mcd.crr_binomial(100, 100, 0.05, 0.0, 0.20, 1.0, 2000, C)
    assert abs(b.price - mcd.black_scholes_merton(100, 100, .05, 0, .2, 1, C)) < 0.05


def test_heston_semi_analytic_lewis():
    p = mcd.HestonParams()
    p.spot, p.rate, p.carry_yield = 100.0, 0.01, 0.02
    p.v0, p.kappa, p.theta, p.xi, p.rho, p.time = 0.04, 4.0, 0.25, 1.0, -0.5, 1.0
    assert abs(mcd.heston_semi_analytic(p, 100.0, mcd.OptionType.Call) - 16.070154917028834278) < 1e-9


def test_default_bumps():
    b = mcd.default_bump_sizes(100, 0.2, 1.0)
    assert (b.spot, b.vol, b.rate, b.time) == (1.0, 0.002, 0.01, 0.01)


def test_engine_call_fails_loudly_without_device():
    import mcdx
    n = mcdx.C.c_int(0)
    mcdx.lib().mcdx_device_count(mcdx.C.byref(n))
    if n.value > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError):
        mcd.monte_carlo_european(100, 100, 0.05, 0.0, 0.2, 1.0, mcd.OptionType.Call, 1000, 42)
    with pytest.raises(ValueError):
        mcd.qmc_sobol_asian(100, 100, .05, .02, .25, 1, mcd.OptionType.Call, mcd.StrikeStyle.Fixed, 8, 10)
