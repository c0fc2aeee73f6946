"""GPU parity of the Heston QE kernel against the CPU oracle (mcdx C ABI).

Bars: FP64 per-path payoffs within 1e-12 relative of the oracle, FP64 prices within 1e-12
relative of the oracle's sequential Welford and of the reference fixture; FP32-generation
prices within 0.25 SE of FP64 on the same paths and within 3 SE of the semi-analytic price
(the reference's acceptance bar, tests/heston_test.cpp:111-140); many strikes from one pass
equal single-strike calls bit for bit.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import mcdx
import oracle_lib as O

pytestmark = pytest.mark.gpu
FP32, FP64 = mcdx.PREC_FP32, mcdx.PREC_FP64
L = O.lib()
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_small.json")))
LEWIS = dict(spot=100.0, rate=0.01, carry_yield=0.02, v0=0.04, kappa=4.0, theta=0.25, xi=1.0,
             rho=-0.5, time=1.0)
C3 = dict(spot=100.0, rate=0.03, carry_yield=0.0, v0=0.04, kappa=2.0, theta=0.04, xi=0.5,
          rho=-0.7, time=1.0)


@pytest.mark.parametrize("pars,strike,typ,steps", [(LEWIS, 100.0, 0, 20), (C3, 90.0, 1, 50),
                                                   (C3, 120.0, 0, 252)])
def test_path_samples_fp64_match_oracle(pars, strike, typ, steps):
    got = mcdx.heston_samples(mcdx.heston(**pars), strike, typ, steps, 7, FP64, 500, 2048)
    h = O.heston(**pars)
    want = np.array([L.orc_heston_path_sample(C.byref(h), strike, typ, steps, 7, 500 + i)
                     for i in range(2048)])
    assert np.max(np.abs(got - want) / np.maximum(1.0, np.abs(want))) <= 1e-12


def test_price_fp64_matches_reference_fixture():
    g = GOLD["heston_qe"]
    (p, se, n), = mcdx.heston_strikes(mcdx.heston(**LEWIS), [g["strike"]], 0, g["steps"],
                                      g["paths"], g["seed"], FP64)
    assert n == g["paths"]
    assert abs(p - float.fromhex(g["price"])) <= 1e-12 * p
    assert abs(se - float.fromhex(g["se"])) <= 1e-9 * se


def test_many_strikes_equal_single_strike_calls():
    ks = [80.0, 85.0, 90.0, 95.0, 100.0, 105.0, 110.0, 115.0, 120.0]
    par = mcdx.heston(**C3)
    multi = mcdx.heston_strikes(par, ks, 0, 30, 100_003, 5, FP32)
    for k, m in zip(ks, multi):
        r = mcdx.Result()
        assert mcdx.lib().mcdx_heston_qe_strikes(par, (C.c_double * 1)(k), 1, 0, 30, 100_003, 5,
                                                 FP32, r) == 0
        assert (r.price, r.standard_error, r.path_count) == m


def test_fp32_bias_and_semi_analytic():
    # C3 parameters (sub-Feller: 2 kappa theta = 0.16 < xi^2 = 0.25), 9 strikes
    ks = [80.0 + 5 * i for i in range(9)]
    par = mcdx.heston(**C3)
    r32 = mcdx.heston_strikes(par, ks, 0, 100, 1 << 18, 42, FP32)
    r64 = mcdx.heston_strikes(par, ks, 0, 100, 1 << 18, 42, FP64)
    h = O.heston(**C3)
    for k, a, b in zip(ks, r32, r64):
        assert abs(a[0] - b[0]) <= 0.25 * b[1], (k, a, b)
        sa = L.orc_heston_semi_analytic(C.byref(h), k, 0)
        assert abs(a[0] - sa) < 3 * a[1] and abs(b[0] - sa) < 3 * b[1], (k, a, b, sa)


def test_lewis_within_3se_via_reference_signature():
    # reference bindings/python/tests/test_smoke.py: 50 steps, 100k paths, seed 42
    r = mcdx.Result()
    assert mcdx.lib().mcdx_heston_qe_european(mcdx.heston(**LEWIS), 100.0, 0, 50, 100_000, 42,
                                              r) == 0
    assert abs(r.price - 16.070154917028834) < 3 * r.standard_error


def test_no_nan_and_reproducible():
    par = mcdx.heston(**C3)
    a = mcdx.heston_strikes(par, [100.0], 0, 252, 50_000, 9, FP32)
    b = mcdx.heston_strikes(par, [100.0], 0, 252, 50_000, 9, FP32)
    assert a == b and np.isfinite(a[0][0])


def test_invalid_heston_arguments():
    r = (mcdx.Result * 40)()
    ks = (C.c_double * 40)(*([100.0] * 40))
    par = mcdx.heston(**C3)
    assert mcdx.lib().mcdx_heston_qe_strikes(par, ks, 33, 0, 10, 100, 1, FP32, r) == mcdx.EINVAL
    assert mcdx.lib().mcdx_heston_qe_strikes(par, ks, 1, 0, 0, 100, 1, FP32, r) == mcdx.EINVAL


def test_c3_full_size_semi_analytic_and_fp64():
    # The C3 workload itself: 2^22 paths x 252 steps, 9 strikes from one pass. Every strike
    # within 3 SE of the semi-analytic price in both modes; FP32 generation within 0.01 SE of
    # the FP64 mode on the same paths.
    ks = [80.0 + 5 * i for i in range(9)]
    par = mcdx.heston(**C3)
    r32 = mcdx.heston_strikes(par, ks, 0, 252, 1 << 22, 42, FP32)
    r64 = mcdx.heston_strikes(par, ks, 0, 252, 1 << 22, 42, FP64)
    h = O.heston(**C3)
    for k, a, b in zip(ks, r32, r64):
        assert a[2] == b[2] == 1 << 22
        assert abs(a[0] - b[0]) <= 0.01 * b[1], (k, a, b)
        sa = L.orc_heston_semi_analytic(C.byref(h), k, 0)
        assert abs(a[0] - sa) < 3 * a[1] and abs(b[0] - sa) < 3 * b[1], (k, a, b, sa)
