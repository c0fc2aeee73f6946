"""GPU parity of Sobol + Brownian-bridge QMC against the CPU oracle (mcdx C ABI).

Bars: Sobol integers bit-exact for every dimension of the extended table (dims 0-6 are the
reference's); FP64 QMC prices within 1e-12 relative of the oracle (the point sum is merged
in a different order); FP32 generation within 2e-6 relative of FP64 on the same points.
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


@pytest.mark.parametrize("dim", [0, 1, 2, 3, 4, 5, 6, 7, 8, 40, 100, 251, 252, 511, 1023])
@pytest.mark.parametrize("index0", [0, 1000, (1 << 20) - 100, (1 << 31) - 7])
def test_sobol_ints_bit_exact(dim, index0):
    got = mcdx.sobol_ints(dim, index0, 2053)
    want = np.array([L.orc_sobol_int(dim, index0 + k) for k in range(2053)], dtype=np.uint32)
    assert np.array_equal(got, want)


def test_sobol_dims_0_to_6_equal_reference_fixture():
    for d, pts in GOLD["sobol"].items():
        got = mcdx.sobol_ints(int(d), 0, len(pts)).astype(np.float64) / 2 ** 32
        assert np.array_equal(got, np.array([float.fromhex(x) for x in pts]))


def test_dimension_caps():
    assert mcdx.lib().mcdx_sobol_max_dimensions() == 7
    assert mcdx.lib().mcdx_sobol_extended_dimensions() == 1024
    price = mcdx.C.c_double()
    assert mcdx.lib().mcdx_qmc_sobol_asian(100, 100, .05, .02, .25, 1, 0, 0, 8, 1000,
                                           mcdx.C.byref(price)) == mcdx.EINVAL


@pytest.mark.parametrize("m,n,ss,typ", [(1, 5000, 0, 0), (7, 20000, 0, 0), (5, 5000, 1, 1),
                                        (12, 4099, 0, 0), (64, 3000, 0, 1), (252, 2048, 0, 0)])
def test_qmc_asian_fp64_matches_oracle(m, n, ss, typ):
    got, _, cnt = mcdx.qmc_asian(100, 95, .05, .01, .3, 1, typ, ss, m, n, FP64)
    want = L.orc_qmc_sobol_asian(100, 95, .05, .01, .3, 1, typ, ss, m, n)
    assert cnt == n
    assert abs(got - want) <= 1e-12 * max(1.0, abs(want))


def test_qmc_fixtures_of_reference():
    q = GOLD["qmc"]
    got = mcdx.qmc_asian(100, 100, .05, .02, .25, 1, 0, 0, 7, 20000, FP64)[0]
    assert abs(got - float.fromhex(q["asian7"])) <= 1e-12 * got
    got = mcdx.qmc_european(100, 100, .05, .02, .2, 1, 0, 20000, FP64)
    assert abs(got - float.fromhex(q["european"])) <= 1e-12 * got


def test_qmc_fp32_close_to_fp64():
    for m in (7, 252):
        a = mcdx.qmc_asian(100, 100, .05, 0, .2, 1, 0, 0, m, 1 << 16, FP32)[0]
        b = mcdx.qmc_asian(100, 100, .05, 0, .2, 1, 0, 0, m, 1 << 16, FP64)[0]
        assert abs(a - b) <= 2e-6 * b, (m, a, b)


def test_qmc_european_near_bsm():
    # reference bindings/python/tests/test_smoke.py:149-156
    bsm = L.orc_bsm(100, 100, .05, .02, .2, 1, 0)
    for prec in (FP32, FP64):
        assert abs(mcdx.qmc_european(100, 100, .05, .02, .2, 1, 0, 200_000, prec) - bsm) < 0.02


def test_qmc_asian_252_converges_to_mc():
    # the 252-dimension QMC Asian agrees with plain MC on the same contract within 3 SE
    q = mcdx.qmc_asian(100, 100, .05, 0, .2, 1, 0, 0, 252, 1 << 20, FP32)[0]
    c = mcdx.contract(mcdx.ASIAN, monitoring_points=252, path_count=1 << 20, seed=3)
    p, se, _ = mcdx.price(c, FP32)
    assert abs(q - p) < 3 * se


def qmc_contract(product, m, n, **kw):
    base = dict(spot=100.0, strike=100.0, rate=0.05, carry_yield=0.0, vol=0.25, time=1.0,
                monitoring_points=m, path_count=n)
    base.update(kw)
    return mcdx.contract(product, **base)


def oracle_qmc(c):
    p = O.product(c.product, spot=c.spot, strike=c.strike, barrier=c.barrier, rebate=c.rebate,
                  rate=c.rate, carry_yield=c.carry_yield, vol=c.vol, time=c.time,
                  monitoring_points=c.monitoring_points, type=c.type, direction=c.direction,
                  knock=c.knock, strike_style=c.strike_style)
    return O.lib().orc_qmc_sobol_path(C.byref(p), c.path_count)


PATH_CASES = [
    (mcdx.BARRIER, dict(barrier=120.0, direction=0, knock=1)),
    (mcdx.BARRIER, dict(barrier=85.0, direction=1, knock=0, type=1, rebate=1.5)),
    (mcdx.LOOKBACK, dict(strike_style=1)),
    (mcdx.LOOKBACK, dict(strike_style=0, strike=95.0, type=1)),
    (mcdx.ASIAN, dict(strike_style=0)),
]


@pytest.mark.parametrize("product,kw", PATH_CASES)
@pytest.mark.parametrize("m,n", [(7, 3000), (50, 1 << 12), (252, 1025)])
def test_qmc_path_products_fp64_match_oracle(product, kw, m, n):
    c = qmc_contract(product, m, n, **kw)
    got = mcdx.qmc_path(c, FP64)
    want = oracle_qmc(c)
    assert abs(got - want) <= 1e-12 * max(1.0, abs(want)), (got, want)


@pytest.mark.parametrize("product,kw", PATH_CASES)
def test_qmc_path_products_fp32_close_and_converge_to_mc(product, kw):
    c = qmc_contract(product, 252, 1 << 18, **kw)
    q32, q64 = mcdx.qmc_path(c, FP32), mcdx.qmc_path(c, FP64)
    assert abs(q32 - q64) <= 2e-4 * max(1.0, abs(q64)), (q32, q64)
    m = mcdx.GbmContract.from_buffer_copy(c)
    m.path_count = 1 << 22
    m.seed = 42
    p, se, _ = mcdx.price(m, FP32)
    assert abs(q64 - p) <= 4 * se + 1e-3, (q64, p, se)
