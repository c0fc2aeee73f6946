"""GPU parity of Sobol + Brownian-bridge QMC against the CPU oracle (mcdx C ABI).

Bars: Sobol integers bit-exact for every dimension of the extended table (dims 0-6 are the
reference's); FP64 QMC prices within 1e-12 relative of the oracle (the point sum is merged
in a different order); FP32 generation within 2e-6 relative of FP64 on the same points.
"""
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
