"""GPU Greeks (pathwise, likelihood ratio, fused CRN finite differences) against the oracle
and the reference fixtures. Arithmetic is FP64 as in the reference; only the Welford merge
tree differs, so values and standard errors agree to 1e-12 relative (1e-10 for FD, whose
reference prices are chunk-merged)."""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

import mcdx
import oracle_lib as O

pytestmark = pytest.mark.gpu
L = O.lib()
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_small.json")))


def close(a, b, tol=1e-12):
    return abs(a - b) <= tol * max(1.0, abs(b))


def test_pathwise_european_vs_fixture_and_oracle():
    got = mcdx.greeks("mcdx_pathwise_european", 100, 100, .05, 0, .2, 1, 0, 20_000, 42, n_out=3)
    want = [float.fromhex(x) for x in GOLD["pathwise_european"]]
    for i in range(3):
        assert close(got[i][0], want[2 * i]) and close(got[i][1], want[2 * i + 1])


@pytest.mark.parametrize("ss,avs,typ", [(0, 0, 0), (1, 1, 1), (0, 1, 0)])
def test_pathwise_asian_vs_oracle(ss, avs, typ):
    got = mcdx.greeks("mcdx_pathwise_asian", 100, 95, .05, .01, .25, 1, typ, ss, avs, 12, 10_007,
                      9, n_out=3)
    g = (O.Greek * 3)()
    L.orc_pathwise_asian(100, 95, .05, .01, .25, 1, typ, ss, avs, 12, 10_007, 9, g)
    for i in range(3):
        assert close(got[i][0], g[i].value) and close(got[i][1], g[i].standard_error)


def test_lr_terminal_vs_fixture_and_oracle():
    got = mcdx.greeks("mcdx_likelihood_ratio_european", 100, 100, .05, 0, .2, 1, 0, 20_000, 42,
                      n_out=5)
    want = [float.fromhex(x) for x in GOLD["lr_european"]]
    for i in range(5):
        assert close(got[i][0], want[2 * i]) and close(got[i][1], want[2 * i + 1])
    got = mcdx.greeks("mcdx_likelihood_ratio_digital", 100, 100, .05, 0, .25, 1, 0, 0, 1.0,
                      20_000, 7, n_out=5)
    g = (O.Greek * 5)()
    L.orc_lr_terminal(100, 100, .05, 0, .25, 1, 0, 1, 0, 1.0, 20_000, 7, g)
    for i in range(5):
        assert close(got[i][0], g[i].value) and close(got[i][1], g[i].standard_error)


def test_lr_barrier_vs_oracle():
    got = mcdx.greeks("mcdx_likelihood_ratio_barrier", 100, 100, 120, .05, 0, .2, 1, 0, 0, 1,
                      0.0, 50, 10_000, 11, n_out=4)
    g = (O.Greek * 4)()
    L.orc_lr_barrier(100, 100, 120, .05, 0, .2, 1, 0, 0, 1, 0.0, 50, 10_000, 11, g)
    for i in range(4):
        assert close(got[i][0], g[i].value) and close(got[i][1], g[i].standard_error)


def _flat(got):
    return [v for g in got for v in g[:2]]


def test_pathwise_asian_vs_reference_fixture():
    for case in GOLD["pathwise_asian"]:
        got = _flat(mcdx.greeks("mcdx_pathwise_asian", *case["args"], n_out=3))
        want = [float.fromhex(x) for x in case["out"]]
        assert all(close(a, b) for a, b in zip(got, want)), case["args"]


def test_lr_digital_vs_reference_fixture():
    for case in GOLD["lr_digital"]:
        got = _flat(mcdx.greeks("mcdx_likelihood_ratio_digital", *case["args"], n_out=5))
        want = [float.fromhex(x) for x in case["out"]]
        assert all(close(a, b) for a, b in zip(got, want)), case["args"]


def test_lr_barrier_vs_reference_fixture():
    for case in GOLD["lr_barrier"]:
        got = _flat(mcdx.greeks("mcdx_likelihood_ratio_barrier", *case["args"], n_out=4))
        want = [float.fromhex(x) for x in case["out"]]
        assert all(close(a, b) for a, b in zip(got, want)), case["args"]


def test_fd_prices_vs_oracle_pricings():
    bumps = (1.0, 0.002, 0.01, 0.01)
    got = mcdx.greeks("mcdx_fd_european_prices", 100, 100, .05, 0, .2, 1, 0, 30_011, 42, *bumps,
                      n_out=9)
    S = [100, 101, 99, 100, 100, 100, 100, 100, 100]
    V = [.2, .2, .2, .202, .198, .2, .2, .2, .2]
    R = [.05] * 5 + [.06, .04, .05, .05]
    T = [1] * 7 + [1.01, .99]
    for i in range(9):
        p = O.product(O.EUROPEAN, spot=S[i], vol=V[i], rate=R[i], time=T[i])
        op, ose, _ = O.price(p, 30_011, 42, chunks=1, threads=1)
        assert close(got[i][0], op, 1e-10) and close(got[i][1], ose, 1e-8)


def test_pathwise_delta_near_bsm_delta():
    got = mcdx.greeks("mcdx_pathwise_european", 100, 100, .05, 0, .2, 1, 0, 1 << 20, 3, n_out=3)
    d1 = (0.05 + 0.02) / 0.2
    delta = 0.5 * math.erfc(-d1 / math.sqrt(2))
    assert abs(got[0][0] - delta) < 3 * got[0][1]
