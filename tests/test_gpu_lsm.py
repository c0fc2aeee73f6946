"""GPU Longstaff-Schwartz against the oracle (mcdx C ABI).

The device regression solves each date's least squares in FP64 by normal equations of rows
preconditioned with a Cholesky-QR factor of the date's in-the-money rows (formed from the store
before the backward induction, two Cholesky steps), so the preconditioned problem is well
conditioned: the same least-squares solution as the reference's sequential Householder sweep
(src/core/linalg.cpp:7-71) with different rounding, so coefficients are compared at 1e-7
relative and prices at 0.02 SE (a decision within rounding of the exercise boundary may flip). FP32-generation prices are compared to
FP64 at 0.5 SE; acceptance against CRR follows the reference's tests/lsm_test.cpp:26-123 at its
bare 3 SE bar.
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


def test_fp64_matches_reference_fixture():
    g = GOLD["lsm"]
    s, k, r, q, v, t, typ, m, n, seed = g["args"]
    p, se, co, has, inc = mcdx.lsm(s, k, r, q, v, t, typ, m, n, seed, FP64)
    want_co = np.array([float.fromhex(x) for x in g["coeffs"]]).reshape(m + 1, 4)
    assert has.tolist() == g["has"] and inc == g["inception"]
    mask = has.astype(bool)
    rel = np.abs(co[mask] - want_co[mask]) / np.maximum(1.0, np.abs(want_co[mask]))
    assert rel.max() < 1e-7, rel.max()
    assert abs(p - float.fromhex(g["price"])) <= 0.02 * float.fromhex(g["se"])


# The last two are deep out of the money with more paths than the 2^16-path preconditioning
# sample: at the middle dates the sample holds a few hundred in-the-money rows, so those dates
# are preconditioned from every path, and the first dates have fewer than four rows. Their
# earliest regressed date is an interpolation: for the put, 4 rows with condition number 9.8e9
# and coefficients near 1e10, where any FP64 least-squares solution is only good to about
# condition x 1e-16 ~ 1e-6 (solved in 60-digit arithmetic: the oracle's sequential Householder
# is 3.1e-7 from the exact coefficients, the device 4.5e-7). Coefficients beyond 1e6 in
# magnitude, which only such near-interpolations produce, are compared at 1e-6.
@pytest.mark.parametrize("typ,s0,vol,m,n", [(1, 100.0, 0.2, 50, 40_000), (1, 90.0, 0.3, 12, 30_011),
                                           (0, 110.0, 0.25, 20, 20_000),
                                           (1, 140.0, 0.2, 20, 100_000), (0, 70.0, 0.25, 20, 100_000)])
def test_fp64_matches_oracle(typ, s0, vol, m, n):
    p, se, co, has, _ = mcdx.lsm(s0, 100.0, 0.06, 0.01, vol, 1.0, typ, m, n, 42, FP64)
    op, ose, oco, ohas, _ = O.lsm(s0, 100.0, 0.06, 0.01, vol, 1.0, typ, m, n, 42)
    assert has.tolist() == ohas.tolist()
    mask = has.astype(bool)
    big = np.abs(oco[mask]).max(axis=1) > 1e6
    assert np.allclose(co[mask][~big], oco[mask][~big], rtol=1e-7, atol=1e-9)
    assert np.allclose(co[mask][big], oco[mask][big], rtol=1e-6, atol=1e-9)
    assert abs(p - op) <= 0.02 * ose


def crr_american(s, k, r, q, v, t, typ, steps=4000):
    return L.orc_crr_american(s, k, r, q, v, t, steps, typ)


def crr_bermudan(s, k, r, q, v, t, typ, dates, steps=10_000):
    """The CRR tree with exercise only at the LSM's `dates` exercise dates (steps must be a
    multiple of dates): the exact value a `dates`-date LSM estimates."""
    return L.orc_crr_bermudan(s, k, r, q, v, t, steps, typ, steps // dates)


def test_c5_full_size_fp64_matches_oracle():
    # The C5 workload itself (2^20 paths x 50 dates, seed 42): FP64 coefficients within 1e-7
    # relative and the price within 0.02 SE of the oracle's sequential Householder LSM; FP32
    # generation within 0.5 SE of FP64; both within 3 SE of CRR-10^4 (the reference's bar,
    # tests/lsm_test.cpp:86) and of the 50-date Bermudan on the same tree.
    args = (100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 1, 50, 1 << 20, 42)
    p64, se64, co, has, _ = mcdx.lsm(*args, FP64)
    op, ose, oco, ohas, _ = O.lsm(*args)
    assert has.tolist() == ohas.tolist()
    mask = has.astype(bool)
    assert np.allclose(co[mask], oco[mask], rtol=1e-7, atol=1e-9)
    assert abs(p64 - op) <= 0.02 * ose, (p64, op, ose)
    p32, se32, _, _, _ = mcdx.lsm(*args, FP32)
    assert abs(p32 - p64) <= 0.5 * se64
    crr = crr_american(100, 100, 0.06, 0, 0.2, 1, 1, 10_000)
    berm = crr_bermudan(100, 100, 0.06, 0, 0.2, 1, 1, 50)
    for p, se in ((p32, se32), (p64, se64)):
        assert abs(p - crr) < 3 * se and abs(p - berm) < 3 * se, (p, se, crr, berm)


# The reference's LSM acceptance matrix (tests/lsm_test.cpp:76-98): N = 1e5 paths, 50 dates,
# seed 909, |LSM - CRR-4000| < 3 SE, in both generation modes.
LSM_MATRIX = [(100, 100, 0.05, 0.00, 0.25, 1.0, 1), (100, 80, 0.05, 0.00, 0.25, 1.0, 1),
              (100, 120, 0.05, 0.00, 0.25, 1.0, 1), (100, 100, 0.05, 0.03, 0.30, 0.5, 1),
              (100, 100, 0.05, 0.03, 0.30, 0.5, 0), (100, 100, 0.03, 0.06, 0.20, 2.0, 1)]


@pytest.mark.parametrize("prec", [FP64, FP32])
@pytest.mark.parametrize("case", LSM_MATRIX)
def test_lsm_vs_binomial_matrix(case, prec):
    p, se, _, _, _ = mcdx.lsm(*case, 50, 100_000, 909, prec)
    crr = crr_american(*case)
    assert abs(p - crr) < 3 * se, (p, se, crr, (p - crr) / se)
    if prec == FP64:
        op, ose, _, _, _ = O.lsm(*case, 50, 100_000, 909)
        assert abs(p - op) <= 0.02 * ose, (p, op)


@pytest.mark.parametrize("prec", [FP64, FP32])
def test_lsm_converges_as_path_count_increases(prec):
    # tests/lsm_test.cpp:103-123: the SE shrinks and every price stays within 3 SE of CRR.
    args = (100.0, 100.0, 0.05, 0.0, 0.25, 1.0, 1)
    crr = crr_american(*args)
    last_se = 1e9
    for n in (10_000, 40_000, 160_000):
        p, se, _, _, _ = mcdx.lsm(*args, 50, n, 909, prec)
        assert se < last_se and abs(p - crr) < 3 * se, (n, p, se, crr)
        last_se = se


@pytest.mark.parametrize("prec", [FP64, FP32])
def test_american_call_equals_european_and_put_bounds(prec):
    # tests/lsm_test.cpp:26-65 at its N = 1e5, seed 909: the call equals BSM within 3 SE; the
    # put is at least the European put and the immediate exercise value, less 3 SE.
    c, cse, _, _, _ = mcdx.lsm(100, 100, 0.05, 0, 0.25, 1, 0, 50, 100_000, 909, prec)
    assert abs(c - L.orc_bsm(100, 100, 0.05, 0, 0.25, 1, 0)) < 3 * cse
    p, pse, _, _, _ = mcdx.lsm(100, 100, 0.05, 0, 0.25, 1, 1, 50, 100_000, 909, prec)
    assert p >= L.orc_bsm(100, 100, 0.05, 0, 0.25, 1, 1) - 3 * pse
    for s in (60.0, 80.0, 100.0, 120.0, 140.0):
        p, pse, _, _, _ = mcdx.lsm(s, 100, 0.05, 0, 0.25, 1, 1, 50, 100_000, 909, prec)
        assert p >= max(100 - s, 0.0) - 3 * pse, (s, p, pse)


def test_frozen_replay_at_bumped_parameters_matches_reference():
    # src/pricers/lsm.cpp:151-213: the reference's fitted policy (fixture) replayed on the
    # device at the base and bumped spot, vol and rate, FP64, against the reference's replay
    # prices (fixture) and the oracle's; a decision on the exercise boundary may flip with a
    # last-bit difference of the path, so the bar is 0.02 SE.
    g = GOLD["lsm"]
    s0, k0, r0, q0, v0, t0, ty, m0, n0, sd = g["args"]
    co = np.array([float.fromhex(x) for x in g["coeffs"]])
    has = np.array(g["has"], np.uint8)
    se = float.fromhex(g["se"])
    for rep in GOLD["lsm_frozen"]["replays"]:
        a = (rep["spot"], k0, rep["rate"], q0, rep["vol"], t0, ty, m0, n0, sd)
        got = mcdx.lsm_reprice(*a, co, has, g["inception"], FP64)
        want = float.fromhex(rep["price"])
        assert abs(got - want) <= 0.02 * se, (rep, got, want)
        orc = L.orc_lsm_reprice_frozen(*a, co.ctypes.data_as(O.P(O.dbl)),
                                       has.ctypes.data_as(O.P(O.u8)), g["inception"])
        assert orc == want
        # FP32 generation replays the same policy within 0.5 SE
        assert abs(mcdx.lsm_reprice(*a, co, has, g["inception"], FP32) - want) <= 0.5 * se
    ic = GOLD["lsm_frozen"]["inception_case"]
    co2 = np.array([float.fromhex(x) for x in ic["coeffs"]])
    got = mcdx.lsm_reprice(ic["replay_spot"], *ic["args"][1:], co2, np.array(ic["has"], np.uint8),
                           1, FP64)
    assert got == float.fromhex(ic["replay_price"])
    # the device's own fitted policy, replayed at bumped spot and vol, equals the oracle's
    # replay of that same policy
    args = (100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 1, 25, 30_011, 5)
    p, se, co, has, inc = mcdx.lsm(*args, FP64)
    cc = np.ascontiguousarray(co.ravel())
    for s, v in ((101.5, 0.2), (98.0, 0.2), (100.0, 0.23)):
        a = (s,) + args[1:4] + (v,) + args[5:]
        got = mcdx.lsm_reprice(*a, co, has, inc, FP64)
        orc = L.orc_lsm_reprice_frozen(*a, cc.ctypes.data_as(O.P(O.dbl)),
                                       has.ctypes.data_as(O.P(O.u8)), inc)
        assert abs(got - orc) <= 0.02 * se, (s, v, got, orc)


def test_fp32_close_to_fp64_and_frozen_replay_exact():
    args = (100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 1, 50, 1 << 17, 42)
    p32, se32, co32, has32, inc32 = mcdx.lsm(*args, FP32)
    p64, se64, _, _, _ = mcdx.lsm(*args, FP64)
    assert abs(p32 - p64) <= 0.5 * se64
    # replaying the fitted policy on the same paths reproduces the fitted price bit for bit
    assert mcdx.lsm_reprice(*args, co32, has32, inc32, FP32) == p32


def test_exercise_at_inception_and_reference_signature():
    p, se, _, _, inc = mcdx.lsm(60, 100, 0.06, 0, 0.2, 1, 1, 10, 20_000, 3, FP32)
    assert inc == 1 and p == 40.0 and se == 0.0
    r = mcdx.Result()
    assert mcdx.lib().mcdx_monte_carlo_lsm_american(100, 100, .05, 0, .25, 1, 1, 20, 50_000, 7, r) == 0
    assert r.price > 0 and r.standard_error > 0


def test_large_path_count():
    """2^22 paths: 32 paths per lane in the backward-induction layout. At this size the SE
    (about 0.0034) is below the gap between 50 exercise dates and continuous exercise
    (about 0.014 on CRR-10^4), so the oracle is the 50-date Bermudan tree."""
    berm = crr_bermudan(100, 100, 0.06, 0, 0.2, 1, 1, 50)
    p, se, _, has, _ = mcdx.lsm(100, 100, 0.06, 0, 0.2, 1, 1, 50, 1 << 22, 7, FP32)
    assert has[1:50].all()
    assert abs(p - berm) < 3 * se, (p, se, berm)


@pytest.mark.parametrize("prec", [FP32, FP64])
def test_american_greeks_are_frozen_policy_differences(prec):
    """One pass over five scenarios equals separate frozen-policy replays
    (reprice_against_frozen_policy) of the fitted policy at the bumped parameters."""
    args = (100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 1, 25)
    n, seed = 30_011, 5
    price, delta, gamma, vega = mcdx.lsm_greeks(*args, n, seed, prec)
    p, se, co, has, inc = mcdx.lsm(*args, n, seed, prec)
    assert price[0] == p and price[2] == n
    hs, hv = 1.0, 0.002

    def frozen(s0=100.0, vol=0.2):
        a = (s0, 100.0, 0.06, 0.0, vol, 1.0, 1, 25)
        return mcdx.lsm_reprice(*a, n, seed, co, has, inc, prec)
    up, dn = frozen(s0=100 + hs), frozen(s0=100 - hs)
    vu, vd = frozen(vol=0.2 + hv), frozen(vol=0.2 - hv)
    assert abs(delta[0] - (up - dn) / (2 * hs)) <= 1e-9 * (1 + abs(up) + abs(dn)) / hs
    assert abs(gamma[0] - (up - 2 * p + dn) / hs ** 2) <= 4e-9 * (1 + abs(up) + abs(dn)) / hs ** 2
    assert abs(vega[0] - (vu - vd) / (2 * hv)) <= 1e-9 * (1 + abs(vu) + abs(vd)) / hv
    assert delta[1] > 0 and vega[1] > 0


def test_american_greeks_near_binomial():
    price, delta, gamma, vega = mcdx.lsm_greeks(100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 1, 50,
                                                1 << 20, 42, FP32)
    crr = lambda s=100.0, v=0.2: L.orc_crr_american(s, 100, 0.06, 0, v, 1, 4000, 1)  # noqa: E731
    d = (crr(s=101.0) - crr(s=99.0)) / 2.0
    g = (crr(s=101.0) - 2 * crr() + crr(s=99.0)) / 1.0
    v = (crr(v=0.202) - crr(v=0.198)) / 0.004
    assert abs(delta[0] - d) < 4 * delta[1] + 0.01, (delta, d)
    assert abs(gamma[0] - g) < 4 * gamma[1] + 0.005, (gamma, g)
    assert abs(vega[0] - v) < 4 * vega[1] + 0.5, (vega, v)


def test_frozen_policy_delta_is_less_noisy_than_refitting():
    """The reference's tests/lsm_test.cpp:138-170 claim, on the device: across 40 seeds the
    frozen-boundary delta varies less than refitting the regression at each bump."""
    args = (100.0, 100.0, 0.05, 0.0, 0.25, 1.0, 1, 20)
    h, n = 0.1, 5000
    frozen, naive = [], []
    for trial in range(40):
        seed = 909 + trial * 1013
        frozen.append(mcdx.lsm_greeks(*args, n, seed, FP64, hs=h)[1][0])
        up = mcdx.lsm(100.0 + h, *args[1:], n, seed, FP64)[0]
        dn = mcdx.lsm(100.0 - h, *args[1:], n, seed, FP64)[0]
        naive.append((up - dn) / (2 * h))
    assert np.std(frozen, ddof=1) < np.std(naive, ddof=1)


# ---- QR-LSM on Sobol + Brownian-bridge paths (plain and randomized) ----
def oracle_lsm_qmc(args, shifted=0, shift_seed=0, replicate=0):
    import ctypes as C
    s, k, r, q, v, t, typ, m, n = args
    co = np.zeros((m + 1) * 4)
    has = np.zeros(m + 1, np.uint8)
    inc = C.c_uint8()
    res = L.orc_lsm_american_qmc(s, k, r, q, v, t, typ, m, n, shifted, shift_seed, replicate,
                                 co.ctypes.data_as(C.POINTER(C.c_double)),
                                 has.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(inc))
    return res.price, res.standard_error


@pytest.mark.parametrize("typ,s0,vol,m,n", [(1, 100.0, 0.2, 50, 20_000), (0, 110.0, 0.25, 12, 16_384)])
def test_qmc_lsm_fp64_matches_oracle(typ, s0, vol, m, n):
    args = (s0, 100.0, 0.06, 0.01, vol, 1.0, typ, m, n)
    p, se, total, _ = mcdx.lsm_qmc(*args, 0, 0, FP64)
    op, ose = oracle_lsm_qmc(args)
    assert total == n
    assert abs(p - op) <= 0.02 * ose, (p, op, ose)
    # randomized runs: every run equals the oracle's run with the same shift
    p, se, total, reps = mcdx.lsm_qmc(*args, 3, 17, FP64)
    assert total == 3 * n
    for r in range(3):
        op, ose = oracle_lsm_qmc(args, 1, 17, r)
        assert abs(reps[r] - op) <= 0.02 * ose, (r, reps[r], op)
    assert abs(p - reps.mean()) <= 1e-14 * p and abs(se - reps.std(ddof=1) / np.sqrt(3)) <= 1e-12


def test_qmc_lsm_replicates_independent_of_batch():
    # The shifted runs share every launch (grid y = run); each run's price must not depend on
    # which runs share its batch: 3 runs equal the first 3 of 5, bit for bit, in both modes.
    args = (100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 1, 20, 12_345)
    for prec in (FP32, FP64):
        _, _, _, r3 = mcdx.lsm_qmc(*args, 3, 99, prec)
        _, _, _, r5 = mcdx.lsm_qmc(*args, 5, 99, prec)
        assert np.array_equal(r3, r5[:3]), (prec, r3, r5)
        assert len(set(r5.tolist())) == 5  # distinct shifts give distinct prices


def test_qmc_lsm_against_crr_and_monte_carlo():
    # 16 shifted runs of 2^16 Sobol paths: the estimate sits where the Philox LSM does (both
    # low-biased by the same regression), with a much smaller error than Monte Carlo's.
    args = (100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 1, 50, 1 << 16)
    q, qse, _, reps = mcdx.lsm_qmc(*args, 16, 2024, FP32)
    p, se, _, _, _ = mcdx.lsm(*args, 42, FP32)
    crr = L.orc_crr_american(100, 100, 0.06, 0, 0.2, 1, 2000, 1)
    assert abs(q - p) <= 4 * np.hypot(qse, se), (q, qse, p, se)
    assert abs(q - crr) <= 4 * qse + 0.05, (q, qse, crr)
    assert qse < 0.5 * se / 4.0, (qse, se)   # Monte Carlo at 16x the paths would give se / 4
    assert len(reps) == 16 and np.all(reps > 4.0)


def test_qmc_lsm_rejects():
    r = mcdx.Result()
    f = mcdx.lib().mcdx_lsm_american_qmc
    args = (100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 1, 50, 1000)
    assert f(*args, 1, 5, FP32, mcdx.C.byref(r), None) == mcdx.EINVAL       # one replicate
    assert f(*args, 4097, 5, FP32, mcdx.C.byref(r), None) == mcdx.EINVAL
    assert f(100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 1, 2000, 1000, 0, 5, FP32, mcdx.C.byref(r),
             None) == mcdx.EINVAL                                           # beyond 1024 dates
