"""GPU parity of the GBM path engine against the CPU oracle, through the mcdx C ABI.

Bars (stated per test):
  * Philox bits: bit-exact, every (seed, path, draw, stream) tested.
  * FP64 mode: per-path samples within 1e-12 relative of the oracle (device libm vs glibc
    differ by at most an ulp per call); prices within 1e-12 relative of the oracle's
    chunked FP64 price (the merge trees differ).
  * FP32 mode: per-path samples agree with the oracle to FP32 generation accuracy; prices
    lie within 3 SE of the closed-form oracle where one exists and within 0.25 SE of the
    FP64 price on the same paths.
  * Two runs of either mode are bitwise identical.
"""
import math

import numpy as np
import pytest

import mcdx
import oracle_lib as O

pytestmark = pytest.mark.gpu

FP32, FP64 = mcdx.PREC_FP32, mcdx.PREC_FP64


def both(kind, **kw):
    """The same contract as an mcdx contract and an oracle product."""
    n = kw.pop("path_count", 4096)
    seed = kw.pop("seed", 42)
    return mcdx.contract(kind, path_count=n, seed=seed, **kw), O.product(kind, **kw), n, seed


# --------------------------------------------------------------------------- RNG
@pytest.mark.parametrize("seed,path0,draw,stream", [
    (42, 0, 0, 0), (42, 0, 251, 0), (42, 12345, 7, 1), (0xDEADBEEFCAFEF00D, 0, 3, 0),
    (7, (1 << 32) - 300, 0, 0), (2024, (1 << 40) + 5, (1 << 31) + 9, 1)])
def test_philox_bits_bit_exact(seed, path0, draw, stream):
    n = 4099
    got = mcdx.philox_bits(seed, path0, n, draw, stream)
    want = O.philox_bits(seed, path0, n, draw, stream)
    assert np.array_equal(got, want)


def test_philox_known_answers_through_stream_layout():
    # Random123 KATs (reference tests/rng_test.cpp:17-39) mapped onto the
    # {path_lo, path_hi, draw, stream} / {seed_lo, seed_hi} layout (rng.cpp:49-57).
    assert mcdx.philox_bits(0, 0, 1, 0, 0)[0] == 0x6627E8D5
    assert mcdx.philox_bits((1 << 64) - 1, (1 << 64) - 1, 1, 0xFFFFFFFF, 0xFFFFFFFF)[0] == 0x408F276D
    assert mcdx.philox_bits(0x299F31D0A4093822, 0x85A308D3243F6A88, 1, 0x13198A2E,
                            0x03707344)[0] == 0xD16CFE09


def test_normal_draws_fp64_match_oracle():
    n = 1 << 16
    got = mcdx.normal_draws(42, 0, n, 3, FP64)
    want = np.array([O.lib().orc_standard_normal_variate(42, i, 3) for i in range(n)])
    assert np.max(np.abs(got - want) / np.maximum(1.0, np.abs(want))) <= 1e-14


def test_fp32_inverse_normal_exhaustive_over_all_words():
    """Every one of the 2^32 Philox words through the FP32 inverse normal, against the FP64
    inverse (the reference's Acklam + Halley, src/core/normal.cpp:70-82) on the same word:
    the largest |z32 - z64| / (1 + |z64|), central and tail words separately. The FP64 side
    is first pinned to the oracle at the extreme and branch-edge words, and the scan's worst
    words are re-evaluated one by one."""
    L = O.lib()
    edge = np.array([0, 1, 2, 2073810, 2073811, 104152956, 104152957, (1 << 31) - 1, 1 << 31,
                     0xFFFFFFFF - 2073811, 0xFFFFFFFF - 2073810, 0xFFFFFFFE, 0xFFFFFFFF],
                    dtype=np.uint32)
    want = np.array([L.orc_inverse_normal_cdf(L.orc_uniform_open01(int(b))) for b in edge])
    z64 = mcdx.inv_normal_words(edge, FP64)
    assert np.all(np.abs(z64 - want) <= 1e-14 * (1 + np.abs(want))), (z64, want)
    z32 = mcdx.inv_normal_words(edge, FP32)
    assert np.all(np.abs(z32 - want) / (1 + np.abs(want)) <= 1.9e-7)
    # signs agree (at u = 1/2 +- 2^-33 the FP32 value is a zero of the right sign)
    assert np.array_equal(np.signbit(z32), np.signbit(want))
    err, bits = mcdx.inv_normal_f32_scan(0, 1 << 32)
    print(f"\nFP32 inverse normal, all 2^32 words: central max {err[0]:.3e} at {bits[0]:#010x}, "
          f"tail max {err[1]:.3e} at {bits[1]:#010x}")
    # measured over every word: central 1.884e-7 (at 0x002619bc), tail 1.863e-7 (0x0000347c)
    assert 0.0 < err[0] <= 1.9e-7 and 0.0 < err[1] <= 1.9e-7, (err, [hex(b) for b in bits])
    w = np.array(bits, dtype=np.uint32)
    a32, a64 = mcdx.inv_normal_words(w, FP32), mcdx.inv_normal_words(w, FP64)
    assert np.allclose(np.abs(a32 - a64) / (1 + np.abs(a64)), err, rtol=0, atol=0)
    oz = np.array([L.orc_inverse_normal_cdf(L.orc_uniform_open01(int(b))) for b in w])
    assert np.all(np.abs(a64 - oz) <= 1e-14 * (1 + np.abs(oz)))
    # halves of the range give the same maxima as the whole (the reduction's order is fixed)
    lo, lb = mcdx.inv_normal_f32_scan(0, 1 << 31)
    hi, hb = mcdx.inv_normal_f32_scan(1 << 31, 1 << 31)
    assert max(lo[0], hi[0]) == err[0] and max(lo[1], hi[1]) == err[1]


def test_normal_draws_fp32_accuracy_and_tails():
    n = 1 << 18
    got = mcdx.normal_draws(7, 0, n, 0, FP32)
    want = np.array([O.lib().orc_standard_normal_variate(7, i, 0) for i in range(n)])
    err = np.abs(got - want) / (1.0 + np.abs(want))
    assert np.all(np.isfinite(got))
    assert err.max() <= 2e-6, err.max()
    # same branch (central vs tail) as the FP64 comparison: signs agree everywhere
    assert np.array_equal(np.sign(got), np.sign(want))


# --------------------------------------------------------------------------- per-path parity
CASES = [
    dict(kind=O.EUROPEAN),
    dict(kind=O.EUROPEAN, type=1, antithetic=1),
    dict(kind=O.DIGITAL, digital_style=0, cash_amount=1.5),
    dict(kind=O.DIGITAL, digital_style=1, type=1),
    dict(kind=O.ASIAN, monitoring_points=12),
    dict(kind=O.ASIAN, monitoring_points=12, average_style=1, strike_style=1, type=1),
    dict(kind=O.ASIAN, monitoring_points=52, control_variate=1, antithetic=1),
    dict(kind=O.BARRIER, barrier=120.0, monitoring_points=20),
    dict(kind=O.BARRIER, barrier=90.0, direction=1, knock=0, rebate=2.0, monitoring_points=20,
         type=1),
    dict(kind=O.BARRIER, barrier=120.0, monitoring_points=50, brownian_bridge=1),
    dict(kind=O.BARRIER, barrier=85.0, direction=1, monitoring_points=50, brownian_bridge=1,
         antithetic=1),
    dict(kind=O.LOOKBACK, strike_style=1, monitoring_points=20),
    dict(kind=O.LOOKBACK, strike_style=0, strike=105.0, monitoring_points=20, type=1),
]


def _ids(c):
    return "-".join(f"{k}={v}" for k, v in c.items())


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_path_samples_fp64_match_oracle(case):
    case = dict(case)
    kind = case.pop("kind")
    c, p, n, seed = both(kind, path_count=2048, **case)
    got = mcdx.path_samples(c, FP64, 1000, 2048)
    want = O.path_samples(p, seed, 1000, 2048)
    assert np.max(np.abs(got - want) / np.maximum(1.0, np.abs(want))) <= 1e-12


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_path_samples_fp32_close_to_oracle(case):
    case = dict(case)
    kind = case.pop("kind")
    c, p, n, seed = both(kind, path_count=4096, **case)
    got = mcdx.path_samples(c, FP32, 0, 4096)
    want = O.path_samples(p, seed, 0, 4096)
    err = np.abs(got - want) / (1.0 + np.abs(want))
    # Discrete decisions (barrier touch, ITM) may flip on paths within FP32 rounding of the
    # boundary; everything else agrees to FP32 generation accuracy.
    assert np.mean(err > 1e-4) <= 0.002, np.sort(err)[-10:]
    assert np.median(err) <= 2e-6


# --------------------------------------------------------------------------- prices
@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_price_fp64_matches_oracle_price(case):
    case = dict(case)
    kind = case.pop("kind")
    c, p, n, seed = both(kind, path_count=50_003, seed=424242, **case)
    gp, gse, gn = mcdx.price(c, FP64)
    op, ose, on = O.price(p, n, seed)
    assert gn == on == n
    assert abs(gp - op) <= 1e-12 * max(1.0, abs(op))
    assert abs(gse - ose) <= 1e-9 * max(1e-12, ose)


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_price_fp32_bias_small_against_fp64(case):
    case = dict(case)
    kind = case.pop("kind")
    c, _, n, _ = both(kind, path_count=200_000, seed=11, **case)
    p32, se32, n32 = mcdx.price(c, FP32)
    p64, se64, n64 = mcdx.price(c, FP64)
    assert n32 == n64 == 200_000
    assert abs(p32 - p64) <= 0.25 * se64 + 1e-12, (p32, p64, se64)


def test_european_c1_within_3se_of_bsm():
    # C1: S=K=100 r=0.05 q=0 sigma=0.2 T=1, 1e6 paths, seed 42; BSM 10.450583572185565.
    bsm = O.lib().orc_bsm(100, 100, 0.05, 0.0, 0.2, 1.0, 0)
    assert abs(bsm - 10.450583572185565) < 1e-12
    for prec in (FP32, FP64):
        c = mcdx.contract(mcdx.EUROPEAN, path_count=1_000_000, seed=42)
        p, se, n = mcdx.price(c, prec)
        assert n == 1_000_000
        assert abs(p - bsm) < 3 * se


def test_digital_and_geometric_asian_within_3se_of_closed_forms():
    c = mcdx.contract(mcdx.DIGITAL, path_count=1 << 20, seed=5)
    p, se, _ = mcdx.price(c, FP32)
    ana = O.lib().orc_digital(100, 100, 0.05, 0, 0.2, 1, 0, 0, 1.0)
    assert abs(p - ana) < 3 * se
    # geometric Asian at fine monitoring vs continuous Kemna-Vorst (reference
    # tests/exotics_mc_test.cpp:24-40 uses 400 steps)
    c = mcdx.contract(mcdx.ASIAN, average_style=1, monitoring_points=400, path_count=1 << 18,
                      seed=9)
    p, se, _ = mcdx.price(c, FP32)
    kv = O.lib().orc_kemna_vorst(100, 100, 0.05, 0, 0.2, 1, 0)
    assert abs(p - kv) < 3 * se + 0.01


def test_in_out_parity_on_gpu():
    # knock-in + knock-out = vanilla path by path (reference exotics_mc_test.cpp:110-137)
    kw = dict(barrier=110.0, monitoring_points=30, path_count=1 << 16, seed=3)
    pin = mcdx.path_samples(mcdx.contract(mcdx.BARRIER, knock=0, **kw), FP32, 0, 1 << 16)
    pout = mcdx.path_samples(mcdx.contract(mcdx.BARRIER, knock=1, **kw), FP32, 0, 1 << 16)
    van = mcdx.path_samples(mcdx.contract(mcdx.BARRIER, knock=1, barrier=1e12,
                                          monitoring_points=30, path_count=1 << 16, seed=3),
                            FP32, 0, 1 << 16)
    assert np.allclose(pin + pout, van, rtol=0, atol=1e-12)


def test_runs_bitwise_reproducible():
    for prec in (FP32, FP64):
        c = mcdx.contract(mcdx.ASIAN, monitoring_points=24, path_count=300_001, seed=77)
        a = mcdx.price(c, prec)
        b = mcdx.price(c, prec)
        assert a == b


@pytest.mark.parametrize("n", [0, 1, 255, 1023, 1024, 1025, 1_000_003])
def test_ragged_path_counts(n):
    c, p, _, seed = both(O.EUROPEAN, path_count=n, seed=1)
    gp, gse, gn = mcdx.price(c, FP64)
    assert gn == n
    if n == 0:
        assert (gp, gse) == (0.0, 0.0)
        return
    if n <= 2000:
        op, ose, on = O.price(p, n, seed)
        assert abs(gp - op) <= 1e-12 * max(1.0, abs(op))


def test_invalid_arguments_rejected():
    L = mcdx.lib()
    r = mcdx.Result()
    c = mcdx.contract(mcdx.ASIAN, monitoring_points=0)
    assert L.mcdx_price_gbm(c, FP32, r) == mcdx.EINVAL
    c = mcdx.contract(7)
    assert L.mcdx_price_gbm(c, FP32, r) == mcdx.EINVAL
    c = mcdx.contract(mcdx.EUROPEAN)
    assert L.mcdx_price_gbm(c, 5, r) == mcdx.EINVAL


def test_reference_signature_entry_points():
    L = mcdx.lib()
    r = mcdx.Result()
    opts = mcdx.Options(0, 0, 0, 0, 8)
    assert L.mcdx_set_precision(FP64) == 0
    try:
        assert L.mcdx_monte_carlo_lookback(100, 100, 0.05, 0, 0.2, 1, 0, 1, 20, 50_003, 424242,
                                           opts, r) == 0
        op, _, _ = O.price(O.product(O.LOOKBACK, strike_style=1, monitoring_points=20), 50_003,
                           424242)
        assert abs(r.price - op) <= 1e-12 * op
    finally:
        L.mcdx_set_precision(FP32)


# --------------------------------------------------------------------------- full sizes
def _reference_fixture():
    import json
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "reference_baseline.json")
    with open(path) as f:
        return json.load(f)


FULL = dict(spot=100.0, strike=100.0, rate=0.05, carry_yield=0.0, vol=0.2, time=1.0, seed=42)


@pytest.mark.parametrize("key,kind,kw,fp64", [
    ("c1_european", mcdx.EUROPEAN, dict(path_count=1_000_000), True),
    ("c2_asian_arith_call", mcdx.ASIAN, dict(monitoring_points=252, path_count=1 << 24), True),
    ("c2_barrier_uo_call", mcdx.BARRIER,
     dict(barrier=120.0, direction=0, knock=1, monitoring_points=252, path_count=1 << 24), True),
    ("c2_barrier_uo_call_bb", mcdx.BARRIER,
     dict(barrier=120.0, direction=0, knock=1, monitoring_points=252, path_count=1 << 24,
          brownian_bridge=1), False),
    ("c2_lookback_float_call", mcdx.LOOKBACK,
     dict(strike_style=1, monitoring_points=252, path_count=1 << 24), True),
])
def test_full_size_against_reference_fp64_prices(key, kind, kw, fp64):
    # The workload sizes themselves (2^24 paths x 252 dates; C1 at 1e6 paths) against the
    # reference's own FP64 prices (tests/golden/reference_baseline.json, the reference built
    # from its sources at 8 chunks): FP64 mode within 1e-12 relative (only the merge order
    # differs), FP32 generation within 0.01 SE, same path counts.
    ref = _reference_fixture()[key]
    rp, rse = float.fromhex(ref["price"]), float.fromhex(ref["se"])
    c = mcdx.contract(kind, **FULL, **kw)
    p32, se32, n32 = mcdx.price(c, FP32)
    assert n32 == ref["paths"]
    assert abs(p32 - rp) <= 0.01 * rse, (key, p32, rp, rse)
    assert abs(se32 - rse) <= 1e-3 * rse, (key, se32, rse)
    if fp64:
        p64, se64, n64 = mcdx.price(c, FP64)
        assert n64 == ref["paths"]
        assert abs(p64 - rp) <= 1e-12 * abs(rp), (key, p64, rp)
        assert abs(se64 - rse) <= 1e-9 * rse, (key, se64, rse)


@pytest.mark.parametrize("avs,ss,cv,anti", [(0, 0, 0, 0), (1, 0, 0, 0), (0, 1, 0, 0), (1, 1, 0, 0),
                                            (0, 0, 1, 0), (0, 0, 0, 1), (1, 0, 0, 1), (0, 0, 1, 1)])
def test_fp32_asian_tile_kernels_equal_their_path_samples(avs, ss, cv, anti):
    # The FP32 Asian tile kernels are instantiated per combination of running sums (and
    # antithetic or not); the per-path sample kernel decides the sums per path. The same
    # per-path arithmetic must come out of both: the tile price equals the mean of the
    # sample kernel's per-path values to summation-order rounding.
    n = 50_000
    c = mcdx.contract(mcdx.ASIAN, monitoring_points=37, path_count=n, seed=17, average_style=avs,
                      strike_style=ss, control_variate=cv, antithetic=anti, strike=98.0)
    p, se, cnt = mcdx.price(c, FP32)
    x = mcdx.path_samples(c, FP32, 0, n)
    assert cnt == n
    assert abs(p - x.mean()) <= 1e-12 * max(1.0, abs(p)), (p, x.mean())
    assert abs(se - x.std(ddof=1) / math.sqrt(n)) <= 1e-9 * se, (se, x.std(ddof=1))
