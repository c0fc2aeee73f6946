"""The one-launch LSM backward induction against the one-launch-per-date schedule.

Runs of at most 2^20 paths per run (whose blocks all fit on the device at once) run every date
in one resident launch; mcdx.lsm_schedule(LSM_PER_DATE) forces one launch per date. Both form
every sum by the same operations in the same order, so prices, standard errors, coefficients
and flags must be bit for bit equal, for fits, frozen-policy replays, shifted-Sobol batches,
the American Greeks and the learned pipeline's captured samples. Runs past 2^20 paths fall
back to the per-date launches by themselves; the oracle parity of both schedules is in
test_gpu_lsm.py.
"""
import numpy as np
import pytest

import mcdx

pytestmark = pytest.mark.gpu
FP32, FP64 = mcdx.PREC_FP32, mcdx.PREC_FP64


def both(fn):
    prev = mcdx.lsm_schedule(mcdx.LSM_AUTO)
    try:
        auto = fn()
        mcdx.lsm_schedule(mcdx.LSM_PER_DATE)
        per_date = fn()
    finally:
        mcdx.lsm_schedule(prev)
    return auto, per_date


def same(a, b):
    if isinstance(a, (tuple, list)):
        assert len(a) == len(b)
        for x, y in zip(a, b):
            same(x, y)
    elif isinstance(a, np.ndarray):
        assert a.dtype == b.dtype and a.shape == b.shape
        assert a.tobytes() == b.tobytes()
    elif isinstance(a, float):
        assert np.float64(a).tobytes() == np.float64(b).tobytes(), (a, b)
    elif isinstance(a, dict):
        assert a.keys() == b.keys()
        for k in a:
            same(a[k], b[k])
    else:
        assert a == b


@pytest.mark.parametrize("prec", [FP32, FP64])
@pytest.mark.parametrize("typ,s0,m,n", [(1, 100.0, 50, 1 << 20), (1, 100.0, 50, 100_003),
                                        (0, 110.0, 20, 20_000), (1, 90.0, 12, 4_097),
                                        (1, 100.0, 2, 50_000), (1, 100.0, 1, 30_000),
                                        (1, 140.0, 20, 100_000)])
def test_fit_bit_equal(prec, typ, s0, m, n):
    args = (s0, 100.0, 0.06, 0.01, 0.2, 1.0, typ, m, n, 42, prec)
    auto, per_date = both(lambda: mcdx.lsm(*args))
    same(auto, per_date)
    assert np.isfinite(auto[0])


def test_large_run_falls_back_and_agrees():
    # 2^20 + 1 paths need 9 paths per lane: the per-date launches run in both schedules.
    args = (100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 1, 10, (1 << 20) + 1, 5, FP32)
    auto, per_date = both(lambda: mcdx.lsm(*args))
    same(auto, per_date)


@pytest.mark.parametrize("prec", [FP32, FP64])
def test_frozen_replay_bit_equal(prec):
    args = (100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 1, 50)
    p, se, co, has, inc = mcdx.lsm(*args, 200_000, 11, prec)
    auto, per_date = both(lambda: [mcdx.lsm_reprice(s, *args[1:], 200_000, 17, co, has, inc, prec)
                                   for s in (95.0, 100.0, 105.0)])
    same(auto, per_date)


def test_shifted_sobol_batch_bit_equal():
    # 16 runs of 2^16 paths share each launch (grid y = run): 16 blocks per run, 256 in all.
    auto, per_date = both(lambda: mcdx.lsm_qmc(100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 1, 50, 1 << 16,
                                               16, 99, FP32))
    same(auto, per_date)


def test_american_greeks_bit_equal():
    auto, per_date = both(lambda: mcdx.lsm_greeks(100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 1, 50,
                                                  1 << 18, 42, FP32))
    same(auto, per_date)


def test_learned_capture_bit_equal():
    # The fit captures every (path, date) sample for the network; the trained weights, the
    # table and both policies' prices follow from those samples.
    def run():
        res, w, co, has = mcdx.lsm_learned(100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 1, 20, 1 << 16, 42,
                                           FP32, steps=200)
        return res, w, co, has
    auto, per_date = both(run)
    same(auto, per_date)


def test_back_to_back_runs_reuse_the_counter():
    # Consecutive fits on one thread reuse the scratch and the run's date counter; each must
    # equal a fresh per-date fit (the counter is cleared before every backward induction).
    args = [(100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 1, 50, 1 << 20, sd, FP32) for sd in (1, 2)]
    args.append((100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 1, 7, 3000, 3, FP32))
    auto, per_date = both(lambda: [mcdx.lsm(*a) for a in args])
    same(auto, per_date)


def test_schedule_switch_rejects_unknown():
    assert mcdx.lib().mcdx_lsm_schedule(7) == -mcdx.EINVAL
    assert mcdx.lsm_schedule(mcdx.LSM_AUTO) == mcdx.LSM_AUTO


def test_concurrent_threads_match_sequential_calls():
    # Threads own separate streams; their resident launches are chained on the device, so
    # concurrent fits (each needing every CU) finish and equal the sequential results.
    import threading
    jobs = [(100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 1, 50, 1 << 20, sd, FP32) for sd in range(4)]
    want = [mcdx.lsm(*j) for j in jobs]
    got = [None] * len(jobs)

    def run(i):
        got[i] = [mcdx.lsm(*jobs[i]) for _ in range(3)]

    ts = [threading.Thread(target=run, args=(i,)) for i in range(len(jobs))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    for i in range(len(jobs)):
        for g in got[i]:
            same(g, want[i])
