#!/usr/bin/env python3
"""Headline benchmark: path-dependent GBM exotics on one MI355X.

Workload (one "step"): price three contracts, each 2^24 Philox paths x 252 monitoring
dates, S=100 r=0.05 q=0 sigma=0.2 T=1, seed 42, FP32 path generation + FP64 accumulation:
  * arithmetic-average Asian call, K=100 (fixed strike)
  * up-and-out barrier call, H=120, K=100, rebate 0 (discrete monitoring)
  * floating-strike lookback call
Each contract is one pricing call of the engine (path kernel + fixed-order merge),
enqueued through the mcdx C ABI on torch's current stream; inputs are scalars, so nothing
crosses PCIe inside the timed region. `value` = contract-paths priced per second.

Also reported (untimed): FP64-mode prices of the same paths (bias and SE), the roofline of
the dominant path kernel from HIP events around each launch, and the reference's own
multithreaded CPU implementation (oracle/_ref) timed on this host on a bounded sample.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "derivative-pricing_amd", "python"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

N_PATHS = 1 << 24
STEPS = 252
SEED = 42

# Algorithmic operation count per path-step in FP32 mode (FMA = 2 flops, every other
# lane operation = 1, each transcendental instruction = 1): Philox4x32-10 word 0 =
# 10 x (2 mul_hi + 2 mul_lo + 4 xor) = 80; inverse normal (central branch, 99.9% of draws)
# = fold u to min(u, 1-u) 2 + convert 1 + 4m 2 + 4m(1-m) 3 + log2 1 + degree-9 Horner 18 +
# |q| 2 + product 1 + sign 2 + tail test 1 = 33; log-space GBM step = 1 FMA + 1 add = 3;
# payoff observation: arithmetic average exp2 + cvt + f64 add = 3, discrete barrier and
# lookback running max + min = 2 (the barrier is decided from the extrema once per path).
OPS_DRAW = 80 + 33
OPS_GBM = 3
OPS_OBSERVE = {"asian": 3, "barrier_uo": 2, "lookback_float": 2}

# HBM bytes per launch of each 2^24 x 252 path kernel: rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE per dispatch, FETCH_SIZE scaled x2 and WRITE_SIZE x1 by the same counters'
# reading of a plain 1 GiB device copy (FETCH 454.8, 224.1, 223.0 KiB; WRITE 768 KiB).
TRAFFIC_BYTES = {"asian": 1717862, "barrier_uo": 1245440, "lookback_float": 1243136}
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector peak (packed FMA)
PEAK_FP64_TFLOPS = 78.6    # MI355X FP64 vector peak

# C1 (European call, FP64 throughout) per path, the same counting rule: one normal draw =
# Philox word 0 (80 integer ops) + uniform map 2 + Acklam central rational 24 + Halley step
# 30 (erfc 20, exp 4, 6 others) = 136, plus the terminal spot, payoff and Welford update 8.
OPS_C1_FP64_PER_PATH = 144

# Secondary path kernels, the same rule (FP32 generation; a draw = OPS_DRAW = 113):
# * bridge-corrected barrier, per path-step: draw 113 + log2-space step 3 + crossing exponent
#   and tests 6 = 122 (the stream-1 uniform is drawn only where the crossing probability can
#   exceed 2^-34; those draws are not counted);
# * Sobol QMC with the Brownian bridge, per point-date: Sobol word (LDS part read, 10 masked
#   XORs) 12 + inverse normal 33 + bridge point 2 FMA = 4 + exp2 with its FMA 3 + observe 1
#   = 53;
# * Heston QE, per path-step: two Philox words 160 + inverse normal of z_s 33 + moment
#   matching m, s^2, m^2, psi with its Newton reciprocal and the switch 12 + quadratic branch
#   (1/psi, b^2, a, b with two square roots, Phi^-1(U) 33, v') 53 + variance and log-spot
#   update 12 = 270;
# * batched book, per shared-draw group path-step: draw 113 + 4 per member evaluated on it
#   (log2 step 3, observation 1); a plain (ungrouped) item 117 per path-step.
OPS_BB = 122
OPS_QMC = 53
OPS_HESTON = 270
OPS_MEMBER = 4

# Lane-issue roofline of the three path kernels. A SIMD issues one wave64 VALU instruction
# per 4 cycles (the issue probe measures the same rate for v_fma_f32, v_mad_u64_u32,
# v_mul_hi_u32, v_bitop3_b32 and v_fma_f64, and half of it for v_exp_f32), so the issue slots
# a kernel needs are rocprofv3's SQ_ACTIVE_INST_VALU (quad-cycles of VALU work, a
# transcendental counting 2) less SQ_ACTIVE_INST_VALU2 (quad-cycles in which gfx950 issued a
# second VALU instruction beside the first). Per wave-path-step of each 2^24 x 252 FP32
# launch (counters / (2^24 x 252 / 64)): (VALU quad-cycles, paired quad-cycles).
ISSUE_QUADS = {"asian": (58.961, 14.827), "barrier_uo": (56.643, 14.812),
               "lookback_float": (56.626, 14.854)}
SIMDS = 1024               # MI355X: 256 CUs x 4 SIMDs
PEAK_CLOCK_HZ = 2.4e9      # MI355X peak engine clock

# The same lane-issue roofline for the secondary path kernels, from one launch of each at its
# configuration (rocprofv3 --pmc SQ_ACTIVE_INST_VALU SQ_ACTIVE_INST_VALU2 ... -- python3
# bench.py --only kernels): (VALU quad-cycles, paired quad-cycles) per wave-unit, i.e. the
# counters over (units / 64): C1 in FP64 per wave-path (2^24 paths: the reference's FP64
# inverse normal with its erfc, exp and IEEE divisions, 531 VALU instructions per path), the
# bridge-corrected barrier per wave-path-step, Sobol + bridge QMC (Asian) per wave-point-date
# and Heston QE (9 strikes) per wave-path-step.
# C5's learned-policy serving kernel per nominal wave-path-date (2^20 x 50 / 64; a wave runs
# until its last path stops), from rocprofv3 --pmc ... -- python3 bench.py --only c5.
ISSUE_QUADS_CONFIG = {"c1_fp64": (564.74, 58.45), "c2_barrier_uo_bb": (91.99, 18.85),
                      "c2_qmc_sobol_bb_asian": (48.56, 13.88), "c3_heston_qe": (175.65, 34.57),
                      "c5_learned_serving": (88.71, 21.00)}


def config_issue(key, units, ms, timing):
    """Lane-issue roofline of one secondary configuration: the issue slots its kernel needs
    for `units` (counter-derived, ISSUE_QUADS_CONFIG) over the slots the SIMDs offer at the peak
    clock in the measured time `ms` (which may include more than the kernel, `timing`)."""
    quads, paired = ISSUE_QUADS_CONFIG[key]
    need_ms = (quads - paired) * 4 * (units / 64) / SIMDS / PEAK_CLOCK_HZ * 1e3
    return {"bound": "valu issue", "issue_limited_ms": need_ms, "ms": ms, "frac": need_ms / ms,
            "quad_cycles_per_wave_unit": quads, "paired_per_wave_unit": paired, "timing": timing}


def contracts(mcdx, n):
    common = dict(spot=100.0, strike=100.0, rate=0.05, carry_yield=0.0, vol=0.2, time=1.0,
                  path_count=n, seed=SEED, monitoring_points=STEPS)
    return {
        "asian": mcdx.contract(mcdx.ASIAN, average_style=0, strike_style=0, **common),
        "barrier_uo": mcdx.contract(mcdx.BARRIER, barrier=120.0, direction=0, knock=1,
                                    rebate=0.0, **common),
        "lookback_float": mcdx.contract(mcdx.LOOKBACK, strike_style=1, **common),
    }


def cpu_baseline(budget_s=12.0):
    """Time the reference's own multithreaded implementation (oracle/_ref) on this host, or
    the C oracle where the reference build is absent, on a bounded sample of the workload."""
    import oracle_lib as O
    cores = len(os.sched_getaffinity(0))
    R = O.ref_o3()
    flags = "-O3 -march=x86-64-v3" if R is not None and R is not O.ref() else "-O2"
    res = O.Result()

    def run(n):
        t0 = time.perf_counter()
        if R is not None:
            R.ref_mc_asian(100, 100, .05, 0, .2, 1, 0, 0, 0, STEPS, n, SEED, 0, 0, cores, C.byref(res))
            R.ref_mc_barrier(100, 100, 120, .05, 0, .2, 1, 0, 0, 1, 0.0, STEPS, n, SEED, 0, 0,
                             cores, C.byref(res))
            R.ref_mc_lookback(100, 100, .05, 0, .2, 1, 0, 1, STEPS, n, SEED, 0, cores, C.byref(res))
        else:
            for prod in (O.product(O.ASIAN, monitoring_points=STEPS),
                         O.product(O.BARRIER, barrier=120.0, monitoring_points=STEPS),
                         O.product(O.LOOKBACK, strike_style=1, monitoring_points=STEPS)):
                O.price(prod, n, SEED, chunks=cores, threads=cores)
        return time.perf_counter() - t0

    n = 1 << 14
    dt = run(n)
    n2 = int(min(N_PATHS, max(n, n * budget_s / max(dt, 1e-3) / 1.2)))
    dt2 = run(n2)
    return {"value": 3 * n2 / dt2, "unit": "paths/s", "cores": cores,
            "kind": "reference" if R is not None else "port",
            "sample": f"3 contracts x {n2} paths x {STEPS} steps (same contracts, seed {SEED}), "
                      f"{dt2:.1f} s, num_threads={cores}, built {flags}"}


def cpu_reference_configs(ex, budget_s=2.5):
    """CPU figures beside the GPU ones for C1, C3 and C5 on bounded samples, each scaled from
    a probe to about budget_s seconds: the reference's own pricers (oracle/_ref, release -O3
    build) -- C1 on its thread pool, Heston QE and LSM single-threaded as the reference runs
    them -- and, for Heston and LSM, the oracle's path-parallel restatements on all cores
    (Heston: the 9 strikes from one path pass, as the GPU prices them; LSM: per-chunk QR blocks
    stacked in a fixed order). Returns None without the reference build."""
    import oracle_lib as O
    R = O.ref_o3()
    if R is None:
        return None
    OL = O.lib_o3()
    flags = "-O3 -march=x86-64-v3" if R is not O.ref() else "-O2"
    cores = len(os.sched_getaffinity(0))
    res = O.Result()

    def scaled(fn, n, n_max):
        while True:  # grow the sample until it takes at least half the budget
            t0 = time.perf_counter()
            fn(n)
            dt = time.perf_counter() - t0
            if dt >= 0.5 * budget_s or n >= n_max:
                return n, dt
            n = int(min(n_max, n * min(16.0, max(2.0, budget_s / max(dt, 1e-4)))))

    out = {}
    n, dt = scaled(lambda n: R.ref_mc_european(100, 100, .05, 0, .2, 1, 0, n, SEED, 0, cores,
                                               C.byref(res)), 1 << 16, 1 << 30)
    out["c1_european"] = {"value": n / dt, "unit": "paths/s", "cores": cores, "kind": "reference",
                          "sample": f"{n} paths, num_threads={cores}, built {flags}, {dt:.2f} s"}
    p9 = (C.c_double * 9)(100.0, 0.03, 0.0, 0.04, 2.0, 0.04, 0.5, -0.7, 1.0)
    n, dt = scaled(lambda n: R.ref_heston_qe(p9, 100.0, 0, STEPS, n, SEED, C.byref(res)),
                   1 << 10, 1 << 22)
    out["c3_heston_qe"] = {"value": n / dt, "unit": "paths/s (one strike)", "cores": 1,
                           "kind": "reference",
                           "sample": f"{n} paths x {STEPS} steps, strike 100, single-threaded "
                                     f"(the reference's Heston pricer), built {flags}, {dt:.2f} s"}
    h = O.Heston(100.0, 0.03, 0.0, 0.04, 2.0, 0.04, 0.5, -0.7, 1.0)
    ks = (C.c_double * 9)(*[80.0 + 5 * i for i in range(9)])
    res9 = (O.Result * 9)()
    n, dt = scaled(lambda n: OL.orc_heston_qe_multi(C.byref(h), ks, 9, 0, STEPS, n, SEED, cores,
                                                    res9), 1 << 12, 1 << 22)
    out["c3_heston_qe_path_parallel"] = {
        "value": n / dt, "unit": "paths/s (9 strikes per path)", "cores": cores, "kind": "port",
        "sample": f"{n} paths x {STEPS} steps x 9 strikes, {cores} threads (the oracle's "
                  f"path-parallel Heston QE), built {flags}, {dt:.2f} s"}
    co = (C.c_double * (51 * 4))()
    has = (C.c_uint8 * 51)()
    inc = C.c_uint8()
    n, dt = scaled(lambda n: R.ref_lsm(100, 100, .06, 0, .2, 1, 1, 50, n, SEED, C.byref(res),
                                       co, has, C.byref(inc)), 1 << 12, 1 << 20)
    out["c5_lsm"] = {"value": n / dt, "unit": "paths/s", "cores": 1, "kind": "reference",
                     "sample": f"{n} paths x 50 dates, single-threaded (the reference's LSM "
                               f"pricer), built {flags}, {dt:.2f} s"}
    n, dt = scaled(lambda n: OL.orc_lsm_american_mt(100, 100, .06, 0, .2, 1, 1, 50, n, SEED,
                                                    cores, co, has, C.byref(inc)),
                   1 << 12, 1 << 20)
    out["c5_lsm_path_parallel"] = {
        "value": n / dt, "unit": "paths/s", "cores": cores, "kind": "port",
        "sample": f"{n} paths x 50 dates, {cores} threads (the oracle's path-parallel LSM: "
                  f"per-chunk QR blocks stacked in thread order), built {flags}, {dt:.2f} s"}
    if ex:
        c3 = ex.get("c3_heston_qe_9_strikes", {}).get("paths_per_s_wall")
        c5 = ex.get("c5_american_put_lsm", {}).get("qr_lsm", {}).get("paths_per_s")
        gpu = {"c1_european": ex.get("c1_european_1000000", {}).get("paths_per_s"),
               "c3_heston_qe": c3, "c3_heston_qe_path_parallel": c3,
               "c5_lsm": c5, "c5_lsm_path_parallel": c5}
        for k, g in gpu.items():
            if g:
                out[k]["gpu_paths_per_s"] = g
                out[k]["gpu_over_cpu"] = g / out[k]["value"]
        out["c3_heston_qe"]["note"] = ("the GPU figure prices 9 strikes per path; the CPU "
                                       "figure one strike")
    return out


def load_reference_fixture():
    path = os.path.join(ROOT, "tests", "golden", "reference_baseline.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def progress(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def timed_events(torch, stream, fn, reps=1):
    """Average ms of fn() over reps, between HIP events recorded on `stream`."""
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(stream)
    for _ in range(reps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def extras(torch, mcdx, L, stream, handle, ref):
    """Secondary configurations: C1 European, C2 bridge-corrected barrier and Sobol QMC,
    C3 Heston 9 strikes, C4 batched book vs per-contract launches, C5 American put by
    QR-LSM and by the learned continuation value. Closed forms (Black-Scholes, Heston
    semi-analytic, CRR) come from the engine's own host library (the `mcd` module)."""
    import mcd
    res = {}
    fh = float.fromhex
    dev = torch.zeros(3, dtype=torch.float64, device="cuda")

    def enq(c, prec=mcdx.PREC_FP32):
        mcdx.check(L.mcdx_price_gbm_enqueue(C.byref(c), prec, handle,
                                            C.c_void_p(dev.data_ptr())), "enqueue")

    # C1: European call, 1e6 paths, seed 42, FP64 throughout (the configuration's precision;
    # launch-latency scale), and 2^24 paths for the steady state. FP32 generation beside it.
    bsm = mcd.black_scholes_merton(100, 100, 0.05, 0.0, 0.2, 1.0, mcd.OptionType.Call)
    for n in (1_000_000, 1 << 24):
        c = mcdx.contract(mcdx.EUROPEAN, path_count=n, seed=42)
        enq(c, mcdx.PREC_FP64)
        ms64 = timed_events(torch, stream, lambda: enq(c, mcdx.PREC_FP64), reps=10)
        enq(c)
        ms32 = timed_events(torch, stream, lambda: enq(c), reps=10)
        p, se, _ = mcdx.price(c, mcdx.PREC_FP32)
        p64, se64, _ = mcdx.price(c, mcdx.PREC_FP64)
        tf = OPS_C1_FP64_PER_PATH * n / (ms64 / 1e3) / 1e12
        entry = {"precision": "fp64", "paths_per_s": n / (ms64 / 1e3), "ms": ms64,
                 "price_fp64": p64, "se_fp64": se64, "bsm": bsm,
                 "dev_in_se": (p64 - bsm) / se64, "within_3se_of_bsm": abs(p64 - bsm) < 3 * se64,
                 "roofline": {"bound": "valu", "ops_per_path": OPS_C1_FP64_PER_PATH,
                              "achieved": tf, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s (FP64)",
                              "frac": tf / PEAK_FP64_TFLOPS,
                              "note": "pricing call (tile kernel + merge) timed by HIP events",
                              "issue": config_issue("c1_fp64", n, ms64,
                                                    "pricing call (tile kernel + merge)")},
                 "fp32_generation": {"paths_per_s": n / (ms32 / 1e3), "ms": ms32, "price": p,
                                     "se": se, "dev_in_se": (p - bsm) / se}}
        if ref and n == 1_000_000:
            rp = fh(ref["c1_european"]["price"])
            entry.update(reference_fp64=rp, bias_vs_reference_in_se=(p64 - rp) / se64)
            entry["fp32_generation"]["bias_vs_reference_in_se"] = (p - rp) / se
        res[f"c1_european_{n}"] = entry
    progress("C1 done")
    # C2: barrier with the Brownian-bridge correction, 2^24 x 252.
    c = mcdx.contract(mcdx.BARRIER, barrier=120.0, monitoring_points=STEPS, path_count=N_PATHS,
                      seed=SEED, brownian_bridge=1)
    enq(c)
    ms = timed_events(torch, stream, lambda: enq(c), reps=2)
    p, se, _ = mcdx.price(c, mcdx.PREC_FP32)
    tf = OPS_BB * N_PATHS * STEPS / (ms / 1e3) / 1e12
    entry = {"paths_per_s": N_PATHS / (ms / 1e3), "ms": ms, "price_fp32": p, "se": se,
             "roofline": {"bound": "valu", "ops_per_path_step": OPS_BB, "achieved": tf,
                          "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": tf / PEAK_FP32_TFLOPS,
                          "timing": "pricing call (tile kernel + merge), HIP events",
                          "issue": config_issue("c2_barrier_uo_bb", N_PATHS * STEPS, ms,
                                                "pricing call (tile kernel + merge)")}}
    if ref:
        rp = fh(ref["c2_barrier_uo_call_bb"]["price"])
        entry.update(reference_fp64=rp, bias_vs_reference_in_se=(p - rp) / se)
    res["c2_barrier_uo_bb"] = entry
    progress("C2 bridge-corrected barrier done")
    # C2 QMC: Sobol 2^20 points x 252 dims + Brownian bridge (synchronous call, wall clock).
    mcdx.qmc_asian(100, 100, .05, 0, .2, 1, 0, 0, STEPS, 1 << 16, mcdx.PREC_FP32)
    t0 = time.perf_counter()
    q32 = mcdx.qmc_asian(100, 100, .05, 0, .2, 1, 0, 0, STEPS, 1 << 20, mcdx.PREC_FP32)[0]
    wall = time.perf_counter() - t0
    q64 = mcdx.qmc_asian(100, 100, .05, 0, .2, 1, 0, 0, STEPS, 1 << 20, mcdx.PREC_FP64)[0]
    # Randomized QMC on the same budget: 16 digitally shifted replicates of 2^16 points, which
    # gives the error estimate plain QMC lacks.
    asian = mcdx.contract(mcdx.ASIAN, strike=100.0, rate=0.05, vol=0.2, time=1.0,
                          monitoring_points=STEPS, path_count=1 << 16)
    mcdx.rqmc_path(asian, mcdx.PREC_FP32, 16, SEED)
    t0 = time.perf_counter()
    rq, rq_se, rq_n, _ = mcdx.rqmc_path(asian, mcdx.PREC_FP32, 16, SEED)
    rq_wall = time.perf_counter() - t0
    # QMC over the bridge for the other two headline payoffs (the reference prices only the
    # Asian by QMC), against the reference's own FP64 Monte Carlo price at 2^24 paths.
    qpaths = {}
    for name, kw, key in (("barrier_uo", dict(barrier=120.0, direction=0, knock=1), "c2_barrier_uo_call"),
                          ("lookback_float", dict(strike_style=1), "c2_lookback_float_call")):
        prod = mcdx.BARRIER if name == "barrier_uo" else mcdx.LOOKBACK
        qc = mcdx.contract(prod, strike=100.0, rate=0.05, vol=0.2, time=1.0,
                           monitoring_points=STEPS, path_count=1 << 20, **kw)
        mcdx.qmc_path(qc, mcdx.PREC_FP32)
        t0 = time.perf_counter()
        qv = mcdx.qmc_path(qc, mcdx.PREC_FP32)
        qms = (time.perf_counter() - t0) * 1e3
        entry = {"ms_wall": qms, "price_fp32": qv}
        if ref and key in ref:
            rp, rse = float.fromhex(ref[key]["price"]), float.fromhex(ref[key]["se"])
            entry.update(reference_mc_fp64=rp, reference_mc_se=rse, dev_in_reference_se=(qv - rp) / rse)
        qpaths[name] = entry
    tq = OPS_QMC * (1 << 20) * STEPS / wall / 1e12
    res["c2_qmc_sobol_bb_asian"] = {"paths_per_s_wall": (1 << 20) / wall, "ms_wall": wall * 1e3,
                                    "roofline": {"bound": "valu", "ops_per_point_date": OPS_QMC,
                                                 "achieved": tq, "peak": PEAK_FP32_TFLOPS,
                                                 "unit": "TFLOP/s", "frac": tq / PEAK_FP32_TFLOPS,
                                                 "timing": "synchronous call, wall clock "
                                                           "(kernel time in profiles/)",
                                                 "issue": config_issue(
                                                     "c2_qmc_sobol_bb_asian", (1 << 20) * STEPS,
                                                     wall * 1e3, "synchronous call, wall clock")},
                                    "other_payoffs_2p20x252": qpaths,
                                    "price_fp32": q32, "price_fp64": q64,
                                    "randomized_16x2p16": {"price": rq, "se": rq_se,
                                                           "points": rq_n,
                                                           "ms_wall": rq_wall * 1e3}}
    progress("C2 QMC done")
    # C3: Heston QE, 2^22 x 252, 9 strikes 80..120 from one pass (wall clock).
    par = mcdx.heston()
    ks = [80.0 + 5 * i for i in range(9)]
    mcdx.heston_strikes(par, ks, 0, STEPS, 1 << 16, SEED, mcdx.PREC_FP32)
    t0 = time.perf_counter()
    r = mcdx.heston_strikes(par, ks, 0, STEPS, 1 << 22, SEED, mcdx.PREC_FP32)
    wall = time.perf_counter() - t0
    h = mcd.HestonParams()
    for k_, v_ in dict(spot=100.0, rate=0.03, carry_yield=0.0, v0=0.04, kappa=2.0, theta=0.04,
                       xi=0.5, rho=-0.7, time=1.0).items():
        setattr(h, k_, v_)
    sa = [mcd.heston_semi_analytic(h, k, mcd.OptionType.Call) for k in ks]
    # FP64 mode on the same paths: the reference's arithmetic (per-path parity with the
    # oracle is tested), so the FP32-generation bias is measured path for path.
    r64 = mcdx.heston_strikes(par, ks, 0, STEPS, 1 << 22, SEED, mcdx.PREC_FP64)
    th = OPS_HESTON * (1 << 22) * STEPS / wall / 1e12
    res["c3_heston_qe_9_strikes"] = {
        "paths_per_s_wall": (1 << 22) / wall, "path_steps_per_s_wall": (1 << 22) * STEPS / wall,
        "roofline": {"bound": "valu", "ops_per_path_step": OPS_HESTON, "achieved": th,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": th / PEAK_FP32_TFLOPS,
                     "timing": "synchronous call, wall clock (kernel time in profiles/)",
                     "issue": config_issue("c3_heston_qe", (1 << 22) * STEPS, wall * 1e3,
                                           "synchronous call, wall clock")},
        "ms_wall": wall * 1e3,
        "prices": [x[0] for x in r], "se": [x[1] for x in r], "semi_analytic": sa,
        "dev_in_se": [(x[0] - a) / x[1] for x, a in zip(r, sa)],
        "prices_fp64": [x[0] for x in r64],
        "bias_fp32_vs_fp64_in_se": [(x[0] - y[0]) / y[1] for x, y in zip(r, r64)]}
    progress("C3 done")
    # C4: 4096-contract book, one persistent graph launch vs 4096 per-contract launches of
    # the same path kernel. Headline case: one Monte Carlo seed for the whole book (common
    # random numbers), so contracts of one payoff family and step count share each path's
    # draws; every contract's result still equals its own launch bit for bit. Also reported:
    # distinct seeds per contract (no sharing possible) and the book with CRN
    # finite-difference delta/gamma/vega against 5 per-contract launches per contract.
    res["c4_book_4096"] = c4_entry(torch, mcdx, L, stream, handle)
    progress("C4 done")
    res["c5_american_put_lsm"] = c5_lsm(mcdx, mcd)
    progress("C5 done")
    return res


def book_ops(mcdx, cs):
    """Algorithmic ops of one price book (shared-draw groups as the engine forms them: FP32
    contracts of one seed, payoff family and step count, path counts descending, up to 8 per
    group; a lone contract is a plain item)."""
    fam = {mcdx.EUROPEAN: 0, mcdx.DIGITAL: 0, mcdx.ASIAN: 1, mcdx.BARRIER: 2, mcdx.LOOKBACK: 4}
    buckets = {}
    ops = 0.0
    for c in cs:
        m = 1 if c.product in (mcdx.EUROPEAN, mcdx.DIGITAL) else c.monitoring_points
        if c.antithetic or (c.product == mcdx.BARRIER and c.brownian_bridge):
            ops += c.path_count * m * (OPS_DRAW + OPS_MEMBER)
            continue
        buckets.setdefault((c.seed, fam[c.product], m), []).append(c.path_count)
    for (_, _, m), counts in buckets.items():
        counts.sort(reverse=True)
        for o in range(0, len(counts), 8):
            grp = counts[o:o + 8]
            if len(grp) < 2:
                ops += grp[0] * m * (OPS_DRAW + OPS_MEMBER)
                continue
            tiles = (grp[0] + 1023) // 1024
            for t in range(tiles):  # members active on tile t share its draws
                paths = min(1024, grp[0] - 1024 * t)
                k = sum(1 for n in grp if n > 1024 * t)
                ops += paths * m * (OPS_DRAW + OPS_MEMBER * k)
    return ops


def c4_entry(torch, mcdx, L, stream, handle):
    def per_contract(cs, outs):
        for i, c in enumerate(cs):
            mcdx.check(L.mcdx_price_gbm_enqueue(C.byref(c), mcdx.PREC_FP32, handle,
                                                C.c_void_p(outs[i].data_ptr())), "enqueue")

    def run(cs, flags=0):
        book = mcdx.Book(cs, mcdx.PREC_FP32, flags)
        tuned = book.autotune()
        book.enqueue(handle.value)
        batched = timed_events(torch, stream, lambda: book.enqueue(handle.value), reps=2)
        res = book.results()
        greeks = (book.greeks(), book.pathwise()) if flags & mcdx.Book.GREEKS else None
        sched = book.schedule2()
        book.close()
        return batched, tuned, res, greeks, sched

    def per(cs):
        outs = torch.zeros((len(cs), 3), dtype=torch.float64, device="cuda")
        per_contract(cs, outs)
        ms = timed_events(torch, stream, lambda: per_contract(cs, outs), reps=1)
        host = outs.cpu().numpy()
        counts = host.view("u8")
        return ms, [mcdx.welford_finish(int(counts[i, 0]), float(host[i, 1]), float(host[i, 2]))
                    for i in range(len(cs))]

    out = {}
    crn = mcdx.c4_book(mc_seed=SEED)
    b_ms, tuned, bres, _, sched = run(crn)
    progress("C4 shared-seed book timed")
    p_ms, pres = per(crn)
    total = sum(c.path_count for c in crn)
    # Closed-form check of the terminal contracts (Black-Scholes European, cash digital).
    import mcd
    ok = tot = 0
    for c, (p, se, _) in zip(crn, bres):
        if c.product not in (mcdx.EUROPEAN, mcdx.DIGITAL):
            continue
        typ = mcd.OptionType.Call if c.type == 0 else mcd.OptionType.Put
        cf = (mcd.black_scholes_merton(c.spot, c.strike, c.rate, c.carry_yield, c.vol, c.time, typ)
              if c.product == mcdx.EUROPEAN else
              mcd.digital(c.spot, c.strike, c.rate, c.carry_yield, c.vol, c.time, typ,
                          mcd.DigitalStyle.CashOrNothing, c.cash_amount))
        tot += 1
        ok += abs(p - cf) <= 3 * se
    out["terminal_contracts_within_3se_of_closed_form"] = [ok, tot]
    steps = sum(c.path_count * (c.monitoring_points if c.product in
                                (mcdx.ASIAN, mcdx.BARRIER, mcdx.LOOKBACK) else 1) for c in crn)
    bops = book_ops(mcdx, crn)
    tb = bops / (b_ms / 1e3) / 1e12
    out["roofline"] = {"bound": "valu", "ops": bops, "achieved": tb, "peak": PEAK_FP32_TFLOPS,
                       "unit": "TFLOP/s", "frac": tb / PEAK_FP32_TFLOPS,
                       "model": "draw 113 per shared-draw path-step + 4 per member evaluated",
                       "timing": "graph replay, HIP events"}
    out.update(seeds="one seed (42) for the book", batched_ms=b_ms, per_contract_ms=p_ms,
               speedup=p_ms / b_ms, paths=total, batched_paths_per_s=total / (b_ms / 1e3),
               contract_path_steps=steps,
               batched_contract_path_steps_per_s=steps / (b_ms / 1e3),
               autotuned_schedule=sched, autotune_best_ms=tuned,
               bitwise_equal_to_per_contract=bres == pres)
    dist = mcdx.c4_book()
    d_ms, _, dres, _, _ = run(dist)
    dp_ms, dpres = per(dist)
    out["distinct_seeds"] = {"batched_ms": d_ms, "per_contract_ms": dp_ms,
                             "speedup": dp_ms / d_ms, "bitwise_equal_to_per_contract": dres == dpres}
    progress("C4 distinct-seed book timed")
    g_ms, _, gres, (greeks, pathwise), g_sched = run(crn, mcdx.Book.GREEKS)
    bumped = []
    for c in crn:
        hs, hv = max(0.01 * c.spot, 1e-4), max(0.01 * c.vol, 1e-4)
        for field, val in (("spot", c.spot + hs), ("spot", c.spot - hs), ("vol", c.vol + hv),
                           ("vol", c.vol - hv)):
            x = mcdx.GbmContract.from_buffer_copy(c)
            setattr(x, field, val)
            bumped.append(x)
    bp_ms, _ = per(bumped)
    # Pathwise delta and vega on the same paths (European and Asian contracts), against the
    # finite differences: the two estimators of the same Greeks agree within their errors.
    agree = tot_pw = 0
    for (d, _, v), (pd, pv) in zip(greeks, pathwise):
        if pd[2] == 0:
            continue
        tot_pw += 1
        agree += (abs(pd[0] - d[0]) <= 4 * (pd[1] ** 2 + d[1] ** 2) ** 0.5 and
                  abs(pv[0] - v[0]) <= 4 * (pv[1] ** 2 + v[1] ** 2) ** 0.5)
    first_pw = next(i for i, (pd, _) in enumerate(pathwise) if pd[2] > 0)
    out["greeks"] = {"what": "price + CRN central-difference delta, gamma, vega per contract "
                             "(default bumps: 1% spot, 1% vol) + pathwise delta, vega "
                             "(European and Asian contracts) on the same paths",
                     "batched_ms": g_ms, "per_contract_ms": p_ms + bp_ms,
                     "per_contract_launches": 5 * len(crn), "speedup": (p_ms + bp_ms) / g_ms,
                     "five_price_books_ms": 5 * b_ms,
                     "speedup_vs_five_price_books": 5 * b_ms / g_ms,
                     "autotuned_schedule": g_sched,
                     "prices_bitwise_equal": gres == pres,
                     "pathwise_contracts": tot_pw,
                     "pathwise_within_4se_of_fd": agree,
                     "example_contract_0": {"delta": greeks[0][0][:2], "gamma": greeks[0][1][:2],
                                            "vega": greeks[0][2][:2]},
                     "example_pathwise_contract": {
                         "index": first_pw, "fd_delta": greeks[first_pw][0][:2],
                         "pathwise_delta": pathwise[first_pw][0][:2],
                         "fd_vega": greeks[first_pw][2][:2],
                         "pathwise_vega": pathwise[first_pw][1][:2]}}
    return out


def crr_bermudan_put(s, k, r, q, v, t, steps, every):
    """CRR tree value of a put exercisable at levels that are multiples of `every` (the value
    a steps / every -date LSM estimates), vectorized over each level's nodes."""
    import numpy as np
    dt = t / steps
    u = math.exp(v * math.sqrt(dt))
    d = 1.0 / u
    pi = (math.exp((r - q) * dt) - d) / (u - d)
    disc = math.exp(-r * dt)
    i = np.arange(steps + 1)
    val = np.maximum(k - s * u ** i * d ** (steps - i), 0.0)
    for level in range(steps - 1, -1, -1):
        val = disc * (pi * val[1:level + 2] + (1.0 - pi) * val[:level + 1])
        if level % every == 0:
            j = np.arange(level + 1)
            val = np.maximum(val, k - s * u ** j * d ** (level - j))
    return float(val[0])


def c5_lsm(mcdx, mcd):
    """C5: American put S=100 K=100 r=0.06 sigma=0.2 T=1, 50 exercise dates, 2^20 paths,
    FP32 store. QR-LSM (store + 49 device regressions) against the learned continuation
    value: the full pipeline (QR fit, capture, training, table, pricing) once, then the
    trained policy serving fresh paths in one streaming kernel (no path store). Wall clock
    of the synchronous calls after a warm-up: median of 15 (QR-LSM), 7 (Greeks) or 3."""
    put = (100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 1, 50)
    n = 1 << 20

    def med(fn, reps=3):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            out = fn()
            ts.append(time.perf_counter() - t0)
        return sorted(ts)[len(ts) // 2] * 1e3, out

    qr_ms, q = med(lambda: mcdx.lsm(*put, n, SEED, mcdx.PREC_FP32), reps=15)
    pipe_ms, (d, w, co, has) = med(lambda: mcdx.lsm_learned(*put, n, SEED, mcdx.PREC_FP32), reps=1)
    serve_ms, sv = med(lambda: mcdx.lsm_learned_price(*put, n, SEED + 1, mcdx.PREC_FP32, w, 1024,
                                                      d["table_lo"], d["table_hi"], co, has,
                                                      judge=0))
    # the QR policy served alone by the same streaming kernel, on the same fresh paths
    qserve_ms, qsv = med(lambda: mcdx.lsm_learned_price(*put, n, SEED + 1, mcdx.PREC_FP32, None,
                                                        1024, d["table_lo"], d["table_hi"], co,
                                                        has, judge=2))
    judged = mcdx.lsm_learned_price(*put, n, SEED + 1, mcdx.PREC_FP32, w, 1024, d["table_lo"],
                                    d["table_hi"], co, has)
    # what the network gains: both policies on eight further fresh path sets
    gains = []
    for k in range(8):
        g = mcdx.lsm_learned_price(*put, n, 5000 + k, mcdx.PREC_FP32, w, 1024, d["table_lo"],
                                   d["table_hi"], co, has)
        gains.append(g["learned"][0] - g["qr_forward"][0])
    gain_mean = sum(gains) / len(gains)
    gain_se = (sum((x - gain_mean) ** 2 for x in gains) / (len(gains) - 1) / len(gains)) ** 0.5
    crr = mcd.crr_binomial_american(100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 10_000, mcd.OptionType.Put)
    berm = crr_bermudan_put(100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 10_000, 200)
    # Frozen-policy American Greeks (one fused pass of 5 bumped scenarios over the QR policy),
    # against central differences of CRR-2000 with the same bumps (1% spot, 1% vol).
    gk_ms, gk = med(lambda: mcdx.lsm_greeks(*put, n, SEED, mcdx.PREC_FP32), reps=7)

    # QR-LSM on Sobol + bridge paths: 16 digitally shifted runs of 2^16 paths (2^20 in all),
    # whose spread gives the error the in-sample Philox estimate does not state.
    qmc_ms, ql = med(lambda: mcdx.lsm_qmc(*put, 1 << 16, 16, SEED, mcdx.PREC_FP32), reps=1)

    def crr_at(s=100.0, v=0.2):
        return mcd.crr_binomial_american(s, 100.0, 0.06, 0.0, v, 1.0, 2000, mcd.OptionType.Put)
    hs, hv = 1.0, 0.002
    c0 = crr_at()
    crr_greeks = {"delta": (crr_at(s=100 + hs) - crr_at(s=100 - hs)) / (2 * hs),
                  "gamma": (crr_at(s=100 + hs) - 2 * c0 + crr_at(s=100 - hs)) / hs ** 2,
                  "vega": (crr_at(v=0.2 + hv) - crr_at(v=0.2 - hv)) / (2 * hv)}
    qp, qse = q[0], q[1]
    lp, lse, _ = d["learned"]
    sp, sse, _ = sv["learned"]
    store_bytes = n * 50 * 4 + n * 49 * 4
    return {
        "paths": n, "exercise_dates": 50, "crr_10000": crr,
        "qr_lsm": {"price": qp, "se": qse, "ms_wall": qr_ms, "paths_per_s": n / (qr_ms / 1e3),
                   "dev_from_crr_in_se": (qp - crr) / qse,
                   "dev_from_bermudan_in_se": (qp - berm) / qse,
                   "store_bytes_algorithmic": store_bytes,
                   # the path store against the HBM roofline over the whole pricing: the
                   # chain of 49 dependent regressions, not HBM, bounds this pass
                   "hbm_roofline": {"achieved": store_bytes / (qr_ms / 1e3) / 1e9,
                                    "peak": 8000.0, "unit": "GB/s",
                                    "frac": store_bytes / (qr_ms / 1e3) / 8e12,
                                    "bound": "latency (dependent regressions)"}},
        "learned_same_paths": {"price": lp, "se": lse, "diff_vs_qr_in_se": (lp - qp) / qse,
                               "qr_forward_replay": d["qr_forward"][0],
                               "disagreement_rate": d["disagreements"] / max(1, d["itm_decisions"]),
                               "fallbacks": d["fallbacks"], "boundary_gap_s_over_k": d["boundary_gap"],
                               "final_loss": d["final_loss"], "pipeline_ms_wall": pipe_ms},
        "learned_serving_fresh_paths": {"price": sp, "se": sse, "ms_wall": serve_ms,
                                        "paths_per_s": n / (serve_ms / 1e3),
                                        "issue": config_issue(
                                            "c5_learned_serving", n * 50, serve_ms,
                                            "wall of the synchronous call: upload, network "
                                            "table, serving kernel, close, read-back"),
                                        "dev_from_crr_in_se": (sp - crr) / sse,
                                        "dev_from_bermudan_in_se": (sp - berm) / sse,
                                        "note": "out-of-sample exercise policy on fresh paths: "
                                                "a low-biased estimator of the option value"},
        "qr_policy_serving_fresh_paths": {"price": qsv["qr_forward"][0],
                                          "se": qsv["qr_forward"][1], "ms_wall": qserve_ms,
                                          "paths_per_s": n / (qserve_ms / 1e3),
                                          "dev_from_bermudan_in_se":
                                              (qsv["qr_forward"][0] - berm) / qsv["qr_forward"][1]},
        "same_fresh_paths": {"learned": judged["learned"][0], "qr_policy": judged["qr_forward"][0],
                             "diff_in_se": (judged["learned"][0] - judged["qr_forward"][0]) / sse,
                             "disagreement_rate": judged["disagreements"] /
                             max(1, judged["itm_decisions"]),
                             "fallbacks": sv["fallbacks"]},
        "network_gain_over_qr_policy": {"mean": gain_mean, "paired_se": gain_se,
                                        "path_sets": len(gains),
                                        "what": "learned minus QR policy value on the same "
                                                "fresh 2^20-path sets (seeds 5000-5007)"},
        "serving_speedup_vs_qr_lsm": qr_ms / serve_ms,
        "serving_speedup_vs_qr_policy_streaming": qserve_ms / serve_ms,
        "bermudan_crr_10000_50_dates": berm,
        "qr_lsm_randomized_qmc_16x2p16": {"price": ql[0], "se": ql[1], "paths": ql[2],
                                          "ms_wall": qmc_ms,
                                          "diff_vs_qr_lsm_in_se": (ql[0] - qp) / math.hypot(ql[1], qse)},
        "american_greeks_frozen_policy": {
            "ms_wall": gk_ms,
            "price": gk[0][:2], "delta": gk[1][:2], "gamma": gk[2][:2], "vega": gk[3][:2],
            "crr_2000_central_differences": crr_greeks},
    }


def issue_roofline(mcdx, per_kernel_ms, paths):
    """Lane-issue roofline of each path launch: the issue slots it needs (counter-derived,
    per wave-path-step) over the slots the SIMDs offer at the peak clock in its measured time;
    with the issue probe's sustained rate per instruction class, measured now."""
    wave_steps = paths * STEPS / 64
    per = {}
    for k, ms in per_kernel_ms.items():
        quads, paired = ISSUE_QUADS[k]
        need_ms = (quads - paired) * 4 * wave_steps / SIMDS / PEAK_CLOCK_HZ * 1e3
        per[k] = {"issue_limited_ms": round(need_ms, 4), "launch_ms": round(ms, 4),
                  "frac": round(need_ms / ms, 4),
                  "single_issue_frac": round(quads * 4 * wave_steps / SIMDS / PEAK_CLOCK_HZ *
                                             1e3 / ms, 4)}
    rates = mcdx.issue_rates(1 << 14)
    slots = SIMDS * PEAK_CLOCK_HZ / 4
    return {"bound": "valu issue", "unit": "wave64 VALU issue slots",
            "peak_slots_per_s": slots, "per_launch": per,
            "quad_cycles_per_wave_path_step": {k: v[0] for k, v in ISSUE_QUADS.items()},
            "paired_per_wave_path_step": {k: v[1] for k, v in ISSUE_QUADS.items()},
            "probe_instructions_per_slot": {k: round(v / slots, 4) for k, v in rates.items()}}


def one_launch_each(torch, mcdx, L, stream, handle):
    """One launch of every kernel family at its configuration's size (the C2 contracts, C1 in
    FP64, the bridge-corrected barrier, Sobol QMC, Heston, the C4 book replay, C5 QR-LSM), for
    per-dispatch counter passes."""
    dev = torch.zeros((4, 3), dtype=torch.float64, device="cuda")
    ran = []
    for k, c in contracts(mcdx, N_PATHS).items():
        mcdx.check(L.mcdx_price_gbm_enqueue(C.byref(c), mcdx.PREC_FP32, handle,
                                            C.c_void_p(dev[0].data_ptr())), "enqueue")
        ran.append(f"c2 {k}")
    c1 = mcdx.contract(mcdx.EUROPEAN, path_count=1 << 24, seed=42)
    mcdx.check(L.mcdx_price_gbm_enqueue(C.byref(c1), mcdx.PREC_FP64, handle,
                                        C.c_void_p(dev[1].data_ptr())), "enqueue")
    bb = mcdx.contract(mcdx.BARRIER, barrier=120.0, monitoring_points=STEPS, path_count=N_PATHS,
                       seed=SEED, brownian_bridge=1)
    mcdx.check(L.mcdx_price_gbm_enqueue(C.byref(bb), mcdx.PREC_FP32, handle,
                                        C.c_void_p(dev[2].data_ptr())), "enqueue")
    torch.cuda.synchronize()
    ran += ["c1 fp64 2^24", "c2 bridge barrier"]
    mcdx.qmc_asian(100, 100, .05, 0, .2, 1, 0, 0, STEPS, 1 << 20, mcdx.PREC_FP32)
    mcdx.heston_strikes(mcdx.heston(), [80.0 + 5 * i for i in range(9)], 0, STEPS, 1 << 22, SEED,
                        mcdx.PREC_FP32)
    book = mcdx.Book(mcdx.c4_book(mc_seed=SEED), mcdx.PREC_FP32)
    book.enqueue(handle.value)
    book.results()
    book.close()
    mcdx.lsm(100.0, 100.0, 0.06, 0.0, 0.2, 1.0, 1, 50, 1 << 20, SEED, mcdx.PREC_FP32)
    ran += ["c2 qmc", "c3 heston", "c4 book", "c5 qr-lsm"]
    return {"ran": ran}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--paths", type=int, default=N_PATHS)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 bias leg")
    ap.add_argument("--no-extras", action="store_true", help="skip the secondary configs")
    ap.add_argument("--only", choices=["c4", "c5", "kernels", "issue"],
                    help="run one secondary config alone, one launch of every kernel family "
                         "(kernels: for counter passes under rocprofv3), or the VALU issue "
                         "probe of every instruction class (issue)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported on this "
              "single-GPU build (multi-GPU sharding is design-only)", file=sys.stderr)
        return 2

    import torch
    import mcdx

    torch.cuda.init()
    L = mcdx.lib()
    cs = contracts(mcdx, args.paths)
    names = list(cs)
    # A dedicated stream (default priority): its handle is non-null, so the engine enqueues
    # on it rather than on its own per-thread stream, and the events below see the kernels.
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    handle = C.c_void_p(stream.cuda_stream)
    assert handle.value, "need a non-default stream handle"
    if args.only == "c5":
        import mcd
        print(json.dumps({"c5_american_put_lsm": c5_lsm(mcdx, mcd)}), flush=True)
        return 0
    if args.only == "c4":
        print(json.dumps({"c4_book_4096": c4_entry(torch, mcdx, L, stream, handle)}), flush=True)
        return 0
    if args.only == "issue":
        rates = mcdx.issue_rates(1 << 14)
        print(json.dumps({"wave64_instructions_per_s": rates}), flush=True)
        return 0
    if args.only == "kernels":
        print(json.dumps(one_launch_each(torch, mcdx, L, stream, handle)), flush=True)
        return 0
    out = torch.zeros((len(names), 3), dtype=torch.float64, device="cuda")

    def step(events=None):
        for i, k in enumerate(names):
            if events is not None:
                events[i].record(stream)
            mcdx.check(L.mcdx_price_gbm_enqueue(C.byref(cs[k]), mcdx.PREC_FP32, handle,
                                                C.c_void_p(out[i].data_ptr())),
                       "mcdx_price_gbm_enqueue")
        if events is not None:
            events[len(names)].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # Timed region: K steps between two events, per-launch events inside.
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(len(names) + 1)]
          for _ in range(args.steps)]
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start.record(stream)
    for s in range(args.steps):
        step(ev[s])
    t_end.record(stream)
    torch.cuda.synchronize()
    ms_per_step = t_start.elapsed_time(t_end) / args.steps
    per_kernel_ms = {k: sum(ev[s][i].elapsed_time(ev[s][i + 1]) for s in range(args.steps)) /
                     args.steps for i, k in enumerate(names)}

    # Each row holds the Welford triple (uint64 n, double mean, double m2).
    host = out.cpu().numpy()
    counts = host.view("u8")
    prices = {k: mcdx.welford_finish(int(counts[i, 0]), float(host[i, 1]), float(host[i, 2]))
              for i, k in enumerate(names)}

    parity = {}
    for k in names:
        p, se, n = prices[k]
        parity[k] = {"price_fp32": p, "se": se, "paths": n}
    if not args.no_fp64:
        for k in names:
            p64, se64, _ = mcdx.price(cs[k], mcdx.PREC_FP64)
            parity[k].update(price_fp64=p64, bias_fp32_vs_fp64=parity[k]["price_fp32"] - p64,
                             bias_in_se=(parity[k]["price_fp32"] - p64) / se64)
    ref = load_reference_fixture()
    ref_keys = {"asian": "c2_asian_arith_call", "barrier_uo": "c2_barrier_uo_call",
                "lookback_float": "c2_lookback_float_call"}
    if ref and args.paths == N_PATHS:
        for k in names:
            rp = float.fromhex(ref[ref_keys[k]]["price"])
            parity[k].update(reference_fp64=rp,
                             bias_vs_reference_in_se=(parity[k]["price_fp32"] - rp) /
                             parity[k]["se"],
                             within_3se_of_reference=abs(parity[k]["price_fp32"] - rp) <
                             3 * parity[k]["se"])

    progress("headline timed; parity legs done")
    dom = max(per_kernel_ms, key=per_kernel_ms.get)
    ops = (OPS_DRAW + OPS_GBM + OPS_OBSERVE[dom]) * args.paths * STEPS
    t_dom = per_kernel_ms[dom] / 1e3
    achieved_tops = ops / t_dom / 1e12
    roofline = {
        "bound": "valu", "kernel": f"gbm_tile_kernel[{dom}, fp32]",
        "achieved": round(achieved_tops, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
        "frac": round(achieved_tops / PEAK_FP32_TFLOPS, 4),
        "ops_per_path_step": OPS_DRAW + OPS_GBM + OPS_OBSERVE[dom],
        "launch_ms": round(per_kernel_ms[dom], 4),
        "per_launch_ms": {k: round(v, 4) for k, v in per_kernel_ms.items()},
        "traffic": TRAFFIC_BYTES[dom] if args.paths == N_PATHS else None,
        "traffic_unit": "bytes per launch",
        "issue": issue_roofline(mcdx, per_kernel_ms, args.paths),
    }

    value = len(names) * args.paths / (ms_per_step / 1e3)
    line = {
        "metric": "paths/sec + % VALU roofline at <=3 SE vs FP64 oracle; batched vs per-contract",
        "value": value, "unit": "paths/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32-gen/f64-acc",
        "data": "synthetic (Philox counter-based draws, seed 42)",
        "config": {"workload": "GBM exotics: arithmetic Asian call K=100, up-and-out call "
                               "H=120 K=100, floating lookback call; S=100 r=0.05 q=0 "
                               "sigma=0.2 T=1", "paths_per_contract": args.paths,
                   "monitoring_points": STEPS, "contracts": len(names), "rng": "philox4x32-10",
                   "parallelism": "single GPU"},
        "path_steps_per_s": value * STEPS,
        "roofline": roofline,
        "parity": parity,
    }
    if not args.no_extras:
        line["extras"] = extras(torch, mcdx, L, stream, handle, ref)
    if not args.no_cpu:
        progress("CPU baseline (reference, bounded sample)")
        cb = cpu_baseline()
        line["cpu_baseline"] = cb
        line["gpu_over_cpu"] = value / cb["value"]
        progress("CPU reference pricers for C1, C3, C5 (bounded samples)")
        line["cpu_reference_configs"] = cpu_reference_configs(line.get("extras"))
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
