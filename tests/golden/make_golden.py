#!/usr/bin/env python3
"""Write golden fixtures from the reference's own implementation.

Runs only where oracle/_ref/libmcdref.so was built from the reference sources (oracle/Makefile
`ref` target); writes numbers only:

  reference_small.json   small cases: draws, MC prices at the reference's chunk count,
                         Sobol points, Heston, LSM, CRR, analytic values, Greeks
  reference_baseline.json  FP64 reference prices at the benchmark configurations
                         (--baseline; minutes of CPU time)

Floats are stored as float.hex() strings so the fixtures are exact.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402

H = float.hex


def mc_case_args():
    """MC cases (reference pricers via the harness); N=50,003 and seed 424242 follow the
    reference's determinism tests (tests/parallelism_test.cpp:36-204)."""
    base = dict(s=100.0, k=100.0, r=0.05, q=0.0, v=0.2, t=1.0)
    return [
        ("european_call", "eur", dict(base, type=0, anti=0)),
        ("european_put_anti", "eur", dict(base, type=1, anti=1)),
        ("digital_cash", "dig", dict(base, type=0, style=0, cash=1.5, anti=0)),
        ("digital_asset_put", "dig", dict(base, type=1, style=1, cash=1.0, anti=0)),
        ("asian_arith_12", "asian", dict(base, type=0, ss=0, avs=0, m=12, anti=0, cv=0)),
        ("asian_geo_float_put", "asian", dict(base, type=1, ss=1, avs=1, m=12, anti=0, cv=0)),
        ("asian_cv_anti_52", "asian", dict(base, type=0, ss=0, avs=0, m=52, anti=1, cv=1)),
        ("barrier_uo_20", "barrier", dict(base, h=120.0, type=0, dir=0, knock=1, rebate=0.0, m=20, anti=0, bb=0)),
        ("barrier_di_put_rebate", "barrier", dict(base, h=90.0, type=1, dir=1, knock=0, rebate=2.0, m=20, anti=0, bb=0)),
        ("barrier_uo_bb_50", "barrier", dict(base, h=120.0, type=0, dir=0, knock=1, rebate=0.0, m=50, anti=0, bb=1)),
        ("barrier_do_bb_anti", "barrier", dict(base, h=85.0, type=0, dir=1, knock=1, rebate=0.0, m=50, anti=1, bb=1)),
        ("lookback_float_20", "lookback", dict(base, type=0, style=1, m=20, anti=0)),
        ("lookback_fixed_put", "lookback", dict(base, k=105.0, type=1, style=0, m=20, anti=0)),
    ]


def run_mc(R, kind, a, n, seed, threads):
    res = O.Result()
    if kind == "eur":
        R.ref_mc_european(a["s"], a["k"], a["r"], a["q"], a["v"], a["t"], a["type"], n, seed, a["anti"], threads, C.byref(res))
    elif kind == "dig":
        R.ref_mc_digital(a["s"], a["k"], a["r"], a["q"], a["v"], a["t"], a["type"], a["style"], a["cash"], n, seed, a["anti"], threads, C.byref(res))
    elif kind == "asian":
        R.ref_mc_asian(a["s"], a["k"], a["r"], a["q"], a["v"], a["t"], a["type"], a["ss"], a["avs"], a["m"], n, seed, a["anti"], a["cv"], threads, C.byref(res))
    elif kind == "barrier":
        R.ref_mc_barrier(a["s"], a["k"], a["h"], a["r"], a["q"], a["v"], a["t"], a["type"], a["dir"], a["knock"], a["rebate"], a["m"], n, seed, a["anti"], a["bb"], threads, C.byref(res))
    else:
        R.ref_mc_lookback(a["s"], a["k"], a["r"], a["q"], a["v"], a["t"], a["type"], a["style"], a["m"], n, seed, a["anti"], threads, C.byref(res))
    return res.price, res.standard_error, res.path_count


LEWIS = dict(spot=100.0, rate=0.01, carry_yield=0.02, v0=0.04, kappa=4.0, theta=0.25, xi=1.0,
             rho=-0.5, time=1.0)


def small(R):
    out = {"chunk_count": R.ref_chunk_count(), "generator": "tests/golden/make_golden.py"}
    # Normal and uniform draws (the reference's FP64 bits, glibc libm).
    draws = []
    for seed, path, draw in [(42, 0, 0), (42, 1, 0), (42, 12345, 7), (7, 999_999, 251),
                             (2024, 1 << 33, 3), (0xDEADBEEF, 5, 1 << 20)]:
        draws.append([seed, path, draw, H(R.ref_standard_normal_variate(seed, path, draw)),
                      H(R.ref_uniform_variate(seed, path, draw))])
    out["draws"] = draws
    out["inverse_normal"] = [[H(u), H(R.ref_inverse_normal_cdf(u))] for u in
                             (1e-9, 0.001, 0.02425, 0.0242500001, 0.3, 0.5, 0.7, 0.97575,
                              0.999, 1 - 1e-9)]
    mc = []
    for name, kind, a in mc_case_args():
        p, se, n = run_mc(R, kind, a, 50_003, 424242, 1)
        mc.append({"name": name, "kind": kind, "args": a, "paths": n, "seed": 424242,
                   "price": H(p), "se": H(se)})
    out["mc"] = mc
    out["sobol"] = {str(d): [H(R.ref_sobol_point(d, i)) for i in range(0, 70)] for d in range(7)}
    out["qmc"] = {
        "european": H(R.ref_qmc_european(100, 100, .05, .02, .2, 1, 0, 20000)),
        "asian7": H(R.ref_qmc_asian(100, 100, .05, .02, .25, 1, 0, 0, 7, 20000)),
        "asian5_float_put": H(R.ref_qmc_asian(100, 95, .05, 0, .3, 1, 1, 1, 5, 5000)),
    }
    p9 = (C.c_double * 9)(*LEWIS.values())
    out["heston_semi_analytic"] = {str(k): H(R.ref_heston_semi_analytic(p9, k, 0))
                                   for k in (80.0, 90.0, 100.0, 110.0, 120.0)}
    res = O.Result()
    R.ref_heston_qe(p9, 100.0, 0, 20, 10_000, 7, C.byref(res))
    out["heston_qe"] = {"strike": 100.0, "steps": 20, "paths": 10_000, "seed": 7,
                        "price": H(res.price), "se": H(res.standard_error)}
    import numpy as np
    co = np.zeros(21 * 4)
    has = np.zeros(21, np.uint8)
    inc = C.c_uint8()
    R.ref_lsm(100, 100, .05, 0, .25, 1, 1, 20, 20_000, 7, C.byref(res),
              co.ctypes.data_as(C.POINTER(C.c_double)), has.ctypes.data_as(C.POINTER(C.c_uint8)),
              C.byref(inc))
    out["lsm"] = {"args": [100, 100, .05, 0, .25, 1, 1, 20, 20_000, 7], "price": H(res.price),
                  "se": H(res.standard_error), "coeffs": [H(x) for x in co],
                  "has": has.tolist(), "inception": inc.value}
    out["crr_american"] = {str(n): H(R.ref_crr_american(100, 100, .06, 0, .2, 1, n, 1))
                           for n in (50, 500, 2000)}
    out["bsm_c1"] = H(R.ref_bsm(100, 100, .05, 0, .2, 1, 0))
    out["reiner_rubinstein"] = [[h, t, d, k, H(R.ref_reiner_rubinstein(100, 100, h, .05, .01, .25, 1, t, d, k, 1.5))]
                                for h in (80.0, 120.0) for t in (0, 1) for d in (0, 1) for k in (0, 1)]
    out["kemna_vorst"] = H(R.ref_kemna_vorst(100, 100, .05, 0, .2, 1, 0))
    out["lookback_fixed"] = H(R.ref_lookback_fixed(100, 90, .05, .01, .2, 1, 0))
    out["lookback_floating"] = H(R.ref_lookback_floating(100, .05, .01, .2, 1, 1))
    o6 = (C.c_double * 6)()
    R.ref_pathwise_european(100, 100, .05, 0, .2, 1, 0, 20_000, 42, o6)
    out["pathwise_european"] = [H(x) for x in o6]
    o10 = (C.c_double * 10)()
    R.ref_lr_european(100, 100, .05, 0, .2, 1, 0, 20_000, 42, o10)
    out["lr_european"] = [H(x) for x in o10]
    # The cases below use the GPU Greeks tests' exact arguments, so the device results can be
    # compared with the reference's own bits as well as with the oracle.
    pw = []
    for ss, avs, typ in [(0, 0, 0), (1, 1, 1), (0, 1, 0)]:
        R.ref_pathwise_asian(100, 95, .05, .01, .25, 1, typ, ss, avs, 12, 10_007, 9, o6)
        pw.append({"args": [100, 95, .05, .01, .25, 1, typ, ss, avs, 12, 10_007, 9],
                   "out": [H(x) for x in o6]})
    out["pathwise_asian"] = pw
    lrd = []
    for a in ([100, 100, .05, 0, .25, 1, 0, 0, 1.0, 20_000, 7],
              [100, 105, .03, .01, .3, .5, 1, 0, 2.5, 15_001, 3],
              [100, 95, .05, .02, .2, 1, 0, 1, 1.0, 12_000, 5],
              [100, 100, .05, 0, .25, 1, 1, 1, 1.0, 9_999, 8]):
        R.ref_lr_digital(*a, o10)
        lrd.append({"args": a, "out": [H(x) for x in o10]})
    out["lr_digital"] = lrd
    lrb = []
    o8 = (C.c_double * 8)()
    for a in ([100, 100, 120, .05, 0, .2, 1, 0, 0, 1, 0.0, 50, 10_000, 11],
              [100, 100, 85, .05, .01, .25, 1, 1, 1, 0, 1.5, 20, 8_001, 4],
              [100, 95, 110, .03, 0, .3, .5, 0, 0, 0, 0.0, 12, 6_000, 9],
              [100, 105, 90, .05, 0, .2, 1, 0, 1, 1, 2.0, 30, 7_000, 13]):
        no_theta = R.ref_lr_barrier(*a, o8)
        lrb.append({"args": a, "out": [H(x) for x in o8], "theta_absent": no_theta})
    out["lr_barrier"] = lrb
    # Frozen-policy replays of the LSM fixture's policy at the base and bumped parameters
    # (src/pricers/lsm.cpp:151-213), plus one replay of a put policy fitted with exercise at
    # inception (deep in the money).
    g = out["lsm"]
    s0, k0, r0, q0, v0, t0, ty, m0, n0, sd = g["args"]
    rep = []
    for s, v, r in [(s0, v0, r0), (s0 + 1.0, v0, r0), (s0 - 1.0, v0, r0), (s0, v0 + .01, r0),
                    (s0, v0, r0 + .01)]:
        p = R.ref_reprice_frozen(s, k0, r, q0, v, t0, ty, m0, n0, sd,
                                 co.ctypes.data_as(C.POINTER(C.c_double)),
                                 has.ctypes.data_as(C.POINTER(C.c_uint8)), inc.value)
        rep.append({"spot": s, "vol": v, "rate": r, "price": H(p)})
    co2 = np.zeros(11 * 4)
    has2 = np.zeros(11, np.uint8)
    inc2 = C.c_uint8()
    R.ref_lsm(60, 100, .05, 0, .2, 1, 1, 10, 5_000, 3, C.byref(res),
              co2.ctypes.data_as(C.POINTER(C.c_double)), has2.ctypes.data_as(C.POINTER(C.c_uint8)),
              C.byref(inc2))
    p_inc = R.ref_reprice_frozen(61, 100, .05, 0, .2, 1, 1, 10, 5_000, 3,
                                 co2.ctypes.data_as(C.POINTER(C.c_double)),
                                 has2.ctypes.data_as(C.POINTER(C.c_uint8)), inc2.value)
    out["lsm_frozen"] = {"replays": rep,
                         "inception_case": {"args": [60, 100, .05, 0, .2, 1, 1, 10, 5_000, 3],
                                            "fit_price": H(res.price), "inception": inc2.value,
                                            "coeffs": [H(x) for x in co2], "has": has2.tolist(),
                                            "replay_spot": 61, "replay_price": H(p_inc)}}
    out["black76"] = [[f, k, r, v, t, ty, H(R.ref_black76(f, k, r, v, t, ty))]
                      for f, k, r, v, t, ty in [(100, 95, .05, .2, 1, 0), (100, 95, .05, .2, 1, 1),
                                                (80, 100, .02, .35, 2, 0), (120, 100, 0, .15, .25, 1),
                                                (100, 100, .05, .2, 1, 0)]]
    return out


def baseline(R, threads):
    """FP64 reference prices at the benchmark configurations (bias anchors)."""
    out = {"chunk_count": R.ref_chunk_count(), "threads": threads}
    t0 = time.time()
    p, se, n = run_mc(R, "eur", dict(s=100, k=100, r=.05, q=0, v=.2, t=1, type=0, anti=0),
                      1_000_000, 42, threads)
    out["c1_european"] = {"price": H(p), "se": H(se), "paths": n}
    N, M = 1 << 24, 252
    a = dict(s=100.0, k=100.0, r=.05, q=0.0, v=.2, t=1.0)
    for key, kind, extra in [
            ("c2_asian_arith_call", "asian", dict(type=0, ss=0, avs=0, m=M, anti=0, cv=0)),
            ("c2_barrier_uo_call", "barrier", dict(h=120.0, type=0, dir=0, knock=1, rebate=0.0, m=M, anti=0, bb=0)),
            ("c2_barrier_uo_call_bb", "barrier", dict(h=120.0, type=0, dir=0, knock=1, rebate=0.0, m=M, anti=0, bb=1)),
            ("c2_lookback_float_call", "lookback", dict(type=0, style=1, m=M, anti=0))]:
        p, se, n = run_mc(R, kind, dict(a, **extra), N, 42, threads)
        out[key] = {"price": H(p), "se": H(se), "paths": n}
        print(key, p, se, f"{time.time() - t0:.0f}s", flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--baseline", action="store_true")
    args = ap.parse_args()
    R = O.ref()
    if R is None:
        print("oracle/_ref/libmcdref.so not built; nothing to do", file=sys.stderr)
        return 1
    with open(os.path.join(HERE, "reference_small.json"), "w") as f:
        json.dump(small(R), f, indent=1)
    if args.baseline:
        with open(os.path.join(HERE, "reference_baseline.json"), "w") as f:
            json.dump(baseline(R, os.cpu_count()), f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
