"""ctypes bindings for the CPU oracle (oracle/liboracle.so) and, where it was built, the
reference harness (oracle/_ref/libmcdref.so). Test infrastructure: only tests, smoke()
and bench.py's cpu_baseline leg import this module."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libmcdref.so")

u8, u32, u64, i32, dbl = C.c_uint8, C.c_uint32, C.c_uint64, C.c_int, C.c_double
P = C.POINTER


class Product(C.Structure):
    """Mirror of orc_product (oracle/mcd_oracle.h)."""
    _fields_ = [(n, dbl) for n in ("spot", "strike", "barrier", "rebate", "rate", "carry_yield",
                                   "vol", "time", "cash_amount")] + [
        ("monitoring_points", C.c_int32)] + [
        (n, u8) for n in ("product", "type", "direction", "knock", "strike_style",
                          "average_style", "digital_style", "antithetic", "control_variate",
                          "brownian_bridge")]


class Result(C.Structure):
    _fields_ = [("price", dbl), ("standard_error", dbl), ("path_count", u64)]


class Heston(C.Structure):
    _fields_ = [(n, dbl) for n in ("spot", "rate", "carry_yield", "v0", "kappa", "theta", "xi",
                                   "rho", "time")]


class Greek(C.Structure):
    _fields_ = [("value", dbl), ("standard_error", dbl)]


EUROPEAN, DIGITAL, ASIAN, BARRIER, LOOKBACK = range(5)


def product(kind, spot=100.0, strike=100.0, barrier=0.0, rebate=0.0, rate=0.05, carry_yield=0.0,
            vol=0.2, time=1.0, cash_amount=1.0, monitoring_points=1, type=0, direction=0,
            knock=1, strike_style=0, average_style=0, digital_style=0, antithetic=0,
            control_variate=0, brownian_bridge=0):
    return Product(spot, strike, barrier, rebate, rate, carry_yield, vol, time, cash_amount,
                   monitoring_points, kind, type, direction, knock, strike_style, average_style,
                   digital_style, antithetic, control_variate, brownian_bridge)


_lib = None
_lib_o3 = None


def fast_isa():
    """True where this host has the x86-64-v3 features (AVX2, FMA, BMI2, MOVBE) that the -O3
    timing builds (liboracle_o3.so, _ref/libmcdref_o3.so) are compiled for."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    have = set(line.split(":", 1)[1].split())
                    return {"avx2", "fma", "bmi2", "movbe", "f16c"} <= have
    except OSError:
        pass
    return False


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        _lib = _bind_oracle(C.CDLL(ORACLE_SO))
    return _lib


def lib_o3():
    """The -O3 x86-64-v3 timing build of the oracle (same bits), or the -O2 one where this
    host lacks those features or the build is absent."""
    global _lib_o3
    if _lib_o3 is None:
        so = ORACLE_SO.replace("liboracle.so", "liboracle_o3.so")
        _lib_o3 = _bind_oracle(C.CDLL(so)) if fast_isa() and os.path.exists(so) else lib()
    return _lib_o3


def _bind_oracle(L):
    sig = {
        "orc_philox4x32_10": (None, [P(u32), P(u32), P(u32)]),
        "orc_philox_bits": (u32, [u64, u64, u32, u32]),
        "orc_philox_bits_range": (None, [u64, u64, u64, u32, u32, P(u32)]),
        "orc_uniform_open01": (dbl, [u32]),
        "orc_standard_normal_variate": (dbl, [u64, u64, u32]),
        "orc_uniform_variate": (dbl, [u64, u64, u32]),
        "orc_normal_pdf": (dbl, [dbl]),
        "orc_normal_cdf": (dbl, [dbl]),
        "orc_inverse_normal_cdf": (dbl, [dbl]),
        "orc_acklam": (dbl, [dbl]),
        "orc_gbm_terminal_spot": (dbl, [dbl] * 6),
        "orc_path_sample": (dbl, [P(Product), u64, u64]),
        "orc_path_samples": (None, [P(Product), u64, u64, u64, P(dbl)]),
        "orc_price": (Result, [P(Product), u64, u64, C.c_uint, C.c_uint]),
        "orc_sobol_int": (u32, [C.c_uint, u64]),
        "orc_sobol_point": (dbl, [C.c_uint, u64]),
        "orc_sobol_max_dims": (C.c_uint, []),
        "orc_sobol_direction_table": (i32, [C.c_uint, P(u32)]),
        "orc_brownian_bridge_path": (None, [dbl] * 5 + [i32, P(dbl), P(dbl)]),
        "orc_qmc_sobol_european": (dbl, [dbl] * 6 + [i32, u64]),
        "orc_qmc_sobol_asian": (dbl, [dbl] * 6 + [i32, i32, i32, u64]),
        "orc_qmc_sobol_path": (dbl, [P(Product), u64]),
        "orc_rqmc_shift": (C.c_uint32, [u64, C.c_uint32, C.c_uint]),
        "orc_rqmc_replicate": (dbl, [P(Product), u64, u64, C.c_uint32]),
        "orc_qe_variance_step": (dbl, [dbl] * 6),
        "orc_heston_path_sample": (dbl, [P(Heston), dbl, i32, i32, u64, u64]),
        "orc_heston_qe_european": (Result, [P(Heston), dbl, i32, i32, u64, u64]),
        "orc_heston_qe_multi": (None, [P(Heston), P(dbl), i32, i32, i32, u64, u64, C.c_uint,
                                       P(Result)]),
        "orc_heston_semi_analytic": (dbl, [P(Heston), dbl, i32]),
        "orc_gauss_legendre_rule": (None, [i32, P(dbl), P(dbl)]),
        "orc_householder_ls": (None, [P(dbl), i32, i32, P(dbl), P(dbl)]),
        "orc_lsm_american": (Result, [dbl] * 6 + [i32, i32, u64, u64, P(dbl), P(u8), P(u8)]),
        "orc_lsm_american_mt": (Result, [dbl] * 6 + [i32, i32, u64, u64, C.c_uint, P(dbl),
                                                    P(u8), P(u8)]),
        "orc_lsm_american_qmc": (Result, [dbl] * 6 + [i32, i32, u64, i32, u64, C.c_uint32,
                                                     P(dbl), P(u8), P(u8)]),
        "orc_lsm_reprice_frozen": (dbl, [dbl] * 6 + [i32, i32, u64, u64, P(dbl), P(u8), u8]),
        "orc_bsm": (dbl, [dbl] * 6 + [i32]),
        "orc_black76": (dbl, [dbl] * 5 + [i32]),
        "orc_kemna_vorst": (dbl, [dbl] * 6 + [i32]),
        "orc_reiner_rubinstein": (dbl, [dbl] * 7 + [i32, i32, i32, dbl]),
        "orc_lookback_fixed": (dbl, [dbl] * 6 + [i32]),
        "orc_lookback_floating": (dbl, [dbl] * 5 + [i32]),
        "orc_digital": (dbl, [dbl] * 6 + [i32, i32, dbl]),
        "orc_crr": (dbl, [dbl] * 6 + [i32, i32, P(dbl), P(dbl), P(dbl)]),
        "orc_crr_american": (dbl, [dbl] * 6 + [i32, i32]),
        "orc_crr_bermudan": (dbl, [dbl] * 6 + [i32, i32, i32]),
        "orc_pathwise_european": (None, [dbl] * 6 + [i32, u64, u64, P(Greek)]),
        "orc_pathwise_asian": (None, [dbl] * 6 + [i32, i32, i32, i32, u64, u64, P(Greek)]),
        "orc_lr_terminal": (None, [dbl] * 6 + [i32, i32, i32, dbl, u64, u64, P(Greek)]),
        "orc_lr_barrier": (None, [dbl] * 7 + [i32, i32, i32, dbl, i32, u64, u64, P(Greek)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def philox_bits(seed, path0, n, draw, stream):
    out = np.empty(n, dtype=np.uint32)
    lib().orc_philox_bits_range(seed, path0, n, draw, stream,
                                out.ctypes.data_as(P(u32)))
    return out


def path_samples(prod, seed, path0, n):
    out = np.empty(n, dtype=np.float64)
    lib().orc_path_samples(C.byref(prod), seed, path0, n, out.ctypes.data_as(P(dbl)))
    return out


def price(prod, n, seed, chunks=8, threads=8):
    r = lib().orc_price(C.byref(prod), n, seed, chunks, threads)
    return r.price, r.standard_error, r.path_count


def sobol_table(dim):
    v = np.empty(32, dtype=np.uint32)
    assert lib().orc_sobol_direction_table(dim, v.ctypes.data_as(P(u32))) == 0
    return v


def heston(spot=100.0, rate=0.03, carry_yield=0.0, v0=0.04, kappa=2.0, theta=0.04, xi=0.5,
           rho=-0.7, time=1.0):
    return Heston(spot, rate, carry_yield, v0, kappa, theta, xi, rho, time)


def heston_multi(h, strikes, type, steps, n, seed, threads=8):
    ks = np.asarray(strikes, dtype=np.float64)
    out = (Result * len(ks))()
    lib().orc_heston_qe_multi(C.byref(h), ks.ctypes.data_as(P(dbl)), len(ks), type, steps, n,
                              seed, threads, out)
    return [(r.price, r.standard_error, r.path_count) for r in out]


def lsm(s, k, r, q, v, t, type, m, n, seed):
    coeffs = np.zeros((m + 1) * 4)
    has = np.zeros(m + 1, dtype=np.uint8)
    inc = u8(0)
    res = lib().orc_lsm_american(s, k, r, q, v, t, type, m, n, seed,
                                 coeffs.ctypes.data_as(P(dbl)), has.ctypes.data_as(P(u8)),
                                 C.byref(inc))
    return res.price, res.standard_error, coeffs.reshape(m + 1, 4), has, inc.value


def lsm_mt(s, k, r, q, v, t, type, m, n, seed, threads):
    coeffs = np.zeros((m + 1) * 4)
    has = np.zeros(m + 1, dtype=np.uint8)
    inc = u8(0)
    res = lib().orc_lsm_american_mt(s, k, r, q, v, t, type, m, n, seed, threads,
                                    coeffs.ctypes.data_as(P(dbl)), has.ctypes.data_as(P(u8)),
                                    C.byref(inc))
    return res.price, res.standard_error, coeffs.reshape(m + 1, 4), has, inc.value


_ref = None
_ref_o3 = None


def ref():
    """The reference harness, or None when oracle/_ref was not built here."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        _ref = _bind_ref(C.CDLL(REF_SO))
    return _ref


def ref_o3():
    """The reference harness built with its release optimisation (-O3) at ISA level x86-64-v3
    (same bits: no contraction), for timing; ref() where this host lacks those features."""
    global _ref_o3
    if _ref_o3 is None:
        so = REF_SO.replace("libmcdref.so", "libmcdref_o3.so")
        _ref_o3 = _bind_ref(C.CDLL(so)) if fast_isa() and os.path.exists(so) else ref()
    return _ref_o3


def _bind_ref(L):
    sig = {
        "ref_chunk_count": (C.c_uint, []),
        "ref_philox": (None, [P(u32), P(u32), P(u32)]),
        "ref_standard_normal_variate": (dbl, [u64, u64, u32]),
        "ref_uniform_variate": (dbl, [u64, u64, u32]),
        "ref_inverse_normal_cdf": (dbl, [dbl]),
        "ref_sobol_point": (dbl, [C.c_uint, u64]),
        "ref_mc_european": (None, [dbl] * 6 + [i32, u64, u64, i32, C.c_uint, P(Result)]),
        "ref_mc_digital": (None, [dbl] * 6 + [i32, i32, dbl, u64, u64, i32, C.c_uint,
                                              P(Result)]),
        "ref_mc_asian": (None, [dbl] * 6 + [i32, i32, i32, i32, u64, u64, i32, i32,
                                            C.c_uint, P(Result)]),
        "ref_mc_barrier": (None, [dbl] * 7 + [i32, i32, i32, dbl, i32, u64, u64, i32, i32,
                                              C.c_uint, P(Result)]),
        "ref_mc_lookback": (None, [dbl] * 6 + [i32, i32, i32, u64, u64, i32, C.c_uint,
                                               P(Result)]),
        "ref_qmc_european": (dbl, [dbl] * 6 + [i32, u64]),
        "ref_qmc_asian": (dbl, [dbl] * 6 + [i32, i32, i32, u64]),
        "ref_heston_qe": (None, [P(dbl), dbl, i32, i32, u64, u64, P(Result)]),
        "ref_heston_semi_analytic": (dbl, [P(dbl), dbl, i32]),
        "ref_lsm": (None, [dbl] * 6 + [i32, i32, u64, u64, P(Result), P(dbl), P(u8), P(u8)]),
        "ref_bsm": (dbl, [dbl] * 6 + [i32]),
        "ref_reiner_rubinstein": (dbl, [dbl] * 7 + [i32, i32, i32, dbl]),
        "ref_kemna_vorst": (dbl, [dbl] * 6 + [i32]),
        "ref_lookback_fixed": (dbl, [dbl] * 6 + [i32]),
        "ref_lookback_floating": (dbl, [dbl] * 5 + [i32]),
        "ref_crr_american": (dbl, [dbl] * 6 + [i32, i32]),
        "ref_pathwise_european": (None, [dbl] * 6 + [i32, u64, u64, P(dbl)]),
        "ref_lr_european": (None, [dbl] * 6 + [i32, u64, u64, P(dbl)]),
        "ref_pathwise_asian": (None, [dbl] * 6 + [i32, i32, i32, i32, u64, u64, P(dbl)]),
        "ref_lr_digital": (None, [dbl] * 6 + [i32, i32, dbl, u64, u64, P(dbl)]),
        "ref_lr_barrier": (i32, [dbl] * 7 + [i32, i32, i32, dbl, i32, u64, u64, P(dbl)]),
        "ref_reprice_frozen": (dbl, [dbl] * 6 + [i32, i32, u64, u64, P(dbl), P(u8), u8]),
        "ref_black76": (dbl, [dbl] * 5 + [i32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L
