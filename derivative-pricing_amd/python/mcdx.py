"""ctypes view of the mcdx C ABI (include/mcdx.h) in libmcd_hip.so.

This is the thin Python handle that tests, smoke() and bench.py use to call the HIP
engine through its C boundary. It loads only the in-tree build and raises if it is
missing; there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "lib", "libmcd_hip.so")

OK, EINVAL, EDEVICE, ENOMEM, ENODEV = range(5)
PREC_FP32, PREC_FP64 = 0, 1
EUROPEAN, DIGITAL, ASIAN, BARRIER, LOOKBACK = range(5)

u8, u32, u64, i32, dbl = C.c_uint8, C.c_uint32, C.c_uint64, C.c_int32, C.c_double
P = C.POINTER


class Result(C.Structure):
    _fields_ = [("price", dbl), ("standard_error", dbl), ("path_count", u64)]


class Options(C.Structure):
    _fields_ = [("antithetic", u8), ("control_variate", u8), ("brownian_bridge", u8),
                ("reserved", u8), ("num_threads", u32)]


class GbmContract(C.Structure):
    _fields_ = [(n, dbl) for n in ("spot", "strike", "barrier", "rebate", "rate", "carry_yield",
                                   "vol", "time", "cash_amount")] + [
        ("path_count", u64), ("seed", u64), ("monitoring_points", i32)] + [
        (n, u8) for n in ("product", "type", "direction", "knock", "strike_style",
                          "average_style", "digital_style", "antithetic", "control_variate",
                          "brownian_bridge")] + [("reserved", u8 * 2)]


def contract(product, spot=100.0, strike=100.0, barrier=0.0, rebate=0.0, rate=0.05,
             carry_yield=0.0, vol=0.2, time=1.0, cash_amount=1.0, path_count=1 << 16, seed=42,
             monitoring_points=1, type=0, direction=0, knock=1, strike_style=0,
             average_style=0, digital_style=0, antithetic=0, control_variate=0,
             brownian_bridge=0):
    c = GbmContract()
    for k, v in dict(spot=spot, strike=strike, barrier=barrier, rebate=rebate, rate=rate,
                     carry_yield=carry_yield, vol=vol, time=time, cash_amount=cash_amount,
                     path_count=path_count, seed=seed, monitoring_points=monitoring_points,
                     product=product, type=type, direction=direction, knock=knock,
                     strike_style=strike_style, average_style=average_style,
                     digital_style=digital_style, antithetic=antithetic,
                     control_variate=control_variate, brownian_bridge=brownian_bridge).items():
        setattr(c, k, v)
    return c


class LearnOptions(C.Structure):
    _fields_ = [("init_seed", C.c_uint64), ("sample_seed", C.c_uint64), ("steps", C.c_int32),
                ("learning_rate", C.c_float), ("final_learning_rate", C.c_float),
                ("grid_points", C.c_uint32), ("average_steps", C.c_uint32),
                ("reserved", C.c_uint32)]


class LearnedResult(C.Structure):
    _fields_ = [("qr_lsm", Result), ("learned", Result), ("qr_forward", Result),
                ("itm_decisions", C.c_uint64), ("disagreements", C.c_uint64),
                ("fallbacks", C.c_uint64), ("final_loss", C.c_double),
                ("boundary_gap", C.c_double), ("boundary_dates", C.c_uint32),
                ("reserved", C.c_uint32), ("table_lo", C.c_double), ("table_hi", C.c_double)]


class HestonParams(C.Structure):
    _fields_ = [(n, dbl) for n in ("spot", "rate", "carry_yield", "v0", "kappa", "theta", "xi",
                                   "rho", "time")]


def heston(spot=100.0, rate=0.03, carry_yield=0.0, v0=0.04, kappa=2.0, theta=0.04, xi=0.5,
           rho=-0.7, time=1.0):
    return HestonParams(spot, rate, carry_yield, v0, kappa, theta, xi, rho, time)


# name -> (restype, argtypes); the exported-symbol test checks this against include/mcdx.h.
SIGNATURES = {
    "mcdx_device_count": (C.c_int, [P(C.c_int)]),
    "mcdx_set_precision": (C.c_int, [C.c_int]),
    "mcdx_get_precision": (C.c_int, []),
    "mcdx_version": (C.c_char_p, []),
    "mcdx_philox_bits": (C.c_int, [u64, u64, u64, u32, u32, P(u32)]),
    "mcdx_normal_draws": (C.c_int, [u64, u64, u64, u32, C.c_int, P(dbl)]),
    "mcdx_inv_normal_f32_scan": (C.c_int, [u64, u64, P(dbl), P(u32)]),
    "mcdx_inv_normal_words": (C.c_int, [P(u32), u64, C.c_int, P(dbl)]),
    "mcdx_issue_probe": (C.c_int, [C.c_int, u32, P(dbl), P(u64)]),
    "mcdx_fault_inject": (C.c_int, [C.c_int, C.c_int32]),
    "mcdx_fault_pending": (C.c_int, [C.c_int]),
    "mcdx_lsm_schedule": (C.c_int, [C.c_int]),
    "mcdx_monte_carlo_european": (C.c_int, [dbl] * 6 + [u8, u64, u64, P(Options), P(Result)]),
    "mcdx_monte_carlo_digital": (C.c_int, [dbl] * 6 + [u8, u8, dbl, u64, u64, P(Options),
                                                       P(Result)]),
    "mcdx_monte_carlo_asian": (C.c_int, [dbl] * 6 + [u8, u8, u8, i32, u64, u64, P(Options),
                                                     P(Result)]),
    "mcdx_monte_carlo_barrier": (C.c_int, [dbl] * 7 + [u8, u8, u8, dbl, i32, u64, u64,
                                                       P(Options), P(Result)]),
    "mcdx_monte_carlo_lookback": (C.c_int, [dbl] * 6 + [u8, u8, i32, u64, u64, P(Options),
                                                        P(Result)]),
    "mcdx_price_gbm": (C.c_int, [P(GbmContract), C.c_int, P(Result)]),
    "mcdx_price_gbm_enqueue": (C.c_int, [P(GbmContract), C.c_int, C.c_void_p, C.c_void_p]),
    "mcdx_price_gbm_enqueue_noalloc": (C.c_int, [P(GbmContract), C.c_int, C.c_void_p,
                                                 C.c_void_p]),
    "mcdx_gbm_path_samples": (C.c_int, [P(GbmContract), C.c_int, u64, u64, P(dbl)]),
    "mcdx_sobol_max_dimensions": (u32, []),
    "mcdx_sobol_extended_dimensions": (u32, []),
    "mcdx_sobol_ints": (C.c_int, [u32, u64, u64, P(u32)]),
    "mcdx_qmc_sobol_european": (C.c_int, [dbl] * 6 + [u8, u64, P(dbl)]),
    "mcdx_qmc_sobol_european_p": (C.c_int, [dbl] * 6 + [u8, u64, C.c_int, P(Result)]),
    "mcdx_qmc_sobol_asian": (C.c_int, [dbl] * 6 + [u8, u8, i32, u64, P(dbl)]),
    "mcdx_qmc_sobol_asian_extended": (C.c_int, [dbl] * 6 + [u8, u8, i32, u64, C.c_int,
                                                            P(Result)]),
    "mcdx_qmc_sobol_path": (C.c_int, [P(GbmContract), C.c_int, P(Result)]),
    "mcdx_rqmc_sobol_path": (C.c_int, [P(GbmContract), C.c_int, u32, u64, P(Result), P(dbl)]),
    "mcdx_heston_qe_european": (C.c_int, [P(HestonParams), dbl, u8, i32, u64, u64, P(Result)]),
    "mcdx_heston_qe_strikes": (C.c_int, [P(HestonParams), P(dbl), i32, u8, i32, u64, u64,
                                         C.c_int, P(Result)]),
    "mcdx_heston_path_samples": (C.c_int, [P(HestonParams), dbl, u8, i32, u64, C.c_int, u64, u64,
                                           P(dbl)]),
    "mcdx_lsm_american": (C.c_int, [dbl] * 6 + [u8, i32, u64, u64, C.c_int, P(Result), P(dbl),
                                                P(u8), P(u8)]),
    "mcdx_lsm_american_qmc": (C.c_int, [dbl] * 6 + [u8, i32, u64, u32, u64, C.c_int, P(Result),
                                                P(dbl)]),
    "mcdx_monte_carlo_lsm_american": (C.c_int, [dbl] * 6 + [u8, i32, u64, u64, P(Result)]),
    "mcdx_lsm_reprice_frozen": (C.c_int, [dbl] * 6 + [u8, i32, u64, u64, P(dbl), P(u8), u8,
                                                      C.c_int, P(Result)]),
    "mcdx_lsm_american_greeks": (C.c_int, [dbl] * 6 + [u8, i32, u64, u64, C.c_int, dbl, dbl,
                                                        P(Result)]),
    "mcdx_pathwise_european": (C.c_int, [dbl] * 6 + [u8, u64, u64, P(Result)]),
    "mcdx_pathwise_asian": (C.c_int, [dbl] * 6 + [u8, u8, u8, i32, u64, u64, P(Result)]),
    "mcdx_likelihood_ratio_european": (C.c_int, [dbl] * 6 + [u8, u64, u64, P(Result)]),
    "mcdx_likelihood_ratio_digital": (C.c_int, [dbl] * 6 + [u8, u8, dbl, u64, u64, P(Result)]),
    "mcdx_likelihood_ratio_barrier": (C.c_int, [dbl] * 7 + [u8, u8, u8, dbl, i32, u64, u64,
                                                            P(Result)]),
    "mcdx_fd_european_prices": (C.c_int, [dbl] * 6 + [u8, u64, u64] + [dbl] * 4 + [P(Result)]),
    "mcdx_mlp_param_count": (u32, []),
    "mcdx_mlp_gradients": (C.c_int, [P(C.c_float)] * 4 + [u32, P(C.c_float), P(C.c_float)]),
    "mcdx_mlp_forward": (C.c_int, [P(C.c_float)] * 3 + [u64, P(C.c_float)]),
    "mcdx_mlp_train": (C.c_int, [P(C.c_float)] * 4 + [u64, i32, C.c_float, C.c_float, u64,
                                                      P(C.c_float)]),
    "mcdx_lsm_learn_default_options": (None, [P(LearnOptions)]),
    "mcdx_lsm_american_learned": (C.c_int, [dbl] * 6 + [u8, i32, u64, u64, C.c_int,
                                                         P(LearnOptions), P(C.c_float), P(dbl),
                                                         P(u8), P(LearnedResult)]),
    "mcdx_lsm_learned_price": (C.c_int, [dbl] * 6 + [u8, i32, u64, u64, C.c_int, P(C.c_float),
                                                      u32, dbl, dbl, P(dbl), P(u8), u8,
                                                      P(LearnedResult)]),
    "mcdx_book_create_ex": (C.c_int, [P(GbmContract), u32, C.c_int, u32, P(C.c_void_p)]),
    "mcdx_book_greeks": (C.c_int, [C.c_void_p, P(Result)]),
    "mcdx_book_create": (C.c_int, [P(GbmContract), u32, C.c_int, P(C.c_void_p)]),
    "mcdx_book_autotune": (C.c_int, [C.c_void_p, P(C.c_float)]),
    "mcdx_book_set_schedule": (C.c_int, [C.c_void_p, u32, u32]),
    "mcdx_book_set_schedule2": (C.c_int, [C.c_void_p, u32, u32, u32, u32]),
    "mcdx_book_get_schedule2": (C.c_int, [C.c_void_p, P(u32), P(u32), P(u32), P(u32), P(u32)]),
    "mcdx_book_get_schedule": (C.c_int, [C.c_void_p, P(u32), P(u32), P(u32)]),
    "mcdx_book_enqueue": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mcdx_book_results": (C.c_int, [C.c_void_p, P(Result)]),
    "mcdx_book_pathwise": (C.c_int, [C.c_void_p, P(Result)]),
    "mcdx_book_price": (C.c_int, [C.c_void_p, P(Result)]),
    "mcdx_book_destroy": (None, [C.c_void_p]),
}

_lib = None


class McdxError(RuntimeError):
    def __init__(self, message, status=None):
        super().__init__(message)
        self.status = status


def lib():
    """Load libmcd_hip.so (built in-tree); raise if it is not there."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise McdxError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status, what=""):
    if status != OK:
        raise McdxError(f"{what} failed with mcdx status {status}", status)


def philox_bits(seed, path0, n, draw, stream):
    out = np.empty(n, dtype=np.uint32)
    check(lib().mcdx_philox_bits(seed, path0, n, draw, stream, out.ctypes.data_as(P(u32))),
          "mcdx_philox_bits")
    return out


def normal_draws(seed, path0, n, draw, precision):
    out = np.empty(n, dtype=np.float64)
    check(lib().mcdx_normal_draws(seed, path0, n, draw, precision, out.ctypes.data_as(P(dbl))),
          "mcdx_normal_draws")
    return out


FAULT_ALLOC, FAULT_LAUNCH = 0, 1


def fault_inject(site, countdown):
    """Arm (countdown >= 1) or disarm (0) an injected failure of this thread's
    countdown-th device allocation (FAULT_ALLOC) or launch check (FAULT_LAUNCH)."""
    check(lib().mcdx_fault_inject(site, countdown), "mcdx_fault_inject")


LSM_AUTO, LSM_PER_DATE = 0, 1


def lsm_schedule(schedule):
    """Select the LSM backward-induction schedule (LSM_AUTO: one resident launch for every
    date when the run fits; LSM_PER_DATE: one launch per date); returns the previous one."""
    prev = lib().mcdx_lsm_schedule(schedule)
    if prev < 0:
        check(-prev, "mcdx_lsm_schedule")
    return prev


ISSUE_CLASSES = {"fma_f32": 0, "mad_u64_u32": 1, "mul_hi_u32": 2, "xor3": 3, "exp2_f32": 4,
                 "fma_f64": 5, "mix_fma_mad": 6, "mix_fma_exp2": 7}


def issue_rates(iterations=1 << 14):
    """Sustained wave64 instructions per second for each probed class, whole GPU."""
    out = {}
    for name, cls in ISSUE_CLASSES.items():
        sec, wi = dbl(), u64()
        check(lib().mcdx_issue_probe(cls, iterations, C.byref(sec), C.byref(wi)), "mcdx_issue_probe")
        out[name] = wi.value / sec.value
    return out


def inv_normal_words(words, precision):
    w = np.ascontiguousarray(words, dtype=np.uint32)
    out = np.empty(len(w), dtype=np.float64)
    check(lib().mcdx_inv_normal_words(w.ctypes.data_as(P(u32)), len(w), precision,
                                      out.ctypes.data_as(P(dbl))), "mcdx_inv_normal_words")
    return out


def inv_normal_f32_scan(begin=0, count=1 << 32):
    """(max_err[central, tail], argmax_bits[central, tail]) of the FP32 inverse normal against
    the FP64 one over Philox words [begin, begin + count)."""
    err = (dbl * 2)()
    bits = (u32 * 2)()
    check(lib().mcdx_inv_normal_f32_scan(begin, count, err, bits), "mcdx_inv_normal_f32_scan")
    return list(err), list(bits)


def price(c, precision=PREC_FP32):
    r = Result()
    check(lib().mcdx_price_gbm(C.byref(c), precision, C.byref(r)), "mcdx_price_gbm")
    return r.price, r.standard_error, r.path_count


def path_samples(c, precision, path0, n):
    out = np.empty(n, dtype=np.float64)
    check(lib().mcdx_gbm_path_samples(C.byref(c), precision, path0, n,
                                      out.ctypes.data_as(P(dbl))), "mcdx_gbm_path_samples")
    return out


def sobol_ints(dim, index0, n):
    out = np.empty(n, dtype=np.uint32)
    check(lib().mcdx_sobol_ints(dim, index0, n, out.ctypes.data_as(P(u32))), "mcdx_sobol_ints")
    return out


def qmc_asian(spot, strike, rate, carry, vol, time, type, strike_style, m, n, precision):
    r = Result()
    check(lib().mcdx_qmc_sobol_asian_extended(spot, strike, rate, carry, vol, time, type,
                                              strike_style, m, n, precision, C.byref(r)),
          "mcdx_qmc_sobol_asian_extended")
    return r.price, r.standard_error, r.path_count


def qmc_european(spot, strike, rate, carry, vol, time, type, n, precision):
    r = Result()
    check(lib().mcdx_qmc_sobol_european_p(spot, strike, rate, carry, vol, time, type, n,
                                          precision, C.byref(r)), "mcdx_qmc_sobol_european_p")
    return r.price


def qmc_path(c, precision):
    r = Result()
    check(lib().mcdx_qmc_sobol_path(C.byref(c), precision, C.byref(r)), "mcdx_qmc_sobol_path")
    return r.price


def rqmc_path(c, precision, replicates, shift_seed):
    """Randomized QMC: (price, standard error, total points, replicate prices)."""
    r = Result()
    reps = np.zeros(replicates, dtype=np.float64)
    check(lib().mcdx_rqmc_sobol_path(C.byref(c), precision, replicates, shift_seed, C.byref(r),
                                     reps.ctypes.data_as(P(dbl))), "mcdx_rqmc_sobol_path")
    return r.price, r.standard_error, r.path_count, reps


def heston_strikes(params, strikes, type, steps, n, seed, precision):
    ks = np.ascontiguousarray(strikes, dtype=np.float64)
    out = (Result * len(ks))()
    check(lib().mcdx_heston_qe_strikes(C.byref(params), ks.ctypes.data_as(P(dbl)), len(ks), type,
                                       steps, n, seed, precision, out), "mcdx_heston_qe_strikes")
    return [(r.price, r.standard_error, r.path_count) for r in out]


def heston_samples(params, strike, type, steps, seed, precision, path0, n):
    out = np.empty(n, dtype=np.float64)
    check(lib().mcdx_heston_path_samples(C.byref(params), strike, type, steps, seed, precision,
                                         path0, n, out.ctypes.data_as(P(dbl))),
          "mcdx_heston_path_samples")
    return out


def lsm(spot, strike, rate, carry, vol, time, type, m, n, seed, precision):
    """(price, se, coefficients[(m+1), 4], has[m+1], exercise_at_inception)"""
    r = Result()
    co = np.zeros((m + 1) * 4)
    has = np.zeros(m + 1, dtype=np.uint8)
    inc = u8(0)
    check(lib().mcdx_lsm_american(spot, strike, rate, carry, vol, time, type, m, n, seed, precision,
                                  C.byref(r), co.ctypes.data_as(P(dbl)), has.ctypes.data_as(P(u8)),
                                  C.byref(inc)), "mcdx_lsm_american")
    return r.price, r.standard_error, co.reshape(m + 1, 4), has, inc.value


def lsm_reprice(spot, strike, rate, carry, vol, time, type, m, n, seed, coeffs, has, inception,
                precision):
    r = Result()
    co = np.ascontiguousarray(coeffs, dtype=np.float64).ravel()
    hs = np.ascontiguousarray(has, dtype=np.uint8)
    check(lib().mcdx_lsm_reprice_frozen(spot, strike, rate, carry, vol, time, type, m, n, seed,
                                        co.ctypes.data_as(P(dbl)), hs.ctypes.data_as(P(u8)),
                                        inception, precision, C.byref(r)), "mcdx_lsm_reprice_frozen")
    return r.price


def lsm_qmc(spot, strike, rate, carry, vol, time, type, m, n, replicates, shift_seed, precision):
    """QR-LSM on Sobol + bridge paths: (price, se, paths, run prices)."""
    r = Result()
    reps = np.zeros(max(replicates, 1), dtype=np.float64)
    check(lib().mcdx_lsm_american_qmc(spot, strike, rate, carry, vol, time, type, m, n, replicates,
                                      shift_seed, precision, C.byref(r),
                                      reps.ctypes.data_as(P(dbl)) if replicates else None),
          "mcdx_lsm_american_qmc")
    return r.price, r.standard_error, r.path_count, reps[:replicates]


def lsm_greeks(spot, strike, rate, carry, vol, time, type, m, n, seed, precision, hs=0.0, hv=0.0):
    """[(value, se, paths)] for price, delta, gamma, vega (frozen-policy central differences)."""
    out = (Result * 4)()
    check(lib().mcdx_lsm_american_greeks(spot, strike, rate, carry, vol, time, type, m, n, seed,
                                         precision, hs, hv, out), "mcdx_lsm_american_greeks")
    return [(r.price, r.standard_error, r.path_count) for r in out]


def greeks(name, *args, n_out):
    out = (Result * n_out)()
    check(getattr(lib(), name)(*args, out), name)
    return [(r.price, r.standard_error, r.path_count) for r in out]


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(P(C.c_float))


def mlp_gradients(w, x, t, y):
    (w_, pw), (x_, px), (t_, pt), (y_, py) = _f32(w), _f32(x), _f32(t), _f32(y)
    g = np.empty(lib().mcdx_mlp_param_count(), dtype=np.float32)
    loss = np.empty(2, dtype=np.float32)
    check(lib().mcdx_mlp_gradients(pw, px, pt, py, len(x_), g.ctypes.data_as(P(C.c_float)),
                                   loss.ctypes.data_as(P(C.c_float))), "mcdx_mlp_gradients")
    return g, loss


def mlp_forward(w, x, t):
    (w_, pw), (x_, px), (t_, pt) = _f32(w), _f32(x), _f32(t)
    out = np.empty(len(x_), dtype=np.float32)
    check(lib().mcdx_mlp_forward(pw, px, pt, len(x_), out.ctypes.data_as(P(C.c_float))),
          "mcdx_mlp_forward")
    return out


def mlp_train(w, x, t, y, steps, lr, lr_decay, seed):
    w = np.array(w, dtype=np.float32)
    (x_, px), (t_, pt), (y_, py) = _f32(x), _f32(t), _f32(y)
    losses = np.empty(steps, dtype=np.float32)
    check(lib().mcdx_mlp_train(w.ctypes.data_as(P(C.c_float)), px, pt, py, len(x_), steps, lr,
                               lr_decay, seed, losses.ctypes.data_as(P(C.c_float))),
          "mcdx_mlp_train")
    return w, losses


def learn_options(**kw):
    o = LearnOptions()
    lib().mcdx_lsm_learn_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _learned_dict(r):
    res = lambda x: (x.price, x.standard_error, x.path_count)  # noqa: E731
    return {"qr_lsm": res(r.qr_lsm), "learned": res(r.learned), "qr_forward": res(r.qr_forward),
            "itm_decisions": r.itm_decisions, "disagreements": r.disagreements,
            "fallbacks": r.fallbacks, "final_loss": r.final_loss,
            "boundary_gap": r.boundary_gap, "boundary_dates": r.boundary_dates,
            "table_lo": r.table_lo, "table_hi": r.table_hi}


def lsm_learned(spot, strike, rate, carry, vol, time, type, m, n, seed, precision, **opts):
    """Fit + train + price on the same paths. Returns (dict, weights, coeffs, has)."""
    o = learn_options(**opts)
    r = LearnedResult()
    w = np.zeros(mlp_param_count(), dtype=np.float32)
    co = np.zeros((m + 1) * 4)
    has = np.zeros(m + 1, dtype=np.uint8)
    check(lib().mcdx_lsm_american_learned(spot, strike, rate, carry, vol, time, type, m, n, seed,
                                          precision, C.byref(o), w.ctypes.data_as(P(C.c_float)),
                                          co.ctypes.data_as(P(dbl)), has.ctypes.data_as(P(u8)),
                                          C.byref(r)), "mcdx_lsm_american_learned")
    return _learned_dict(r), w, co, has


def lsm_learned_price(spot, strike, rate, carry, vol, time, type, m, n, seed, precision, w,
                      grid, lo, hi, coeffs, has, judge=True):
    """judge True/1: both policies on the same paths; False/0: the network only; 2: the QR
    policy alone through the same streaming kernel."""
    r = LearnedResult()
    w_, pw = _f32(w) if w is not None else (None, None)
    co = np.ascontiguousarray(coeffs, dtype=np.float64).ravel()
    hs = np.ascontiguousarray(has, dtype=np.uint8)
    check(lib().mcdx_lsm_learned_price(spot, strike, rate, carry, vol, time, type, m, n, seed,
                                       precision, pw, grid, lo, hi, co.ctypes.data_as(P(dbl)),
                                       hs.ctypes.data_as(P(u8)), int(judge), C.byref(r)),
          "mcdx_lsm_learned_price")
    return _learned_dict(r)


def mlp_param_count():
    return lib().mcdx_mlp_param_count()


class Book:
    """Owning handle over an mcdx_book (batched contracts, one persistent launch)."""

    NO_SHARING, GREEKS = 1, 2

    def __init__(self, contracts, precision=PREC_FP32, flags=0):
        arr = (GbmContract * len(contracts))(*contracts)
        self.n = len(contracts)
        self.h = C.c_void_p()
        check(lib().mcdx_book_create_ex(arr, self.n, precision, flags, C.byref(self.h)),
              "mcdx_book_create_ex")

    def greeks(self):
        """[(delta, gamma, vega)] per contract, each (value, se, paths)."""
        out = (Result * (3 * self.n))()
        check(lib().mcdx_book_greeks(self.h, out), "mcdx_book_greeks")
        r = [(x.price, x.standard_error, x.path_count) for x in out]
        return [tuple(r[3 * i:3 * i + 3]) for i in range(self.n)]

    def pathwise(self):
        """[(delta, vega)] pathwise per contract, each (value, se, paths); NaN where the
        reference has no pathwise estimator."""
        out = (Result * (2 * self.n))()
        check(lib().mcdx_book_pathwise(self.h, out), "mcdx_book_pathwise")
        r = [(x.price, x.standard_error, x.path_count) for x in out]
        return [tuple(r[2 * i:2 * i + 2]) for i in range(self.n)]

    def price(self):
        out = (Result * self.n)()
        check(lib().mcdx_book_price(self.h, out), "mcdx_book_price")
        return [(r.price, r.standard_error, r.path_count) for r in out]

    def enqueue(self, stream=None):
        check(lib().mcdx_book_enqueue(self.h, C.c_void_p(stream) if stream else None),
              "mcdx_book_enqueue")

    def results(self):
        out = (Result * self.n)()
        check(lib().mcdx_book_results(self.h, out), "mcdx_book_results")
        return [(r.price, r.standard_error, r.path_count) for r in out]

    def autotune(self):
        ms = C.c_float()
        check(lib().mcdx_book_autotune(self.h, C.byref(ms)), "mcdx_book_autotune")
        return ms.value

    def set_schedule(self, blocks_per_cu, order):
        check(lib().mcdx_book_set_schedule(self.h, blocks_per_cu, order), "mcdx_book_set_schedule")

    def set_schedule2(self, blocks_per_cu, order, width=0, tiles_per_item=0):
        check(lib().mcdx_book_set_schedule2(self.h, blocks_per_cu, order, width, tiles_per_item),
              "mcdx_book_set_schedule2")

    def schedule(self):
        b, o, g = u32(), u32(), u32()
        check(lib().mcdx_book_get_schedule(self.h, C.byref(b), C.byref(o), C.byref(g)),
              "mcdx_book_get_schedule")
        return b.value, o.value, g.value

    def schedule2(self):
        """(blocks per CU, order, grid, fusion width, tiles per item)"""
        v = [u32() for _ in range(5)]
        check(lib().mcdx_book_get_schedule2(self.h, *[C.byref(x) for x in v]),
              "mcdx_book_get_schedule2")
        return tuple(x.value for x in v)

    def close(self):
        if self.h:
            lib().mcdx_book_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def c4_book(n_contracts=4096, seed=2026, log2_min=14, log2_max=20, scale_paths=1.0,
            mc_seed=None):
    """The batched-book workload: K ~ U[70,130], T in {0.25,0.5,1,2}, sigma ~ U[0.1,0.5],
    S=100 r=0.05 q=0; type uniform over European, cash digital, arithmetic Asian and the 8
    barriers (H = K*(1 +/- 0.2) by direction, rebate 0); N = floor(2^U(log2_min, log2_max));
    M = 252*T for path-dependent products. Parameters come from Philox (key `seed`, stream
    2) so the book is reproducible without any library RNG. Monte Carlo seeds are 42 + i per
    contract, or `mc_seed` for every contract (common random numbers across the book)."""
    import math
    out = []
    for i in range(n_contracts):
        u = [(int(b) + 0.5) / 2 ** 32 for b in philox_bits_host(seed, i, 6)]
        k = 70.0 + 60.0 * u[0]
        t = (0.25, 0.5, 1.0, 2.0)[min(3, int(u[1] * 4))]
        vol = 0.1 + 0.4 * u[2]
        typ = min(10, int(u[3] * 11))
        n = int(math.floor(2.0 ** (log2_min + (log2_max - log2_min) * u[4]) * scale_paths))
        m = int(round(252 * t))
        common = dict(spot=100.0, strike=k, rate=0.05, carry_yield=0.0, vol=vol, time=t,
                      path_count=max(n, 1), seed=42 + i if mc_seed is None else mc_seed)
        if typ == 0:
            out.append(contract(EUROPEAN, **common))
        elif typ == 1:
            out.append(contract(DIGITAL, digital_style=0, cash_amount=1.0, **common))
        elif typ == 2:
            out.append(contract(ASIAN, monitoring_points=m, **common))
        else:
            b = typ - 3  # 8 barriers: type x direction x knock
            opt, direction, knock = b & 1, (b >> 1) & 1, (b >> 2) & 1
            h = k * (1.2 if direction == 0 else 0.8)
            out.append(contract(BARRIER, barrier=h, direction=direction, knock=knock, type=opt,
                                rebate=0.0, monitoring_points=m, **common))
    return out


def philox_bits_host(seed, index, n):
    """Word 0 of Philox4x32-10 at counter {index, 0, j, 2}, key seed, j < n (host side)."""
    M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
    res = []
    for j in range(n):
        c = [index & 0xFFFFFFFF, index >> 32, j, 2]
        k0, k1 = seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF
        for _ in range(10):
            p0, p1 = M0 * c[0], M1 * c[2]
            c = [((p1 >> 32) ^ c[1] ^ k0) & 0xFFFFFFFF, p1 & 0xFFFFFFFF,
                 ((p0 >> 32) ^ c[3] ^ k1) & 0xFFFFFFFF, p0 & 0xFFFFFFFF]
            k0, k1 = (k0 + W0) & 0xFFFFFFFF, (k1 + W1) & 0xFFFFFFFF
        res.append(c[0])
    return res


def welford_finish(n, mean, m2):
    """price, SE, count from a device Welford triple (stats.hpp:35-41)."""
    import math
    var = 0.0 if n < 2 else m2 / (n - 1)
    se = 0.0 if n == 0 else math.sqrt(var / n)
    return mean, se, n
