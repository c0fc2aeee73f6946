"""Error paths of the C ABI on the device, forced from the host through the fault-injection
seam (mcdx_fault_inject): an allocation that reports out-of-memory, or a launch check that
reports a device error, at every position a call reaches. Each failing call must return its
status (never abort or hang), and the next call on the same thread must return exactly the
result of an undisturbed call. Kernels queued before an injected launch failure still run to
completion, so no GPU fault is involved anywhere."""
import threading

import numpy as np
import pytest

import mcdx

pytestmark = pytest.mark.gpu
L = mcdx.lib()


def _ops():
    rng = np.random.default_rng(3)
    xs = rng.uniform(-1, 1, 4096).astype(np.float32)
    ts = rng.uniform(-1, 1, 4096).astype(np.float32)
    ys = (xs * xs).astype(np.float32)
    w0 = (rng.standard_normal(mcdx.mlp_param_count()) * 0.1).astype(np.float32)
    asian = mcdx.contract(mcdx.ASIAN, monitoring_points=24, path_count=50_000, seed=5)
    eur64 = mcdx.contract(mcdx.EUROPEAN, path_count=30_001, seed=9)
    book_cs = [mcdx.contract(mcdx.ASIAN, strike=90 + i, monitoring_points=12, path_count=4096,
                             seed=42 + i) for i in range(6)]

    def book_price():
        b = mcdx.Book(book_cs)
        try:
            return b.price()
        finally:
            b.close()

    return {
        "gbm_fp32": lambda: mcdx.price(asian, mcdx.PREC_FP32),
        "gbm_fp64": lambda: mcdx.price(eur64, mcdx.PREC_FP64),
        "qmc": lambda: mcdx.qmc_asian(100, 100, .05, 0, .2, 1, 0, 0, 16, 4096, mcdx.PREC_FP32),
        "heston": lambda: mcdx.heston_strikes(mcdx.heston(), [90.0, 100.0], 0, 16, 8192, 7,
                                              mcdx.PREC_FP32),
        "lsm": lambda: mcdx.lsm(100, 100, .06, 0, .2, 1, 1, 10, 20_000, 4, mcdx.PREC_FP64)[:2],
        "greeks": lambda: mcdx.greeks("mcdx_pathwise_european", 100, 100, .05, 0, .2, 1, 0,
                                      20_000, 42, n_out=3),
        "mlp_train": lambda: mcdx.mlp_train(w0, xs, ts, ys, 5, 1e-3, 1.0, 3)[0].tobytes(),
        "book": book_price,
    }


OPS = _ops()


def _status(fn):
    try:
        return mcdx.OK, fn()
    except mcdx.McdxError as e:
        return e.status, None


@pytest.mark.parametrize("name", sorted(OPS))
def test_injected_launch_failure_at_every_check(name):
    fn = OPS[name]
    clean = fn()
    fired = 0
    for k in range(1, 64):
        mcdx.fault_inject(mcdx.FAULT_LAUNCH, k)
        st, got = _status(fn)
        pending = L.mcdx_fault_pending(mcdx.FAULT_LAUNCH)
        mcdx.fault_inject(mcdx.FAULT_LAUNCH, 0)
        if pending > 0:         # the call made fewer than k launch checks: nothing fired
            assert st == mcdx.OK and got == clean
            break
        fired += 1
        assert st == mcdx.EDEVICE, (name, k, st)
        assert fn() == clean, (name, k)  # the engine recovers on the same thread
    else:
        pytest.fail(f"{name}: more than 63 launch checks")
    assert fired >= 1


@pytest.mark.parametrize("name", sorted(OPS))
def test_injected_allocation_failure_in_a_fresh_thread(name):
    # A fresh thread owns no scratch yet, so its first call allocates; arm each allocation
    # position of that first call in turn.
    fn = OPS[name]
    clean = fn()
    fired = 0
    for k in range(1, 16):
        out = {}

        def run():
            mcdx.fault_inject(mcdx.FAULT_ALLOC, k)
            out["first"] = _status(fn)
            out["pending"] = L.mcdx_fault_pending(mcdx.FAULT_ALLOC)
            mcdx.fault_inject(mcdx.FAULT_ALLOC, 0)
            out["second"] = _status(fn)

        t = threading.Thread(target=run)
        t.start()
        t.join()
        assert out["second"] == (mcdx.OK, clean), (name, k)
        if out["pending"] > 0:
            assert out["first"] == (mcdx.OK, clean)
            break
        fired += 1
        assert out["first"][0] == mcdx.ENOMEM, (name, k, out["first"])
    assert fired >= 1, name


def test_enqueue_launch_failure_keeps_scratch_ordered():
    # An enqueue whose merge launch reports a failure has already queued its tile kernel on
    # the caller's stream; the scratch it used must still be ordered before the next user.
    torch = pytest.importorskip("torch")
    import ctypes as C
    c = mcdx.contract(mcdx.ASIAN, monitoring_points=64, path_count=1 << 20, seed=11)
    ref = mcdx.price(c, mcdx.PREC_FP32)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    out = torch.zeros((2, 3), dtype=torch.float64, device="cuda")
    mcdx.fault_inject(mcdx.FAULT_LAUNCH, 2)  # the final merge's check
    st = L.mcdx_price_gbm_enqueue(C.byref(c), mcdx.PREC_FP32, C.c_void_p(s1.cuda_stream),
                                  C.c_void_p(out[0].data_ptr()))
    mcdx.fault_inject(mcdx.FAULT_LAUNCH, 0)
    assert st == mcdx.EDEVICE
    mcdx.check(L.mcdx_price_gbm_enqueue(C.byref(c), mcdx.PREC_FP32, C.c_void_p(s2.cuda_stream),
                                        C.c_void_p(out[1].data_ptr())), "enqueue")
    torch.cuda.synchronize()
    host = out.cpu().numpy()
    got = mcdx.welford_finish(int(host[1].view("u8")[0]), float(host[1, 1]), float(host[1, 2]))
    assert got == ref


def test_module_raises_memory_error_and_runtime_error():
    mcd = pytest.importorskip("mcd")
    C_ = mcd.OptionType.Call
    clean = mcd.monte_carlo_european(100, 100, .05, 0, .2, 1, C_, 10_000, 3).price
    out = {}

    def run():
        mcdx.fault_inject(mcdx.FAULT_ALLOC, 1)
        try:
            mcd.monte_carlo_european(100, 100, .05, 0, .2, 1, C_, 10_000, 3)
        except MemoryError:
            out["alloc"] = "MemoryError"
        mcdx.fault_inject(mcdx.FAULT_ALLOC, 0)
        mcdx.fault_inject(mcdx.FAULT_LAUNCH, 1)
        try:
            mcd.monte_carlo_european(100, 100, .05, 0, .2, 1, C_, 10_000, 3)
        except RuntimeError as e:
            out["launch"] = str(e)
        mcdx.fault_inject(mcdx.FAULT_LAUNCH, 0)
        out["after"] = mcd.monte_carlo_european(100, 100, .05, 0, .2, 1, C_, 10_000, 3).price

    t = threading.Thread(target=run)
    t.start()
    t.join()
    assert out.get("alloc") == "MemoryError"
    assert "device error" in out.get("launch", "")
    assert out["after"] == clean
