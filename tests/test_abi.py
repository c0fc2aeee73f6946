"""CPU checks of the C-ABI boundary: the in-tree library loads without a GPU and exports
every entry point include/mcdx.h declares, and the Python handle binds each of them."""
import ctypes as C
import os
import re

import mcdx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mcdx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mcdx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "mcdx_price_gbm" in syms and len(syms) >= 10


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(mcdx.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_handle_binds_every_declared_symbol():
    assert sorted(mcdx.SIGNATURES) == declared_symbols()
    mcdx.lib()


def test_engine_calls_without_device_report_status_not_crash():
    # Without a visible device every compute entry point returns a status (never aborts).
    L = mcdx.lib()
    n = C.c_int(-1)
    st = L.mcdx_device_count(C.byref(n))
    if n.value > 0:
        return  # a GPU is visible: covered by the gpu tests
    r = mcdx.Result()
    assert L.mcdx_price_gbm(mcdx.contract(mcdx.EUROPEAN), 0, r) == mcdx.ENODEV
    assert L.mcdx_price_gbm(None, 0, r) == mcdx.EINVAL
    assert L.mcdx_set_precision(9) == mcdx.EINVAL
    assert L.mcdx_get_precision() in (0, 1)
    assert L.mcdx_version() == b"0.1.0"


def test_fault_injection_seam_arms_per_thread():
    # The seam only sets thread-local countdowns: it validates its arguments and reports what
    # is pending without touching a device.
    import threading
    L = mcdx.lib()
    assert L.mcdx_fault_inject(7, 1) == mcdx.EINVAL
    assert L.mcdx_fault_inject(mcdx.FAULT_ALLOC, -1) == mcdx.EINVAL
    assert L.mcdx_fault_pending(7) < 0
    mcdx.fault_inject(mcdx.FAULT_ALLOC, 3)
    mcdx.fault_inject(mcdx.FAULT_LAUNCH, 2)
    seen = []
    t = threading.Thread(target=lambda: seen.append((L.mcdx_fault_pending(mcdx.FAULT_ALLOC),
                                                     L.mcdx_fault_pending(mcdx.FAULT_LAUNCH))))
    t.start()
    t.join()
    assert seen == [(0, 0)]  # another thread is not armed
    assert L.mcdx_fault_pending(mcdx.FAULT_ALLOC) == 3
    assert L.mcdx_fault_pending(mcdx.FAULT_LAUNCH) == 2
    mcdx.fault_inject(mcdx.FAULT_ALLOC, 0)
    mcdx.fault_inject(mcdx.FAULT_LAUNCH, 0)
    assert L.mcdx_fault_pending(mcdx.FAULT_ALLOC) == 0 == L.mcdx_fault_pending(mcdx.FAULT_LAUNCH)
