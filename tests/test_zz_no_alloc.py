"""The allocation-free enqueue entry point (mcdx_price_gbm_enqueue_noalloc) against the existing
enqueue entry point (mcdx_price_gbm_enqueue), through the C ABI: the same kernels, so the
results must agree bit for bit, on consecutive calls whose second input is smaller than the
first (the shared per-device array keeps the first call's partials; every tile partial the
second call merges is rewritten by it)."""
import ctypes as C

import pytest

import mcdx

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _call(name, c, prec, stream, row):
    f = getattr(mcdx.lib(), name)
    return f(C.byref(c), prec, C.c_void_p(stream.cuda_stream if stream is not None else 0),
             C.c_void_p(row.data_ptr()))


def _read(out):
    host = out.cpu().numpy()
    counts = host.view("u8")
    return [(int(counts[i, 0]), float(host[i, 1]), float(host[i, 2]))
            for i in range(host.shape[0])]


@pytest.mark.parametrize("prec", [mcdx.PREC_FP32, mcdx.PREC_FP64])
def test_noalloc_equals_enqueue_on_consecutive_calls(prec):
    big = mcdx.contract(mcdx.ASIAN, monitoring_points=52, path_count=(1 << 21) + 777, seed=21)
    small = mcdx.contract(mcdx.BARRIER, barrier=115.0, monitoring_points=20, path_count=5_003,
                          seed=22)
    s = torch.cuda.Stream()
    out = torch.zeros((4, 3), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert _call("mcdx_price_gbm_enqueue", big, prec, s, out[0]) == mcdx.OK
    assert _call("mcdx_price_gbm_enqueue", small, prec, s, out[1]) == mcdx.OK
    assert _call("mcdx_price_gbm_enqueue_noalloc", big, prec, s, out[2]) == mcdx.OK
    assert _call("mcdx_price_gbm_enqueue_noalloc", small, prec, s, out[3]) == mcdx.OK
    torch.cuda.synchronize()
    got = _read(out)
    assert got[2] == got[0] and got[3] == got[1]
    assert got[2][0] == big.path_count and got[3][0] == small.path_count
    assert mcdx.welford_finish(*got[2]) == mcdx.price(big, prec)


def test_noalloc_null_stream_empty_and_limits():
    c = mcdx.contract(mcdx.EUROPEAN, path_count=1_000_000, seed=42)
    out = torch.full((2, 3), 7.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert _call("mcdx_price_gbm_enqueue_noalloc", c, mcdx.PREC_FP64, None, out[0]) == mcdx.OK
    empty = mcdx.contract(mcdx.EUROPEAN, path_count=0, seed=42)
    assert _call("mcdx_price_gbm_enqueue_noalloc", empty, mcdx.PREC_FP32, None, out[1]) == mcdx.OK
    torch.cuda.synchronize()
    got = _read(out)
    assert mcdx.welford_finish(*got[0]) == mcdx.price(c, mcdx.PREC_FP64)
    assert got[1] == (0, 0.0, 0.0)
    # the largest size the array holds is accepted, one tile more is rejected before any launch
    top = mcdx.contract(mcdx.EUROPEAN, path_count=1 << 26, seed=3)
    over = mcdx.contract(mcdx.EUROPEAN, path_count=(1 << 26) + 1, seed=3)
    assert _call("mcdx_price_gbm_enqueue_noalloc", over, mcdx.PREC_FP32, None, out[0]) == mcdx.EINVAL
    assert _call("mcdx_price_gbm_enqueue_noalloc", top, mcdx.PREC_FP32, None, out[0]) == mcdx.OK
    torch.cuda.synchronize()
    assert mcdx.welford_finish(*_read(out)[0]) == mcdx.price(top, mcdx.PREC_FP32)
    assert mcdx.lib().mcdx_price_gbm_enqueue_noalloc(C.byref(c), mcdx.PREC_FP32, None,
                                                     None) == mcdx.EINVAL
