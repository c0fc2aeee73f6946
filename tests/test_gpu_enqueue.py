"""The enqueue-only pricing entry point (mcdx_price_gbm_enqueue) on caller streams, interleaved
with synchronous calls on the same thread and with enqueues on other streams, with no host
synchronisation in between: every result must equal the synchronous call's bit for bit (the
tile partials of an enqueue live in scratch ordered across streams; the synchronous entry
points use their own)."""
import ctypes as C

import pytest

import mcdx

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _enqueue(c, stream, out_row, prec=mcdx.PREC_FP32):
    mcdx.check(mcdx.lib().mcdx_price_gbm_enqueue(C.byref(c), prec, C.c_void_p(stream.cuda_stream),
                                                  C.c_void_p(out_row.data_ptr())),
               "mcdx_price_gbm_enqueue")


def _read(out):
    host = out.cpu().numpy()
    counts = host.view("u8")
    return [mcdx.welford_finish(int(counts[i, 0]), float(host[i, 1]), float(host[i, 2]))
            for i in range(host.shape[0])]


def test_enqueue_interleaved_with_sync_calls_and_other_streams():
    big = mcdx.contract(mcdx.ASIAN, monitoring_points=64, path_count=1 << 22, seed=11)
    mid = mcdx.contract(mcdx.BARRIER, barrier=125.0, monitoring_points=32, path_count=3 << 19,
                        seed=12)
    small = mcdx.contract(mcdx.EUROPEAN, path_count=50_003, seed=13)
    want = {k: mcdx.price(c) for k, c in (("big", big), ("mid", mid), ("small", small))}
    a, b = torch.cuda.Stream(), torch.cuda.Stream()
    out = torch.zeros((6, 3), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    for _ in range(2):
        _enqueue(big, a, out[0])
        sync_mid = mcdx.price(mid)            # thread stream, while `a` still runs
        _enqueue(mid, b, out[1])              # another caller stream, same scratch
        _enqueue(small, a, out[2])
        sync_big = mcdx.price(big)
        _enqueue(big, b, out[3])              # grows nothing, waits for `a`'s use
        _enqueue(mid, a, out[4])
        _enqueue(small, b, out[5])
        assert sync_mid == want["mid"] and sync_big == want["big"]
        torch.cuda.synchronize()
        got = _read(out)
        assert got[0] == want["big"] and got[3] == want["big"]
        assert got[1] == want["mid"] and got[4] == want["mid"]
        assert got[2] == want["small"] and got[5] == want["small"]
        out.zero_()


def test_enqueue_scratch_growth_after_use_on_another_stream():
    # A small enqueue on one stream, then a larger one on another: the buffer grows only after
    # the first use has ended, and both results stand.
    small = mcdx.contract(mcdx.ASIAN, monitoring_points=252, path_count=1 << 20, seed=5)
    huge = mcdx.contract(mcdx.EUROPEAN, path_count=(1 << 27) + 12_345, seed=6)
    want_s, want_h = mcdx.price(small), mcdx.price(huge)
    a, b = torch.cuda.Stream(), torch.cuda.Stream()
    out = torch.zeros((2, 3), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _enqueue(small, a, out[0])
    _enqueue(huge, b, out[1])
    torch.cuda.synchronize()
    got = _read(out)
    assert got[0] == want_s and got[1] == want_h
