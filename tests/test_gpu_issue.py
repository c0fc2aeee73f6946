"""The VALU issue probe (mcdx_issue_probe): one wave64 instruction of every probed class per
SIMD per 4 cycles, except the transcendental exp2 at half of that, and two classes mixed
issue at the rate their separate rates imply (no overlap). These are the per-class peaks the
path kernels' lane-issue roofline is priced against; the bounds leave room for the engine
clock to sit below its 2.4 GHz peak."""
import ctypes as C

import pytest

import mcdx

pytestmark = pytest.mark.gpu

SLOTS = 1024 * 2.4e9 / 4  # wave64 issue slots per second: 1024 SIMDs, one per 4 cycles


def test_issue_rates_per_class():
    r = {k: v / SLOTS for k, v in mcdx.issue_rates(1 << 14).items()}
    print({k: round(v, 3) for k, v in r.items()})
    for k in ("fma_f32", "mad_u64_u32", "mul_hi_u32", "xor3", "fma_f64"):
        assert 0.7 < r[k] < 1.1, (k, r[k])
    # integer multiplies issue at the FMA rate on gfx950, not at a fraction of it
    assert r["mad_u64_u32"] > 0.85 * r["fma_f32"]
    assert r["mul_hi_u32"] > 0.85 * r["fma_f32"]
    assert 0.4 < r["exp2_f32"] / r["fma_f32"] < 0.6
    # alternating FMA and exp2 chains: the harmonic combination, within 10%
    harmonic = 2 / (1 / r["fma_f32"] + 1 / r["exp2_f32"])
    assert abs(r["mix_fma_exp2"] / harmonic - 1) < 0.1
    assert abs(r["mix_fma_mad"] / r["fma_f32"] - 1) < 0.1


def test_issue_probe_rejects_bad_arguments():
    sec, wi = C.c_double(), C.c_uint64()
    L = mcdx.lib()
    assert L.mcdx_issue_probe(99, 16, C.byref(sec), C.byref(wi)) == mcdx.EINVAL
    assert L.mcdx_issue_probe(0, 0, C.byref(sec), C.byref(wi)) == mcdx.EINVAL
    assert L.mcdx_issue_probe(0, (1 << 24) + 1, C.byref(sec), C.byref(wi)) == mcdx.EINVAL
