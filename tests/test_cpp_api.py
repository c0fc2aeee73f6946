"""Runs the C++ tests of the drop-in API (tests/cpp/api_test.cpp, built by build.py)."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "api_test")


def run(mode):
    p = subprocess.run([BIN, mode], capture_output=True, timeout=300)
    out = p.stdout.decode()
    assert p.returncode == 0, out
    assert "0 failures" in out, out
    return out


def test_host_functions_and_engine_error_convention():
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible: the engine-error half of host mode needs no device")
    run("host")


@pytest.mark.gpu
def test_device_pricers_against_closed_forms():
    run("device")
