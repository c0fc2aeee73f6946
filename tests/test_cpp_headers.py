"""The C++ header surface a reference caller uses (include/mcd/core, models, payoffs):
tests/cpp/ref_caller_test.cpp includes what the reference's benchmark and core tests include
and restates their checks against libmcd.so; here its host functions are also pinned bit for
bit to the reference's own draws (committed fixture) and to the oracle's bridge and least
squares, which are themselves pinned to the reference (tests/test_oracle.py)."""
import ctypes as C
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "ref_caller_test")
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_small.json")))
L = O.lib()


def run(*args):
    p = subprocess.run([BIN, *map(str, args)], capture_output=True, timeout=300)
    assert p.returncode == 0, p.stdout.decode() + p.stderr.decode()
    return p.stdout.decode()


def test_reference_style_caller_host_checks():
    out = run("host")
    assert " 0 failures" in out, out


@pytest.mark.gpu
def test_reference_style_caller_device_checks():
    out = run("device")
    assert " 0 failures" in out, out


def test_host_draws_equal_reference_fixture_bits():
    args = [x for seed, path, draw, _, _ in GOLD["draws"] for x in (seed, path, draw)]
    lines = run("draws", *args).split()
    for k, (seed, path, draw, z, u) in enumerate(GOLD["draws"]):
        assert float.fromhex(lines[2 * k]) == float.fromhex(z)
        assert float.fromhex(lines[2 * k + 1]) == float.fromhex(u)


@pytest.mark.parametrize("n", [1, 2, 7, 50, 252])
def test_host_bridge_equals_oracle_bits(n):
    got = [float.fromhex(x) for x in run("bridge", n, 5, 77).split()]
    z = np.array([L.orc_standard_normal_variate(5, 77, i) for i in range(n)])
    want = np.zeros(n)
    L.orc_brownian_bridge_path(100.0, 0.05, 0.01, 0.25, 1.5, n, z.ctypes.data_as(C.POINTER(C.c_double)),
                               want.ctypes.data_as(C.POINTER(C.c_double)))
    assert got == want.tolist()


@pytest.mark.parametrize("rows,cols", [(3, 2), (50, 4), (1000, 4), (7, 7)])
def test_host_least_squares_equals_oracle_bits(rows, cols):
    got = [float.fromhex(x) for x in run("lsq", rows, cols, 13).split()]
    a = np.array([[L.orc_standard_normal_variate(13, i, j) for j in range(cols)] for i in range(rows)])
    y = np.array([L.orc_standard_normal_variate(13, i, 99) for i in range(rows)])
    beta = np.zeros(cols)
    L.orc_householder_ls(a.ravel().copy().ctypes.data_as(C.POINTER(C.c_double)), rows, cols,
                         y.copy().ctypes.data_as(C.POINTER(C.c_double)),
                         beta.ctypes.data_as(C.POINTER(C.c_double)))
    assert got == beta.tolist()
    assert np.allclose(got, np.linalg.lstsq(a, y, rcond=None)[0], atol=1e-12)
