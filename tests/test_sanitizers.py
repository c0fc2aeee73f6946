"""Host code under sanitizers (CPU only; built by build.py into derivative-pricing_amd/_san):
the C++ host layer (the reference-style caller's host checks, the API test's host mode and the
JSON CLI over good and malformed requests) under AddressSanitizer + UndefinedBehaviorSanitizer,
and the oracle driven over every function and its edge cases under ASan + UBSan and under
ThreadSanitizer (its pricers run worker threads). Any sanitizer report fails the test: the
builds abort on the first error, and the output is searched for report headers as well. The
device side is not instrumented: GPU sanitizers are not available on this hardware pool."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "derivative-pricing_amd", "_san")
REPORTS = ("ERROR: AddressSanitizer", "ERROR: LeakSanitizer", "runtime error:",
           "WARNING: ThreadSanitizer")


def run(name, *args, stdin=None):
    path = os.path.join(SAN, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not built (run __graft_entry__.build())")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", TSAN_OPTIONS="halt_on_error=1")
    p = subprocess.run([path, *args], input=stdin, capture_output=True, timeout=300, env=env)
    out, err = p.stdout.decode(), p.stderr.decode()
    for r in REPORTS:
        assert r not in out + err, out + err
    return p.returncode, out


@pytest.mark.parametrize("name", ["oracle_asan", "oracle_tsan"])
def test_oracle_driver_clean(name):
    rc, out = run(name)
    assert rc == 0 and " 0 failures" in out, out


def test_reference_style_caller_host_checks_clean():
    rc, out = run("ref_caller_test", "host")
    assert rc == 0 and " 0 failures" in out, out


def test_api_test_host_mode_clean():
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible: the engine-error half of host mode needs no device")
    rc, out = run("api_test", "host")
    assert rc == 0 and " 0 failures" in out, out


EURO = dict(product="european", spot=100, strike=100, rate=0.05, carry_yield=0.0, vol=0.2,
            time=1.0, type="call")


@pytest.mark.parametrize("req", [
    json.dumps(dict(product="forward", spot=100, rate=0.05, carry_yield=0.02, time=1.0)),
    json.dumps(dict(product="binomial_american", spot=100, strike=100, rate=0.05,
                    carry_yield=0.0, vol=0.2, time=1.0, type="put", steps=200)),
    json.dumps(dict(product="heston", request="semi_analytic", spot=100, strike=100, rate=0.01,
                    carry_yield=0.02, v0=0.04, kappa=4.0, theta=0.25, xi=1.0, rho=-0.5,
                    time=1.0, type="call")),
    json.dumps(dict(EURO, path_count=1000, seed=7)),
    json.dumps(dict(EURO, path_count=1000, seed=1e20)),
    json.dumps(dict(EURO, path_count=-5, seed=1)),
    json.dumps({"product": "not_a_real_product"}),
    "", "this is not json", "{", '{"a":1,}', '{"a":[1]}', '{"a":1} trailing', '{"a":"\\u0041"}',
    '{"product":"european","spot":"1e999999"}', '{"' + "k" * 5000 + '":1}',
    '{"product":"european","spot":1e308,"strike":1e-308,"rate":0,"carry_yield":0,"vol":0.2,'
    '"time":1,"type":"call","path_count":10,"seed":1}',
])
def test_cli_requests_clean(req):
    rc, out = run("mcd_cli", stdin=req.encode())
    assert rc in (0, 1)
    json.loads(out.strip())  # one JSON object, success or error
