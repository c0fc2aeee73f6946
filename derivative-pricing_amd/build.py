"""Build recipe for the MI355X engine (run by __graft_entry__.build()).

Produces, in-tree:
  derivative-pricing_amd/lib/libmcd_hip.so   HIP kernels + the mcdx C ABI (include/mcdx.h)
  derivative-pricing_amd/python/mcd*.so       pybind11 `mcd` module (drop-in Python API)
  derivative-pricing_amd/bin/mcd_cli          JSON command-line front end
  derivative-pricing_amd/_san/                host-only sanitizer builds for the CPU tests
and asks oracle/Makefile for the CPU oracle (test infrastructure) and, where the reference
tree exists, oracle/_ref.
"""
import glob
import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIBDIR = os.path.join(PKG, "lib")
PYDIR = os.path.join(PKG, "python")
LIB = os.path.join(LIBDIR, "libmcd_hip.so")
ARCH = os.environ.get("MCDX_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++20",
    # FP64 mode must round exactly as the reference's scalar code: no contraction unless a
    # function opts in (the FP32 generation code uses explicit fmaf).
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd, **kw)


def build_hip(jobs=8):
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objs = []
    procs = []
    for s in srcs:
        o = os.path.join(objdir, os.path.basename(s) + ".o")
        objs.append(o)
        cmd = [HIPCC] + HIP_FLAGS + ["-c", s, "-o", o]
        print("+", " ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd)))
        if len(procs) >= jobs:
            s0, p0 = procs.pop(0)
            if p0.wait() != 0:
                raise RuntimeError(f"hipcc failed on {s0}")
    for s0, p0 in procs:
        if p0.wait() != 0:
            raise RuntimeError(f"hipcc failed on {s0}")
    tmp = LIB + ".tmp"
    run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs +
        ["-Wl,-soname,libmcd_hip.so"])
    os.replace(tmp, LIB)


def build_host_lib():
    """libmcd.so: the drop-in C++20 API (include/mcd/**) over libmcd_hip.so."""
    srcs = [s for s in sorted(glob.glob(os.path.join(HOST, "*.cpp")))
            if not s.endswith("module.cpp")]
    out = os.path.join(LIBDIR, "libmcd.so")
    run(["g++", "-O2", "-std=c++20", "-fPIC", "-shared", "-ffp-contract=off",
         "-I", os.path.join(ROOT, "include"), "-I", CSRC] + srcs +
        ["-L", LIBDIR, "-lmcd_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-soname,libmcd.so",
         "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)


def build_pybind():
    hosts = sorted(glob.glob(os.path.join(HOST, "*.cpp")))
    if not hosts:
        return
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(PYDIR, "mcd" + suffix)
    cmd = ["g++", "-O2", "-std=c++20", "-fPIC", "-shared", "-ffp-contract=off",
           "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-I", pybind11.get_include(),
           "-I", sysconfig.get_paths()["include"]] + hosts + [
           "-L", LIBDIR, "-lmcd_hip", "-Wl,-rpath,$ORIGIN/../lib", "-o", out + ".tmp"]
    run(cmd)
    os.replace(out + ".tmp", out)


def build_cli():
    """bin/mcd_cli: the JSON command-line front end over libmcd.so."""
    bindir = os.path.join(PKG, "bin")
    os.makedirs(bindir, exist_ok=True)
    out = os.path.join(bindir, "mcd_cli")
    run(["g++", "-O2", "-std=c++20", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
         os.path.join(PKG, "cli", "mcd_cli.cpp"), "-L", LIBDIR, "-lmcd", "-lmcd_hip",
         "-Wl,-rpath,$ORIGIN/../lib", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)


def build_cpp_tests():
    """tests/cpp/*.cpp: C++ callers of the drop-in API, each linked against libmcd.so."""
    for src in sorted(glob.glob(os.path.join(ROOT, "tests", "cpp", "*.cpp"))):
        out = src[:-len(".cpp")]
        run(["g++", "-O2", "-std=c++20", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
             src, "-L", LIBDIR, "-lmcd", "-lmcd_hip",
             "-Wl,-rpath,$ORIGIN/../../derivative-pricing_amd/lib", "-o", out + ".tmp"])
        os.replace(out + ".tmp", out)


def build_sanitized():
    """_san/: host-only sanitizer builds for the CPU tests (g++ and gcc, no HIP source on these
    lines): the C++ host layer compiled into the reference-style caller, the API test and the
    CLI under AddressSanitizer + UndefinedBehaviorSanitizer (linked against the normal
    libmcd_hip.so), and the oracle's driver under ASan + UBSan and under ThreadSanitizer."""
    san = os.path.join(PKG, "_san")
    os.makedirs(san, exist_ok=True)
    asan = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
            "-fno-sanitize-recover=all", "-ffp-contract=off"]
    hosts = [s for s in sorted(glob.glob(os.path.join(HOST, "*.cpp")))
             if not s.endswith("module.cpp")]
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    link = ["-L", LIBDIR, "-lmcd_hip", "-Wl,-rpath,$ORIGIN/../lib"]
    mains = {"ref_caller_test": os.path.join(ROOT, "tests", "cpp", "ref_caller_test.cpp"),
             "api_test": os.path.join(ROOT, "tests", "cpp", "api_test.cpp"),
             "mcd_cli": os.path.join(PKG, "cli", "mcd_cli.cpp")}
    for name, main_src in mains.items():
        out = os.path.join(san, name)
        run(["g++", "-std=c++20"] + asan + inc + hosts + [main_src] + link + ["-o", out + ".tmp"])
        os.replace(out + ".tmp", out)
    odir = os.path.join(ROOT, "oracle")
    drv = os.path.join(ROOT, "tests", "cpp", "oracle_san_driver.c")
    for name, flags in (("oracle_asan", asan),
                        ("oracle_tsan", ["-O1", "-g", "-fsanitize=thread",
                                         "-fno-omit-frame-pointer", "-ffp-contract=off"])):
        out = os.path.join(san, name)
        run(["gcc", "-std=c11"] + flags + ["-I", odir, drv, os.path.join(odir, "mcd_oracle.c"),
                                           "-lm", "-lpthread", "-o", out + ".tmp"])
        os.replace(out + ".tmp", out)


def build_oracle():
    run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    ref_root = os.environ.get("MCDX_REFERENCE_ROOT", "/root/reference")
    if os.path.isdir(os.path.join(ref_root, "src")):
        run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", f"REF_ROOT={ref_root}"])


def main():
    build_hip()
    build_host_lib()
    build_pybind()
    build_cli()
    build_cpp_tests()
    build_oracle()
    build_sanitized()


if __name__ == "__main__":
    sys.exit(main())
