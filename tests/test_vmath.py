"""Accuracy of the viscosity solves' FP64 quotient, logarithm and exponential
(velo2cycles_amd/csrc/v2c_vmath.hpp) against long-double references: the
host build (IEEE reciprocal) on the CPU, the gfx950 build on the GPU.  The
references are the x86 long-double logl/expl and a/b, exact to far below a
double's ulp, so these bars pin the atanh/Taylor reductions, the ln 2 splits
and the reciprocal refinement; out-of-range arguments must give NaN (the
callers then recompute with the library functions)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
PREBUILT = os.path.join(HERE, "native", "vmath_check")


def _parse(out):
    f = out.split()
    return {f[i]: float(f[i + 1]) for i in range(0, len(f) - 1, 2)}


@pytest.fixture(scope="module")
def host_checker(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("vmath") / "vmath_check"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-o", str(out),
                    os.path.join(HERE, "native", "vmath_check.hip")], check=True, capture_output=True)
    return str(out)


def test_vmath_host(host_checker):
    r = subprocess.run([host_checker, "host", "3000000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    s = _parse(r.stdout)
    assert s["n"] == 3000000
    assert s["range_violations"] == 0
    assert s["log_max_ulp"] <= 2.5 and s["log_mean_ulp"] <= 0.3, s
    assert s["exp_max_ulp"] <= 1.5 and s["exp_mean_ulp"] <= 0.1, s
    # the host's reciprocal is IEEE 1/b, so the corrected quotient is a/b
    assert s["quot_exact"] == 1.0 and s["quot_max_ulp"] <= 0.5, s


@pytest.mark.gpu
def test_vmath_device():
    assert os.path.exists(PREBUILT), "tests/native/vmath_check is not built (make -C tests/native)"
    print("vmath_check device 3000000", flush=True)
    r = subprocess.run([PREBUILT, "device", "3000000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    s = _parse(r.stdout)
    assert s["n"] == 3000000
    assert s["range_violations"] == 0
    assert s["log_max_ulp"] <= 2.5 and s["log_mean_ulp"] <= 0.3, s
    assert s["exp_max_ulp"] <= 1.5 and s["exp_mean_ulp"] <= 0.1, s
    # the refined hardware reciprocal: faithful, and nearly always exact
    assert s["quot_max_ulp"] <= 1.0 and s["quot_exact"] >= 0.999, s
