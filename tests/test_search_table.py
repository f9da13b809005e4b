"""The table-driven cell search returns exactly what the reference's bisection
(interpolation.c:1536-1573) returns, on uniform, refined, random and extreme
axes, at gridlines, their neighbouring doubles, bin edges, random points,
infinities and NaN.  Host-side build of the same header the kernels use."""
import os
import subprocess

import numpy as np
import pytest

from velo2cycles_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _build(tmp_path_factory, name):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp(name) / name
    src = os.path.join(HERE, "native", name + ".hip")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-o", str(out), src],
                   check=True, capture_output=True)
    return str(out)


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    return _build(tmp_path_factory, "search_check")


@pytest.fixture(scope="module")
def divrcp(tmp_path_factory):
    return _build(tmp_path_factory, "divrcp_check")


def _axes():
    rng = np.random.default_rng(0)
    axes = {
        "uniform": synth.uniform_axis(1501, 1500e3),
        "refined": synth.refined_axis(1501, 1500e3, 450e3, 800e3, 500.0),
        "random": np.cumsum(np.concatenate([[0.0], rng.uniform(0.1, 10.0, 700)])),
        "extreme": np.cumsum(np.concatenate([[-5.0], [1e-6, 1.0, 1e-6, 1e-6, 3.0, 1e-9, 2.0]])),
        "tiny": np.array([0.0, 1.0, 2.0, 3.0]),
    }
    return axes


def test_table_search_matches_bisection(checker, tmp_path):
    files = []
    for name, g in _axes().items():
        g = np.ascontiguousarray(g, dtype=np.float64)
        assert np.all(np.diff(g) > 0), name
        p = tmp_path / f"{name}.bin"
        g.tofile(p)
        files.append(str(p))
    r = subprocess.run([checker] + files, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    for line in r.stdout.strip().splitlines():
        probes, bad = map(int, line.split())
        assert probes > 1000 and bad == 0


def test_division_by_reciprocal_is_exact(divrcp, tmp_path):
    """div_rcp(a, b, RN(1/b)) == a / b bit for bit: on every stencil
    numerator/denominator pair of the test axes and on 20 M random pairs.
    The kernels use it for every weight that decides a marker position."""
    files = []
    for name, g in _axes().items():
        p = tmp_path / f"{name}.bin"
        np.ascontiguousarray(g, dtype=np.float64).tofile(p)
        files.append(str(p))
    r = subprocess.run([divrcp] + files, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    checks, bad = map(int, r.stdout.strip().splitlines()[-1].split())
    assert checks > 20_000_000 and bad == 0
