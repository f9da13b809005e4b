"""marker2nodes' sticky-air marks (v2c_sticky_marks, interpolation.c:861-916)
and calc_timestep's sticky-air statistics (v2c_calc_sticky_stats,
calc_RSF.c:364-374, 505-550) on the GPU, against the reference's outputs
(tests/golden/sticky.npz) and the oracle.

Bar: the marks bit-exact; counts exact; extrema, P values and sums within
1e-12 relative (tests/util.py sticky_mismatches: the device takes a square
root and log10 where the reference calls pow and log10, and adds in a fixed
tree order where it adds serially)."""
import os

import numpy as np
import pytest

from velo2cycles_amd import synth
from velo2cycles_amd.abi import Context
from oracle.pyoracle import Oracle
from tests.util import compare, load_fixture, run_gpu, sticky_mismatches

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_sticky_fixture_matches_reference(gpu):
    cfg, ops, exp, z = load_fixture(os.path.join(GOLDEN, "sticky.npz"))
    print("case sticky fixture", ops, cfg.n, cfg.grid.nodenum, flush=True)
    _, gn, ret, ctx = run_gpu(gpu, cfg, ops)
    ctx.close()
    assert np.array_equal(gn["sticky01"], exp["nodal"]["sticky01"])
    bad = compare(None, None, {"nu": gn["nu"]}, {"nu": exp["nodal"]["nu"]}, nodal_fields=("nu",))
    assert not bad, bad
    ref = {k: float(z["g_" + k][0]) for k in ret["sticky"]}
    bad = sticky_mismatches(ret["sticky"], ref)
    assert not bad, bad


def _state(nx, ny, seed, per=1):
    """A grid with random marker types (1 and 2 among them), nodal stresses
    and positive viscosities, and a type-1 marker as the last marker of the
    cloud marking each of nodes (50, 50) and (500, 50) that exist."""
    cfg = synth.config_grid(nx, ny, (nx - 1) * 1e3, (ny - 1) * 1e3, per, refine=True, seed=seed)
    rng = np.random.default_rng(seed)
    m = cfg.markers
    m.t[:] = rng.choice(np.array([0, 1, 2, 4, 7], np.int8), m.n, p=[0.1, 0.35, 0.35, 0.1, 0.1])
    g = cfg.grid
    k = m.n - 1
    for m1 in (50, 500):
        if m1 < g.xnumx - 1 and 50 < g.ynumy - 1:
            # cell (m1, 49), lower-left quadrant: m3 = m1*ynumy + 49 + 1
            m["x"][k] = g.gx[m1] + 0.25 * (g.gx[m1 + 1] - g.gx[m1])
            m["y"][k] = g.gy[49] + 0.25 * (g.gy[50] - g.gy[49])
            m.t[k] = 1
            k -= 1
    nn = g.nodenum
    cfg.nodal["sxxs"] = rng.normal(0, 2e7, nn)
    cfg.nodal["sxy"] = rng.normal(0, 2e7, nn)
    cfg.nodal["nu"] = 10.0 ** rng.uniform(18, 25, nn)
    return cfg


def _device_vs_oracle(gpu, cfg):
    o = Oracle(cfg)
    o.sticky_marks()
    ost = o.sticky_stats()
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=max(cfg.n, 1), lib=gpu)
    try:
        ctx.set_all_nodal(cfg.nodal)
        ctx.upload_markers(cfg.markers)
        ctx.sticky_marks()
        marks = ctx.get_nodal("sticky01")
        gst = ctx.sticky_stats()
    finally:
        ctx.close()
    nd = int(np.sum(marks != o.nodal["sticky01"]))
    assert nd == 0, f"{nd} marks differ"
    bad = sticky_mismatches(gst, ost)
    assert not bad, bad
    return ost


@pytest.mark.parametrize("nx,ny,per", [(502, 53, 1), (1501, 301, 6)])
def test_sticky_large_matches_oracle(gpu, nx, ny, per):
    """P1 and P2 defined; at 1501 x 301 with 6 x 6 markers per cell (16.2 M
    markers) many markers mark each node, so the highest-index rule decides
    most nodes, and the statistics run over more than 2^16 nodes."""
    cfg = _state(nx, ny, seed=nx, per=per)
    print("case sticky large", nx, ny, cfg.n, flush=True)
    st = _device_vs_oracle(gpu, cfg)
    assert st["sigin_P1"] != 0 and st["visco_P1"] != 0
    assert st["sigin_P2"] != 0 and st["visco_P2"] != 0
    assert st["nr_1"] > 0 and st["nr_0"] > 0


def test_sticky_without_markers(gpu):
    """No markers: no node marked, the reference's initial statistics."""
    cfg = _state(41, 21, seed=3)
    cfg.nodal["sticky01"] = np.full(cfg.grid.nodenum, 2.0)
    cfg.markers = synth.MarkerArrays(0)
    st = _device_vs_oracle(gpu, cfg)
    assert st["nr_1"] == 0 and st["nr_0"] == 0 and st["min_sigin_1"] == 1e30 and st["max_visc_0"] == 0.0


def test_sticky_nan_and_marks_from_nodal_input(gpu):
    """Statistics over marks the host wrote (calc_timestep reads whatever
    sticky01 holds) with NaN stresses and viscosities at some marked nodes:
    the serial MAXV/MINV folds pass NaN on in the reference's way."""
    cfg = _state(301, 101, seed=8)
    rng = np.random.default_rng(9)
    nn = cfg.grid.nodenum
    mk = rng.integers(0, 3, nn).astype(float)
    idx = rng.choice(nn, 40, replace=False)
    cfg.nodal["sxxs"][idx[:20]] = np.nan
    cfg.nodal["nu"][idx[20:]] = np.nan
    o = Oracle(cfg)
    o.nodal["sticky01"][:] = mk
    ost = o.sticky_stats()
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=max(cfg.n, 1), lib=gpu)
    try:
        ctx.set_all_nodal(cfg.nodal)
        ctx.set_nodal("sticky01", mk)
        gst = ctx.sticky_stats()
    finally:
        ctx.close()
    bad = sticky_mismatches(gst, ost)
    assert not bad, bad
