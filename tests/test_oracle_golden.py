"""The oracle against the reference's own outputs (golden vectors written by
tests/golden/make_golden.py from the in-place reference build, one thread).
Bar: bit for bit, every marker and nodal field."""
import glob
import os

import numpy as np
import pytest

from velo2cycles_amd.abi import MARKER_FIELDS, NODAL_FIELDS
from oracle.pyoracle import Oracle
from tests.util import STICKY_KEYS, load_fixture, run_oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
STEP_FIXTURES = sorted(p for p in glob.glob(os.path.join(GOLDEN, "edge_*.npz")))


@pytest.mark.parametrize("path", STEP_FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_matches_reference_bitwise(path):
    cfg, ops, exp, z = load_fixture(path)
    om, on, ret = run_oracle(cfg, ops)
    for f in MARKER_FIELDS:
        assert np.array_equal(om[f], exp["markers"][f], equal_nan=True), f"marker {f}"
    assert np.array_equal(om.t, exp["markers"].t)
    for k in NODAL_FIELDS:
        assert np.array_equal(on[k], exp["nodal"][k], equal_nan=True), f"nodal {k}"
    if any(o.startswith("after") for o in ops):
        assert ret["after"][0] == z["ret"][1]


def test_ductilevisc_vectors():
    cfg, _, _, z = load_fixture(os.path.join(GOLDEN, "ductilevisc.npz"))
    o = Oracle(cfg)
    out = np.array([o.ductilevisc(t, a, b, c, d)[0] for t, a, b, c, d in
                    zip(z["dv_type"], z["dv_sxy"], z["dv_sxx"], z["dv_tk"], z["dv_pb"])])
    assert np.array_equal(out, z["dv_out"], equal_nan=True)


def test_allinteri_vectors():
    cfg, _, _, z = load_fixture(os.path.join(GOLDEN, "allinteri.npz"))
    o = Oracle(cfg)
    got = np.array([o.allinteri(x, y) for x, y in zip(z["ai_x"], z["ai_y"])])
    for j, k in enumerate(("ai_vx", "ai_vy", "ai_esp", "ai_ee")):
        assert np.array_equal(got[:, j], z[k]), k


POST_FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "post_*.npz")))


@pytest.mark.parametrize("path", POST_FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_post_routines_match_reference_bitwise(path):
    from velo2cycles_amd.abi import TimestepParams
    from tests.util import load_post_fixture
    cfg, runs, maxtmstep = load_post_fixture(path)
    ops, out, scal = runs["ops"]
    _, on, ret = run_oracle(cfg, ops)
    for k in NODAL_FIELDS:
        exp = out.get(k, cfg.nodal[k])
        assert np.array_equal(on[k], exp, equal_nan=True), f"nodal {k}"
    assert ret["stress"]["maxdphi"] == scal["maxdphi"] and ret["stress"]["maxnvp"] == scal["maxnvp"]
    for tag, oo in (("tstep0", 0), ("tstep1", 1)):
        _, _, scal = runs[tag]
        rc, r = Oracle(cfg).calc_timestep(oo, TimestepParams.default(maxtmstep=maxtmstep))
        assert rc == 0
        assert r["dt"] == scal["dt"], tag
        for k in ("dtnvp", "dtd", "dtphi", "dtL", "maxvel"):
            assert r[k] == scal[k], (tag, k)
        # the sticky-air statistics of the same calc_timestep (sticky01 is an
        # input of these fixtures)
        st = Oracle(cfg).sticky_stats()
        for k in STICKY_KEYS:
            assert st[k] == scal[k] or (np.isnan(st[k]) and np.isnan(scal[k])), (tag, k, st[k], scal[k])
        assert st["nr_1"] + st["nr_0"] > 0


def test_oracle_sticky_marks_and_stats_match_reference_bitwise():
    """marker2nodes' sticky-air marks and calc_timestep's statistics
    (interpolation.c:861-916, calc_RSF.c:364-374, 505-550) on the fixture's
    repeated marks, midline markers, wrapped and dropped indices, markers
    outside the grid and at NaN: marks, nu and every statistic bit for bit."""
    cfg, ops, exp, z = load_fixture(os.path.join(GOLDEN, "sticky.npz"))
    _, on, ret = run_oracle(cfg, ops)
    assert np.array_equal(on["sticky01"], exp["nodal"]["sticky01"])
    assert np.array_equal(on["nu"], exp["nodal"]["nu"], equal_nan=True)
    st = ret["sticky"]
    for k in STICKY_KEYS:
        assert st[k] == float(z["g_" + k][0]), k
    assert st["nr_1"] > 0 and st["nr_0"] > 0 and st["sigin_P1"] != 0 and st["visco_P1"] != 0
    assert np.all(np.isfinite([st[k] for k in STICKY_KEYS]))


BC_FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "bc_*.npz")))


def load_bc_fixture(path):
    import json
    from velo2cycles_amd import synth
    from velo2cycles_amd.abi import GridArrays, Params
    z = np.load(path)
    cfg = synth.config_grid(len(z["gx"]), len(z["gy"]), float(z["xsize"]), float(z["ysize"]), 1)
    cfg.grid = GridArrays(z["gx"], z["gy"], float(z["xsize"]), float(z["ysize"]))
    p = Params()
    for k, v in json.loads(str(z["params"])).items():
        setattr(p, k, v)
    cfg.params = p
    cfg.nodal = {k: np.zeros(cfg.grid.nodenum) for k in NODAL_FIELDS}
    for k in ("mvx", "mvy", "sppe", "tk"):
        cfg.nodal[k] = np.array(z["in_" + k])
    return cfg, (z["bondm"], z["bondn"], z["bondv"]), z


@pytest.mark.parametrize("path", BC_FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_boundary_tail_matches_reference_bitwise(path):
    cfg, bc, z = load_bc_fixture(path)
    for which, op, fields in (("vxvy", "bcvxvy", ("mvx", "mvy")), ("P", "bcp", ("sppe",)), ("tk", "bctk", ("tk",))):
        o = Oracle(cfg)
        assert o.boundary(which, *bc) == 0
        for k in fields:
            assert np.array_equal(o.nodal[k], z[f"out_{op}_{k}"]), (which, k)
