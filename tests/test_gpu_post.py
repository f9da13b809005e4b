"""Post-solve nodal routines on the GPU (recompute_stress_strainrate,
interpolate_center2basic(_old), interpolate_basic2center, copyarrays,
calc_timestep) against the reference's golden outputs and the oracle.

Bar: every nodal field within 1e-12 of the scaled oracle / reference value
(tests/util.py); calc_timestep's results within rel 1e-12 (its running
minima select among per-node candidates computed with device libm)."""
import glob
import math
import os

import numpy as np
import pytest

from velo2cycles_amd import synth
from velo2cycles_amd.abi import NODAL_FIELDS, Context, TimestepParams, V2CError
from oracle.pyoracle import Oracle
from tests.util import compare, load_post_fixture, run_gpu, run_oracle, sticky_mismatches

pytestmark = pytest.mark.gpu
DT = synth.YEAR
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
POST_OPS = [f"stress:{DT}:0", f"c2b:{DT}", f"c2bold:{DT}", "b2c", "copy:0"]
TS_KEYS = ("dt", "dtnvp", "dtd", "dtphi", "dtL", "maxvel")


def _close(a, b, rel=1e-12):
    if math.isnan(a) or math.isnan(b):
        return math.isnan(a) and math.isnan(b)
    return a == b or abs(a - b) <= rel * max(abs(a), abs(b))


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "post_*.npz"))),
                         ids=lambda p: os.path.basename(p)[:-4])
def test_post_routines_match_reference_golden(gpu, path):
    cfg, runs, maxtmstep = load_post_fixture(path)
    ops, out, scal = runs["ops"]
    print("case", os.path.basename(path), flush=True)
    _, gn, ret, ctx = run_gpu(gpu, cfg, ops)
    exp = {k: out.get(k, cfg.nodal[k]) for k in NODAL_FIELDS}
    bad = compare(None, None, gn, exp)
    assert not bad, bad
    assert _close(ret["stress"]["maxdphi"], scal["maxdphi"]) and _close(ret["stress"]["maxnvp"], scal["maxnvp"])
    for tag, oo in (("tstep0", 0), ("tstep1", 1)):
        ctx.set_all_nodal(cfg.nodal)
        r = ctx.calc_timestep(oo, TimestepParams.default(maxtmstep=maxtmstep))
        _, _, s = runs[tag]
        for k in TS_KEYS:
            assert _close(r[k], s[k]), (tag, k, r[k], s[k])
        # the sticky-air statistics of the same calc_timestep
        bad = sticky_mismatches(ctx.sticky_stats(), s)
        assert not bad, (tag, bad)
    ctx.close()


def _large(pmode=0, rsf=1, seed=21, long_nan=False, nx=401, ny=101):
    cfg = synth.config_grid(nx, ny, (nx - 1) * 1e3, (ny - 1) * 1e3, 1, refine=True, seed=seed)
    cfg.params.pmode = pmode
    cfg.params.RSFmode = rsf
    cfg.nodal = synth.post_solve_nodal(cfg, seed=seed + 1)
    if long_nan:
        # a = 0 (chi NaN: lambda carried over) across several 256-node blocks,
        # b <= 0 resets and NaN velocities near block edges
        cfg.nodal["as"][5000:6100] = 0.0
        cfg.nodal["bs"][[255, 256, 511, 20000]] = 0.0
        cfg.nodal["mvx"][[767, 768, 30001]] = np.nan
        cfg.nodal["vx"][[1023, 1024]] = np.nan
        if cfg.grid.nodenum > 300_000:
            # more than 256 blocks of 256 nodes: the final compositions and
            # the lambda carry run over several blocks per thread; a = 0 over
            # 16 blocks so lambda is carried across those runs
            cfg.nodal["as"][100_000:104_100] = 0.0
            cfg.nodal["bs"][[1791, 1792, 262_143]] = 0.0
            cfg.nodal["mvx"][[262_144, 300_001]] = np.nan
    return cfg


@pytest.mark.parametrize("pmode,rsf,long_nan,nx,ny", [(0, 1, False, 401, 101), (1, 1, True, 401, 101),
                                                      (0, 0, False, 401, 101), (1, 1, True, 1501, 301),
                                                      (0, 1, False, 1501, 301)])
def test_post_routines_match_oracle_large(gpu, pmode, rsf, long_nan, nx, ny):
    cfg = _large(pmode, rsf, long_nan=long_nan, nx=nx, ny=ny)
    print("case large", pmode, rsf, long_nan, cfg.grid.nodenum, flush=True)
    _, gn, gret, ctx = run_gpu(gpu, cfg, POST_OPS)
    _, on, oret = run_oracle(cfg, POST_OPS)
    bad = compare(None, None, gn, on)
    assert not bad, bad
    for k, v in oret["stress"].items():
        assert _close(gret["stress"][k], v), k
    for oo in (0, 1):
        for lam0 in (float("nan"), 0.05):
            p = TimestepParams.default(maxtmstep=DT)
            p.lambda0 = lam0
            ctx.set_all_nodal(cfg.nodal)
            r = ctx.calc_timestep(oo, p)
            rc, o = Oracle(cfg).calc_timestep(oo, p)
            assert rc == 0
            for k in TS_KEYS + ("maxVx",):
                assert _close(r[k], o[k]), (oo, lam0, k, r[k], o[k])
    ctx.close()


def test_timestep_all_chi_nan_uses_lambda0(gpu):
    """a = 0 everywhere: chi is NaN at every node, so the reference never sets
    lambda; the value before the first node is what counts (capped at 1/5)."""
    cfg = _large(seed=33)
    cfg.nodal["as"][:] = 0.0
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=1, lib=gpu)
    ctx.set_time(cfg.timesum, cfg.timestep)
    ctx.set_all_nodal(cfg.nodal)
    for lam0 in (0.01, 0.5, float("nan")):
        p = TimestepParams.default(maxtmstep=DT)
        p.lambda0 = lam0
        r = ctx.calc_timestep(0, p)
        rc, o = Oracle(cfg).calc_timestep(0, p)
        assert rc == 0
        assert _close(r["dtL"], o["dtL"]), (lam0, r["dtL"], o["dtL"])
    ctx.close()


def test_timestep_unknown_pmode_is_an_error(gpu):
    cfg = _large(seed=44)
    cfg.params.pmode = 2
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=1, lib=gpu)
    ctx.set_all_nodal(cfg.nodal)
    with pytest.raises(V2CError, match="pmode"):
        ctx.calc_timestep(0)
    assert Oracle(cfg).calc_timestep(0)[0] == -2
    ctx.close()


def _heat_case(seed=51, per=4):
    cfg = synth.config_grid(401, 101, 400e3, 100e3, per, refine=True, seed=seed)
    rng = np.random.default_rng(seed)
    nn, m = cfg.grid.nodenum, cfg.markers
    cfg.nodal["tk"] = rng.uniform(300, 1600, nn)
    cfg.nodal["tk2"] = cfg.nodal["tk"] + rng.normal(0, 5, nn)
    cfg.nodal["kt"] = rng.uniform(1.5, 3.5, nn)
    cfg.nodal["cp"] = rng.uniform(800, 1300, nn)
    cfg.nodal["ro"] = rng.uniform(2600, 3400, nn)
    m["k"][:] = rng.uniform(300, 1600, m.n)
    k = rng.permutation(m.n)
    m["k"][k[:500]] = -1.0                 # newly incoming
    m.t[k[500:3000]] = rng.integers(50, 100, 2500).astype(np.int8)   # fluid
    m.t[k[3000:3500]] = 110                # immobile kinds: untouched
    m["x"][k[3500:3600]] = -1.0            # outside
    m["y"][k[3600:3700]] = cfg.grid.ysize + 1.0
    # 300 markers in one cell: its segment spans several 64-marker batches
    gx, gy = cfg.grid.gx, cfg.grid.gy
    m["x"][k[3700:4000]] = rng.uniform(gx[200], gx[201], 300)
    m["y"][k[3700:4000]] = rng.uniform(gy[50], gy[51], 300)
    return cfg


@pytest.mark.parametrize("heatdif", [0.0, 1e-3, 1.0, 1e6])
def test_marker_temperatures_match_oracle(gpu, heatdif):
    cfg = _heat_case()
    print("case heat", heatdif, cfg.n, flush=True)
    gm, gn, _, ctx = run_gpu(gpu, cfg, [f"heat:{heatdif}"])
    om, on, _ = run_oracle(cfg, [f"heat:{heatdif}"])
    bad = compare(gm, om, gn, on)
    assert not bad, bad
    ctx.close()


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "bc_*.npz"))),
                         ids=lambda p: os.path.basename(p)[:-4])
def test_boundary_tail_matches_reference_golden(gpu, path):
    from tests.test_oracle_golden import load_bc_fixture
    cfg, bc, z = load_bc_fixture(path)
    print("case", os.path.basename(path), flush=True)
    for which, op, fields in (("vxvy", "bcvxvy", ("mvx", "mvy")), ("P", "bcp", ("sppe",)), ("tk", "bctk", ("tk",))):
        ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=1, lib=gpu)
        ctx.set_all_nodal(cfg.nodal)
        ctx.set_boundary(*bc)
        {"vxvy": ctx.vxvy_boundarycorrection, "P": ctx.P_boundarycorrection, "tk": ctx.tkrecalc}[which]()
        for k in fields:
            got, exp = ctx.get_nodal(k), z[f"out_{op}_{k}"]
            scale = max(np.max(np.abs(exp)), 1e-300)
            assert np.max(np.abs(got - exp)) <= 1e-12 * scale, (which, k)
        ctx.close()


def test_boundary_tail_matches_oracle_large(gpu):
    cfg = synth.config_grid(401, 101, 400e3, 100e3, 1, refine=True, seed=61)
    rng = np.random.default_rng(61)
    nn = cfg.grid.nodenum
    for k, s in (("mvx", 1e-9), ("mvy", 1e-9), ("sppe", 1e6)):
        cfg.nodal[k] = rng.normal(0, s, nn)
    cfg.nodal["tk"] = rng.uniform(273, 1600, nn)
    bc = synth.boundary_equations(cfg, 61)
    print("case bc large", nn, len(bc[1]), flush=True)
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=1, lib=gpu)
    ctx.set_all_nodal(cfg.nodal)
    ctx.set_boundary(*bc)
    ctx.vxvy_boundarycorrection()
    ctx.P_boundarycorrection()
    ctx.tkrecalc()
    o = Oracle(cfg)
    for which in ("vxvy", "P", "tk"):
        assert o.boundary(which, *bc) == 0
    for k in ("mvx", "mvy", "sppe", "tk"):
        got, exp = ctx.get_nodal(k), o.nodal[k]
        assert np.max(np.abs(got - exp)) <= 1e-12 * np.max(np.abs(exp)), k
    ctx.close()


def test_boundary_tail_errors(gpu):
    cfg = synth.config_grid(25, 15, 24e3, 14e3, 1, refine=True, seed=62)
    ny = cfg.grid.ynumy
    bm, bn, bv = synth.boundary_equations(cfg, 62, extra=False)
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=1, lib=gpu)
    ctx.set_all_nodal(cfg.nodal)
    with pytest.raises(V2CError, match="no boundary"):
        ctx.tkrecalc()
    # a cycle that never settles: vx(3, 4) = 1 + vx(3, 5) and vx(3, 5) = 1 + vx(3, 4)
    a, b = 3 * ny + 4, 3 * ny + 5
    bn2 = np.vstack([bn, [[3 * b + 1 + 1, 0, 0], [3 * a + 1 + 1, 0, 0]]])
    bv2 = np.vstack([bv, [[1.0, 1.0, 0.0, 0.0], [1.0, 1.0, 0.0, 0.0]]])
    bm2 = bm.copy()
    bm2[3 * a + 1], bm2[3 * b + 1] = len(bn2) - 2, len(bn2) - 1
    ctx.set_boundary(bm2, bn2, bv2)
    with pytest.raises(V2CError, match="cycle"):
        ctx.vxvy_boundarycorrection()
    # a row tied to the vx of (2, ny-1), which has no row (the reference exits)
    bn3 = np.vstack([bn, [[3 * (2 * ny + ny - 1) + 1 + 1, 0, 0]]])
    bv3 = np.vstack([bv, [[0.0, 1.0, 0.0, 0.0]]])
    bm3 = bm.copy()
    bm3[3 * (3 * ny + 4) + 1] = len(bn3) - 1
    ctx.set_boundary(bm3, bn3, bv3)
    with pytest.raises(V2CError, match="absent"):
        ctx.vxvy_boundarycorrection()
    assert Oracle(cfg).boundary("vxvy", bm3, bn3, bv3) == -1
    ctx.close()
