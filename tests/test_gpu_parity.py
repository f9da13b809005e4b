"""HIP path (libv2c through the C-ABI) against the FP64 oracle.

Bar: marker positions and types bit-exact; every FP64 marker and nodal field
within 1e-12 of the oracle, scaled by max(|oracle|, rms(field)) (tests/util.py).
"""
import numpy as np
import pytest

from velo2cycles_amd import synth
from velo2cycles_amd.abi import NODAL_FIELDS, default_params
from tests.util import compare, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

DT = synth.YEAR
STEP = ["m2n", f"before:{DT}:0", f"after:{DT}:0", "n2m", "move"]


def _check(lib, cfg, ops, gops=None, skip=()):
    om, on, oret = run_oracle(cfg, ops)
    gm, gn, gret, ctx = run_gpu(lib, cfg, gops or ops)
    ctx.close()
    bad = compare(gm, om, gn, on, skip=skip)
    assert not bad, f"{cfg.name} {ops}:\n" + "\n".join(bad)
    for k in oret:
        if k == "before":
            assert gret[k] == oret[k]
        else:
            d_o, e_o = oret[k]
            d_g, e_g = gret[k]
            assert e_g == e_o
            if np.isnan(d_o):       # e.g. emptied cells: 0/0 in the misfit on both sides
                assert np.isnan(d_g)
            else:
                assert abs(d_g - d_o) <= 1e-12 * max(abs(d_o), 1e-300)


@pytest.mark.parametrize("op", ["m2n", "n2m", "move", f"before:{DT}:0", f"after:{DT}:0"])
def test_edge_single_ops(gpu, op):
    _check(gpu, synth.config_edge(), [op])


@pytest.mark.parametrize("timesum", [0.0, DT])
def test_edge_full_step(gpu, timesum):
    cfg = synth.config_edge()
    cfg.timesum = timesum
    _check(gpu, cfg, STEP)


@pytest.mark.parametrize("antidiffinter", [1, 0])
def test_edge_fused_n2m_advect(gpu, antidiffinter):
    """The fused update gathers 9 nodal interpolants with antidiffusion and
    18 without; both against the oracle."""
    cfg = synth.config_edge()
    cfg.timesum = DT
    cfg.params.antidiffinter = antidiffinter
    ops = ["m2n", f"before:{DT}:0", f"after:{DT}:0", "n2m_move"]
    _check(gpu, cfg, ops)


@pytest.mark.parametrize("flags", [
    dict(markmod=1), dict(markmod=0), dict(stoksmod=2), dict(outgrid=0), dict(outgrid=1),
    dict(densimod=1), dict(antidiffinter=0), dict(RSFmode=0), dict(power_law_RSF=1),
    dict(inertyn=1), dict(pmode=1), dict(compressiblemode=1), dict(Dphis=1.0), dict(tempmod=1),
    dict(n2m_phantom=0), dict(stickymod=0)])
def test_edge_flag_variants(gpu, flags):
    cfg = synth.config_edge()
    cfg.timesum = DT
    for k, v in flags.items():
        setattr(cfg.params, k, v)
    _check(gpu, cfg, STEP)


@pytest.mark.parametrize("rkmod,rk", [(1, 1), (2, 1), (3, 1), (3, 2), (3, 3)])
def test_edge_rk_variants(gpu, rkmod, rk):
    cfg = synth.config_edge()
    cfg.timesum = DT
    cfg.params.rkmod = rkmod
    _check(gpu, cfg, ["m2n", f"before:{DT / 2}:{rk}", f"after:{DT / 2}:{rk}"])


def _clustered():
    """Edge cloud with sub-cells holding several 64-marker batches (runs that
    are carried from batch to batch), runs of exactly 64 and 65, and the
    emptied sub-cells they were taken from."""
    cfg = synth.config_edge()
    cfg.timesum = DT
    g = cfg.grid
    rng = np.random.default_rng(5)
    x, y = cfg.markers["x"], cfg.markers["y"]
    idx = rng.permutation(cfg.n)
    j = 0
    for cnt, cx, cy, qx, qy in ((300, 3, 2, 0, 0), (130, 3, 3, 1, 1), (65, 4, 2, 0, 1), (64, 5, 5, 1, 0)):
        sel = idx[j:j + cnt]
        j += cnt
        x[sel] = g.gx[cx] + (g.gx[cx + 1] - g.gx[cx]) * (0.5 * qx + rng.uniform(0.05, 0.45, cnt))
        y[sel] = g.gy[cy] + (g.gy[cy + 1] - g.gy[cy]) * (0.5 * qy + rng.uniform(0.05, 0.45, cnt))
    return cfg


def test_clustered_runs_full_step(gpu):
    _check(gpu, _clustered(), STEP)


def test_clustered_runs_two_steps(gpu):
    cfg = _clustered()
    ops = ["m2n", f"before:{DT}:0", f"after:{DT}:0", "n2m_move", f"time:{2 * DT}:{DT}", "m2n",
           f"before:{DT}:0", f"after:{DT}:0", "n2m_move"]
    _check(gpu, cfg, ops)


def test_medium_refined_two_steps(gpu):
    cfg = synth.config_grid(201, 101, 200e3, 100e3, 4, refine=True)
    ops = STEP + [f"time:{DT}:{DT}"] + STEP
    _check(gpu, cfg, ops)


def test_small_config_step(gpu):
    """401x101 nodes, 16 markers per cell (640,000 markers), one full step."""
    cfg = synth.config_small()
    cfg.timesum = DT
    _check(gpu, cfg, STEP)


def test_determinism(gpu):
    cfg = synth.config_grid(101, 51, 100e3, 50e3, 5, refine=True)
    cfg.timesum = DT
    ops = ["m2n", f"before:{DT}:0", f"after:{DT}:0", "n2m_move", "m2n"]
    m1, n1, _, c1 = run_gpu(gpu, cfg, ops)
    c1.close()
    m2, n2, _, c2 = run_gpu(gpu, cfg, ops)
    c2.close()
    for f in m1.arrays:
        assert np.array_equal(m1[f], m2[f], equal_nan=True), f
    for k in NODAL_FIELDS:
        assert np.array_equal(n1[k], n2[k], equal_nan=True), k


def test_ductilevisc_batch(gpu):
    cfg = synth.config_edge()
    from oracle.pyoracle import Oracle
    from velo2cycles_amd.abi import Context
    rng = np.random.default_rng(7)
    n = 4000
    t = rng.integers(0, 9, n).astype(np.int8)
    sxy = rng.normal(0, 3e7, n)
    sxx = rng.normal(0, 3e7, n)
    sxx[:50] = 0.0
    sxy[:50] = 0.0
    tk = rng.uniform(273, 1900, n)
    pb = rng.uniform(1, 4e4, n)
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=1, lib=gpu)
    ctx.set_time(cfg.timesum, cfg.timestep)     # G*dt enters the iteration
    nu, it = ctx.ductilevisc_batch(t, sxy, sxx, np.zeros(n), np.zeros(n), tk, pb)
    ctx.close()
    o = Oracle(cfg)
    ref = np.array([o.ductilevisc(t[i], sxy[i], sxx[i], tk[i], pb[i]) for i in range(n)])
    assert np.array_equal(it, ref[:, 1].astype(np.int32))
    err = np.abs(nu - ref[:, 0]) / np.abs(ref[:, 0])
    assert err.max() <= 1e-12


def test_unsupported_immobile_spawning(gpu):
    """Types >= 100 (immobile-marker spawning, moveandrotatemarkers.c:90-136)
    return an error code instead of corrupting state."""
    from velo2cycles_amd.abi import Context, V2CError
    cfg = synth.config_edge()
    cfg.markers.t[5] = 101
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=cfg.n, lib=gpu)
    ctx.set_all_nodal(cfg.nodal)
    ctx.upload_markers(cfg.markers)
    with pytest.raises(V2CError, match="EUNSUPPORTED"):
        ctx.moveandrotatemarkers()
    ctx.close()


def test_capacity_error(gpu):
    from velo2cycles_amd.abi import Context, V2CError
    cfg = synth.config_edge()
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=10, lib=gpu)
    with pytest.raises(V2CError, match="ECAPACITY"):
        ctx.upload_markers(cfg.markers)
    ctx.close()


def test_production_full_step_parity(gpu):
    """The production workload itself: 1501x301 nodes refined to 500 m, 36
    markers per cell (16.2 M markers), one full step (marker2nodes,
    compute_before_solve, compute_after_solve, nodes2markers + advection)
    against the oracle on every marker: positions and types bit-exact, every
    FP64 marker and nodal field at the 1e-12 bar; then a second device run of
    marker2nodes gives the same bits."""
    from velo2cycles_amd.abi import Context
    cfg = synth.config_production()
    cfg.timesum = DT
    print("case production full step", cfg.n, flush=True)
    ops = ["m2n", f"before:{DT}:0", f"after:{DT}:0", "n2m", "move"]
    om, on, oret = run_oracle(cfg, ops)
    gm, gn, gret, ctx = run_gpu(gpu, cfg, ["m2n", f"before:{DT}:0", f"after:{DT}:0", "n2m_move"])
    ctx.close()
    bad = compare(gm, om, gn, on)
    assert not bad, "\n".join(bad)
    assert gret["before"] == oret["before"]
    (d_g, e_g), (d_o, e_o) = gret["after"], oret["after"]
    assert e_g == e_o and abs(d_g - d_o) <= 1e-12 * abs(d_o)
    del om, on, gm
    # determinism at full size
    ctx2 = Context(cfg.grid, cfg.params, cfg.mats, capacity=cfg.n, lib=gpu)
    ctx2.set_time(cfg.timesum, cfg.timestep)
    ctx2.set_all_nodal(cfg.nodal)
    ctx2.upload_markers(cfg.markers)
    ctx2.marker2nodes()
    ctx3 = Context(cfg.grid, cfg.params, cfg.mats, capacity=cfg.n, lib=gpu)
    ctx3.set_time(cfg.timesum, cfg.timestep)
    ctx3.set_all_nodal(cfg.nodal)
    ctx3.upload_markers(cfg.markers)
    ctx3.marker2nodes()
    for k in NODAL_FIELDS:
        assert np.array_equal(ctx2.get_nodal(k), ctx3.get_nodal(k), equal_nan=True), k
    ctx2.close()
    ctx3.close()


def _oracle_keys(cfg, x, y):
    """Sub-cell key of every marker from the oracle's nxsearch / nysearch
    (the reference bisection, interpolation.c:1536-1573) at the position
    clamped to the domain: cell nx*(ynumy-1)+ny, times 4, plus the quadrant
    (2 if past the cell's x midline, 1 if past its y midline)."""
    from oracle.pyoracle import Oracle
    g = cfg.grid
    xc = np.clip(x, 0.0, g.xsize)
    yc = np.clip(y, 0.0, g.ysize)
    nx, ny = Oracle(cfg).cells(xc, yc)
    cx = (xc - g.gx[nx]) / (g.gx[nx + 1] - g.gx[nx])
    cy = (yc - g.gy[ny]) / (g.gy[ny + 1] - g.gy[ny])
    q = np.where(cx > 0.5, 2, 0) + np.where(cy > 0.5, 1, 0)
    return ((nx * (g.ynumy - 1) + ny) * 4 + q).astype(np.int64)


@pytest.mark.parametrize("case", ["edge", "clustered", "refined", "small"])
def test_subcell_keys_match_oracle_cells(gpu, case):
    """The device's cell of every marker (its table-driven search, kept as the
    sort key) equals the oracle's bisection cell, before and after an
    advection step; the edge cloud puts markers on gridlines, midlines, the
    domain boundary and outside it."""
    from velo2cycles_amd.abi import Context
    cfg = {"edge": synth.config_edge, "clustered": _clustered,
           "refined": lambda: synth.config_grid(201, 101, 200e3, 100e3, 4, refine=True, seed=41),
           "small": synth.config_small}[case]()
    cfg.timesum = DT
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=cfg.n, lib=gpu)
    ctx.set_time(cfg.timesum, cfg.timestep)
    ctx.set_all_nodal(cfg.nodal)
    m = cfg.markers.copy()
    ctx.upload_markers(m)
    for stage in ("uploaded", "advected"):
        order, skey = ctx.order_get()
        key = np.empty(cfg.n, np.int64)
        key[order.astype(np.int64)] = skey
        ctx.download_markers(m)
        ref = _oracle_keys(cfg, m["x"], m["y"])
        bad = np.nonzero(key != ref)[0]
        assert bad.size == 0, f"{stage}: {bad.size} keys differ, first marker {bad[:5]}"
        assert np.all(np.diff(skey.astype(np.int64)) >= 0)
        if stage == "uploaded":
            ctx.marker2nodes()
            ctx.nodes2markers_advect()
    ctx.close()


@pytest.mark.parametrize("dt", [1e9, 1e10])
def test_incremental_order_equals_full_sort(gpu, dt):
    """After advection, the incremental merge of movers into the kept order
    gives exactly the (key, id) order a full radix sort of the new keys gives."""
    from velo2cycles_amd.abi import Context
    cfg = synth.config_small()
    cfg.timesum = DT
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=cfg.n, lib=gpu)
    ctx.set_time(cfg.timesum, dt)
    ctx.set_all_nodal(cfg.nodal)
    ctx.upload_markers(cfg.markers)
    ctx.order_get()                       # full sort of the initial keys
    ctx.nodes2markers_advect()
    o_inc, k_inc = ctx.order_get()        # incremental merge
    st = ctx.order_stats()
    assert st["incremental"] == 1 and st["full_sorts"] == 1 and 0 < st["last_movers"] <= cfg.n // 4
    m = cfg.markers.copy()
    ctx.download_markers(m)
    ctx.close()
    ctx2 = Context(cfg.grid, cfg.params, cfg.mats, capacity=cfg.n, lib=gpu)
    ctx2.upload_markers(m)
    o_full, k_full = ctx2.order_get()
    ctx2.close()
    assert np.array_equal(k_inc, k_full)
    assert np.array_equal(o_inc, o_full)
    assert np.all(np.diff(k_full.astype(np.int64)) >= 0)


def _coseismic_flow(seed=83):
    """201x101 refined grid, 16 markers per cell, the corner flow scaled from
    5 cm/yr to ~5 m/s (velocities and strain rates x 3e9), slip rates up to
    10 m/s on the nodes."""
    cfg = synth.config_grid(201, 101, 200e3, 100e3, 4, refine=True, seed=seed)
    s = 3e9
    for k in ("vx", "vy", "mvx", "mvy", "exx", "exy", "esp", "exxpl"):
        cfg.nodal[k] = cfg.nodal[k] * s
    rng = np.random.default_rng(seed)
    nn = cfg.grid.nodenum
    cfg.nodal["RK_sliprs_final"] = 10.0 ** rng.uniform(-3.0, 1.0, nn)
    cfg.nodal["RK_sliprs_solved"] = cfg.nodal["RK_sliprs_final"] * np.exp(rng.uniform(-0.5, 0.5, nn))
    return cfg


@pytest.mark.parametrize("markmod,stoksmod", [(2, 1), (0, 1), (1, 2)])
def test_coseismic_rates_adaptive_steps(gpu, markmod, stoksmod):
    """Four steps at coseismic rates with dt growing from 1 ms to 1000 s:
    displacements up to ~10 cells per step, rotation angles up to ~0.5 rad,
    slip rates 1e-3..10 m/s; positions and types bit-exact after the last."""
    cfg = _coseismic_flow()
    cfg.params.markmod = markmod
    cfg.params.stoksmod = stoksmod
    ops, ts = [], 0.0
    for dt in (1e-3, 1e-1, 10.0, 1e3):
        ops += [f"time:{ts}:{dt}", "m2n", f"before:{dt}:0", f"after:{dt}:0", "n2m_move"]
        ts += dt
    print("case coseismic steps", markmod, stoksmod, cfg.n, flush=True)
    _check(gpu, cfg, ops)


def test_async_and_timed_runs_match_sync(gpu):
    """Two full steps queued without host waits (asynchronous mode, with the
    per-kernel timing events on both of the context's streams) give the same
    bits as the synchronous calls: the side-stream work of marker2nodes and
    nodes2markers is joined before anything that reads it."""
    cfg = synth.config_grid(201, 101, 200e3, 100e3, 4, refine=True, seed=29)
    cfg.timesum = DT
    ops = ["m2n", f"before:{DT}:0", f"after:{DT}:0", "n2m", "move", f"time:{2 * DT}:{DT}",
           "m2n", f"before:{DT}:0", f"after:{DT}:0", "n2m_move"]
    ms, ns, _, cs = run_gpu(gpu, cfg, ops)
    cs.close()
    from velo2cycles_amd.abi import Context
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=cfg.n, lib=gpu)
    ctx.set_async(True)
    ctx.enable_timing(True)
    ma, na, _, ca = run_gpu(gpu, cfg, ops, ctx=ctx)
    ca.synchronize()
    ma2 = ma.copy()
    ca.download_markers(ma2)
    for f in ms.arrays:
        assert np.array_equal(ms[f], ma2[f], equal_nan=True), f
    assert np.array_equal(ms.t, ma2.t)
    for k in NODAL_FIELDS:
        assert np.array_equal(ns[k], ca.get_nodal(k), equal_nan=True), k
    ms_props, n_props = ca.kernel_time("m2n_props")
    assert n_props == 2 and ms_props > 0
    ca.close()


def test_many_field_transfers_roundtrip(gpu):
    """v2c_nodal_set_many / get_many (staged, 24 fields per copy): 30 fields
    in, read back one by one and all at once, bit for bit; fields not named
    are left alone."""
    from velo2cycles_amd.abi import Context
    cfg = synth.config_grid(61, 31, 60e3, 30e3, 2, seed=3)
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=1, lib=gpu)
    nn = cfg.grid.nodenum
    rng = np.random.default_rng(3)
    names = list(NODAL_FIELDS[:30])
    untouched = NODAL_FIELDS[30]
    ctx.set_nodal(untouched, np.full(nn, 7.0))
    vals = {k: rng.standard_normal(nn) for k in names}
    print("case many-field transfers", len(names), flush=True)
    ctx.set_nodal_many(vals)
    for k in names:
        assert np.array_equal(ctx.get_nodal(k), vals[k]), k
    got = ctx.get_nodal_many(names[::-1])
    for k in names:
        assert np.array_equal(got[k], vals[k]), k
    assert np.array_equal(ctx.get_nodal(untouched), np.full(nn, 7.0))
    ctx.close()
