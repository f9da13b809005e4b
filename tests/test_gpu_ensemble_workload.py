"""The batched-ensemble workload itself: 64 independent models with grids of
201x51 to 801x201 nodes and 4 to 36 markers per cell (89.8 M markers), a-b
and convergence rate drawn per model (synth.config_ensemble).

Two graph-replayed batched steps must equal, bit for bit, the same two steps
through the per-model entry points, for every model; the per-model path is in
turn checked against the oracle: marker-local results of nodes2markers +
advection on a seeded 100,000-marker subset of every model (each marker's
update depends only on its own state and the nodal fields, which are taken
from the device after that step's marker2nodes and solves), and every marker
and nodal field of the three smallest models over the full step (the
reference step i2evps.c:89-113)."""
import numpy as np
import pytest

from velo2cycles_amd import synth
from velo2cycles_amd.abi import NODAL_FIELDS, Context, Ensemble, MarkerArrays
from tests.util import compare, run_oracle

pytestmark = pytest.mark.gpu

SUBSET = 100_000
# marker fields nodes2markers + advection write (phi also takes marker2nodes'
# friction reset, which the subset replay does not repeat)
FIELDS_N2M = ("xx", "xy", "exx", "exy", "p", "vx", "vy", "e", "k")


def _ctx(lib, cfg):
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=cfg.n, lib=lib)
    ctx.set_time(cfg.timesum, cfg.timestep)
    ctx.set_all_nodal(cfg.nodal)
    ctx.upload_markers(cfg.markers)
    return ctx


def _subset(m, idx):
    f = {k: m[k][idx] for k in m.arrays}
    f["t"] = m.t[idx]
    return MarkerArrays(len(idx), f)


def _check_subset(cfg, before, nodal, after, idx, what):
    """Oracle nodes2markers + advection of the markers `idx` of state
    `before`, on the device's nodal fields of that step, against `after`."""
    sub = _subset(before, idx)
    prm = type(cfg.params).from_buffer_copy(cfg.params)
    prm.n2m_phantom = 0               # the phantom marker only feeds the node sums
    c = synth.Config(cfg.name, cfg.grid, prm, cfg.mats, sub, nodal, cfg.timestep, cfg.timesum)
    om, _, _ = run_oracle(c, ["n2m", "move"])
    for f in ("x", "y"):
        assert np.array_equal(after[f][idx], om[f]), (what, f)
    assert np.array_equal(after.t[idx], om.t), what
    for f in FIELDS_N2M:
        a, b = after[f][idx], om[f]
        scale = np.maximum(np.abs(b), np.sqrt(np.mean(b * b)))
        scale[scale == 0] = 1.0
        err = np.max(np.abs(a - b) / scale)
        assert err <= 1e-12, (what, f, err)


@pytest.mark.timeout(900)
def test_ensemble_64_models_workload(gpu):
    dt = synth.YEAR
    batch, single, cfgs_small, subsets, sizes = [], [], {}, [], []
    for k, cfg in synth.config_ensemble(64):
        cfg.timesum = dt
        batch.append(_ctx(gpu, cfg))
        single.append(_ctx(gpu, cfg))
        sizes.append(cfg.n)
        rng = np.random.default_rng(100 + k)
        idx = np.sort(rng.choice(cfg.n, min(SUBSET, cfg.n), replace=False))
        subsets.append((cfg, idx))
        if k % 8 == 0:
            print(f"model {k}: {cfg.grid.xnumx}x{cfg.grid.ynumy} nodes, {cfg.n} markers", flush=True)
    total = sum(sizes)
    print(f"64 models, {total} markers", flush=True)
    assert total > 80_000_000
    smallest = set(np.argsort(sizes)[:3].tolist())
    ens = Ensemble(batch)
    for _ in range(2):
        ens.step(graph=True)
    print("batched steps done", flush=True)
    for k, (cfg, idx) in enumerate(subsets):
        ctx = single[k]
        state = cfg.markers.copy()
        for s in range(2):
            ts = cfg.timesum + s * cfg.timestep
            ctx.set_time(ts, cfg.timestep)
            ctx.marker2nodes()
            ctx.compute_before_solve(cfg.timestep, 0)
            ctx.compute_after_solve(cfg.timestep, 0)
            nod = ctx.get_all_nodal()
            ctx.nodes2markers_advect()
            after = cfg.markers.copy()
            ctx.download_markers(after)
            c_step = synth.Config(cfg.name, cfg.grid, cfg.params, cfg.mats, state, cfg.nodal, cfg.timestep, ts)
            _check_subset(c_step, state, nod, after, idx, f"model {k} step {s}")
            state = after
        if k in smallest:
            # every field of the full two steps against the oracle
            ops = []
            for s in range(2):
                ts = cfg.timesum + s * cfg.timestep
                ops += [f"time:{ts}:{cfg.timestep}", "m2n", f"before:{cfg.timestep}:0",
                        f"after:{cfg.timestep}:0", "n2m", "move"]
            om, on, _ = run_oracle(cfg, ops)
            # the batched step, like v2c_marker2nodes, leaves out the
            # sticky-air marks that the oracle's m2n op adds
            bad = compare(state, om, ctx.get_all_nodal(), on, skip=("sticky01",))
            assert not bad, f"model {k}:\n" + "\n".join(bad)
        # batched == per-model, bit for bit
        mb = cfg.markers.copy()
        batch[k].download_markers(mb)
        for f in mb.arrays:
            assert np.array_equal(mb[f], state[f], equal_nan=True), (k, f)
        assert np.array_equal(mb.t, state.t), k
        nb, ns = batch[k].get_all_nodal(), ctx.get_all_nodal()
        for f in NODAL_FIELDS:
            assert np.array_equal(nb[f], ns[f], equal_nan=True), (k, f)
        ctx.close()
        single[k] = None
        subsets[k] = None
        if k % 8 == 0:
            print(f"model {k} checked", flush=True)
    ens.close()
    for c in batch:
        c.close()
