"""Batched ensembles: one launch per stage over all models (persistent grid,
HIP graph replay) gives bit for bit the per-model results, whatever the tile
shapes; the autotuner runs on the ensemble and keeps valid shapes."""
import numpy as np
import pytest

from velo2cycles_amd import synth
from velo2cycles_amd.abi import NODAL_FIELDS, Context, Ensemble

pytestmark = pytest.mark.gpu
DT = synth.YEAR


def _models():
    """Uneven models: grid sizes, marker densities, refinement and flags differ."""
    cfgs = [synth.config_grid(61, 31, 60e3, 30e3, 2, refine=False, seed=1),
            synth.config_grid(81, 41, 80e3, 40e3, 3, refine=True, seed=2),
            synth.config_grid(121, 61, 120e3, 60e3, 4, refine=True, seed=3),
            synth.config_edge()]
    cfgs[1].params.stoksmod = 2
    cfgs[2].params.markmod = 1
    for c in cfgs:
        c.timesum = DT
    return cfgs


def _ctx(lib, cfg):
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=cfg.n, lib=lib)
    ctx.set_time(cfg.timesum, cfg.timestep)
    ctx.set_all_nodal(cfg.nodal)
    ctx.upload_markers(cfg.markers)
    return ctx


def _single_step(ctx, cfg, ts):
    ctx.set_time(ts, cfg.timestep)
    ctx.marker2nodes()
    ctx.compute_before_solve(cfg.timestep, 0)
    ctx.compute_after_solve(cfg.timestep, 0)
    ctx.nodes2markers_advect()


def _state(ctx, cfg):
    m = cfg.markers.copy()
    ctx.download_markers(m)
    return m, ctx.get_all_nodal()


def _assert_same(a, b, what):
    (ma, na), (mb, nb) = a, b
    for f in ma.arrays:
        assert np.array_equal(ma[f], mb[f], equal_nan=True), (what, f)
    assert np.array_equal(ma.t, mb.t)
    for k in NODAL_FIELDS:
        assert np.array_equal(na[k], nb[k], equal_nan=True), (what, k)


@pytest.mark.parametrize("graph,branches", [(True, 1), (False, 1), (True, 2), (True, 3), (False, 4)])
def test_ensemble_equals_single_models(gpu, graph, branches):
    """Batched step (graph replay or direct launches; the models split into
    1..4 concurrent branches) == per-model entry points, bit for bit."""
    cfgs = _models()
    singles = [_ctx(gpu, c) for c in cfgs]
    batch = [_ctx(gpu, c) for c in cfgs]
    ens = Ensemble(batch)
    ens.set_branches(branches)
    assert ens.branches() == min(branches, len(cfgs))
    print("case ensemble graph", graph, "branches", branches, flush=True)
    for step in range(2):
        for ctx, cfg in zip(singles, cfgs):
            _single_step(ctx, cfg, cfg.timesum + step * cfg.timestep)
        ens.step(graph=graph)
    for k, cfg in enumerate(cfgs):
        _assert_same(_state(batch[k], cfg), _state(singles[k], cfg), f"model {k}")
    ens.close()
    for c in singles + batch:
        c.close()


def test_ensemble_shapes_do_not_change_results(gpu):
    cfgs = _models()
    a = [_ctx(gpu, c) for c in cfgs]
    b = [_ctx(gpu, c) for c in cfgs]
    ea, eb = Ensemble(a), Ensemble(b)
    for cls in range(3):
        eb.set_shape(cls, 8 if cls != 1 else 37, 256, 256)
    eb.set_blocks_per_cu(2)
    for _ in range(2):
        ea.step()
        eb.step()
    for k, cfg in enumerate(cfgs):
        _assert_same(_state(a[k], cfg), _state(b[k], cfg), f"model {k}")
    assert np.array_equal(ea.rms(), eb.rms(), equal_nan=True)
    ea.close()
    eb.close()
    for c in a + b:
        c.close()


def test_ensemble_looped_baseline_and_autotune(gpu):
    cfgs = _models()
    a = [_ctx(gpu, c) for c in cfgs]
    b = [_ctx(gpu, c) for c in cfgs]
    ea, eb = Ensemble(a), Ensemble(b)
    ea.step()
    eb.step_looped()
    for k, cfg in enumerate(cfgs):
        _assert_same(_state(a[k], cfg), _state(b[k], cfg), f"model {k}")
    ms = ea.autotune(reps=2)
    assert ms > 0
    for cls in range(3):
        cpg, mpt, npt = ea.get_shape(cls)
        assert cpg >= 1 and mpt % 256 == 0 and npt % 256 == 0
    ea.close()
    eb.close()
    for c in a + b:
        c.close()


def test_ensemble_batched_reordering_matches_single_models(gpu):
    """Timesteps from 1 to 3000 years: some models re-order incrementally in
    the batched pass (one mover sort for all), some fall back to a full sort;
    orders, markers and nodes equal the per-model path after every step."""
    cfgs = _models()
    for c, years in zip(cfgs, (1.0, 30.0, 300.0, 3000.0)):
        c.timestep = years * DT
    singles = [_ctx(gpu, c) for c in cfgs]
    batch = [_ctx(gpu, c) for c in cfgs]
    ens = Ensemble(batch)
    for step in range(3):
        for ctx, cfg in zip(singles, cfgs):
            _single_step(ctx, cfg, cfg.timesum + step * cfg.timestep)
        ens.step(graph=True)
        ens.order()                       # the batched pass the next step would run
        for k, cfg in enumerate(cfgs):
            oa, sa = batch[k].order_get()
            ob, sb = singles[k].order_get()
            assert np.array_equal(sa, sb) and np.array_equal(oa, ob), (step, k)
            assert batch[k].order_stats()["last_movers"] == singles[k].order_stats()["last_movers"], (step, k)
    for k, cfg in enumerate(cfgs):
        _assert_same(_state(batch[k], cfg), _state(singles[k], cfg), f"model {k}")
    stats = [c.order_stats() for c in batch]
    assert any(s["incremental"] > 0 for s in stats), stats
    ens.close()
    for c in singles + batch:
        c.close()


def test_batched_and_single_steps_interleave(gpu):
    """Contexts stepped alternately by the batched ensemble and by their own
    entry points (the batched re-ordering recomputes the sub-cell starts from
    the keys; a single-model re-ordering after it rebuilds the per-sub-cell
    counts it keeps, and scans them on the steps after) give, step for step,
    the bits of contexts stepped by their entry points alone."""
    cfgs = _models()
    for c, years in zip(cfgs, (1.0, 10.0, 100.0, 1.0)):
        c.timestep = years * DT
    singles = [_ctx(gpu, c) for c in cfgs]
    mixed = [_ctx(gpu, c) for c in cfgs]
    ens = Ensemble(mixed)
    for step, how in enumerate(("single", "batched", "single", "single", "batched", "batched", "single",
                                "single")):
        for ctx, cfg in zip(singles, cfgs):
            _single_step(ctx, cfg, cfg.timesum + step * cfg.timestep)
        if how == "batched":
            for ctx, cfg in zip(mixed, cfgs):
                ctx.set_time(cfg.timesum + step * cfg.timestep, cfg.timestep)
            ens.step(graph=True)
        else:
            for ctx, cfg in zip(mixed, cfgs):
                _single_step(ctx, cfg, cfg.timesum + step * cfg.timestep)
        for k, cfg in enumerate(cfgs):
            _assert_same(_state(mixed[k], cfg), _state(singles[k], cfg), f"step {step} ({how}) model {k}")
    stats = [c.order_stats() for c in mixed]
    assert any(s["incremental"] > 1 for s in stats), stats
    ens.close()
    for c in singles + mixed:
        c.close()
