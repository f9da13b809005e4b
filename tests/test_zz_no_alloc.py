"""v2c_ductilevisc_batch_noalloc: the allocation-free variant of
v2c_ductilevisc_batch (static device staging of at most
V2C_DUCTILEVISC_NOALLOC_MAX inputs).  Two calls in a row through the C-ABI,
the second smaller, each bit-identical to v2c_ductilevisc_batch on the same
inputs (viscosities and iteration counts); a size above the limit is
refused on the host with V2C_EINVAL before anything is launched."""
import numpy as np
import pytest

from velo2cycles_amd import synth
from velo2cycles_amd.abi import NOALLOC_MAX, Context, V2CError

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    t = rng.integers(0, 9, n).astype(np.int8)
    sxy = rng.normal(0, 3e7, n)
    sxx = rng.normal(0, 3e7, n)
    sxx[:50] = 0.0
    sxy[:50] = 0.0
    tk = rng.uniform(273, 1900, n)
    pb = rng.uniform(1, 4e4, n)
    z = np.zeros(n)
    return t, sxy, sxx, z, z, tk, pb


def test_noalloc_equals_allocating_entry_point(gpu):
    cfg = synth.config_edge()
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=1, lib=gpu)
    ctx.set_time(cfg.timesum, cfg.timestep)
    for n, seed in ((NOALLOC_MAX, 1), (5000, 2)):
        print("case noalloc ductilevisc_batch n =", n, flush=True)
        args = _inputs(n, seed)
        nu_a, it_a = ctx.ductilevisc_batch(*args)
        nu_b, it_b = ctx.ductilevisc_batch(*args, noalloc=True)
        assert np.array_equal(nu_a, nu_b, equal_nan=True), n
        assert np.array_equal(it_a, it_b), n
        assert np.isfinite(nu_b).all() and (it_b > 0).any()
    with pytest.raises(V2CError, match="static staging limit"):
        ctx.ductilevisc_batch(*_inputs(NOALLOC_MAX + 1, 3), noalloc=True)
    # a second context on the same device shares the staging
    ctx2 = Context(cfg.grid, cfg.params, cfg.mats, capacity=1, lib=gpu)
    ctx2.set_time(cfg.timesum, cfg.timestep)
    args = _inputs(777, 4)
    assert np.array_equal(ctx2.ductilevisc_batch(*args, noalloc=True)[0], ctx2.ductilevisc_batch(*args)[0])
    ctx2.close()
    ctx.close()
