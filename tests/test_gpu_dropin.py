"""The reference-side binding (velo2cycles_amd/csrc/dropin/v2c_dropin.c) linked
into the reference's own routines: oracle/_ref/dropin_driver is the reference
driver with the hot-path routines renamed and the binding's definitions live,
so each op below runs exactly as the reference's call sites would run it
through libv2c on the GPU.  Its outputs are compared with the plain reference
driver (oracle/_ref/ref_driver) on the same input.

Both drivers are built here from the reference sources (make -C oracle) and
carry no sources at run time; without them the tests skip.  Bar as
tests/util.py: positions and types bit-exact, FP64 within 1e-12 scaled."""
import math
import os

import numpy as np
import pytest

from oracle import refrun
from velo2cycles_amd import synth
from velo2cycles_amd.abi import MarkerArrays
from tests.util import compare, load_post_fixture, run_oracle

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (refrun.built(refrun.DRIVER) and refrun.built(refrun.DROPIN)),
                                 reason="oracle/_ref drivers not built (make -C oracle with the reference sources)")]

DT = synth.YEAR
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
STEP = ["m2n", f"before:{DT}:0", f"after:{DT}:0", "n2m", "move"]
SCALARS = ("maxvel", "maxdphi", "maxnvp", "dtL", "dtphi", "dtd", "dtnvp")
# nodes2markers' per-node sums (interpolation.c:934-1048, mrx0 / mry0 among
# them): marker2nodes clears them (interpolation.c:76-138, 793) and
# nodes2markers accumulates into them and reads them; nothing else does, so
# the binding leaves them on the device
N2M_SCRATCH = tuple("wtphisum phisum sxysum exysum esum wtsxxsum psum sxxsum exxsum wtvxsum vxsum "
                    "wtvysum vysum mrx0 mry0".split())


def _bond_extra(bc):
    bm, bn, bv = bc
    return {"bondm": ("l", bm), "bondn": ("l", bn.ravel()), "bondv": ("d", bv.ravel()),
            "bondnum": ("l", np.array([len(bn)], np.int64))}


def _both(cfg, ops, extra):
    print("case dropin", cfg.name, ops, flush=True)
    return (refrun.run(cfg, ops, extra=extra),
            refrun.run(cfg, ops, extra=extra, driver=refrun.DROPIN))


def _close(a, b, rel=1e-12):
    if math.isnan(a) or math.isnan(b):
        return math.isnan(a) and math.isnan(b)
    return a == b or abs(a - b) <= rel * max(abs(a), abs(b))


@pytest.mark.parametrize("timesum,flags", [(0.0, {}), (DT, {}), (DT, dict(tempmod=1)),
                                           (DT, dict(markmod=1, pmode=1))])
def test_step_through_the_binding(gpu, timesum, flags):
    cfg = synth.config_edge()
    cfg.timesum = timesum
    for k, v in flags.items():
        setattr(cfg.params, k, v)
    extra = _bond_extra(synth.boundary_equations(cfg, 5))
    (rm, rn, rraw), (dm, dn, draw) = _both(cfg, STEP, extra)
    bad = compare(MarkerArrays(cfg.n, dm), MarkerArrays(cfg.n, rm), dn, rn, skip=N2M_SCRATCH)
    assert not bad, "\n".join(bad)
    assert _close(draw["drv_ret"][1], rraw["drv_ret"][1])        # compute_after_solve's misfit
    assert draw["errcheck"][0] == rraw["errcheck"][0]


@pytest.mark.parametrize("name", ["post_rsf_p0", "post_rsf_p1", "post_dp"])
@pytest.mark.parametrize("oldornew", [0, 1])
def test_post_routines_through_the_binding(gpu, name, oldornew):
    cfg, _, maxtmstep = load_post_fixture(os.path.join(GOLDEN, name + ".npz"))
    ops = [f"stress:{DT}:0", f"c2b:{DT}", f"c2bold:{DT}", "b2c", f"tstep:{oldornew}", "copy:0"]
    extra = {"maxtmstep": ("d", np.array([maxtmstep]))}
    (_, rn, rraw), (_, dn, draw) = _both(cfg, ops, extra)
    bad = compare(None, None, dn, rn)
    assert not bad, "\n".join(bad)
    assert _close(draw["drv_ret"][2], rraw["drv_ret"][2])        # calc_timestep's return
    for k in SCALARS:
        assert _close(draw[k][0], rraw[k][0]), k


@pytest.mark.parametrize("name", ["bc_walls_chains", "bc_compressible"])
def test_boundary_tail_through_the_binding(gpu, name):
    from tests.test_oracle_golden import load_bc_fixture
    cfg, bc, _ = load_bc_fixture(os.path.join(GOLDEN, name + ".npz"))
    (_, rn, _), (_, dn, _) = _both(cfg, ["bcvxvy", "bcp", "bctk"], _bond_extra(bc))
    bad = compare(None, None, dn, rn)
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("heatdif", [0.0, 1.0])
def test_heat_markers_through_the_binding(gpu, heatdif):
    """The binding's marker part of solve_temperature against the oracle (the
    reference has it only inside solve_temperature)."""
    cfg = synth.config_grid(41, 21, 40e3, 20e3, 3, refine=True, seed=57)
    rng = np.random.default_rng(57)
    nn = cfg.grid.nodenum
    cfg.nodal["tk"] = rng.uniform(300, 1600, nn)
    cfg.nodal["tk2"] = cfg.nodal["tk"] + rng.normal(0, 5, nn)
    for k, lo, hi in (("kt", 1.5, 3.5), ("cp", 800, 1300), ("ro", 2600, 3400)):
        cfg.nodal[k] = rng.uniform(lo, hi, nn)
    cfg.markers["k"][:] = rng.uniform(300, 1600, cfg.n)
    cfg.markers["k"][::37] = -1.0
    cfg.markers.t[::11] = 60
    print("case dropin heat", heatdif, flush=True)
    dm, dn, _ = refrun.run(cfg, [f"heat:{heatdif}"], driver=refrun.DROPIN)
    om, on, _ = run_oracle(cfg, [f"heat:{heatdif}"])
    bad = compare(MarkerArrays(cfg.n, dm), om, dn, on)
    assert not bad, "\n".join(bad)
