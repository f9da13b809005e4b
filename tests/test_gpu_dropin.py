"""The reference-side binding (velo2cycles_amd/csrc/dropin/v2c_dropin.c) as a C
host calls it: tests/native/dropin_harness defines the reference globals the
binding declares, reads a driver input file, calls the binding's routines by
the reference names (marker2nodes(), compute_after_solve(dt, rk), ...) and
writes the globals back.  The binding forwards every call to libv2c on the GPU.

Each case replays a fixture that the reference's own routines wrote from the
same input file (tests/golden/dropin_*.npz, made by
tests/golden/make_dropin_golden.py in a container holding the reference
sources); nothing built from the reference runs here.  The loop cases run
host stand-ins (tests/native/host_stand_in.h) where the reference runs host
code, between the binding's hand-overs, so the binding's residency tracking
(what crosses PCIe when) is checked against the reference's plain globals.  The marker part of the
heat step has no stand-alone reference routine and is compared with the
oracle.  Bar as tests/util.py: positions and types bit-exact, FP64 within
1e-12 scaled."""
import glob
import json
import math
import os
import tempfile

import numpy as np
import pytest

from oracle import refrun
from velo2cycles_amd import synth
from velo2cycles_amd.abi import MARKER_FIELDS, NODAL_FIELDS, MarkerArrays
from tests.util import check_deltavs, compare, run_oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
HARNESS = os.path.join(HERE, "native", "dropin_harness")
FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "dropin_*.npz")))

pytestmark = [pytest.mark.gpu]

SCALARS = ("maxvel", "maxdphi", "maxnvp", "dtL", "dtphi", "dtd", "dtnvp",
           # calc_timestep's sticky-air statistics (counts exact: _close of ints)
           "nr_1", "nr_0", "max_sigin_1", "max_sigin_0", "min_sigin_1", "min_sigin_0", "sum_sigin_1",
           "sum2_sigin_1", "sum_sigin_0", "sum2_sigin_0", "max_visc_1", "max_visc_0", "min_visc_1",
           "min_visc_0", "sum_visc_1", "sum2_visc_1", "sum_visc_0", "sum2_visc_0", "sigin_P1", "sigin_P2",
           "visco_P1", "visco_P2")
# nodes2markers' per-node sums (interpolation.c:934-1048, mrx0 / mry0 among
# them): marker2nodes clears them (interpolation.c:76-138, 793) and
# nodes2markers accumulates into them and reads them; nothing else does, so
# the binding leaves them on the device
N2M_SCRATCH = tuple("wtphisum phisum sxysum exysum esum wtsxxsum psum sxxsum exxsum wtvxsum vxsum "
                    "wtvysum vysum mrx0 mry0".split())


@pytest.fixture(scope="module")
def harness():
    if not os.path.exists(HARNESS):
        pytest.fail("tests/native/dropin_harness is not built (run __graft_entry__.build())")
    return HARNESS


def _close(a, b, rel=1e-12):
    if math.isnan(a) or math.isnan(b):
        return math.isnan(a) and math.isnan(b)
    return a == b or abs(a - b) <= rel * max(abs(a), abs(b))


def test_fixtures_present():
    assert len(FIXTURES) == 15, FIXTURES


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[7:-4])
def test_binding_matches_reference_fixture(gpu, harness, path):
    z = np.load(path, allow_pickle=False)
    ops = json.loads(str(z["ops"]))
    n, nn = int(z["n"]), int(z["nodenum"])
    print("case dropin", os.path.basename(path), ops, flush=True)
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        z["in_bin"].tofile(fin)
        raw = refrun.run_file(harness, fin, ops)
    dm, dn = refrun.split_output(raw, n, nn)
    rm = {f: z["out_m_" + f] for f in MARKER_FIELDS}
    rm["t"] = z["out_m_t"]
    rn = {k: z["out_n_" + k] for k in NODAL_FIELDS}
    with_markers = n > 0 and any(o in ("m2n", "n2m", "move") for o in ops)
    skip = N2M_SCRATCH if "n2m" in ops else ()
    copies = [o for o in ops if o.startswith("copy:")]
    if copies:
        skip = skip + ("deltaVs",)
    bad = compare(MarkerArrays(n, dm) if with_markers else None, MarkerArrays(n, rm) if with_markers else None,
                  dn, rn, skip=skip)
    if copies:
        # the one copyarrays of the case reads the input's RK_sliprs_final
        # and the RK_sliprs_solved it leaves
        assert len(copies) == 1
        with tempfile.TemporaryDirectory() as td:
            fin = os.path.join(td, "in.bin")
            z["in_bin"].tofile(fin)
            inp = refrun.read_records(fin)
        final = inp["RK_sliprs_final"][:nn]
        bad += check_deltavs(dn["deltaVs"], rn["deltaVs"], (dn["RK_sliprs_solved"], rn["RK_sliprs_solved"]),
                             (final, final), float(inp["timestep"][0]))
    assert not bad, "\n".join(bad)
    if any(o.startswith("after:") for o in ops):
        assert _close(raw["drv_ret"][1], z["drv_ret"][1])      # compute_after_solve's misfit
    if any(o.startswith("tstep:") for o in ops):
        assert _close(raw["drv_ret"][2], z["drv_ret"][2])      # calc_timestep's return
        for k in SCALARS:
            assert _close(raw[k][0], z["s_" + k][0]), k
    if "hostout" in ops:
        assert _close(raw["drv_ret"][3], z["drv_ret"][3])      # what host output code read
    assert raw["errcheck"][0] == z["errcheck"][0]


@pytest.mark.parametrize("heatdif", [0.0, 1.0])
def test_heat_markers_through_the_binding(gpu, harness, heatdif):
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
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        refrun.write_input(fin, cfg)
        raw = refrun.run_file(harness, fin, [f"heat:{heatdif}"])
    dm, dn = refrun.split_output(raw, cfg.n, nn)
    om, on, _ = run_oracle(cfg, [f"heat:{heatdif}"])
    bad = compare(MarkerArrays(cfg.n, dm), om, dn, on)
    assert not bad, "\n".join(bad)
