"""The oracle against the live in-place reference build (oracle/_ref), at
reference-legal sizes beyond the committed fixtures.  Skipped where the
reference sources are absent."""
import numpy as np
import pytest

from velo2cycles_amd import synth
from velo2cycles_amd.abi import MARKER_FIELDS, NODAL_FIELDS
from oracle import refrun
from tests.util import run_oracle

pytestmark = pytest.mark.skipif(not refrun.available(), reason="reference sources not mounted")
DT = synth.YEAR
STEP = ["m2n", f"before:{DT}:0", f"after:{DT}:0", "n2m", "move"]


def _bitwise(cfg, ops):
    mk, nd, raw = refrun.run(cfg, ops)
    om, on, _ = run_oracle(cfg, ops)
    bad = [f for f in MARKER_FIELDS if not np.array_equal(om[f], mk[f], equal_nan=True)]
    bad += [k for k in NODAL_FIELDS if not np.array_equal(on[k], nd[k], equal_nan=True)]
    assert np.array_equal(om.t, mk["t"])
    assert not bad, bad


@pytest.mark.parametrize("timesum", [0.0, DT])
def test_medium_refined_step(timesum):
    """201x101 refined grid, 16 markers per cell (320,000 markers)."""
    cfg = synth.config_grid(201, 101, 200e3, 100e3, 4, refine=True)
    cfg.timesum = timesum
    _bitwise(cfg, STEP)


def test_reference_shipped_grid_size():
    """251x101 nodes over 50x20 km with 16 markers per cell (init_RSF.t3c:2-7),
    band geometry of the shipped set-up inside the domain."""
    cfg = synth.config_grid(251, 101, 50e3, 20e3, 4)
    cfg.timesum = DT
    _bitwise(cfg, STEP)


@pytest.mark.parametrize("flags", [dict(markmod=1, stoksmod=2), dict(densimod=1, outgrid=1),
                                   dict(RSFmode=0, antidiffinter=0), dict(power_law_RSF=1, inertyn=1)])
def test_flag_variants_small(flags):
    cfg = synth.config_grid(61, 41, 60e3, 40e3, 3, refine=True, seed=3)
    cfg.timesum = DT
    for k, v in flags.items():
        setattr(cfg.params, k, v)
    _bitwise(cfg, STEP)
