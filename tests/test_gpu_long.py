"""Long runs of the HIP path against the FP64 oracle: drift over the run
length of the production configuration (100 steps)."""
import numpy as np
import pytest

from velo2cycles_amd import synth
from velo2cycles_amd.abi import MARKER_FIELDS, NODAL_FIELDS
from tests.util import compare, rel_errors, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(600)
def test_hundred_steps_small_config(gpu):
    """401x101 nodes, 16 markers per cell (640,000 markers), 100 full steps
    at dt = 1 yr (marker2nodes, compute_before_solve, compute_after_solve,
    nodes2markers + advection), the i2evps.c:82-132 loop without the solves:
    positions and types still bit-exact after the last step, every marker and
    nodal field within 1e-12 (scaled) of the oracle."""
    cfg = synth.config_small()
    dt = cfg.timestep
    cfg.timesum = dt
    ops, gops = [], []
    for s in range(100):
        step = [f"time:{cfg.timesum + s * dt}:{dt}", "m2n", f"before:{dt}:0", f"after:{dt}:0"]
        ops += step + ["n2m", "move"]
        gops += step + ["n2m_move"]
    print("case 100 steps", cfg.name, cfg.n, flush=True)
    om, on, _ = run_oracle(cfg, ops)
    gm, gn, _, ctx = run_gpu(gpu, cfg, gops)
    ctx.close()
    bad = compare(gm, om, gn, on)
    worst = max(rel_errors(gm[f], om[f])[0] for f in MARKER_FIELDS if f not in ("x", "y"))
    worst = max(worst, max(rel_errors(gn[k], on[k])[0] for k in NODAL_FIELDS))
    print(f"100 steps: worst scaled error {worst:.2e}", flush=True)
    assert not bad, "\n".join(bad)
    moved = np.sum(gm["x"] != cfg.markers["x"])
    assert moved > 0.9 * cfg.n
