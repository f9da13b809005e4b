"""Per-field error of the HIP path against the FP64 oracle, printed as a table.

    python tests/parity_report.py [--config small|edge] [--steps 2]

Runs the full step (marker2nodes, compute_before_solve, compute_after_solve,
nodes2markers + advection) the given number of times through libv2c on the GPU
and through the oracle on the same seeded input, then prints, for every
marker and nodal field that either side changed: the largest and the rms error
scaled by max(|oracle|, rms(field)) (the bar of the parity tests, 1e-12), the
number of entries that differ at all, and NaN / Inf mismatches.  Positions,
types and cell keys are integer-exact or bit-exact and are reported as counts
of differing entries.  Needs a GPU (it is a report, not a test).
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from velo2cycles_amd import synth  # noqa: E402
from velo2cycles_amd.abi import MARKER_FIELDS, NODAL_FIELDS, load_lib  # noqa: E402
from tests.util import rel_errors, run_gpu, run_oracle  # noqa: E402


def report(cfg, steps: int):
    dt = cfg.timestep
    ops, gops = [], []
    for s in range(steps):
        ts = cfg.timesum + s * dt
        step = [f"time:{ts}:{dt}", "m2n", f"before:{dt}:0", f"after:{dt}:0"]
        ops += step + ["n2m", "move"]
        gops += step + ["n2m_move"]
    om, on, _ = run_oracle(cfg, ops)
    gm, gn, _, ctx = run_gpu(load_lib(), cfg, gops)
    ctx.close()
    rows = []
    rows.append(("marker", "t (type)", None, None, int(np.sum(gm.t != om.t)), 0))
    for f in MARKER_FIELDS:
        a, b = gm[f], om[f]
        if np.array_equal(b, cfg.markers[f], equal_nan=True) and np.array_equal(a, b, equal_nan=True):
            continue
        mx, rms, nm = rel_errors(a, b)
        diff = int(np.sum(~((a == b) | (np.isnan(a) & np.isnan(b)))))
        rows.append(("marker", f, mx, rms, diff, nm))
    for k in NODAL_FIELDS:
        a, b = gn[k], on[k]
        if np.array_equal(b, cfg.nodal[k], equal_nan=True) and np.array_equal(a, b, equal_nan=True):
            continue
        mx, rms, nm = rel_errors(a, b)
        diff = int(np.sum(~((a == b) | (np.isnan(a) & np.isnan(b)))))
        rows.append(("nodal", k, mx, rms, diff, nm))
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="small", choices=["small", "edge"])
    ap.add_argument("--steps", type=int, default=2)
    args = ap.parse_args()
    cfg = synth.config_small() if args.config == "small" else synth.config_edge()
    cfg.timesum = synth.YEAR
    rows = report(cfg, args.steps)
    print(f"# {cfg.name}: {cfg.n:,} markers, {cfg.grid.nodenum:,} nodes, {args.steps} full steps, dt = {cfg.timestep:g} s")
    print(f"# {'kind':6s} {'field':18s} {'max scaled err':>14s} {'rms scaled err':>14s} {'#differ':>9s} {'NaN/Inf mism.':>13s}")
    for kind, f, mx, rms, diff, nm in rows:
        smx = "-" if mx is None else f"{mx:.2e}"
        srms = "-" if rms is None else f"{rms:.2e}"
        print(f"  {kind:6s} {f:18s} {smx:>14s} {srms:>14s} {diff:9d} {nm:13d}")
    worst = max((r[2] for r in rows if r[2] is not None), default=0.0)
    print(f"# worst scaled error {worst:.2e} (bar 1e-12); positions {'bit-exact' if all(r[4] == 0 for r in rows if r[1] in ('x', 'y', 't (type)')) else 'DIFFER'}")


if __name__ == "__main__":
    main()
