"""Time the reference's own routines (oracle/_ref/ref_driver, built in this
container from the reference sources) beside the FP64 oracle on the same
workload, and write oracle/reference_timing.json.

Test infrastructure, run only where the reference is built (the built
driver never leaves this container).  The benchmark reads the JSON for its
CPU baseline: the reference's rate on this host, and the oracle/reference
rate ratio at 1 thread and at all cores, which calibrates the oracle rate
the benchmark measures on the GPU host's cores.

Workload: the largest window of the production grid the reference's
capacity macros admit (synth.config_production_window: 250 x 120 nodes of
the 500 m refined part, 36 markers per cell, 1.07 M markers), dt = 1 yr,
`steps` full steps of marker2nodes, compute_before_solve,
compute_after_solve, nodes2markers, moveandrotatemarkers (i2evps.c:89-113
without the solves); the reference's time is its driver's own clock around
the routines (OpenMP with OMP_NUM_THREADS threads), the oracle's the wall
time of its threads (oracle/cpu_rate.py); per routine, ns per marker-step
(the reference's clock around each routine call; the oracle's slowest
thread).

    python -m oracle.time_reference [--steps 4]
"""
from __future__ import annotations

import argparse
import json
import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

from velo2cycles_amd import synth  # noqa: E402
from oracle import cpu_rate, refrun  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_timing.json")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--repeats", type=int, default=3)
    args = ap.parse_args()
    if not refrun.built(refrun.DRIVER):
        sys.exit("oracle/_ref/ref_driver is not built (make -C oracle with the reference present)")
    win = synth.config_production_window()
    dt = win.timestep
    win.timesum = dt
    ops = ["m2n", f"before:{dt}:0", f"after:{dt}:0", "n2m", "move"] * args.steps
    cores = cpu_rate.usable_threads()
    all_t = cores["threads"]
    res = {"workload": f"production window {win.grid.xnumx}x{win.grid.ynumy} nodes, {win.n} markers, "
                       f"{args.steps} steps, dt = 1 yr",
           "markers": win.n, "steps": args.steps, "host_cpu": cpu_model_line(), **cores,
           "reference": {}, "oracle": {}, "oracle_over_reference": {}}
    res["ns_per_marker_step"] = {"reference": {}, "oracle": {}}
    per_ms = 1e9 / (win.n * args.steps)
    for th in sorted({1, all_t}):
        best_r = best_o = 0.0
        for _ in range(args.repeats):
            _, _, raw = refrun.run(win, ops, timesum=dt, threads=th)
            r = win.n * args.steps / float(raw["drv_time"][0])
            if r > best_r:
                best_r = r
                res["ns_per_marker_step"]["reference"][str(th)] = {
                    k: float(v) * per_ms for k, v in zip(cpu_rate.ROUTINES, raw["drv_optime"])}
            o = cpu_rate.step_rate(win, win.n, args.steps, th)
            if o["rate"] > best_o:
                best_o = o["rate"]
                res["ns_per_marker_step"]["oracle"][str(th)] = {
                    k: v * per_ms for k, v in o["routine_seconds"].items()}
        res["reference"][str(th)] = best_r
        res["oracle"][str(th)] = best_o
        res["oracle_over_reference"][str(th)] = best_o / best_r
        print(f"{th} threads: reference {best_r:.4g}, oracle {best_o:.4g} marker-steps/s", flush=True)
    with open(OUT, "w") as f:
        json.dump(res, f, indent=1)
        f.write("\n")
    print(json.dumps(res))


def cpu_model_line():
    return cpu_rate.cpu_model()


if __name__ == "__main__":
    main()
