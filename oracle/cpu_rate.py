"""CPU throughput of the FP64 oracle -- test infrastructure: it times the
checker as the benchmark's CPU baseline and is never the thing measured on
the GPU nor part of the product.

The oracle's routines keep all state in their or_state argument and ctypes
releases the GIL around each call, so T Python threads, each with its own
Oracle over a contiguous slice of the markers (x-major cell order) and the
full nodal grid, run the routines on T cores at once.  Each thread also runs
the node loops over the whole grid, T-fold redundant work, so the T-thread
figure is a lower bound of what a shared-grid parallel oracle would reach.
"""
from __future__ import annotations

import os
import subprocess
import threading
import time

from velo2cycles_amd import synth
from velo2cycles_amd.abi import MarkerArrays
from oracle.pyoracle import Oracle


ROUTINES = ("marker2nodes", "compute_before_solve", "compute_after_solve", "nodes2markers",
            "moveandrotatemarkers")


def physical_cores() -> int | None:
    """Physical cores (distinct core, socket pairs) that lscpu reports."""
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        return None
    pairs = {ln.strip() for ln in out.splitlines() if ln and not ln.startswith("#")}
    return len(pairs) or None


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def usable_threads() -> dict:
    """Threads for an all-core run: the physical cores lscpu reports, capped
    by the CPUs this process may run on and by OMP_NUM_THREADS when set (the
    host's CPU share for this job)."""
    phys = physical_cores()
    aff = len(os.sched_getaffinity(0))
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    t = min(x for x in (phys, aff, omp or None) if x)
    return {"threads": t, "physical_cores_lscpu": phys, "affinity_cpus": aff, "omp_num_threads": omp or None}


def _slice(cfg, lo, hi):
    f = {k: cfg.markers[k][lo:hi] for k in cfg.markers.arrays}
    f["t"] = cfg.markers.t[lo:hi]
    return synth.Config(cfg.name, cfg.grid, cfg.params, cfg.mats, MarkerArrays(hi - lo, f), cfg.nodal,
                        cfg.timestep, cfg.timesum)


def step_rate(cfg, sample: int, steps: int, threads: int) -> dict:
    """Full steps (marker2nodes, compute_before_solve, compute_after_solve,
    nodes2markers, moveandrotatemarkers) of the first `sample` markers of
    cfg on `threads` threads; marker-steps per second of wall time."""
    sample = min(sample, cfg.n)
    cuts = [sample * i // threads for i in range(threads + 1)]
    oracles = [Oracle(_slice(cfg, cuts[i], cuts[i + 1])) for i in range(threads)]
    dt = cfg.timestep
    optime = [[0.0] * len(ROUTINES) for _ in range(threads)]

    def work(o, acc):
        for s in range(steps):
            o.set_time(cfg.timesum + s * dt, dt)
            for r, fn in enumerate((o.marker2nodes, lambda: o.compute_before_solve(dt, 0),
                                    lambda: o.compute_after_solve(dt, 0), o.nodes2markers,
                                    o.moveandrotatemarkers)):
                t = time.perf_counter()
                fn()
                acc[r] += time.perf_counter() - t

    pool = [threading.Thread(target=work, args=(o, optime[i])) for i, o in enumerate(oracles)]
    t0 = time.perf_counter()
    for th in pool:
        th.start()
    for th in pool:
        th.join()
    el = time.perf_counter() - t0
    # per routine: the slowest thread's time
    per = {name: max(acc[r] for acc in optime) for r, name in enumerate(ROUTINES)}
    return {"rate": sample * steps / el, "seconds": el, "markers": sample, "steps": steps, "threads": threads,
            "routine_seconds": per}
