"""Throughput of the per-marker hot path on one MI355X.

One step = marker2nodes -> compute_before_solve -> compute_after_solve ->
nodes2markers + moveandrotatemarkers (fused), i.e. the reference time loop
(i2evps.c:89-113) minus the host Stokes/heat solves, whose nodal outputs are
synthetic fields already resident in HBM.  Workload: 1501x301 nodes over
1500x300 km refined to 500 m near the megathrust, 36 markers per cell
(16.2 M markers, FP64 SoA), dt = 1 yr.

Prints ONE JSON line (rank 0).  The cpu_baseline leg runs the FP64 oracle on
a bounded sample of the same marker cloud.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

import numpy as np  # noqa: E402

from velo2cycles_amd import synth  # noqa: E402
from velo2cycles_amd.abi import KERNELS, Context, load_lib  # noqa: E402

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes (FP64 = 8 B, type = 1 B): what each kernel must read and
# write for its function, excluding the library's own intermediates (sort
# order, keys, per-cell partials, per-marker property hand-over).
#   m2n_props   reads x y T sxx sxy phi + type; writes the calc_frictionpara
#               outputs (6 fields for the markers outside the fault band)
#               + nodal pr                                   -> 49 R + 48 W
#   m2n_cells   reads x y T P sxx sxy exx exy e phi vx vy C + type   -> 105 R
#   n2m_scatter reads x y + type + 9 nodal fields                    ->  17 R
#   n2m_update  reads x y T + type; writes phi exx exy sxx sxy P vx vy e x y
#               + 19 nodal fields                            -> 25 R + 88 W
# Per node: the unique nodal arrays each kernel touches.
ALG = {
    "m2n_props": (49 + 48, 1 * 8),
    "m2n_cells": (105, 0),
    "n2m_cells": (17, 9 * 8),
    "n2m_update": (25 + 88, 19 * 8),
}
# Whole step: marker2nodes 161 B and nodes2markers + advection 113 B per
# marker; ~360 B per node over all passes (each nodal array touched once).
B_STEP_MARKER = 161 + 113
B_STEP_NODE = 360

# Nodal fields the reference-side binding moves per step when the host keeps
# the nodal arrays (marker2nodes 24 in / 43 out, nodes2markers 28 / 9,
# moveandrotatemarkers 5 / 0, before solve 22 / 6, after solve 19 / 3).
PCIE_H2D_FIELDS = 24 + 28 + 5 + 22 + 19
PCIE_D2H_FIELDS = 43 + 9 + 6 + 3

# FP64 vector issue: one 64-lane wavefront instruction per SIMD every 4
# cycles (16 FP64 lanes per SIMD); 256 CUs x 4 SIMDs at 2.4 GHz.  Quarter-rate
# transcendentals (rcp, rsq, sqrt of f64) take 16 cycles.
SIMDS, CLOCK_HZ = 256 * 4, 2.4e9


def cpu_baseline(cfg, sample: int, steps: int = 2):
    """The reference's own OpenMP routines (oracle/_ref/ref_driver, built from
    its sources) timed on this host's cores for `steps` full steps on the
    largest window of the production grid its capacity macros allow
    (synth.config_production_window: 250 x 120 nodes of the 500 m refined
    part, 36 markers per cell, 1.07 M markers), with OMP_NUM_THREADS threads
    and with one; plus the FP64 oracle on a sample of the production cloud
    itself.  Without the built driver, the oracle figure alone ("port")."""
    port = cpu_port(cfg, sample, steps)
    from oracle import refrun
    if not refrun.built(refrun.DRIVER):
        return port
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
    win = synth.config_production_window()
    dt = cfg.timestep
    ops = ["m2n", f"before:{dt}:0", f"after:{dt}:0", "n2m", "move"] * steps
    rates = {}
    for th in sorted({threads, 1}):
        _, _, raw = refrun.run(win, ops, timesum=dt, threads=th)
        rates[th] = win.n * steps / float(raw["drv_time"][0])
    return {"value": rates[threads], "unit": "marker-steps/s", "cores": threads, "kind": "reference",
            "sample": f"reference routines (marker2nodes, compute_before_solve, compute_after_solve, "
                      f"nodes2markers, moveandrotatemarkers; OpenMP, {threads} threads) on "
                      f"{win.grid.xnumx}x{win.grid.ynumy} nodes of the production grid's 500 m part with "
                      f"{win.n:,} markers (the reference's capacity), {steps} steps, op time only",
            "value_1thread": rates[1], "port_1thread": port}


def cpu_port(cfg, sample: int, steps: int = 2):
    """FP64 oracle, one thread, full step on the first `sample` markers of the
    same cloud (x-major cell order) with the full nodal grid."""
    from oracle.pyoracle import Oracle
    from velo2cycles_amd.abi import MarkerArrays
    fields = {f: cfg.markers[f][:sample] for f in cfg.markers.arrays}
    fields["t"] = cfg.markers.t[:sample]
    sub = synth.Config(cfg.name, cfg.grid, cfg.params, cfg.mats, MarkerArrays(sample, fields),
                       cfg.nodal, cfg.timestep, cfg.timesum)
    o = Oracle(sub)
    dt = cfg.timestep
    t0 = time.perf_counter()
    for s in range(steps):
        o.set_time(cfg.timesum + s * dt, dt)
        o.marker2nodes()
        o.compute_before_solve(dt, 0)
        o.compute_after_solve(dt, 0)
        o.nodes2markers()
        o.moveandrotatemarkers()
    el = time.perf_counter() - t0
    return {"value": sample * steps / el, "unit": "marker-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle (FP64 C restatement, 1 thread) on the first {sample:,} markers of the "
                      f"same cloud with the full {cfg.grid.nodenum:,}-node grid, {steps} full steps, "
                      f"{el:.1f} s"}


def pcie_inclusive(ctx, cfg, ms_step):
    """Step time if the host keeps the nodal arrays (reference-side binding):
    the device step plus the per-step nodal transfers, each field timed as a
    synchronous copy of the reference's pageable host array."""
    names = list(cfg.nodal)
    host = {k: np.ascontiguousarray(cfg.nodal[k]) for k in names}
    ctx.synchronize()
    t0 = time.perf_counter()
    for k in names:
        ctx.set_nodal(k, host[k])
    t_h2d = (time.perf_counter() - t0) / len(names)
    t0 = time.perf_counter()
    for k in names:
        ctx.get_nodal(k)
    t_d2h = (time.perf_counter() - t0) / len(names)
    ms = ms_step + (PCIE_H2D_FIELDS * t_h2d + PCIE_D2H_FIELDS * t_d2h) * 1e3
    return {"ms_per_step": ms, "value": cfg.n / (ms * 1e-3), "unit": "marker-steps/s",
            "h2d_fields": PCIE_H2D_FIELDS, "d2h_fields": PCIE_D2H_FIELDS,
            "ms_per_field_h2d": t_h2d * 1e3, "ms_per_field_d2h": t_d2h * 1e3,
            "field_mb": cfg.grid.nodenum * 8 / 1e6}


# Algorithmic bytes of the routines after the solves, per node (and per
# marker, per boundary row): the unique arrays each routine must read and write.
#   recompute_stress_strainrate  reads vx vy nd gd sxxe nu gg sxye phis phis_it,
#                                writes sxx exx deltasxx sxy exy esp deltasxy  17 x 8
#   interpolate_center2basic     reads exx sxx sxxe gd nd_m2c pr, writes exxs
#                                exxpl sxxs prs exxel                          11 x 8
#   interpolate_center2basic_old reads exxe sxxe sppe, writes exxes sxxes sppes 6 x 8
#   interpolate_basic2center     reads nu, writes nd                            2 x 8
#   copyarrays                   reads RK_sliprs_solved RK_sliprs_final
#                                RK_phidts_solved, writes deltaVs sliprs
#                                RK_sliprs_final RK_phidts_final                7 x 8
#   calc_timestep(0)             reads bm gg Ls as bs V0s RK_sliprs_final
#                                RK_sliprs_solved nu_m2c sxxes sxye nu mvx mvy
#                                phis                                          15 x 8
#   temperature markers          per marker x y k type + k written (33 B); per
#                                node tk tk2 kt cp ro read, tol0 tol1 written
#   boundary tail                per node the field(s) read and written and the
#                                bondm entries; per row bondn + bondv (56 B)
NEXT_ALG = {
    "recompute_stress_strainrate": (0, 17 * 8),
    "interpolate_center2basic": (0, 11 * 8),
    "interpolate_center2basic_old": (0, 6 * 8),
    "interpolate_basic2center": (0, 2 * 8),
    "copyarrays": (0, 7 * 8),
    "calc_timestep": (0, 15 * 8),
    "temperature_markers": (33, 7 * 8),
    "vxvy_boundarycorrection": (0, 2 * 8 * 3),
    "P_boundarycorrection": (0, 8 * 3),
    "tkrecalc": (0, 8 * 3),
}


def next_rows(ctx, cfg, reps: int = 20):
    """The routines after the solves on the same workload: per call, the
    HIP-event kernel time (all kernels of the call), the wall time including
    its host synchronisations, and the algorithmic bytes over the kernel time.
    Also the .prn writer and reader on the grid plus the first 1 M markers."""
    import tempfile
    from velo2cycles_amd import prn
    from velo2cycles_amd.abi import TimestepParams
    n, nn = cfg.n, cfg.grid.nodenum
    rng = np.random.default_rng(5)
    nod = synth.post_solve_nodal(cfg, seed=5, degenerate=False)
    nod["tk2"] = nod["tk"] + rng.normal(0.0, 5.0, nn)
    bc = synth.boundary_equations(cfg, 5, extra=False)
    rows = len(bc[1])
    ctx.set_all_nodal(nod)
    ctx.set_boundary(*bc)
    dt = cfg.timestep
    calls = {
        "recompute_stress_strainrate": lambda: ctx.recompute_stress_strainrate(dt, 0),
        "interpolate_center2basic": lambda: ctx.interpolate_center2basic(dt),
        "interpolate_center2basic_old": ctx.interpolate_center2basic_old,
        "interpolate_basic2center": ctx.interpolate_basic2center,
        "copyarrays": ctx.copyarrays,
        "calc_timestep": lambda: ctx.calc_timestep(0, TimestepParams.default(maxtmstep=dt)),
        "temperature_markers": lambda: ctx.temperature_markers(1.0),
        "vxvy_boundarycorrection": ctx.vxvy_boundarycorrection,
        "P_boundarycorrection": ctx.P_boundarycorrection,
        "tkrecalc": ctx.tkrecalc,
    }
    out = {}
    for name, fn in calls.items():
        fn()                                   # warm
        ctx.synchronize()
        ctx.enable_timing(True)
        ctx.reset_timing()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        ctx.synchronize()
        wall = (time.perf_counter() - t0) / reps * 1e3
        ctx.enable_timing(False)
        ms, _ = ctx.kernel_time("post")
        ms /= reps
        per_m, per_n = NEXT_ALG[name]
        alg = n * per_m + nn * per_n + (rows * 56 if name.endswith(("correction", "recalc")) else 0)
        out[name] = {"kernel_ms": ms, "wall_ms": wall, "alg_bytes": alg,
                     "achieved_gbs": alg / (ms * 1e-3) / 1e9 if ms > 0 else None}
    d = synth.prn_model(cfg, 5, bonds=False)
    k = min(n, 1_000_000)
    d["marknum"] = k
    d["mark"] = {c: v[:k] for c, v in d["mark"].items()}
    d["t"] = d["t"][:k]
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "m.prn")
        t0 = time.perf_counter()
        prn.write(path, d)
        t_w = time.perf_counter() - t0
        size = os.path.getsize(path)
        t0 = time.perf_counter()
        prn.read(path)
        t_r = time.perf_counter() - t0
    out["prn"] = {"markers": k, "file_mb": size / 1e6, "write_mb_s": size / t_w / 1e6,
                  "read_mb_s": size / t_r / 1e6}
    return out


def valu_issue(workload, kern):
    """Per-kernel FP64-VALU issue bound from the committed counter profile
    (profiles/valu.json: wavefront VALU and f64-transcendental instructions
    per launch, from rocprofv3 --pmc SQ_INSTS_VALU SQ_INSTS_VALU_TRANS_F64 on
    this workload) against the kernel times measured here."""
    path = os.path.join(REPO, "profiles", "valu.json")
    if not os.path.exists(path):
        return None
    prof = json.load(open(path)).get(workload, {})
    out = {}
    for k, ent in prof.items():
        if k not in kern:
            continue
        cycles = 4.0 * ent["valu_per_launch"] + 12.0 * ent.get("trans_f64_per_launch", 0.0)
        bound_ms = cycles / SIMDS / CLOCK_HZ * 1e3
        ms = kern[k]["ms_total"] / max(kern[k]["launches"], 1)
        out[k] = {"issue_bound_ms": bound_ms, "avg_launch_ms": ms, "frac": bound_ms / ms}
    return out or None


def _warm_regime(lib, dt: float, maker, label: str, seeds=(11, 12, 13, 14), keep_other=1.0, epochs=8):
    from velo2cycles_amd.warmstart import WarmStart, evaluate, generate, init_params
    cfg = maker(seed=synth.DEFAULT_SEED)
    cfg.timestep = dt
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=cfg.n, lib=lib)
    t0 = time.perf_counter()
    X, M = generate(ctx, list(seeds), dt, maker=maker, max_rows=2_000_000, keep_other=keep_other)
    ws = WarmStart(ctx)
    ws.fit_norm(X, M)
    ws.load(init_params(0))
    loss = ws.train(X, M, epochs=epochs, batch=4096, lr=2e-3)
    t_train = time.perf_counter() - t0
    modes = ("plain", "warm", "warm_rsf")
    res = evaluate(ctx, cfg, ws, dt, tol=1e-10, repeats=3, modes=modes)
    ctx.close()
    out = {"workload": label, "dt_s": dt, "residual": 1e-10, "training_rows": int(len(X)),
           "slip_rate_rows": int((M[:, 1] > 0).sum()), "epochs": epochs,
           "train_s": round(t_train, 2), "final_loss": float(loss[-1])}
    n = cfg.n
    for mode in modes:
        st = res[mode]["stats"]
        out[mode] = {
            "iters_per_marker": (st["ductile_iters"] + st["rsf_iters"]) / n,
            "ductile_iters_per_marker": st["ductile_iters"] / max(st["ductile_markers"], 1),
            "rsf_iters_per_fault_marker": st["rsf_iters"] / max(st["fault_markers"], 1),
            "fallbacks": st["fallbacks"], "nonfinite": st["nonfinite"],
            "ms_ductile": res[mode]["ms_ductile"], "ms_rsf": res[mode]["ms_rsf"], "ms_mlp": res[mode]["ms_mlp"],
            "ms_wall": res[mode]["ms_wall"],
            "ms_total": res[mode]["ms_total"]}
    p = out["plain"]
    for mode in modes[1:]:
        w = out[mode]
        out[mode]["iters_ratio"] = p["iters_per_marker"] / max(w["iters_per_marker"], 1e-30)
        out[mode]["rsf_iters_ratio"] = p["rsf_iters_per_fault_marker"] / max(w["rsf_iters_per_fault_marker"], 1e-30)
        # whole local update (ductile solve of every marker + slip-rate solve)
        out[mode]["wall_ratio"] = p["ms_total"] / max(w["ms_total"], 1e-30)
        # the slip-rate solve alone, with whatever inference it needs
        out[mode]["rsf_wall_ratio"] = p["ms_rsf"] / max(w["ms_rsf"] + w["ms_mlp"], 1e-30)
    return out


def warm_start_eval(lib, dt: float):
    """Newton iterations per marker and kernel time of the per-marker local
    update, plain vs learned warm start (both guesses, or the slip-rate guess
    only), at residual 1e-10 on the 401x101 / 16-per-cell cloud (640,000
    markers, 2 % fault markers), in two regimes: interseismic (dt = 1 yr, the
    synthetic RSF state) and coseismic (dt = 0.01 s, fault markers slipping at
    1e-2..10 m/s, synth.coseismic).  Per regime the MLP is trained here, on
    the GPU, from plain solves on other seeds: 2 M rows of four clouds
    (interseismic); every fault-marker row of 64 clouds and 2 % of the other
    rows, 32 epochs (coseismic, where only the slip-rate guess matters)."""
    inter = _warm_regime(lib, dt, synth.config_small,
                         "401x101 uniform, 16 markers/cell, 2 % fault markers, interseismic")
    dtc = 1e-2
    cos = _warm_regime(lib, dtc, lambda seed: synth.coseismic(synth.config_small(seed=seed), seed=seed, dt=dtc),
                       "the same cloud, fault markers coseismic (V 1e-3..10 m/s, roots 1e-2..10 m/s)",
                       seeds=range(11, 75), keep_other=0.02, epochs=32)
    out = dict(inter)
    out["iters_ratio"] = inter["warm_rsf"]["iters_ratio"]
    out["wall_ratio"] = inter["warm_rsf"]["wall_ratio"]
    out["coseismic"] = cos
    return out


def device_memory_used():
    """Bytes of device memory in use (hipMemGetInfo), or None."""
    try:
        hip = C.CDLL("libamdhip64.so")
        free, total = C.c_size_t(0), C.c_size_t(0)
        if hip.hipMemGetInfo(C.byref(free), C.byref(total)) != 0:
            return None
        return total.value - free.value
    except OSError:
        return None


def ensemble_bench(args):
    """Batched ensemble of 64 uneven models: one step = marker2nodes,
    before/after solve, nodes2markers + advection of every model.  Times the
    graph-replayed batched step with default tile shapes, after autotuning, and
    the same step as per-model launches."""
    from velo2cycles_amd.abi import Ensemble
    lib = load_lib()
    ctxs, n_total, nodes_total, sizes = [], 0, 0, []
    t0 = time.perf_counter()
    prod = args.production_models > 0
    models = (synth.config_ensemble_production(args.production_models) if prod
              else synth.config_ensemble(args.models, small=args.small_models))
    for k, cfg in models:
        ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=cfg.n, lib=lib)
        ctx.set_time(cfg.timestep, cfg.timestep)
        ctx.set_all_nodal(cfg.nodal)
        ctx.upload_markers(cfg.markers)
        ctxs.append(ctx)
        n_total += cfg.n
        nodes_total += cfg.grid.nodenum
        sizes.append(cfg.n)
        del cfg
    t_gen = time.perf_counter() - t0
    ens = Ensemble(ctxs)

    def timed(fn, steps):
        for c in ctxs:
            c.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        for c in ctxs:
            c.synchronize()
        return (time.perf_counter() - t) / steps

    resident = device_memory_used()
    for _ in range(args.warmup):
        ens.step(graph=True)
    t_default = timed(lambda: ens.step(graph=True), args.steps)
    t_looped = timed(ens.step_looped, args.steps)
    if args.no_tune:
        t_graph_only, t_tuned = None, t_default
    else:
        t_graph_only = ens.autotune(reps=2)
        t_tuned = timed(lambda: ens.step(graph=True), args.steps)
    shapes = {str(c): ens.get_shape(c) for c in range(3)}
    ens.close()
    for c in ctxs:
        c.close()
    best = min(t_default, t_tuned)
    out = {"metric": "marker-steps/sec", "value": n_total / best, "unit": "marker-steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": best * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": ("synthetic (one seeded production cloud, per-model convergence rate and a-b, "
                    "analytic corner-flow nodal fields)" if prod else
                    "synthetic (seeded uneven models, analytic corner-flow nodal fields)"),
           "config": {"workload": (f"ensemble-{len(ctxs)}-production" if prod else
                                   f"ensemble-{args.models}" + ("-small" if args.small_models else "")),
                      "models": len(ctxs), "markers": n_total,
                      "nodes": nodes_total, "markers_min": min(sizes), "markers_max": max(sizes),
                      "grids": ("1501x301 refined to 500 m, 36 markers per cell, one cloud, 2-8 cm/yr"
                                if prod else
                                ("41x21 .. 161x61" if args.small_models else "201x51 .. 801x201")
                                + " nodes, 4-36 markers per cell"), "dt_s": synth.YEAR},
           "device_memory_used_gb": None if resident is None else round(resident / 1e9, 1),
           "batched_graph_default_ms": t_default * 1e3, "batched_graph_tuned_ms": t_tuned * 1e3,
           "looped_single_model_ms": t_looped * 1e3,
           "tuned_graph_only_ms": t_graph_only, "speedup_vs_looped": t_looped / best,
           "tuned_shapes": shapes, "setup_s": round(t_gen, 1)}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="production", choices=["production", "small"])
    ap.add_argument("--cpu-sample", type=int, default=3_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-warm", action="store_true", help="skip the warm-start evaluation")
    ap.add_argument("--no-next", action="store_true",
                    help="skip the post-solve / boundary / heat / .prn measurements")
    ap.add_argument("--ensemble", action="store_true", help="batched ensemble workload instead")
    ap.add_argument("--models", type=int, default=64)
    ap.add_argument("--no-tune", action="store_true", help="ensemble: skip the autotuner")
    ap.add_argument("--small-models", action="store_true", help="ensemble: 41x21 .. 161x61-node models")
    ap.add_argument("--production-models", type=int, default=0,
                    help="ensemble: this many production-size models (16.2 M markers each) instead")
    ap.add_argument("--calib", action="store_true",
                    help="also run a plain 1 GiB device copy (calibrates rocprof byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; multi-GPU runs are not part of it "
                                   "(ensemble sharding across GPUs is a design note only)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.ensemble:
        ensemble_bench(args)
        return
    cfg = synth.config_production() if args.config == "production" else synth.config_small()
    lib = load_lib()
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=cfg.n, lib=lib)
    dt = cfg.timestep
    ctx.set_time(0.0, dt)
    ctx.set_all_nodal(cfg.nodal)
    ctx.upload_markers(cfg.markers)
    h = ctx.h
    if args.calib:
        ms = C.c_double(0)
        lib.v2c_copy_calibration(h, 1 << 30, C.byref(ms))
    ctx.set_async(True)

    def step(ts):
        lib.v2c_set_time(h, ts, dt)
        for rc in (lib.v2c_marker2nodes(h), lib.v2c_compute_before_solve(h, dt, 0, None),
                   lib.v2c_compute_after_solve(h, dt, 0, None, None), lib.v2c_nodes2markers_advect(h)):
            if rc != 0:
                raise RuntimeError(lib.v2c_last_error(h).decode())

    ts = 0.0
    for _ in range(args.warmup):
        step(ts)
        ts += dt
    ctx.synchronize()
    ctx.enable_timing(True)
    ctx.reset_timing()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(ts)
        ts += dt
    ctx.synchronize()
    el = time.perf_counter() - t0
    ctx.enable_timing(False)
    kern = {}
    for k in KERNELS:
        ms, nl = ctx.kernel_time(k)
        if nl:
            kern[k] = {"ms_total": ms, "launches": nl, "ms_per_step": ms / args.steps}
    n, nn = cfg.n, cfg.grid.nodenum
    ms_step = el / args.steps * 1e3
    bytes_step = n * B_STEP_MARKER + nn * B_STEP_NODE
    dom = max((k for k in ALG if k in kern), key=lambda k: kern[k]["ms_total"])
    per_launch_bytes = n * ALG[dom][0] + nn * ALG[dom][1]
    avg_ms = kern[dom]["ms_total"] / max(kern[dom]["launches"], 1)
    achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9
    pcie = pcie_inclusive(ctx, cfg, ms_step)
    valu = valu_issue(cfg.name, kern)
    traffic = None
    tpath = os.path.join(REPO, "profiles", "traffic.json")
    if os.path.exists(tpath):
        tj = json.load(open(tpath))
        ent = tj.get(cfg.name, {}).get(dom)
        if ent:
            traffic = ent["hbm_bytes_per_launch"]
    out = {
        "metric": "marker-steps/sec", "value": n * args.steps / el, "unit": "marker-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded marker cloud, analytic corner-flow nodal fields)",
        "config": {"workload": cfg.name, "markers": n, "nodes": nn, "markers_per_cell": 36
                   if args.config == "production" else 16, "dt_s": dt, "grid": "1501x301 refined to 500 m"
                   if args.config == "production" else "401x101 uniform"},
        "hbm_frac_step": bytes_step / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_launch": per_launch_bytes, "avg_launch_ms": avg_ms},
        "kernels_ms_per_step": {k: v["ms_per_step"] for k, v in kern.items()},
        # m2n_props runs on a second stream beside the re-ordering ("sort"):
        # their event times include each other's share of the GPU
        "overlapped_kernels": ["m2n_props", "sort"],
        "pcie_inclusive": pcie,
    }
    if valu:
        out["valu_issue"] = valu
    if not args.no_next:
        ctx.set_async(False)
        out["next_rows"] = next_rows(ctx, cfg)
    ctx.close()
    if not args.no_warm:
        out["warm_start"] = warm_start_eval(lib, synth.YEAR)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, min(args.cpu_sample, cfg.n))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
