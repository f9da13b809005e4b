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


def log(msg):
    """Progress of the phases on stderr (the JSON line alone goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


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
# Whole step, the path's byte model: 256 B per marker (marker2nodes 116 R +
# 16 W, nodes2markers + advection 28 R + 96 W), 80 B per fault marker (its
# rate-and-state parameters read in each pass) and 360 B per node.  The sum
# of the per-kernel figures above (161 + 113 B per marker) is the same path
# counted field by field for this build's kernels; both are reported.
B_STEP_MARKER = 256
B_STEP_FAULT = 80
B_STEP_NODE = 360
B_STEP_MARKER_KERNELS = 161 + 113

# Nodal arrays the reference-side binding moves per step of the reference loop
# (one Picard iteration): the host Stokes solve's results to the device, and
# back to the host what the Stokes assembly reads of the device's outputs and
# what the per-step diagnostics and postproc read (v2c_dropin.c's
# STOKES_READS / OUTPUT_READS, those a device routine wrote).
PCIE_H2D = ("pr", "vx", "vy")
PCIE_D2H = ("ro", "nu", "nd", "gg", "gd", "sxxe", "sxye", "sppe", "mvx", "mvy", "mrx", "mry", "bm",
            "sxxs", "sxy", "sliprs", "phis_it", "syields", "sticky01")
# Before residency tracking: every array each routine reads, sent, and every
# array it writes, fetched (marker2nodes 24 / 43, nodes2markers 28 / 9,
# moveandrotatemarkers 5 / 0, before solve 22 / 6, after solve 19 / 3).
PCIE_H2D_EVERY_CALL = 24 + 28 + 5 + 22 + 19
PCIE_D2H_EVERY_CALL = 43 + 9 + 6 + 3

# FP64 vector issue: one 64-lane wavefront instruction per SIMD every 4
# cycles (16 FP64 lanes per SIMD); 256 CUs x 4 SIMDs at 2.4 GHz.  Quarter-rate
# transcendentals (rcp, rsq, sqrt of f64) take 16 cycles.
SIMDS, CLOCK_HZ = 256 * 4, 2.4e9


def cpu_baseline(cfg, sample: int, steps: int = 2):
    """The FP64 oracle timed on this host's cores ("port"): full steps on all
    usable physical cores (lscpu's count, capped by this job's CPU share)
    over the whole cloud, and on one core over its first `sample` markers.
    The reference's own routines are never run here: their rate comes from
    oracle/reference_timing.json, measured where the reference is built
    (oracle/time_reference.py: the reference and the oracle on the same
    1.07 M-marker production window), which also gives the oracle/reference
    rate ratio that calibrates the oracle figures."""
    from oracle import cpu_rate
    cores = cpu_rate.usable_threads()
    th = cores["threads"]
    multi = cpu_rate.step_rate(cfg, cfg.n, steps, th)
    one = cpu_rate.step_rate(cfg, sample, steps, 1)
    out = {"value": multi["rate"], "unit": "marker-steps/s", "cores": th, "kind": "port",
           "sample": (f"oracle (FP64 C restatement), {steps} full steps of all {cfg.n:,} markers of the "
                      f"same cloud with the full {cfg.grid.nodenum:,}-node grid on {th} threads, one "
                      f"marker slice each ({multi['seconds']:.1f} s wall; each thread also runs the grid's "
                      f"node loops); one core: the first {one['markers']:,} markers ({one['seconds']:.1f} s)"),
           "value_1core": one["rate"], "host_cpu": cpu_rate.cpu_model(), **{k: v for k, v in cores.items()
                                                                              if k != "threads"}}
    path = os.path.join(REPO, "oracle", "reference_timing.json")
    if os.path.exists(path):
        with open(path) as f:
            ref = json.load(f)
        r_all = ref["oracle_over_reference"][str(ref["threads"])]
        out["reference_elsewhere"] = {
            "what": "the reference's routines (its OpenMP build) timed where it is built; never run here",
            "workload": ref["workload"], "host_cpu": ref["host_cpu"], "threads": ref["threads"],
            "rate": ref["reference"][str(ref["threads"])], "rate_1thread": ref["reference"]["1"],
            "oracle_over_reference": ref["oracle_over_reference"]}
        out["reference_equivalent_here"] = {"value": multi["rate"] / r_all, "value_1core": one["rate"] /
                                            ref["oracle_over_reference"]["1"],
                                            "how": "this host's oracle rates divided by the oracle/reference "
                                                   "ratios measured beside the reference (estimate)"}
    return out


def pcie_inclusive(ctx, cfg, ms_step, reps: int = 5):
    """Step time with the host keeping its nodal arrays (the reference-side
    binding): the device step plus the binding's per-step hand-overs, timed
    here as the staged many-field transfers it makes (v2c_nodal_set_many /
    get_many) between ordinary pageable host arrays; best of `reps`.  The
    figure of the binding without residency tracking (every array a routine
    touches, one synchronous pageable copy each) is given beside it."""
    nn = cfg.grid.nodenum
    host = {k: np.ascontiguousarray(cfg.nodal.get(k, np.zeros(nn)), dtype=np.float64).copy()
            for k in set(PCIE_H2D) | set(PCIE_D2H)}
    ctx.synchronize()
    t_h2d = t_d2h = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        ctx.set_nodal_many({k: host[k] for k in PCIE_H2D})
        t_h2d = min(t_h2d, time.perf_counter() - t0)
        t0 = time.perf_counter()
        ctx.get_nodal_many(PCIE_D2H, out=host)
        t_d2h = min(t_d2h, time.perf_counter() - t0)
    ms = ms_step + (t_h2d + t_d2h) * 1e3
    # per-field synchronous pageable copies, as the binding made them before
    names = list(cfg.nodal)
    t0 = time.perf_counter()
    for k in names:
        ctx.set_nodal(k, np.ascontiguousarray(cfg.nodal[k]))
    f_h2d = (time.perf_counter() - t0) / len(names)
    t0 = time.perf_counter()
    for k in names:
        ctx.get_nodal(k)
    f_d2h = (time.perf_counter() - t0) / len(names)
    ms_every = ms_step + (PCIE_H2D_EVERY_CALL * f_h2d + PCIE_D2H_EVERY_CALL * f_d2h) * 1e3
    return {"ms_per_step": ms, "value": cfg.n / (ms * 1e-3), "unit": "marker-steps/s",
            "h2d_fields": len(PCIE_H2D), "d2h_fields": len(PCIE_D2H), "ms_h2d": t_h2d * 1e3,
            "ms_d2h": t_d2h * 1e3, "field_mb": nn * 8 / 1e6,
            "h2d_gbs": len(PCIE_H2D) * nn * 8 / t_h2d / 1e9, "d2h_gbs": len(PCIE_D2H) * nn * 8 / t_d2h / 1e9,
            "every_call_binding": {"ms_per_step": ms_every, "h2d_fields": PCIE_H2D_EVERY_CALL,
                                   "d2h_fields": PCIE_D2H_EVERY_CALL, "ms_per_field_h2d": f_h2d * 1e3,
                                   "ms_per_field_d2h": f_d2h * 1e3}}


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
    # reads every marker type and the positions of the markers of types 1
    # and 2 (added per call below), writes sticky01
    "sticky_marks": (1, 8),
    "calc_sticky_stats": (0, 4 * 8),
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
        "sticky_marks": ctx.sticky_marks,
        "calc_sticky_stats": ctx.sticky_stats,
    }
    n12 = int(np.count_nonzero((cfg.markers.t == 1) | (cfg.markers.t == 2)))
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
        alg = n * per_m + nn * per_n + (rows * 56 if name.endswith(("correction", "recalc")) else 0) + \
            (n12 * 16 if name == "sticky_marks" else 0)
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
    # 21 repeats interleaved across the modes; times are each mode's median
    res = evaluate(ctx, cfg, ws, dt, tol=1e-10, repeats=21, modes=modes)
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
            "ms_total": res[mode]["ms_total"], "ms_total_best": res[mode]["ms_total_best"],
            "repeats": res[mode]["repeats"]}
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


COSEISMIC_DT = 1e-3


def coseismic_network(lib):
    """Warm-start network trained on plain coseismic solves of eight small
    clouds (101x51 nodes, 16 markers per cell, 3 % fault markers in the
    coseismic regime of synth.coseismic_run)."""
    from velo2cycles_amd.warmstart import WarmStart, generate, init_params

    def maker(seed):
        return synth.coseismic_run(synth.config_grid(101, 51, 100e3, 50e3, 4, seed=seed, fault_frac=0.03),
                                   seed=seed, dt=COSEISMIC_DT)
    cfg = maker(101)
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=cfg.n, lib=lib)
    X, M = generate(ctx, range(101, 109), COSEISMIC_DT, maker=maker, keep_other=0.05)
    ws = WarmStart(ctx)
    ws.fit_norm(X, M)
    ws.load(init_params(0))
    ws.train(X, M, epochs=8, batch=4096, lr=2e-3)
    net = (ws.fetch().copy(), ws.norm.copy())
    ctx.close()
    return net


def ensemble_bench(args):
    """Batched ensemble of 64 uneven models: one step = marker2nodes,
    before/after solve, nodes2markers + advection of every model.  Times the
    graph-replayed batched step with default tile shapes, after autotuning, and
    the same step as per-model launches.  --warm (with --production-models):
    the combined workload, every model's cloud in the coseismic regime, the
    per-marker local update with the network's slip-rate guess inside the
    solve (residual 1e-10) and adaptive steps from each model's
    calc_timestep (1e-3 s .. 1 yr) added to every step."""
    from velo2cycles_amd.abi import Ensemble, TimestepParams
    from velo2cycles_amd.warmstart import WarmStart, upload_cloud
    lib = load_lib()
    ctxs, n_total, nodes_total, sizes = [], 0, 0, []
    t0 = time.perf_counter()
    prod = args.production_models > 0
    warm = args.warm and prod
    net = coseismic_network(lib) if warm else None
    models = (synth.config_ensemble_production(args.production_models,
                                               coseismic_dt=COSEISMIC_DT if warm else None) if prod
              else synth.config_ensemble(args.models, small=args.small_models))
    n_fault = 0
    for k, cfg in models:
        ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=cfg.n, lib=lib)
        ctx.set_all_nodal(cfg.nodal)
        if warm:
            upload_cloud(ctx, cfg)
            WarmStart(ctx).load(*net)
            n_fault += int(cfg.rsf["fault"].sum())
        else:
            ctx.set_time(cfg.timestep, cfg.timestep)
            ctx.upload_markers(cfg.markers)
        ctxs.append(ctx)
        n_total += cfg.n
        nodes_total += cfg.grid.nodenum
        sizes.append(cfg.n)
        del cfg
    t_gen = time.perf_counter() - t0
    log(f"{len(ctxs)} models set up, {n_total} markers ({t_gen:.0f} s)")
    ens = Ensemble(ctxs)
    tsp = TimestepParams.default(maxtmstep=synth.YEAR)
    if warm:
        ens.set_local(warm=2, tol=1e-10)
        ens.set_adaptive(tsp, 0, COSEISMIC_DT, synth.YEAR)

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
    log(f"batched graph step {t_default * 1e3:.1f} ms")
    t_looped = timed(ens.step_looped, args.steps)
    log(f"looped step {t_looped * 1e3:.1f} ms")
    # the same graph as 2..4 concurrent branches of model groups
    t_branch = {1: t_default}
    for b in range(2, min(4, len(ctxs)) + 1):
        ens.set_branches(b)
        ens.step(graph=True)
        t_branch[b] = timed(lambda: ens.step(graph=True), args.steps)
        log(f"{b} branches: {t_branch[b] * 1e3:.1f} ms")
    best_b = min(t_branch, key=t_branch.get)
    ens.set_branches(best_b)
    t_default = t_branch[best_b]
    if args.no_tune:
        t_graph_only, t_tuned = None, t_default
    else:
        t_graph_only = ens.autotune(reps=2)
        t_tuned = timed(lambda: ens.step(graph=True), args.steps)
        log(f"tuned batched step {t_tuned * 1e3:.1f} ms")
    shapes = {str(c): ens.get_shape(c) for c in range(3)}
    ens_branches = ens.branches()
    extra = {}
    if warm:
        times = ens.times()
        st = [ens.local_stats(k) for k in range(len(ctxs))]
        fm = sum(x["fault_markers"] for x in st)
        extra["combined"] = {
            "fault_markers": n_fault, "residual": 1e-10, "dt_range_s": [COSEISMIC_DT, synth.YEAR],
            "last_dt_s": {"min": float(times[:, 1].min()), "max": float(times[:, 1].max())},
            "rsf_iters_per_fault_marker": sum(x["rsf_iters"] for x in st) / max(fm, 1),
            "ductile_iters_per_marker": sum(x["ductile_iters"] for x in st) / max(sum(x["ductile_markers"]
                                                                                   for x in st), 1),
            "nonfinite": sum(x["nonfinite"] for x in st), "fallbacks": sum(x["fallbacks"] for x in st)}
        # the same batched step with the plain slip-rate start (no network)
        ens.set_local(warm=0, tol=1e-10)
        t_plain = timed(lambda: ens.step(graph=True), args.steps)
        log(f"plain-start batched step {t_plain * 1e3:.1f} ms")
        st = [ens.local_stats(k) for k in range(len(ctxs))]
        extra["combined"]["plain_start_batched_ms"] = t_plain * 1e3
        extra["combined"]["plain_rsf_iters_per_fault_marker"] = (sum(x["rsf_iters"] for x in st) /
                                                                 max(sum(x["fault_markers"] for x in st), 1))
        extra["combined"]["warm_over_plain_wall"] = t_plain / min(t_default, t_tuned)
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
           "config": {"workload": (f"ensemble-{len(ctxs)}-production" + ("-coseismic-warm" if warm else "")
                                   if prod else
                                   f"ensemble-{args.models}" + ("-small" if args.small_models else "")),
                      "models": len(ctxs), "markers": n_total,
                      "nodes": nodes_total, "markers_min": min(sizes), "markers_max": max(sizes),
                      "grids": ("1501x301 refined to 500 m, 36 markers per cell, one cloud, 2-8 cm/yr"
                                if prod else
                                ("41x21 .. 161x61" if args.small_models else "201x51 .. 801x201")
                                + " nodes, 4-36 markers per cell"),
                      "dt_s": "adaptive 1e-3 s .. 1 yr" if warm else synth.YEAR,
                      "step": ("marker2nodes, before/after solve, local update (network warm start), "
                               "nodes2markers + advection, copyarrays, calc_timestep" if warm else
                               "marker2nodes, before/after solve, nodes2markers + advection")},
           "device_memory_used_gb": None if resident is None else round(resident / 1e9, 1),
           "batched_graph_default_ms": t_default * 1e3, "batched_graph_tuned_ms": t_tuned * 1e3,
           "looped_single_model_ms": t_looped * 1e3,
           "branches_ms": {str(b): v * 1e3 for b, v in t_branch.items()}, "branches": ens_branches,
           "tuned_graph_only_ms": t_graph_only, "speedup_vs_looped": t_looped / best,
           "tuned_shapes": shapes, "setup_s": round(t_gen, 1), **extra}
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
    ap.add_argument("--warm", action="store_true",
                    help="ensemble with --production-models: the combined coseismic workload (local update "
                         "with the network warm start, adaptive steps)")
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
    # Per-kernel breakdown in a pass of its own, with HIP events around every
    # kernel (each event pair adds a few microseconds of stream time).
    prof_steps = max(2, min(args.steps, 10))
    ctx.enable_timing(True)
    ctx.reset_timing()
    ctx.synchronize()
    for _ in range(prof_steps):
        step(ts)
        ts += dt
    ctx.synchronize()
    ctx.enable_timing(False)
    kern = {}
    for k in KERNELS:
        ms, nl = ctx.kernel_time(k)
        if nl:
            kern[k] = {"ms_total": ms, "launches": nl, "ms_per_step": ms / prof_steps}
    dom = max((k for k in ALG if k in kern), key=lambda k: kern[k]["ms_total"])
    # The timed steps: events around the dominant kernel only.
    ctx.timing_select([dom])
    ctx.enable_timing(True)
    ctx.reset_timing()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(ts)
        ts += dt
    ctx.synchronize()
    el = time.perf_counter() - t0
    log(f"{args.steps} steps: {el / args.steps * 1e3:.3f} ms per step")
    ctx.enable_timing(False)
    dom_ms, dom_launches = ctx.kernel_time(dom)
    ctx.timing_select(None)
    n, nn = cfg.n, cfg.grid.nodenum
    ms_step = el / args.steps * 1e3
    f_fault = float(np.mean(cfg.rsf["fault"])) if cfg.rsf is not None else 0.0
    bytes_step = n * (B_STEP_MARKER + B_STEP_FAULT * f_fault) + nn * B_STEP_NODE
    bytes_step_kernels = n * B_STEP_MARKER_KERNELS + nn * B_STEP_NODE
    per_launch_bytes = n * ALG[dom][0] + nn * ALG[dom][1]
    avg_ms = dom_ms / max(dom_launches, 1)
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
    # what limits the dominant kernel: its FP64 VALU issue bound over its time
    # against its counted HBM traffic over its time at peak (both counter
    # figures from the committed profiles of this workload)
    valu_frac = valu[dom]["frac"] if valu and dom in valu else None
    hbm_frac_traffic = traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None
    limiter = None
    if valu_frac is not None and hbm_frac_traffic is not None:
        limiter = "fp64-valu" if valu_frac > hbm_frac_traffic else "hbm"
    out = {
        "metric": "marker-steps/sec", "value": n * args.steps / el, "unit": "marker-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded marker cloud, analytic corner-flow nodal fields)",
        "config": {"workload": cfg.name, "markers": n, "nodes": nn, "markers_per_cell": 36
                   if args.config == "production" else 16, "dt_s": dt, "grid": "1501x301 refined to 500 m"
                   if args.config == "production" else "401x101 uniform"},
        "hbm_frac_step": bytes_step / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "bytes_step": bytes_step, "fault_fraction": f_fault,
        "hbm_frac_step_kernel_bytes": bytes_step_kernels / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        # bound: the roofline `achieved` is priced against (algorithmic bytes
        # over HBM peak; the path has no MFMA work); limiter: what the counter
        # evidence says limits the kernel
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_launch": per_launch_bytes, "avg_launch_ms": avg_ms,
                     "limiter": limiter, "valu_issue_frac": valu_frac, "hbm_traffic_frac": hbm_frac_traffic,
                     "counter_source": "committed rocprofv3 --pmc runs of this workload: profiles/traffic.json "
                                       "(FETCH_SIZE, WRITE_SIZE, copy-calibrated) and profiles/valu.json "
                                       "(SQ_INSTS_VALU, SQ_INSTS_VALU_TRANS_F64); commands in profiles/*.txt"},
        # from a separate pass of `kernel_breakdown_steps` steps with events
        # around every kernel; the timed steps carry events around `kernel`
        # (the dominant one) only
        "kernels_ms_per_step": {k: v["ms_per_step"] for k, v in kern.items()},
        "kernel_breakdown_steps": prof_steps,
        # m2n_props runs on a second stream beside the re-ordering ("sort"):
        # their event times include each other's share of the GPU
        "overlapped_kernels": ["m2n_props", "sort"],
        "pcie_inclusive": pcie,
    }
    if valu:
        out["valu_issue"] = valu
    if not args.no_next:
        ctx.set_async(False)
        log("post-solve routines")
        out["next_rows"] = next_rows(ctx, cfg)
    ctx.close()
    if not args.no_warm:
        log("warm-start evaluation")
        out["warm_start"] = warm_start_eval(lib, synth.YEAR)
    if not args.no_cpu:
        log("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(cfg, min(args.cpu_sample, cfg.n))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
