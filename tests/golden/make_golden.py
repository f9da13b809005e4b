"""Write the golden vectors of tests/golden/ by running the reference's own
hot-path routines, compiled in place from /root/reference/src
(oracle/_ref/ref_driver, built by oracle/Makefile), single-threaded.

    python tests/golden/make_golden.py            # step fixtures and pure functions
    python tests/golden/make_golden.py post       # post-solve nodal routines
    python tests/golden/make_golden.py bc         # boundary-condition tail
    python tests/golden/make_golden.py prn        # a .prn written by save_model
    python tests/golden/make_golden.py sticky     # sticky-air marks and statistics
    python tests/golden/make_golden.py addmarks   # add the marks to older step fixtures

Each fixture holds its complete input (grid, flags, rock table, markers, nodal
arrays) and the reference's output after the listed ops, so the tests never
need the reference or the generator that produced the inputs.  The committed
edge_*.npz step fixtures hold inputs from an earlier version of
velo2cycles_amd/synth.py; `addmarks` re-runs the reference on those stored
inputs (so the default run would write equally valid fixtures from today's
synthetic inputs rather than reproduce those files).
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from velo2cycles_amd import synth  # noqa: E402
from velo2cycles_amd.abi import MARKER_FIELDS, NODAL_FIELDS, Params  # noqa: E402
from oracle import refrun  # noqa: E402

DT = synth.YEAR
MAT = "nu dh dv ss mm gg n0 n1 ro bb aa cp kt kf kp ht".split()
# calc_timestep's scalars: the time-step limits and the sticky-air statistics
TSTEP_SCALARS = ("maxvel", "maxdphi", "maxnvp", "dtL", "dtphi", "dtd", "dtnvp")
STICKY_SCALARS = ("nr_1", "nr_0", "max_sigin_1", "max_sigin_0", "min_sigin_1", "min_sigin_0", "sum_sigin_1",
                  "sum2_sigin_1", "sum_sigin_0", "sum2_sigin_0", "max_visc_1", "max_visc_0", "min_visc_1",
                  "min_visc_0", "sum_visc_1", "sum2_visc_1", "sum_visc_0", "sum2_visc_0", "sigin_P1",
                  "sigin_P2", "visco_P1", "visco_P2")


def pack(cfg, ops, out_markers, out_nodal, raw, extra=None):
    d = {"gx": cfg.grid.gx, "gy": cfg.grid.gy,
         "xsize": np.array(cfg.grid.xsize), "ysize": np.array(cfg.grid.ysize),
         "timesum": np.array(cfg.timesum), "timestep": np.array(cfg.timestep),
         "params": np.array(json.dumps(cfg.params.as_dict())),
         "ops": np.array(json.dumps(list(ops))),
         "mat_ntypes": np.array(cfg.mats.ntypes),
         "mat_im": np.array(cfg.mats.im[:], dtype=np.int32)}
    for k in MAT:
        d["mat_" + k] = np.array(getattr(cfg.mats, k)[:])
    for f in MARKER_FIELDS:
        d["in_m_" + f] = cfg.markers[f]
        d["out_m_" + f] = out_markers[f]
    d["in_m_t"] = cfg.markers.t
    d["out_m_t"] = out_markers["t"]
    for k in NODAL_FIELDS:
        d["in_n_" + k] = cfg.nodal[k]
        d["out_n_" + k] = out_nodal[k]
    d["ret"] = raw["drv_ret"][:2]
    d["errcheck"] = raw["errcheck"]
    if extra:
        d.update(extra)
    return d


def case(name, cfg, ops):
    mk, nd, raw = refrun.run(cfg, ops)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **pack(cfg, ops, mk, nd, raw))
    print("wrote", name, cfg.n, "markers", cfg.grid.nodenum, "nodes", ops)


def post_cases():
    """Post-solve nodal routines on a synthetic post-solve state (with the
    degenerate nodes of synth.post_solve_nodal).  Markers are not used; the
    fixture keeps the nonzero input fields, the fields the ops changed and
    the scalars the reference leaves behind."""
    post = [f"stress:{DT}:0", f"c2b:{DT}", f"c2bold:{DT}", "b2c", "copy:0"]
    for name, pmode, rsf, refine, seed in (("post_rsf_p0", 0, 1, True, 3), ("post_rsf_p1", 1, 1, False, 4),
                                           ("post_dp", 0, 0, True, 5)):
        cfg = synth.config_grid(25, 15, 24e3, 14e3, 2, refine=refine, seed=seed)
        cfg.params.pmode = pmode
        cfg.params.RSFmode = rsf
        cfg.nodal = synth.post_solve_nodal(cfg, seed=seed + 7)
        # sticky-air marks for calc_timestep's statistics
        cfg.nodal["sticky01"] = np.random.default_rng(seed + 11).integers(0, 3, cfg.grid.nodenum).astype(float)
        runs = {"ops": post}
        for oo in (0, 1):
            runs[f"tstep{oo}"] = [f"tstep:{oo}"]
        d = {k: v for k, v in pack(cfg, [], {f: cfg.markers[f] for f in MARKER_FIELDS} | {"t": cfg.markers.t},
                                   cfg.nodal, refrun.run(cfg, [])[2]).items()
             if not k.startswith(("in_m_", "out_m_", "in_n_", "out_n_", "g_", "ret", "errcheck", "ops"))}
        for k in NODAL_FIELDS:
            if np.any(cfg.nodal[k] != 0):
                d["in_n_" + k] = cfg.nodal[k]
        extra = {"maxtmstep": ("d", np.array([DT]))}
        for tag, ops in runs.items():
            _, nd, raw = refrun.run(cfg, ops, extra=extra)
            pre = "" if tag == "ops" else tag + "_"
            d[pre + "opslist"] = np.array(json.dumps(ops))
            for k in NODAL_FIELDS:
                if not np.array_equal(nd[k], cfg.nodal[k], equal_nan=True):
                    d[pre + "out_n_" + k] = nd[k]
            d[pre + "ret"] = raw["drv_ret"][:3]
            for k in TSTEP_SCALARS + (STICKY_SCALARS if tag != "ops" else ()):
                d[pre + "g_" + k] = raw[k]
        d["maxtmstep"] = np.array(DT)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
        print("wrote", name, cfg.grid.nodenum, "nodes")


def bc_cases():
    """Boundary-condition tail (vxvy_boundarycorrection, P_boundarycorrection,
    tkrecalc) with the equations of synth.boundary_equations."""
    for name, comp, seed in (("bc_walls_chains", 0, 4), ("bc_compressible", 1, 5)):
        cfg = synth.config_grid(25, 15, 24e3, 14e3, 2, refine=True, seed=seed)
        cfg.params.compressiblemode = comp
        rng = np.random.default_rng(seed)
        nn = cfg.grid.nodenum
        inp = {"mvx": rng.normal(0, 1e-9, nn), "mvy": rng.normal(0, 1e-9, nn),
               "sppe": rng.normal(0, 1e6, nn), "tk": rng.uniform(273, 1600, nn)}
        cfg.nodal.update(inp)
        bm, bn, bv = synth.boundary_equations(cfg, seed)
        extra = {"bondm": ("l", bm), "bondn": ("l", bn.ravel()), "bondv": ("d", bv.ravel())}
        d = {"gx": cfg.grid.gx, "gy": cfg.grid.gy, "xsize": np.array(cfg.grid.xsize),
             "ysize": np.array(cfg.grid.ysize), "params": np.array(json.dumps(cfg.params.as_dict())),
             "bondm": bm, "bondn": bn, "bondv": bv}
        for k, v in inp.items():
            d["in_" + k] = v
        for op, fields in (("bcvxvy", ("mvx", "mvy")), ("bcp", ("sppe",)), ("bctk", ("tk",))):
            nd = refrun.run(cfg, [op], extra=extra)[1]
            for k in fields:
                d[f"out_{op}_{k}"] = nd[k]
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
        print("wrote", name, nn, "nodes", len(bn), "equations")


def sticky_cases():
    """sticky.npz: marker2nodes, compute_before_solve and calc_timestep(0)
    on a 53 x 53 refined grid (node (50, 50) is interior, so sigin_P1 and
    visco_P1 are set), 2 x 2 markers per cell of types 1, 2, 4, 7, 8, plus
    markers that mark one node several times (the last one wins), markers on
    cell midlines (no mark), in the last cell row (the mark wraps into the
    next column) and the last corner cell (the mark runs past nodenum),
    outside the grid and at NaN.  Keeps the input, the marks, nu (which the
    statistics read) and the statistics."""
    cfg = synth.config_grid(53, 53, 52e3, 52e3, 2, refine=True, seed=77)
    cfg.params.RSFmode = 0
    g = cfg.grid
    rng = np.random.default_rng(77)
    base = cfg.markers
    n0 = base.n
    extra_n = 600
    fields = {f: np.concatenate([base[f], base[f][rng.integers(0, n0, extra_n)]]) for f in MARKER_FIELDS}
    t = np.concatenate([base.t, np.zeros(extra_n, np.int8)])
    t[:] = rng.choice(np.array([1, 2, 4, 7, 8], np.int8), t.size, p=[0.35, 0.35, 0.1, 0.1, 0.1])
    x, y = fields["x"], fields["y"]
    j = n0
    # repeated marks: copies of earlier markers, jittered within their quadrant
    x[j:j + 300] += rng.uniform(-1.0, 1.0, 300)
    y[j:j + 300] += rng.uniform(-1.0, 1.0, 300)
    j += 300
    # cell midlines in x, in y, and both
    for k in range(60):
        cx, cy = rng.integers(0, g.xnumx - 1), rng.integers(0, g.ynumy - 1)
        mx, my = (g.gx[cx] + g.gx[cx + 1]) / 2, (g.gy[cy] + g.gy[cy + 1]) / 2
        x[j], y[j] = (mx, y[j]) if k % 3 == 0 else (x[j], my) if k % 3 == 1 else (mx, my)
        j += 1
    # the last cell row (celdy > 0.5) and the last corner cell
    for k in range(80):
        cx = rng.integers(0, g.xnumx - 1) if k < 60 else g.xnumx - 2
        cy = g.ynumy - 2
        x[j] = g.gx[cx] + rng.uniform(0.05, 0.95) * (g.gx[cx + 1] - g.gx[cx])
        y[j] = g.gy[cy] + rng.uniform(0.6, 0.95) * (g.gy[cy + 1] - g.gy[cy])
        j += 1
    # outside the grid and NaN
    for k in range(40):
        if k % 4 == 0:
            x[j] = g.gx[0] - rng.uniform(1.0, 5e3)
        elif k % 4 == 1:
            x[j] = g.gx[-1] + rng.uniform(1.0, 5e3)
        elif k % 4 == 2:
            y[j] = g.gy[0] - rng.uniform(1.0, 5e3)
        else:
            x[j] = np.nan
        j += 1
    # node (50, 50) marked 1 by the last marker of the cloud
    x[-1] = g.gx[50] + 0.25 * (g.gx[51] - g.gx[50])
    y[-1] = g.gy[49] + 0.25 * (g.gy[50] - g.gy[49])
    t[-1] = 1
    fields["t"] = t
    cfg.markers = synth.MarkerArrays(t.size, fields)
    nn = g.nodenum
    cfg.nodal["sxxs"] = rng.normal(0, 2e7, nn)
    cfg.nodal["sxy"] = rng.normal(0, 2e7, nn)
    # compute_before_solve sets the nu the statistics read (marker2nodes
    # clears it, interpolation.c:72)
    ops = ["m2n", f"before:{DT}:0", "tstep:0"]
    mk, nd, raw = refrun.run(cfg, ops, extra={"maxtmstep": ("d", np.array([DT]))})
    d = {k: v for k, v in pack(cfg, ops, mk, nd, raw).items() if not k.startswith(("out_m_", "out_n_"))}
    d = {k: v for k, v in d.items() if not (k.startswith("in_n_") and not np.any(v != 0))}
    d["out_n_sticky01"] = nd["sticky01"]
    d["out_n_nu"] = nd["nu"]
    for k in STICKY_SCALARS:
        d["g_" + k] = raw[k]
    np.savez_compressed(os.path.join(HERE, "sticky.npz"), **d)
    print("wrote sticky", cfg.n, "markers", nn, "nodes", int(np.sum(nd["sticky01"] != 0)), "marked")


def add_marks():
    """Add out_n_sticky01 to the step fixtures written before the marks were
    recorded: the reference re-runs each fixture's own stored input and ops,
    every stored output must come back bit for bit, and the marks its
    marker2nodes leaves are added."""
    from tests.util import load_fixture
    for name in ("edge_step_t0", "edge_step_t1", "edge_euler_jaumann_densimod", "edge_dp_noantidiff",
                 "edge_rk3_powerlaw"):
        path = os.path.join(HERE, name + ".npz")
        cfg, ops, _, z = load_fixture(path)
        d = {k: z[k] for k in z.files}
        if "out_n_sticky01" in d:
            continue
        mk, nd, raw = refrun.run(cfg, ops)
        for k in d:
            if k.startswith("out_m_"):
                got = mk["t"] if k == "out_m_t" else mk[k[6:]]
            elif k.startswith("out_n_"):
                got = nd[k[6:]]
            else:
                continue
            assert np.array_equal(got, d[k], equal_nan=True), f"{name}: {k} does not reproduce"
        d["out_n_sticky01"] = nd["sticky01"]
        np.savez_compressed(path, **d)
        print("added marks to", name, int(np.sum(nd["sticky01"] != 0)), "marked")


def prn_cases():
    """prn_small.prn: the reference's save_model output for a seeded model
    (synth.prn_model of a 9 x 7 grid, 2 x 2 markers per cell, boundary
    equations), plus prn_small.npz listing some of its values."""
    import shutil
    import tempfile
    cfg = synth.config_grid(9, 7, 8e3, 6e3, 2, refine=True, seed=91)
    d = synth.prn_model(cfg, 91)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "r.prn")
        refrun.run(cfg, [f"save:{d['cycle']}:{path}"], extra=refrun.prn_extra(d))
        shutil.copyfile(path, os.path.join(HERE, "prn_small.prn"))
    np.savez_compressed(os.path.join(HERE, "prn_small.npz"), gx=d["gx"], gy=d["gy"], markx=d["mark"]["x"],
                        markC0=d["mark"]["C0"], RK_phidts_final=d["node"]["RK_phidts_final"], markt=d["t"],
                        cycle=np.array(d["cycle"]))
    print("wrote prn_small", cfg.grid.nodenum, "nodes", d["marknum"], "markers")


def main():
    assert refrun.available(), "build oracle/_ref first (make -C oracle)"
    if sys.argv[1:] == ["prn"]:
        prn_cases()
        return
    if sys.argv[1:] == ["post"]:
        post_cases()
        return
    if sys.argv[1:] == ["bc"]:
        bc_cases()
        return
    if sys.argv[1:] == ["sticky"]:
        sticky_cases()
        return
    if sys.argv[1:] == ["addmarks"]:
        add_marks()
        return
    step = ["m2n", f"before:{DT}:0", f"after:{DT}:0", "n2m", "move"]
    cfg = synth.config_edge()
    case("edge_step_t0", cfg, step)
    cfg = synth.config_edge()
    cfg.timesum = DT
    case("edge_step_t1", cfg, step)
    cfg = synth.config_edge(seed=99)
    cfg.timesum = DT
    cfg.params.markmod = 1
    cfg.params.stoksmod = 2
    cfg.params.densimod = 1
    cfg.params.outgrid = 0
    case("edge_euler_jaumann_densimod", cfg, step)
    cfg = synth.config_edge(seed=5)
    cfg.timesum = DT
    cfg.params.RSFmode = 0
    cfg.params.antidiffinter = 0
    cfg.params.tempmod = 1
    case("edge_dp_noantidiff", cfg, step)
    cfg = synth.config_edge(seed=6)
    cfg.timesum = DT
    cfg.params.power_law_RSF = 1
    cfg.params.rkmod = 3
    cfg.params.inertyn = 1
    cfg.params.pmode = 1
    cfg.params.Dphis = 1.0
    cfg.params.compressiblemode = 1
    case("edge_rk3_powerlaw", cfg, ["m2n", f"before:{DT / 2}:1", f"after:{DT / 2}:1",
                                    f"before:{DT / 2}:2", f"after:{DT / 2}:2", f"before:{DT}:3"])

    # pure-function vectors: ductilevisc and allinteri
    cfg = synth.config_edge()
    rng = np.random.default_rng(2026)
    n = 3000
    t = rng.integers(0, 9, n).astype(np.int8)
    sxy = rng.normal(0, 3e7, n)
    sxx = rng.normal(0, 3e7, n)
    sxx[:40] = 0.0
    sxy[:40] = 0.0
    tk = rng.uniform(273, 1900, n)
    pb = rng.uniform(1, 4e4, n)
    _, _, raw = refrun.run(cfg, ["ductile"], extra={
        "drv_n": ("l", np.array([n], np.int64)), "dv_type": ("c", t), "dv_sxy": ("d", sxy),
        "dv_sxx": ("d", sxx), "dv_tk": ("d", tk), "dv_pb": ("d", pb)})
    base = pack(cfg, [], {f: cfg.markers[f] for f in MARKER_FIELDS} | {"t": cfg.markers.t},
                cfg.nodal, raw)
    keep = {k: v for k, v in base.items() if not k.startswith(("in_m_", "out_m_", "out_n_"))}
    keep.update(dv_type=t, dv_sxy=sxy, dv_sxx=sxx, dv_tk=tk, dv_pb=pb, dv_out=raw["dv_out"][:n])
    np.savez_compressed(os.path.join(HERE, "ductilevisc.npz"), **keep)
    print("wrote ductilevisc", n)

    g = cfg.grid
    x = rng.uniform(-0.05 * g.xsize, 1.05 * g.xsize, n)
    y = rng.uniform(-0.05 * g.ysize, 1.05 * g.ysize, n)
    x[:13] = g.gx
    y[13:22] = g.gy
    _, _, raw = refrun.run(cfg, ["allinteri"], extra={
        "drv_n": ("l", np.array([n], np.int64)), "ai_x": ("d", x), "ai_y": ("d", y)})
    keep = {k: v for k, v in base.items() if not k.startswith(("in_m_", "out_m_", "out_n_"))}
    keep.update(ai_x=x, ai_y=y, ai_vx=raw["ai_vx"][:n], ai_vy=raw["ai_vy"][:n],
                ai_esp=raw["ai_esp"][:n], ai_ee=raw["ai_ee"][:n])
    np.savez_compressed(os.path.join(HERE, "allinteri.npz"), **keep)
    print("wrote allinteri", n)


if __name__ == "__main__":
    main()
