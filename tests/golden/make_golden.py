"""Write the golden vectors of tests/golden/ by running the reference's own
hot-path routines, compiled in place from /root/reference/src
(oracle/_ref/ref_driver, built by oracle/Makefile), single-threaded.

    python tests/golden/make_golden.py            # step fixtures and pure functions
    python tests/golden/make_golden.py post       # post-solve nodal routines
    python tests/golden/make_golden.py bc         # boundary-condition tail
    python tests/golden/make_golden.py prn        # a .prn written by save_model

Each fixture holds its complete input (grid, flags, rock table, markers, nodal
arrays) and the reference's output after the listed ops, so the tests never
need the reference or the generator that produced the inputs.
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
            for k in ("maxvel", "maxdphi", "maxnvp", "dtL", "dtphi", "dtd", "dtnvp"):
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
