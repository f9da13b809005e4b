"""Shared helpers of the parity tests: running a Config through the oracle and
through libv2c, and the field-by-field comparison bar."""
from __future__ import annotations

import os

import numpy as np

from velo2cycles_amd.abi import MARKER_FIELDS, NODAL_FIELDS, Context
from oracle.pyoracle import Oracle

# FP64 tolerance: |gpu - oracle| <= REL * max(|oracle|, rms(field)) elementwise
# (the rms floor keeps signed fields with near-zero entries meaningful).
REL = 1e-12
# positions and integer state must match bit for bit
EXACT_MARKER = ("x", "y")


# calc_timestep's sticky-air statistics (v2c_sticky_stats): counts, extrema
# and P values, then the sums
STICKY_EXACT = ("nr_1", "nr_0")
STICKY_KEYS = ("nr_1", "nr_0", "max_sigin_1", "max_sigin_0", "min_sigin_1", "min_sigin_0", "sum_sigin_1",
               "sum2_sigin_1", "sum_sigin_0", "sum2_sigin_0", "max_visc_1", "max_visc_0", "min_visc_1",
               "min_visc_0", "sum_visc_1", "sum2_visc_1", "sum_visc_0", "sum2_visc_0", "sigin_P1", "sigin_P2",
               "visco_P1", "visco_P2")


def sticky_mismatches(got: dict, exp: dict, rel: float = 1e-12) -> list:
    """Statistics that differ: counts exactly, the rest within rel (the
    device's square root, log10 and summation order against the serial
    loop's pow and order), NaN only against NaN."""
    bad = []
    for k in STICKY_KEYS:
        a, b = float(got[k]), float(exp[k])
        if k in STICKY_EXACT:
            ok = a == b
        elif np.isnan(a) or np.isnan(b):
            ok = np.isnan(a) and np.isnan(b)
        else:
            ok = a == b or abs(a - b) <= rel * max(abs(a), abs(b))
        if not ok:
            bad.append(f"{k}: {a!r} vs {b!r}")
    return bad


def rel_errors(a: np.ndarray, b: np.ndarray) -> tuple[float, float, int]:
    """(max scaled error, rms scaled error, #nan mismatches) of a against b."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nan_mis = int(np.sum(np.isnan(a) != np.isnan(b)))
    ok = ~(np.isnan(a) | np.isnan(b))
    if not ok.any():
        return 0.0, 0.0, nan_mis
    a, b = a[ok], b[ok]
    inf = np.isinf(a) | np.isinf(b)
    nan_mis += int(np.sum(a[inf] != b[inf]))
    a, b = a[~inf], b[~inf]
    if a.size == 0:
        return 0.0, 0.0, nan_mis
    rms = np.sqrt(np.mean(b * b))
    scale = np.maximum(np.abs(b), rms)
    scale[scale == 0] = 1.0
    e = np.abs(a - b) / scale
    return float(e.max()), float(np.sqrt(np.mean(e * e))), nan_mis


def _field(z, name, n):
    """A nodal array of a golden file; fields newer than the file (which its
    operations never touch) are zero, as the reference's globals start."""
    return z[name] if name in z.files else np.zeros(n)


def compare(gpu_m, ora_m, gpu_n, ora_n, rel=REL, nodal_fields=NODAL_FIELDS, skip=()):
    """Return a list of failure strings (empty == parity); markers None: nodal only."""
    bad = []
    if gpu_m is not None and not np.array_equal(gpu_m.t, ora_m.t):
        bad.append("marker types differ")
    for f in (MARKER_FIELDS if gpu_m is not None else ()):
        if f in skip:
            continue
        a, b = gpu_m[f], ora_m[f]
        if f in EXACT_MARKER:
            if not np.array_equal(a, b, equal_nan=True):
                d = np.nanmax(np.abs(a - b))
                bad.append(f"marker {f}: not bit-exact (max |d| {d:.3e}, {int(np.sum(a != b))} differ)")
            continue
        mx, _, nm = rel_errors(a, b)
        if mx > rel or nm:
            bad.append(f"marker {f}: max scaled err {mx:.3e}, nan mismatches {nm}")
    for k in nodal_fields:
        if k in skip:
            continue
        mx, _, nm = rel_errors(gpu_n[k], ora_n[k])
        if mx > rel or nm:
            bad.append(f"nodal {k}: max scaled err {mx:.3e}, nan mismatches {nm}")
    return bad


def run_oracle(cfg, ops, markers=None, nodal=None):
    o = Oracle(cfg, markers, nodal)
    ret = {}
    for op in ops:
        name, *args = op.split(":")
        if name == "m2n":
            # the reference's marker2nodes ends with the sticky-air marks
            o.marker2nodes()
            o.sticky_marks()
        elif name == "n2m":
            o.nodes2markers()
        elif name == "move":
            o.moveandrotatemarkers()
        elif name == "n2m_move":
            o.nodes2markers()
            o.moveandrotatemarkers()
        elif name == "before":
            ret["before"] = o.compute_before_solve(float(args[0]), int(args[1]))
        elif name == "after":
            ret["after"] = o.compute_after_solve(float(args[0]), int(args[1]))
        elif name == "time":
            o.set_time(float(args[0]), float(args[1]))
        elif name == "stress":
            ret["stress"] = o.recompute_stress_strainrate(float(args[0]), int(args[1]))
        elif name == "c2b":
            o.interpolate_center2basic(float(args[0]))
        elif name == "c2bold":
            o.interpolate_center2basic_old()
        elif name == "b2c":
            o.interpolate_basic2center()
        elif name == "copy":
            o.copyarrays(int(args[0]))
        elif name == "tstep":
            rc, r = o.calc_timestep(int(args[0]))
            assert rc == 0, rc
            ret["tstep"] = r
            ret["sticky"] = o.sticky_stats()
        elif name == "heat":
            o.temperature_markers(float(args[0]))
        else:
            raise ValueError(op)
    return o.markers, o.nodal, ret


def run_gpu(lib, cfg, ops, markers=None, nodal=None, ctx=None):
    ctx = ctx or Context(cfg.grid, cfg.params, cfg.mats, capacity=max(cfg.n, 1), lib=lib)
    ctx.set_time(cfg.timesum, cfg.timestep)
    ctx.set_all_nodal(nodal or cfg.nodal)
    m = (markers or cfg.markers).copy()
    ctx.upload_markers(m)
    ret = {}
    for op in ops:
        name, *args = op.split(":")
        if name == "m2n":
            ctx.marker2nodes()
            ctx.sticky_marks()
        elif name == "n2m":
            ctx.nodes2markers()
        elif name == "move":
            ctx.moveandrotatemarkers()
        elif name == "n2m_move":
            ctx.nodes2markers_advect()
        elif name == "before":
            ret["before"] = ctx.compute_before_solve(float(args[0]), int(args[1]))
        elif name == "after":
            ret["after"] = ctx.compute_after_solve(float(args[0]), int(args[1]))
        elif name == "time":
            ctx.set_time(float(args[0]), float(args[1]))
        elif name == "stress":
            ret["stress"] = ctx.recompute_stress_strainrate(float(args[0]), int(args[1]))
        elif name == "c2b":
            ctx.interpolate_center2basic(float(args[0]))
        elif name == "c2bold":
            ctx.interpolate_center2basic_old()
        elif name == "b2c":
            ctx.interpolate_basic2center()
        elif name == "copy":
            ctx.copyarrays(int(args[0]))
        elif name == "tstep":
            ret["tstep"] = ctx.calc_timestep(int(args[0]))
            ret["sticky"] = ctx.sticky_stats()
        elif name == "heat":
            ctx.temperature_markers(float(args[0]))
        else:
            raise ValueError(op)
    ctx.download_markers(m)
    nod = ctx.get_all_nodal()
    return m, nod, ret, ctx


def load_fixture(path):
    """Rebuild (cfg, ops, expected markers, expected nodal, ret) from a golden .npz."""
    import json
    from velo2cycles_amd import synth
    from velo2cycles_amd.abi import GridArrays, MarkerArrays, Materials, Params
    z = np.load(path, allow_pickle=False)
    grid = GridArrays(z["gx"], z["gy"], float(z["xsize"]), float(z["ysize"]))
    p = Params()
    for k, v in json.loads(str(z["params"])).items():
        setattr(p, k, v)
    mats = Materials()
    mats.ntypes = int(z["mat_ntypes"])
    for k in "nu dh dv ss mm gg n0 n1 ro bb aa cp kt kf kp ht".split():
        getattr(mats, k)[:] = list(z["mat_" + k])
    mats.im[:] = [int(v) for v in z["mat_im"]]
    out = {}
    if "in_m_x" in z:
        fields = {f: z["in_m_" + f] for f in MARKER_FIELDS}
        fields["t"] = z["in_m_t"]
        markers = MarkerArrays(len(z["in_m_x"]), fields)
        if "out_m_x" in z.files:
            ef = {f: z["out_m_" + f] for f in MARKER_FIELDS}
            ef["t"] = z["out_m_t"]
            out["markers"] = MarkerArrays(len(z["in_m_x"]), ef)
            out["nodal"] = {k: _field(z, "out_n_" + k, grid.nodenum) for k in NODAL_FIELDS}
        else:
            # a fixture that keeps some outputs only
            out["nodal"] = {k[6:]: z[k] for k in z.files if k.startswith("out_n_")}
    else:
        markers = MarkerArrays(0)
    nodal = {k: np.array(_field(z, "in_n_" + k, grid.nodenum)) for k in NODAL_FIELDS}
    cfg = synth.Config(os.path.basename(path), grid, p, mats, markers, nodal,
                       float(z["timestep"]), float(z["timesum"]))
    ops = json.loads(str(z["ops"]))
    return cfg, ops, out, z


def load_post_fixture(path):
    """(cfg, runs) from a post-solve golden .npz: cfg holds the grid, flags
    and nodal input (fields absent from the file are zero); runs maps a tag
    ('ops', 'tstep0', 'tstep1') to (ops, {field: expected output} for the
    fields the ops changed, reference scalars)."""
    import json
    from velo2cycles_amd import synth
    from velo2cycles_amd.abi import GridArrays, Materials, Params, MarkerArrays
    z = np.load(path)
    grid = GridArrays(z["gx"], z["gy"], float(z["xsize"]), float(z["ysize"]))
    p = Params()
    for k, v in json.loads(str(z["params"])).items():
        setattr(p, k, v)
    mats = Materials()
    mats.ntypes = int(z["mat_ntypes"])
    for k in "nu dh dv ss mm gg n0 n1 ro bb aa cp kt kf kp ht".split():
        getattr(mats, k)[:] = list(z["mat_" + k])
    mats.im[:] = [int(v) for v in z["mat_im"]]
    nodal = {k: np.array(_field(z, "in_n_" + k, grid.nodenum)) for k in NODAL_FIELDS}
    cfg = synth.Config(os.path.basename(path), grid, p, mats, MarkerArrays(0), nodal,
                       float(z["timestep"]), float(z["timesum"]))
    runs = {}
    for tag, pre in (("ops", ""), ("tstep0", "tstep0_"), ("tstep1", "tstep1_")):
        ops = json.loads(str(z[pre + "opslist"]))
        out = {k[len(pre) + 6:]: z[k] for k in z.files if k.startswith(pre + "out_n_")}
        scal = {k[len(pre) + 2:]: float(z[k][0]) for k in z.files if k.startswith(pre + "g_")}
        scal["dt"] = float(z[pre + "ret"][2])
        runs[tag] = (ops, out, scal)
    return cfg, runs, float(z["maxtmstep"])


def check_deltavs(got, ora, solved, final, dt):
    """Failures of copyarrays' deltaVs = (log(solved) - log(final)) / dt.
    It subtracts two nearly equal logarithms, so last-bit differences of its
    inputs show undamped (one bit of a logarithm is |log V| eps / dt).
    Allowed per node: the first-order effect of the inputs' relative
    differences, (r_solved + r_final) / dt, with solved = (device, oracle)
    and final = (device, oracle) the two slip-rate arrays copyarrays read,
    plus two last bits of each logarithm, 8 (|log V| + 1) eps / dt, plus
    1e-15 relative for the step; equal values (infinities of zero slip
    rates among them) and both NaN (negative ones) pass."""
    def rel(a, b):
        return np.abs(a - b) / np.where(np.abs(b) > 0, np.abs(b), 1.0)

    r = rel(solved[0], solved[1]) + rel(final[0], final[1])
    v = np.asarray(solved[1], dtype=np.float64)
    lv = np.abs(np.log(np.where(v > 0, v, 1.0)))
    allow = (r + 8.0 * np.finfo(np.float64).eps * (lv + 1.0)) / dt + 1e-15 * np.abs(ora)
    both_nan = np.isnan(got) & np.isnan(ora)
    with np.errstate(invalid="ignore"):
        ok = both_nan | (got == ora) | (np.abs(got - ora) <= allow)
    if ok.all():
        return []
    with np.errstate(invalid="ignore"):
        worst = float(np.nanmax(np.where(ok, 0.0, np.abs(got - ora) / allow)))
    return [f"nodal deltaVs: {int((~ok).sum())} nodes beyond the conditioning bound (worst {worst:.3g}x)"]
