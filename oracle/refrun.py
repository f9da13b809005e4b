"""Run the reference's own hot-path routines (oracle/_ref/ref_driver, compiled
in place from /root/reference/src by oracle/Makefile) on a Config.

TEST INFRASTRUCTURE ONLY, and only in the container that holds the reference
sources: it writes golden vectors (tests/golden/) and pins the oracle.  The
binaries under oracle/_ref never leave that container.  The same input/output
record format drives tests/native/dropin_harness (the builder's stand-in for
the reference program around the reference-side binding), which is what runs
on a GPU: `write_input`, `run_file` and `split_output` serve both.
"""
from __future__ import annotations

import os
import struct
import subprocess
import tempfile

import numpy as np

from velo2cycles_amd.abi import MARKER_FIELDS, NODAL_FIELDS

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(ORACLE_DIR, "_ref", "ref_driver")
# the binding compiled into the reference translation unit (a link check of
# the binding's declarations against the reference globals; not run)
DROPIN = os.path.join(ORACLE_DIR, "_ref", "dropin_driver")
REF_SRC = "/root/reference/src"

# reference capacity (headevps.c:10-20)
MAXXLN, MAXYLN, MAXNOD, MAXMRK = 300, 120, 36000, 1080000

_MARK_NAME = {f: ("mark" + f) for f in MARKER_FIELDS}
_MARK_NAME.update(a="mark_a", b="mark_b", L="mark_L", L0="mark_L0", V0="mark_V0",
                  fric_static="fric_static", C0="markC0", C="markC", t="markt")

_PARAM_INT = ("markmod stoksmod outgrid densimod interpolationmod antidiffinter rkmod RSFmode "
              "power_law_RSF reloadfromDP compressiblemode stickymod inertyn pmode tempmod").split()
_PARAM_DBL = ("nubeg nuend tdeep zdeep nstress Dphis half_range start_updip end_updip start_downdip "
              "end_downdip a_vw b_vw L_vw V0_vw a_vs b_vs L_vs V0_vs a_hr b_hr L_hr V0_hr").split()
_MAT = "nu dh dv ss mm gg n0 n1 ro bb aa cp kt kf kp ht".split()


def available() -> bool:
    if os.path.exists(REF_SRC) and not os.path.exists(DRIVER):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=False)
    return os.path.exists(DRIVER) and os.path.exists(REF_SRC)


def built(driver=DRIVER) -> bool:
    return os.path.exists(driver)


def legal(cfg) -> bool:
    g = cfg.grid
    return g.xnumx <= MAXXLN and g.ynumy <= MAXYLN and g.nodenum <= MAXNOD and cfg.n < MAXMRK


def _rec(f, name, kind, arr):
    arr = np.ascontiguousarray(arr)
    f.write(name.encode().ljust(32, b"\0"))
    f.write(kind.encode())
    f.write(struct.pack("<q", arr.size))
    f.write(arr.tobytes())


def _read(path):
    out = {}
    dt = {"d": np.float64, "l": np.int64, "i": np.int32, "c": np.int8}
    with open(path, "rb") as f:
        while True:
            h = f.read(32)
            if len(h) < 32:
                break
            name = h.rstrip(b"\0").decode()
            kind = f.read(1).decode()
            (cnt,) = struct.unpack("<q", f.read(8))
            a = np.frombuffer(f.read(cnt * np.dtype(dt[kind]).itemsize), dtype=dt[kind]).copy()
            out[name] = a
    return out


def read_records(path):
    """The records of a driver input or output file (name -> array)."""
    return _read(path)


def write_input(path, cfg, ops=None, markers=None, nodal=None, timesum=None, timestep=None, extra=None):
    """Write cfg (+ overrides and `extra` records) as a driver input file."""
    m = markers or cfg.markers
    nod = nodal or cfg.nodal
    p = cfg.params
    g = cfg.grid
    with open(path, "wb") as f:
        _rec(f, "gx", "d", g.gx)
        _rec(f, "gy", "d", g.gy)
        _rec(f, "xnumx", "l", np.array([g.xnumx], np.int64))
        _rec(f, "ynumy", "l", np.array([g.ynumy], np.int64))
        _rec(f, "marknum", "l", np.array([m.n], np.int64))
        _rec(f, "xsize", "d", np.array([g.xsize]))
        _rec(f, "ysize", "d", np.array([g.ysize]))
        _rec(f, "timesum", "d", np.array([cfg.timesum if timesum is None else timesum]))
        _rec(f, "timestep", "d", np.array([cfg.timestep if timestep is None else timestep]))
        for k in _PARAM_INT:
            _rec(f, k, "i", np.array([getattr(p, k)], np.int32))
        for k in _PARAM_DBL:
            _rec(f, k, "d", np.array([getattr(p, k)]))
        eps = np.zeros(100)
        eps[2] = p.incoming_T
        _rec(f, "eps", "d", eps)
        for k in _MAT:
            _rec(f, "mark" + k, "d", np.array(getattr(cfg.mats, k)[:200]))
        _rec(f, "markim", "i", np.array(cfg.mats.im[:200], np.int32))
        for fld in MARKER_FIELDS:
            _rec(f, _MARK_NAME[fld], "d", m[fld])
        _rec(f, "markt", "c", m.t)
        for k in NODAL_FIELDS:
            _rec(f, k, "d", nod[k])
        for k, (kind, arr) in (extra or {}).items():
            _rec(f, k, kind, arr)


def run_file(driver, fin, ops, threads=1):
    """Run `driver IN OUT ops...` on an input file; returns the raw output
    records (name -> array)."""
    with tempfile.TemporaryDirectory() as td:
        fout = os.path.join(td, "out.bin")
        # one thread by default: the reference's merge order
        # (interpolation.c:631-674) and its racy antidiffusion scatter (:942)
        # make multithreaded runs vary (threads > 1 is for timing only)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        with open(os.devnull, "w") as dn:
            subprocess.run([driver, os.path.abspath(fin), fout] + list(ops), check=True, stdout=dn, env=env,
                           cwd=td)
        return _read(fout)


def split_output(raw, n, nodenum):
    """(markers dict by library field names, nodal dict by reference names)."""
    mk = {fld: raw[_MARK_NAME[fld]][:n].copy() for fld in MARKER_FIELDS}
    mk["t"] = raw["markt"][:n].copy()
    nd = {k: raw[k][:nodenum].copy() for k in NODAL_FIELDS}
    return mk, nd


def run(cfg, ops, markers=None, nodal=None, timesum=None, timestep=None, extra=None, driver=DRIVER,
        threads=1):
    """Write cfg (+ overrides) into the reference globals, run `ops`, return
    (markers dict, nodal dict, raw dict).  Markers are returned by the library's
    field names, nodal arrays by reference names (length nodenum)."""
    assert built(driver) or (driver == DRIVER and available()), "reference driver unavailable"
    assert legal(cfg), "size exceeds the reference capacity macros"
    m = markers or cfg.markers
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        write_input(fin, cfg, markers=markers, nodal=nodal, timesum=timesum, timestep=timestep, extra=extra)
        raw = run_file(driver, fin, ops, threads=threads)
    mk, nd = split_output(raw, m.n, cfg.grid.nodenum)
    return mk, nd, raw


# .prn marker columns -> reference globals (load_save.c:494-509, input_RSF.c:97-101)
PRN_MARK_GLOBAL = {k: ("mark" + k) for k in "x y k e xx v xy p exx exy d w vx vy phi C C0".split()}
PRN_MARK_GLOBAL.update(L="mark_L", L0="mark_L0")


def prn_extra(d):
    """Driver input records that set the reference's globals to the .prn model
    `d` (velo2cycles_amd.prn.read's dict); pass as run(..., extra=...)."""
    ex = {"mnumx": ("l", np.array([d["mnumx"]], np.int64)), "mnumy": ("l", np.array([d["mnumy"]], np.int64)),
          "pinit": ("d", np.array([d["pinit"]])), "pkf": ("d", np.asarray(d["pkf"], np.float64)),
          "GXKOEF": ("d", np.array([d["GXKOEF"]])), "GYKOEF": ("d", np.array([d["GYKOEF"]])),
          "rocknum": ("i", np.array([d["rocknum"]], np.int32)),
          "bondnum": ("l", np.array([d["bondnum"]], np.int64)),
          "timesave": ("d", np.asarray(d["timesave"], np.float64)),
          "markim": ("i", np.asarray(d["markim"], np.int32)),
          "bondm": ("l", np.asarray(d["bondm"], np.int64)),
          "bondn": ("l", np.asarray(d["bondn"], np.int64).ravel()),
          "bondv": ("d", np.asarray(d["bondv"], np.float64).ravel()),
          "markt": ("c", np.asarray(d["t"], np.int8))}
    for k, v in d["rock"].items():
        ex["mark" + k] = ("d", np.asarray(v, np.float64))
    for k, v in d["node"].items():
        if k != "RK_sliprs_final_a":
            ex[k] = ("d", np.asarray(v, np.float64))
    for k, v in d["mark"].items():
        ex[PRN_MARK_GLOBAL[k]] = ("d", np.asarray(v, np.float64))
    return ex
