"""The reference's binary model file (.prn): read and write through the C-ABI
(v2c_prn_read / v2c_prn_write, include/v2c.h), and map its contents onto the
marker and nodal arrays a Context takes.

Layout (load_save.c:164-347 reader, 353-527 writer; input_RSF.c:94-132 the
extra per-node and per-marker values): 4 size bytes (int, long, float,
double), the grid scalars, the rock table (markim + 27 double columns), one
200-byte record per node (19 doubles, 4 longs of bondm, RK_sliprs_final and
RK_phidts_final), gx, gy, boundary equations 1..bondnum-1, timesave[5], one
153-byte record per marker (14 doubles, the type byte, 5 doubles).
"""
import ctypes as C

import numpy as np

from .abi import (ERRORS, MARKER_FIELDS, NODAL_FIELDS, V2C_MAX_TYPES, GridArrays, MarkerArrays, Materials,
                  V2CError, load_lib)

PRN_ROCK = "n0 n1 s0 s1 nu dh dv ss mm gg ll a0 a1 b0 b1 e0 e1 f0 f1 ro bb aa cp kt kf kp ht".split()
PRN_NODE = ("sxxee sxx exxee exx sxyee sxy exyee exy sppee pr phie phis_it mvxe vx mvye vy pes pee "
            "RK_sliprs_final_a RK_sliprs_final RK_phidts_final").split()
PRN_MARK = "x y k e xx v xy p exx exy d w vx vy phi L L0 C C0".split()


class Prn(C.Structure):
    """struct v2c_prn (include/v2c.h): the reference's .prn model file."""
    _fields_ = ([(n, C.c_long) for n in ("xnumx", "ynumy", "mnumx", "mnumy", "marknum")]
                + [(n, C.c_double) for n in ("xsize", "ysize", "pinit")] + [("pkf", C.c_double * 4)]
                + [(n, C.c_double) for n in ("timestep", "timesum", "GXKOEF", "GYKOEF")]
                + [("rocknum", C.c_int), ("bondnum", C.c_long), ("cycle", C.c_int),
                   ("markim", C.POINTER(C.c_int)), ("rock", C.POINTER(C.c_double) * len(PRN_ROCK)),
                   ("node", C.POINTER(C.c_double) * len(PRN_NODE)), ("bondm", C.POINTER(C.c_long)),
                   ("gx", C.POINTER(C.c_double)), ("gy", C.POINTER(C.c_double)),
                   ("bondv", C.POINTER(C.c_double)), ("bondn", C.POINTER(C.c_long)),
                   ("timesave", C.c_double * 5), ("mark", C.POINTER(C.c_double) * len(PRN_MARK)),
                   ("t", C.POINTER(C.c_char))])


_PRN_SCALARS = ("xnumx ynumy mnumx mnumy marknum xsize ysize pinit timestep timesum GXKOEF GYKOEF "
                "rocknum bondnum cycle").split()


def _bind_prn(lib):
    if getattr(lib, "_prn_bound", False):
        return
    lib.v2c_prn_read.argtypes = [C.c_char_p, C.POINTER(Prn)]
    lib.v2c_prn_read.restype = C.c_int
    lib.v2c_prn_write.argtypes = [C.c_char_p, C.POINTER(Prn)]
    lib.v2c_prn_write.restype = C.c_int
    lib.v2c_prn_free.argtypes = [C.POINTER(Prn)]
    lib.v2c_prn_free.restype = None
    lib.v2c_prn_last_error.argtypes = []
    lib.v2c_prn_last_error.restype = C.c_char_p
    lib._prn_bound = True


def read(path: str, lib=None) -> dict:
    """Read a .prn file into a dict of scalars and numpy arrays: 'rock',
    'node', 'mark' are dicts by column name; 'bondm', 'bondn' (bondnum x 3),
    'bondv' (bondnum x 4), 'gx', 'gy', 'markim', 't', 'timesave', 'pkf'."""
    lib = lib or load_lib()
    _bind_prn(lib)
    p = Prn()
    rc = lib.v2c_prn_read(path.encode(), C.byref(p))
    if rc:
        raise V2CError(f"prn_read: {ERRORS.get(rc, rc)}: {lib.v2c_prn_last_error().decode()}")
    try:
        nn, nm, nr, nb = p.xnumx * p.ynumy, p.marknum, p.rocknum, p.bondnum

        def arr(ptr, n, dt):
            return np.ctypeslib.as_array(ptr, shape=(max(n, 1),))[:n].astype(dt, copy=True)
        out = {k: getattr(p, k) for k in _PRN_SCALARS}
        out["pkf"] = np.array(p.pkf[:])
        out["timesave"] = np.array(p.timesave[:])
        out["markim"] = arr(p.markim, nr, np.int32)
        out["rock"] = {k: arr(p.rock[i], nr, np.float64) for i, k in enumerate(PRN_ROCK)}
        out["node"] = {k: arr(p.node[i], nn, np.float64) for i, k in enumerate(PRN_NODE)}
        out["mark"] = {k: arr(p.mark[i], nm, np.float64) for i, k in enumerate(PRN_MARK)}
        out["bondm"] = arr(p.bondm, 4 * nn, np.int64)
        out["gx"] = arr(p.gx, p.xnumx, np.float64)
        out["gy"] = arr(p.gy, p.ynumy, np.float64)
        out["bondv"] = arr(p.bondv, 4 * nb, np.float64).reshape(nb, 4)
        out["bondn"] = arr(p.bondn, 3 * nb, np.int64).reshape(nb, 3)
        out["t"] = np.frombuffer(C.string_at(p.t, nm), dtype=np.int8).copy()
    finally:
        lib.v2c_prn_free(C.byref(p))
    return out


def write(path: str, d: dict, lib=None):
    """Write a dict shaped like read()'s result to a .prn file."""
    lib = lib or load_lib()
    _bind_prn(lib)
    p = Prn()
    for k in _PRN_SCALARS:
        setattr(p, k, d[k])
    p.pkf[:] = list(d["pkf"])
    p.timesave[:] = list(d["timesave"])
    keep = []

    def ptr(a, dt, ct):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(ct))
    p.markim = ptr(d["markim"], np.int32, C.c_int)
    for i, k in enumerate(PRN_ROCK):
        p.rock[i] = ptr(d["rock"][k], np.float64, C.c_double)
    for i, k in enumerate(PRN_NODE):
        p.node[i] = ptr(d["node"][k], np.float64, C.c_double)
    for i, k in enumerate(PRN_MARK):
        p.mark[i] = ptr(d["mark"][k], np.float64, C.c_double)
    p.bondm = ptr(d["bondm"], np.int64, C.c_long)
    p.gx = ptr(d["gx"], np.float64, C.c_double)
    p.gy = ptr(d["gy"], np.float64, C.c_double)
    p.bondv = ptr(np.asarray(d["bondv"]).ravel(), np.float64, C.c_double)
    p.bondn = ptr(np.asarray(d["bondn"]).ravel(), np.int64, C.c_long)
    p.t = ptr(d["t"], np.int8, C.c_char)
    rc = lib.v2c_prn_write(path.encode(), C.byref(p))
    if rc:
        raise V2CError(f"prn_write: {ERRORS.get(rc, rc)}: {lib.v2c_prn_last_error().decode()}")


# .prn marker columns that are marker fields of the library (markv, the
# marker's viscosity, only travels through the file).
_MARK_TO_FIELD = {k: k for k in PRN_MARK if k in MARKER_FIELDS}


def materials(d: dict) -> Materials:
    """The rock table of a read() result as the Materials a Context takes
    (columns outside the library's material table, such as the brittle koefficients
    ll, a0 ... f1, stay in the file only)."""
    nr = int(d["rocknum"])
    if nr > V2C_MAX_TYPES:
        raise V2CError(f"prn: {nr} rock types, the library holds {V2C_MAX_TYPES}")
    mats = Materials()
    mats.ntypes = nr
    for f, _ in Materials._fields_:
        if f == "im":
            mats.im[:nr] = [int(v) for v in d["markim"]]
        elif f != "ntypes":
            getattr(mats, f)[:nr] = [float(v) for v in d["rock"][f]]
    return mats


def to_state(d: dict):
    """(GridArrays, Materials, MarkerArrays, nodal dict) from a read() result,
    ready for Context(grid, params, mats, ...), upload_markers and
    set_all_nodal.  The nodal dict holds the file's node columns that are
    nodal fields of the library (of the two RK_sliprs_final columns, the second,
    the one load_model keeps: input_RSF.c:126)."""
    grid = GridArrays(d["gx"], d["gy"], d["xsize"], d["ysize"])
    fields = {f: d["mark"][k] for k, f in _MARK_TO_FIELD.items()}
    fields["t"] = d["t"]
    markers = MarkerArrays(int(d["marknum"]), fields)
    nodal = {k: d["node"][k].copy() for k in PRN_NODE if k in NODAL_FIELDS}
    return grid, materials(d), markers, nodal


def from_state(d: dict, markers: MarkerArrays, nodal: dict, timesum=None, timestep=None) -> dict:
    """A copy of the read() result `d` with its marker columns, node columns
    and times replaced by a Context's state (download_markers /
    get_all_nodal), for write()."""
    out = dict(d)
    out["marknum"] = markers.n
    out["mark"] = {k: (markers[_MARK_TO_FIELD[k]].copy() if k in _MARK_TO_FIELD
                       else np.resize(d["mark"][k], markers.n)) for k in PRN_MARK}
    out["t"] = markers.t.copy()
    out["node"] = dict(d["node"])
    for k in PRN_NODE:
        if k in nodal:
            out["node"][k] = np.asarray(nodal[k], dtype=np.float64).copy()
    if "RK_sliprs_final" in nodal:
        out["node"]["RK_sliprs_final_a"] = out["node"]["RK_sliprs_final"].copy()
    if timesum is not None:
        out["timesum"] = float(timesum)
    if timestep is not None:
        out["timestep"] = float(timestep)
    return out
