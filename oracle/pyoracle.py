"""ctypes front end of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product path never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from velo2cycles_amd.abi import (NF, NODAL_FIELDS, Grid, Markers, Materials, Params, StickyStats,
                                 TimestepParams, TimestepResult, _dptr)

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(ORACLE_DIR, "liboracle.so")


class OrState(C.Structure):
    _fields_ = [("grid", Grid), ("prm", Params), ("mat", C.POINTER(Materials)),
                ("timesum", C.c_double), ("timestep", C.c_double), ("m", Markers),
                ("n", C.c_long), ("nod", C.POINTER(C.c_double) * NF)]


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.POINTER
        L.or_nxsearch.argtypes = [P(Grid), C.c_double]
        L.or_nxsearch.restype = C.c_long
        L.or_nysearch.argtypes = [P(Grid), C.c_double]
        L.or_nysearch.restype = C.c_long
        L.or_cells.argtypes = [P(Grid), C.c_long, P(C.c_double), P(C.c_double), P(C.c_long), P(C.c_long)]
        L.or_cells.restype = None
        L.or_allinterpomp.argtypes = [P(OrState), C.c_double, C.c_double]
        L.or_allinterpomp.restype = C.c_double
        L.or_ductilevisc.argtypes = [P(OrState), C.c_int] + [C.c_double] * 6 + [P(C.c_int)]
        L.or_ductilevisc.restype = C.c_double
        L.or_ductilevisc_newton.argtypes = ([P(OrState), C.c_int] + [C.c_double] * 6
                                            + [P(C.c_int), P(C.c_double)])
        L.or_ductilevisc_newton.restype = C.c_double
        L.or_marker2nodes.argtypes = [P(OrState)]
        L.or_marker2nodes.restype = None
        L.or_nodes2markers.argtypes = [P(OrState)]
        L.or_nodes2markers.restype = None
        L.or_moveandrotatemarkers.argtypes = [P(OrState)]
        L.or_moveandrotatemarkers.restype = C.c_long
        L.or_compute_before_solve.argtypes = [P(OrState), C.c_double, C.c_int]
        L.or_compute_before_solve.restype = C.c_long
        L.or_compute_after_solve.argtypes = [P(OrState), C.c_double, C.c_int, P(C.c_long)]
        L.or_compute_after_solve.restype = C.c_double
        L.or_allinteri.argtypes = [P(OrState), C.c_double, C.c_double] + [P(C.c_double)] * 4
        L.or_allinteri.restype = None
        L.or_local_update.argtypes = [P(OrState), C.c_double, C.c_char_p, P(C.c_double), C.c_int,
                                      C.c_double, P(C.c_long), P(C.c_int), P(C.c_int)]
        L.or_local_update.restype = None
        L.or_recompute_stress_strainrate.argtypes = [P(OrState), C.c_double, C.c_int, P(C.c_double)]
        L.or_recompute_stress_strainrate.restype = None
        L.or_interpolate_center2basic.argtypes = [P(OrState), C.c_double]
        L.or_interpolate_center2basic.restype = None
        for name in ("or_interpolate_center2basic_old", "or_interpolate_basic2center"):
            getattr(L, name).argtypes = [P(OrState)]
            getattr(L, name).restype = None
        L.or_copyarrays.argtypes = [P(OrState), C.c_int]
        L.or_copyarrays.restype = None
        L.or_calc_timestep.argtypes = [P(OrState), C.c_int, P(TimestepParams), P(TimestepResult)]
        L.or_calc_timestep.restype = C.c_int
        L.or_sticky_marks.argtypes = [P(OrState)]
        L.or_sticky_marks.restype = None
        L.or_calc_sticky_stats.argtypes = [P(OrState), P(StickyStats)]
        L.or_calc_sticky_stats.restype = None
        L.or_temperature_markers.argtypes = [P(OrState), C.c_double]
        L.or_temperature_markers.restype = None
        for name in ("or_vxvy_boundarycorrection", "or_P_boundarycorrection", "or_tkrecalc"):
            getattr(L, name).argtypes = [P(OrState), P(C.c_long), P(C.c_long), P(C.c_double)]
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


class Oracle:
    """Oracle state over caller-visible numpy arrays (a copy of a Config)."""

    def __init__(self, cfg, markers=None, nodal=None):
        self.cfg = cfg
        self.grid = cfg.grid
        self.mats = cfg.mats
        self.markers = (markers or cfg.markers).copy()
        self.nodal = {k: np.array(v, dtype=np.float64, copy=True)
                      for k, v in (nodal or cfg.nodal).items()}
        for k in NODAL_FIELDS:
            self.nodal.setdefault(k, np.zeros(cfg.grid.nodenum))
        self.st = OrState()
        self.st.grid = cfg.grid.struct
        self.st.prm = cfg.params
        self.st.mat = C.pointer(cfg.mats)
        self.st.timesum = cfg.timesum
        self.st.timestep = cfg.timestep
        self.st.m = self.markers.struct
        self.st.n = self.markers.n
        for i, k in enumerate(NODAL_FIELDS):
            self.st.nod[i] = _dptr(self.nodal[k])

    def set_time(self, timesum, timestep):
        self.st.timesum = timesum
        self.st.timestep = timestep

    def marker2nodes(self):
        lib().or_marker2nodes(C.byref(self.st))

    def nodes2markers(self):
        lib().or_nodes2markers(C.byref(self.st))

    def moveandrotatemarkers(self):
        return lib().or_moveandrotatemarkers(C.byref(self.st))

    def compute_before_solve(self, dt, rk):
        return lib().or_compute_before_solve(C.byref(self.st), dt, rk)

    def compute_after_solve(self, dt, rk):
        nerr = C.c_long(0)
        d = lib().or_compute_after_solve(C.byref(self.st), dt, rk, C.byref(nerr))
        return d, nerr.value

    def recompute_stress_strainrate(self, dt, rk):
        st = (C.c_double * 5)()
        lib().or_recompute_stress_strainrate(C.byref(self.st), dt, rk, st)
        return dict(zip(("maxds", "minnu", "maxnu", "maxdphi", "maxnvp"), list(st)))

    def interpolate_center2basic(self, dt):
        lib().or_interpolate_center2basic(C.byref(self.st), dt)

    def interpolate_center2basic_old(self):
        lib().or_interpolate_center2basic_old(C.byref(self.st))

    def interpolate_basic2center(self):
        lib().or_interpolate_basic2center(C.byref(self.st))

    def copyarrays(self, rk=0):
        lib().or_copyarrays(C.byref(self.st), rk)

    def calc_timestep(self, oldornew, params=None):
        """(return code, result dict); code -1 where chi == 0, -2 for an unknown pmode."""
        p = params or TimestepParams.default()
        r = TimestepResult()
        rc = lib().or_calc_timestep(C.byref(self.st), oldornew, C.byref(p), C.byref(r))
        return rc, r.as_dict()

    def sticky_marks(self):
        lib().or_sticky_marks(C.byref(self.st))

    def sticky_stats(self) -> dict:
        r = StickyStats()
        lib().or_calc_sticky_stats(C.byref(self.st), C.byref(r))
        return r.as_dict()

    def boundary(self, which, bondm, bondn, bondv):
        """which: 'vxvy', 'P' or 'tk'; returns the oracle's code (0, or -1 / -2
        where the reference exits)."""
        fn = {"vxvy": "or_vxvy_boundarycorrection", "P": "or_P_boundarycorrection", "tk": "or_tkrecalc"}[which]
        bm = np.ascontiguousarray(bondm, dtype=np.int64)
        bn = np.ascontiguousarray(bondn, dtype=np.int64)
        bv = np.ascontiguousarray(bondv, dtype=np.float64)
        return getattr(lib(), fn)(C.byref(self.st), bm.ctypes.data_as(C.POINTER(C.c_long)),
                                  bn.ctypes.data_as(C.POINTER(C.c_long)), _dptr(bv))

    def temperature_markers(self, heatdif):
        lib().or_temperature_markers(C.byref(self.st), heatdif)

    def ductilevisc(self, t, sxy, sxx, tk, pb):
        it = C.c_int(0)
        v = lib().or_ductilevisc(C.byref(self.st), int(t), sxy, sxx, 0.0, 0.0, tk, pb, C.byref(it))
        return v, it.value

    def ductilevisc_newton(self, t, sxy, sxx, tk, pb, guess=0.0, tol=1e-10):
        it = C.c_int(0)
        r = C.c_double(0)
        v = lib().or_ductilevisc_newton(C.byref(self.st), int(t), sxy, sxx, tk, pb, guess, tol,
                                        C.byref(it), C.byref(r))
        return v, it.value, r.value

    def allinteri(self, x, y):
        o = [C.c_double() for _ in range(4)]
        lib().or_allinteri(C.byref(self.st), x, y, *[C.byref(v) for v in o])
        return tuple(v.value for v in o)

    def allinterpomp(self, x, y):
        return lib().or_allinterpomp(C.byref(self.st), x, y)

    def nxsearch(self, x):
        return lib().or_nxsearch(C.byref(self.grid.struct), x)

    def nysearch(self, y):
        return lib().or_nysearch(C.byref(self.grid.struct), y)

    def cells(self, x, y):
        """(nxsearch(x_i), nysearch(y_i)) of every point (interpolation.c:1536-1573)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        nx = np.zeros(len(x), np.int64)
        ny = np.zeros(len(x), np.int64)
        lib().or_cells(C.byref(self.grid.struct), len(x), _dptr(x), _dptr(y),
                       nx.ctypes.data_as(C.POINTER(C.c_long)), ny.ctypes.data_as(C.POINTER(C.c_long)))
        return nx, ny


MX_FIELDS = ("V", "phi", "a", "b", "L", "V0", "mu0", "eta", "eta_vp", "sy", "geta", "gV",
             "root_eta", "root_V")


def local_update(oracle: "Oracle", dt, fault, mx: dict, warm=False, tol=1e-10):
    """Per-marker local update on the oracle's markers.  mx maps MX_FIELDS to
    arrays (missing ones are zero); returns (outputs dict, stats, iteration arrays)."""
    n = oracle.markers.n
    buf = np.zeros((len(MX_FIELDS), n))
    for i, k in enumerate(MX_FIELDS):
        if k in mx:
            buf[i] = mx[k]
    buf = np.ascontiguousarray(buf)
    stats = (C.c_long * 11)()
    di = np.zeros(n, dtype=np.int32)
    ri = np.zeros(n, dtype=np.int32)
    f8 = np.ascontiguousarray(fault, dtype=np.uint8)
    lib().or_local_update(C.byref(oracle.st), dt, C.cast(f8.ctypes.data, C.c_char_p),
                          buf.ctypes.data_as(C.POINTER(C.c_double)), int(warm), tol, stats,
                          di.ctypes.data_as(C.POINTER(C.c_int)), ri.ctypes.data_as(C.POINTER(C.c_int)))
    out = {k: buf[i].copy() for i, k in enumerate(MX_FIELDS)}
    names = ("markers ductile_markers ductile_iters fault_markers rsf_iters max_ductile_iters "
             "max_rsf_iters fallbacks nonfinite ductile_evals rsf_evals").split()
    return out, dict(zip(names, list(stats))), di, ri
