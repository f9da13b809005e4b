"""ctypes mirror of include/v2c.h and the Python handle around libv2c.so.

The structures here are byte-for-byte the C structures of the C-ABI; the same
definitions serve the HIP library (libv2c.so) and the oracle.  Nodal fields are
addressed by the reference's global names (headevps.c:89-196); markers by the
reference's marker array names without the ``mark`` prefix (headevps.c:208-246).
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libv2c.so")
HEADER_PATH = os.path.join(REPO_DIR, "include", "v2c.h")

V2C_MAX_TYPES = 256

# Order must equal `enum v2c_nodal_field` in include/v2c.h (checked by a test).
NODAL_FIELDS = (
    "pr vx vy exx exy esp sxx sxy sxxs sppes exxpl exxel "
    "bm ro et nu nd gg gd sppe phis phin sxyen sxxen exyen exxen sppen mus_s as bs Ls V0s "
    "sxxe sxye exxe exye cohesion dro drp cp kt ht tk mrx mry mvx mvy pe pen nu_m2c nd_m2c "
    "mvxn mvyn "
    "phie sxyee exyee pee sppee sxxee exxee mvxe mvye "
    "wtphisum phisum sxysum exysum esum wtsxxsum psum sxxsum exxsum wtvxsum vxsum wtvysum vysum "
    "mrx0 mry0 "
    "pes phis_it syields nuvp_it sliprs RK_sliprs_final RK_phidts_final RK_sliprs_solved "
    "RK_phidts_solved RK_deltas RK_sliprs0 RK_sliprs1 RK_phidts0 RK_phidts1 "
    "deltasxx deltasxy deltaVs exxs exxes sxxes prs "
    "tk2 tol0 tol1 "
    "sticky01"
).split()
NF = len(NODAL_FIELDS)
FIELD_ID = {n: i for i, n in enumerate(NODAL_FIELDS)}

MARKER_FIELDS = ("x y k p xx xy exx exy e phi vx vy C d w a b L L0 V0 fric_static C0").split()

KERNELS = ("m2n_cells", "m2n_nodes", "n2m_cells", "n2m_nodes", "advect", "sort",
           "before", "after", "local", "mlp", "m2n_props", "n2m_update", "post", "local_rsf",
           "local_wall")

ERRORS = {0: "ok", -1: "EINVAL", -2: "ENOMEM", -3: "EHIP", -4: "ECAPACITY",
          -5: "EUNSUPPORTED", -6: "ENOGPU"}


def header_enum_names(path: str = HEADER_PATH) -> list[str]:
    """Names of `enum v2c_nodal_field`, parsed from the header."""
    text = open(path).read()
    body = text[text.index("enum v2c_nodal_field"):]
    body = body[body.index("{") + 1: body.index("V2C_NF")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    return [n[len("V2C_N_"):] for n in re.findall(r"V2C_N_\w+", body)]


class Grid(C.Structure):
    _fields_ = [("xnumx", C.c_long), ("ynumy", C.c_long),
                ("gx", C.POINTER(C.c_double)), ("gy", C.POINTER(C.c_double)),
                ("xsize", C.c_double), ("ysize", C.c_double)]


_PARAM_INTS = ("markmod stoksmod outgrid densimod interpolationmod antidiffinter rkmod RSFmode "
               "power_law_RSF reloadfromDP compressiblemode stickymod inertyn pmode tempmod "
               "n2m_phantom").split()
_PARAM_DBLS = ("nubeg nuend tdeep zdeep incoming_T nstress Dphis half_range start_updip end_updip "
               "start_downdip end_downdip a_vw b_vw L_vw V0_vw a_vs b_vs L_vs V0_vs a_hr b_hr "
               "L_hr V0_hr xstpx faultwidth rad_damping bulk_modulus band_y band_halfwidth").split()


class Params(C.Structure):
    _fields_ = [(n, C.c_int) for n in _PARAM_INTS] + [(n, C.c_double) for n in _PARAM_DBLS]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


_MAT_DBLS = "nu dh dv ss mm gg n0 n1 ro bb aa cp kt kf kp ht".split()


class Materials(C.Structure):
    _fields_ = ([("ntypes", C.c_int)] + [(n, C.c_double * V2C_MAX_TYPES) for n in _MAT_DBLS]
                + [("im", C.c_int * V2C_MAX_TYPES)])


class LUStats(C.Structure):
    _fields_ = ([(n, C.c_long) for n in ("markers ductile_markers ductile_iters fault_markers rsf_iters "
                                         "max_ductile_iters max_rsf_iters fallbacks nonfinite "
                                         "ductile_evals rsf_evals").split()]
                + [("hist_ductile", C.c_long * 32), ("hist_rsf", C.c_long * 32)])

    def as_dict(self) -> dict:
        d = {n: getattr(self, n) for n, _ in self._fields_[:11]}
        d["hist_ductile"] = list(self.hist_ductile)
        d["hist_rsf"] = list(self.hist_rsf)
        return d


class Markers(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in MARKER_FIELDS] + [("t", C.c_char_p)]


def default_params(xsize: float = 0.0, xnumx: int = 2) -> Params:
    """Reference defaults (input_RSF.c:13-59, mode_RSF.t3c flags), markmod=2.

    Kept equal to v2c_params_default() in the library (a test checks both)."""
    p = Params()
    vals = dict(markmod=2, stoksmod=1, outgrid=2, densimod=0, interpolationmod=2, antidiffinter=1,
                rkmod=0, RSFmode=1, power_law_RSF=0, reloadfromDP=0, compressiblemode=0,
                stickymod=1, inertyn=0, pmode=0, tempmod=0, n2m_phantom=1,
                nubeg=1e1, nuend=1.667e27, tdeep=1873.0, zdeep=3000000.0, incoming_T=0.0,
                nstress=10.0e6, Dphis=0.0, half_range=2.0e3,
                a_vw=0.009, b_vw=0.017, L_vw=0.0100, V0_vw=4.0e-09,
                a_vs=0.020, b_vs=0.001, L_vs=0.0100, V0_vs=4.0e-09,
                a_hr=0.009, b_hr=0.017, L_hr=0.0100, V0_hr=4.0e-09,
                faultwidth=400.0, rad_damping=4.3303e+06, bulk_modulus=50e9,
                band_y=10.000e3, band_halfwidth=100.0)
    start_sez, end_sez = 10.0e3, 40.0e3
    vals["start_updip"] = start_sez - vals["half_range"]
    vals["end_updip"] = start_sez + vals["half_range"]
    vals["start_downdip"] = end_sez - vals["half_range"]
    vals["end_downdip"] = end_sez + vals["half_range"]
    vals["xstpx"] = xsize / float(xnumx - 1) if xnumx > 1 else 0.0
    for k, v in vals.items():
        setattr(p, k, v)
    return p


P_DOUBLE = C.POINTER(C.c_double)


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


class GridArrays:
    """Owns the numpy gridline arrays behind a Grid struct."""

    def __init__(self, gx: np.ndarray, gy: np.ndarray, xsize: float | None = None,
                 ysize: float | None = None):
        self.gx = np.ascontiguousarray(gx, dtype=np.float64)
        self.gy = np.ascontiguousarray(gy, dtype=np.float64)
        self.xsize = float(self.gx[-1]) if xsize is None else float(xsize)
        self.ysize = float(self.gy[-1]) if ysize is None else float(ysize)
        self.struct = Grid(len(self.gx), len(self.gy), _dptr(self.gx), _dptr(self.gy),
                           self.xsize, self.ysize)

    @property
    def xnumx(self):
        return len(self.gx)

    @property
    def ynumy(self):
        return len(self.gy)

    @property
    def nodenum(self):
        return len(self.gx) * len(self.gy)


class MarkerArrays:
    """Numpy SoA of markers plus the Markers pointer struct that views it."""

    def __init__(self, n: int, fields: dict | None = None):
        self.n = n
        self.arrays = {}
        for f in MARKER_FIELDS:
            src = None if fields is None else fields.get(f)
            self.arrays[f] = (np.zeros(n) if src is None
                              else np.ascontiguousarray(src, dtype=np.float64).copy())
        t = None if fields is None else fields.get("t")
        self.t = np.zeros(n, dtype=np.int8) if t is None else np.ascontiguousarray(t, np.int8).copy()
        self._make_struct()

    def _make_struct(self):
        ptrs = [_dptr(self.arrays[f]) for f in MARKER_FIELDS]
        self.struct = Markers(*ptrs, C.cast(self.t.ctypes.data, C.c_char_p))

    def __getitem__(self, k):
        return self.t if k == "t" else self.arrays[k]

    def copy(self) -> "MarkerArrays":
        d = dict(self.arrays)
        d["t"] = self.t
        return MarkerArrays(self.n, d)


class StressStats(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("maxds", "minnu", "maxnu", "maxdphi", "maxnvp")]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class TimestepParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("maxtmstep", "tcnvp", "tcd", "tcphih", "dtfactor", "lambda0")]

    @classmethod
    def default(cls, maxtmstep: float = 3.15576e7):
        """input_RSF.c:17-20 values; lambda0 NaN (uninitialised in the reference)."""
        return cls(maxtmstep, 0.2, 1e-3, 0.2, 1.0, float("nan"))


class TimestepResult(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("dt", "dtnvp", "dtd", "dtphi", "dtL", "maxvel", "maxVx")]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class StickyStats(C.Structure):
    """v2c_sticky_stats: calc_timestep's sticky-air statistics (calc_RSF.c:364-374, 505-550)."""
    _fields_ = [("nr_1", C.c_int), ("nr_0", C.c_int)] + [
        (n, C.c_double) for n in ("max_sigin_1", "max_sigin_0", "min_sigin_1", "min_sigin_0",
                                  "sum_sigin_1", "sum2_sigin_1", "sum_sigin_0", "sum2_sigin_0",
                                  "max_visc_1", "max_visc_0", "min_visc_1", "min_visc_0",
                                  "sum_visc_1", "sum2_visc_1", "sum_visc_0", "sum2_visc_0",
                                  "sigin_P1", "sigin_P2", "visco_P1", "visco_P2")]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


def load_lib(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise RuntimeError(f"libv2c.so not built at {path}: run __graft_entry__.build()")
    lib = C.CDLL(path)
    P = C.POINTER
    lib.v2c_params_default.argtypes = [P(Params)]
    lib.v2c_params_default.restype = None
    lib.v2c_nodal_name.argtypes = [C.c_int]
    lib.v2c_nodal_name.restype = C.c_char_p
    lib.v2c_create.argtypes = [P(C.c_void_p), P(Grid), P(Params), P(Materials), C.c_long, C.c_int]
    lib.v2c_destroy.argtypes = [C.c_void_p]
    lib.v2c_destroy.restype = None
    lib.v2c_last_error.argtypes = [C.c_void_p]
    lib.v2c_last_error.restype = C.c_char_p
    for name, args in [("v2c_set_async", [C.c_int]), ("v2c_synchronize", []),
                       ("v2c_set_time", [C.c_double, C.c_double]),
                       ("v2c_markers_upload", [P(Markers), C.c_long]),
                       ("v2c_markers_download", [P(Markers), C.c_long]),
                       ("v2c_nodal_set", [C.c_int, P(C.c_double)]),
                       ("v2c_nodal_get", [C.c_int, P(C.c_double)]),
                       ("v2c_nodal_set_many", [P(C.c_int), P(P(C.c_double)), C.c_int]),
                       ("v2c_nodal_get_many", [P(C.c_int), P(P(C.c_double)), C.c_int]),
                       ("v2c_marker2nodes", []), ("v2c_nodes2markers", []),
                       ("v2c_moveandrotatemarkers", []), ("v2c_nodes2markers_advect", []),
                       ("v2c_compute_before_solve", [C.c_double, C.c_int, P(C.c_long)]),
                       ("v2c_compute_after_solve", [C.c_double, C.c_int, P(C.c_double), P(C.c_long)]),
                       ("v2c_ductilevisc_batch", [C.c_long, C.c_char_p, P(C.c_double), P(C.c_double),
                                                  P(C.c_double), P(C.c_double), P(C.c_double),
                                                  P(C.c_double), P(C.c_double), P(C.c_int)]),
                       ("v2c_ductilevisc_batch_noalloc", [C.c_long, C.c_char_p, P(C.c_double), P(C.c_double),
                                                          P(C.c_double), P(C.c_double), P(C.c_double),
                                                          P(C.c_double), P(C.c_double), P(C.c_int)]),
                       ("v2c_enable_timing", [C.c_int]),
                       ("v2c_timing_select", [C.c_uint]),
                       ("v2c_set_friction_fields", [C.c_int]),
                       ("v2c_copy_calibration", [C.c_long, P(C.c_double)]),
                       ("v2c_order_stats", [P(C.c_long), P(C.c_long), P(C.c_long)]),
                       ("v2c_order_get", [P(C.c_uint), P(C.c_uint), C.c_long]),
                       ("v2c_rsf_upload", [C.c_char_p] + [P(C.c_double)] * 7 + [C.c_long]),
                       ("v2c_local_update", [C.c_double, C.c_int, C.c_double, P(LUStats)]),
                       ("v2c_local_download", [P(C.c_double)] * 5 + [C.c_long]),
                       ("v2c_local_set_guess", [P(C.c_double), P(C.c_double), C.c_long]),
                       ("v2c_local_get_guess", [P(C.c_double), P(C.c_double), C.c_long]),
                       ("v2c_kernel_time", [C.c_int, P(C.c_double), P(C.c_long)]),
                       ("v2c_recompute_stress_strainrate", [C.c_double, C.c_int, P(StressStats)]),
                       ("v2c_interpolate_center2basic", [C.c_double]),
                       ("v2c_interpolate_center2basic_old", []), ("v2c_interpolate_basic2center", []),
                       ("v2c_copyarrays", [C.c_int]), ("v2c_temperature_markers", [C.c_double]),
                       ("v2c_set_boundary", [P(C.c_long), P(C.c_long), P(C.c_double), C.c_long]),
                       ("v2c_vxvy_boundarycorrection", []), ("v2c_P_boundarycorrection", []), ("v2c_tkrecalc", []),
                       ("v2c_calc_timestep", [C.c_int, P(TimestepParams), P(TimestepResult)]),
                       ("v2c_sticky_marks", []), ("v2c_calc_sticky_stats", [P(StickyStats)]),
                       ("v2c_reset_timing", [])]:
        fn = getattr(lib, name)
        fn.argtypes = [C.c_void_p] + args
        fn.restype = C.c_int
    lib.v2c_marker_count.argtypes = [C.c_void_p]
    lib.v2c_marker_count.restype = C.c_long
    return lib


class V2CError(RuntimeError):
    pass


class Context:
    """One device context of libv2c (one HIP device, one marker cloud)."""

    def __init__(self, grid: GridArrays, params: Params, mats: Materials, capacity: int,
                 device: int = 0, lib: C.CDLL | None = None):
        self.lib = lib or load_lib()
        self.grid = grid
        self.params = params
        self.mats = mats
        self.nodenum = grid.nodenum
        h = C.c_void_p()
        rc = self.lib.v2c_create(C.byref(h), C.byref(grid.struct), C.byref(params),
                                 C.byref(mats), capacity, device)
        if rc != 0:
            raise V2CError(f"v2c_create failed: {ERRORS.get(rc, rc)}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.v2c_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.v2c_last_error(self.h)
            raise V2CError(f"{what}: {ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")

    def set_time(self, timesum, timestep):
        self._check(self.lib.v2c_set_time(self.h, timesum, timestep), "set_time")

    def set_async(self, on: bool):
        self._check(self.lib.v2c_set_async(self.h, int(on)), "set_async")

    def synchronize(self):
        self._check(self.lib.v2c_synchronize(self.h), "synchronize")

    def upload_markers(self, m: MarkerArrays):
        self._check(self.lib.v2c_markers_upload(self.h, C.byref(m.struct), m.n), "markers_upload")

    def download_markers(self, m: MarkerArrays):
        self._check(self.lib.v2c_markers_download(self.h, C.byref(m.struct), m.n), "markers_download")

    def set_nodal(self, name: str, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.size == self.nodenum, (name, a.size)
        self._check(self.lib.v2c_nodal_set(self.h, FIELD_ID[name], _dptr(a)), f"nodal_set {name}")

    def get_nodal(self, name: str) -> np.ndarray:
        out = np.empty(self.nodenum)
        self._check(self.lib.v2c_nodal_get(self.h, FIELD_ID[name], _dptr(out)), f"nodal_get {name}")
        return out

    def set_nodal_many(self, nod: dict):
        """Several fields in one staged transfer (v2c_nodal_set_many)."""
        names = list(nod)
        arrs = [np.ascontiguousarray(nod[k], dtype=np.float64) for k in names]
        for k, a in zip(names, arrs):
            assert a.size == self.nodenum, (k, a.size)
        ids = (C.c_int * len(names))(*[FIELD_ID[k] for k in names])
        ptrs = (P_DOUBLE * len(names))(*[_dptr(a) for a in arrs])
        self._check(self.lib.v2c_nodal_set_many(self.h, ids, ptrs, len(names)), "nodal_set_many")

    def get_nodal_many(self, names, out: dict | None = None) -> dict:
        """Several fields in one staged transfer (v2c_nodal_get_many); `out`
        may hold the destination arrays."""
        out = out if out is not None else {k: np.empty(self.nodenum) for k in names}
        ids = (C.c_int * len(names))(*[FIELD_ID[k] for k in names])
        ptrs = (P_DOUBLE * len(names))(*[_dptr(out[k]) for k in names])
        self._check(self.lib.v2c_nodal_get_many(self.h, ids, ptrs, len(names)), "nodal_get_many")
        return out

    def set_all_nodal(self, nod: dict):
        for k, v in nod.items():
            self.set_nodal(k, v)

    def get_all_nodal(self) -> dict:
        return {k: self.get_nodal(k) for k in NODAL_FIELDS}

    def marker2nodes(self):
        self._check(self.lib.v2c_marker2nodes(self.h), "marker2nodes")

    def recompute_stress_strainrate(self, dt, rk=0) -> dict:
        st = StressStats()
        self._check(self.lib.v2c_recompute_stress_strainrate(self.h, dt, rk, C.byref(st)),
                    "recompute_stress_strainrate")
        return st.as_dict()

    def interpolate_center2basic(self, dt):
        self._check(self.lib.v2c_interpolate_center2basic(self.h, dt), "interpolate_center2basic")

    def interpolate_center2basic_old(self):
        self._check(self.lib.v2c_interpolate_center2basic_old(self.h), "interpolate_center2basic_old")

    def interpolate_basic2center(self):
        self._check(self.lib.v2c_interpolate_basic2center(self.h), "interpolate_basic2center")

    def copyarrays(self, rk=0):
        self._check(self.lib.v2c_copyarrays(self.h, rk), "copyarrays")

    def set_boundary(self, bondm, bondn, bondv):
        """Boundary equations: bondm over 4*nodenum parameters, bondn (nbond x 3)
        and bondv (nbond x 4) rows, row 0 unused (include/v2c.h)."""
        bm = np.ascontiguousarray(bondm, dtype=np.int64)
        bn = np.ascontiguousarray(bondn, dtype=np.int64)
        bv = np.ascontiguousarray(bondv, dtype=np.float64)
        assert bm.size == 4 * self.nodenum and bn.shape[1:] == (3,) and bv.shape == (bn.shape[0], 4)
        self._bond = (bm, bn, bv)
        self._check(self.lib.v2c_set_boundary(self.h, bm.ctypes.data_as(C.POINTER(C.c_long)),
                                              bn.ctypes.data_as(C.POINTER(C.c_long)), _dptr(bv), bn.shape[0]),
                    "set_boundary")

    def vxvy_boundarycorrection(self):
        self._check(self.lib.v2c_vxvy_boundarycorrection(self.h), "vxvy_boundarycorrection")

    def P_boundarycorrection(self):
        self._check(self.lib.v2c_P_boundarycorrection(self.h), "P_boundarycorrection")

    def tkrecalc(self):
        self._check(self.lib.v2c_tkrecalc(self.h), "tkrecalc")

    def temperature_markers(self, heatdif):
        self._check(self.lib.v2c_temperature_markers(self.h, heatdif), "temperature_markers")

    def calc_timestep(self, oldornew, params=None) -> dict:
        p = params or TimestepParams.default()
        r = TimestepResult()
        self._check(self.lib.v2c_calc_timestep(self.h, oldornew, C.byref(p), C.byref(r)), "calc_timestep")
        return r.as_dict()

    def sticky_marks(self):
        self._check(self.lib.v2c_sticky_marks(self.h), "sticky_marks")

    def sticky_stats(self) -> dict:
        r = StickyStats()
        self._check(self.lib.v2c_calc_sticky_stats(self.h, C.byref(r)), "calc_sticky_stats")
        return r.as_dict()

    def nodes2markers(self):
        self._check(self.lib.v2c_nodes2markers(self.h), "nodes2markers")

    def moveandrotatemarkers(self):
        self._check(self.lib.v2c_moveandrotatemarkers(self.h), "moveandrotatemarkers")

    def nodes2markers_advect(self):
        self._check(self.lib.v2c_nodes2markers_advect(self.h), "nodes2markers_advect")

    def compute_before_solve(self, dt, rk) -> int:
        nerr = C.c_long(0)
        self._check(self.lib.v2c_compute_before_solve(self.h, dt, rk, C.byref(nerr)),
                    "compute_before_solve")
        return nerr.value

    def compute_after_solve(self, dt, rk):
        d = C.c_double(0)
        nerr = C.c_long(0)
        self._check(self.lib.v2c_compute_after_solve(self.h, dt, rk, C.byref(d), C.byref(nerr)),
                    "compute_after_solve")
        return d.value, nerr.value

    def ductilevisc_batch(self, t, sxy, sxx, exy, exx, tk, pb, noalloc=False):
        """noalloc: v2c_ductilevisc_batch_noalloc (static staging, at most
        NOALLOC_MAX inputs)."""
        n = len(t)
        t8 = np.ascontiguousarray(t, dtype=np.int8)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (sxy, sxx, exy, exx, tk, pb)]
        out = np.empty(n)
        it = np.empty(n, dtype=np.int32)
        fn = self.lib.v2c_ductilevisc_batch_noalloc if noalloc else self.lib.v2c_ductilevisc_batch
        self._check(fn(self.h, n, C.cast(t8.ctypes.data, C.c_char_p), *[_dptr(a) for a in arrs], _dptr(out),
                       it.ctypes.data_as(C.POINTER(C.c_int))),
                    "ductilevisc_batch_noalloc" if noalloc else "ductilevisc_batch")
        return out, it

    def enable_timing(self, on=True):
        self._check(self.lib.v2c_enable_timing(self.h, int(on)), "enable_timing")

    def timing_select(self, names=None):
        """Time only the named kernels (KERNELS entries) while timing is on;
        None restores all."""
        mask = 0xffffffff if names is None else sum(1 << KERNELS.index(k) for k in names)
        self._check(self.lib.v2c_timing_select(self.h, mask), "timing_select")

    def reset_timing(self):
        self._check(self.lib.v2c_reset_timing(self.h), "reset_timing")

    def kernel_time(self, name: str):
        ms = C.c_double(0)
        n = C.c_long(0)
        self._check(self.lib.v2c_kernel_time(self.h, KERNELS.index(name), C.byref(ms), C.byref(n)),
                    "kernel_time")
        return ms.value, n.value


    def order_get(self):
        """(order, skey) of the current cell ordering, brought up to date."""
        n = self.lib.v2c_marker_count(self.h)
        order = np.zeros(n, dtype=np.uint32)
        skey = np.zeros(n, dtype=np.uint32)
        self._check(self.lib.v2c_order_get(self.h, order.ctypes.data_as(C.POINTER(C.c_uint)),
                                           skey.ctypes.data_as(C.POINTER(C.c_uint)), n), "order_get")
        return order, skey

    def order_stats(self):
        v = [C.c_long(0) for _ in range(3)]
        self._check(self.lib.v2c_order_stats(self.h, *[C.byref(x) for x in v]), "order_stats")
        return {"last_movers": v[0].value, "full_sorts": v[1].value, "incremental": v[2].value}

    def rsf_upload(self, fault, V, phi, a, b, L, V0, mu0):
        n = len(fault)
        f8 = np.ascontiguousarray(fault, dtype=np.uint8)
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (V, phi, a, b, L, V0, mu0)]
        self._check(self.lib.v2c_rsf_upload(self.h, C.cast(f8.ctypes.data, C.c_char_p),
                                            *[_dptr(x) for x in arrs], n), "rsf_upload")

    def local_update(self, dt, warm=0, tol=1e-10):
        """warm: 0 default starts, 1 stored guesses, 2 network guess for the
        slip rate inside the solve."""
        st = LUStats()
        self._check(self.lib.v2c_local_update(self.h, dt, int(warm), tol, C.byref(st)), "local_update")
        return st.as_dict()

    def local_download(self, n):
        out = {k: np.empty(n) for k in ("eta", "V", "phi", "eta_vp", "sy")}
        self._check(self.lib.v2c_local_download(self.h, *[_dptr(out[k]) for k in
                                                           ("eta", "V", "phi", "eta_vp", "sy")], n),
                    "local_download")
        return out

    def local_set_guess(self, eta=None, V=None):
        n = len(eta) if eta is not None else len(V)
        e = None if eta is None else np.ascontiguousarray(eta, dtype=np.float64)
        v = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
        self._check(self.lib.v2c_local_set_guess(self.h, _opt(e), _opt(v), n), "local_set_guess")

    def local_get_guess(self, n):
        eta = np.zeros(n)
        V = np.zeros(n)
        self._check(self.lib.v2c_local_get_guess(self.h, _dptr(eta), _dptr(V), n), "local_get_guess")
        return eta, V

    def set_friction_fields(self, on):
        self._check(self.lib.v2c_set_friction_fields(self.h, int(on)), "set_friction_fields")

class TileShape(C.Structure):
    _fields_ = [("cells_per_group", C.c_int), ("markers_per_tile", C.c_int), ("nodes_per_tile", C.c_int)]


ENS_CLASSES = 3
NOALLOC_MAX = 1 << 18          # V2C_DUCTILEVISC_NOALLOC_MAX


def _bind_ens(lib):
    if getattr(lib, "_ens_bound", False):
        return
    P = C.POINTER
    lib.v2c_ens_create.argtypes = [P(C.c_void_p), P(C.c_void_p), C.c_int]
    lib.v2c_ens_destroy.argtypes = [C.c_void_p]
    lib.v2c_ens_destroy.restype = None
    lib.v2c_ens_last_error.argtypes = [C.c_void_p]
    lib.v2c_ens_last_error.restype = C.c_char_p
    for name, args in [("v2c_ens_size_class", [C.c_int]), ("v2c_ens_set_shape", [C.c_int, P(TileShape)]),
                       ("v2c_ens_get_shape", [C.c_int, P(TileShape)]), ("v2c_ens_set_blocks_per_cu", [C.c_int]),
                       ("v2c_ens_step", [C.c_int]), ("v2c_ens_step_looped", []), ("v2c_ens_order", []),
                       ("v2c_ens_set_branches", [C.c_int]),
                       ("v2c_ens_rms", [P(C.c_double)]),
                       ("v2c_ens_set_local", [C.c_int, C.c_double]),
                       ("v2c_ens_local_stats", [C.c_int, P(LUStats)]),
                       ("v2c_ens_set_adaptive", [P(TimestepParams), C.c_int, C.c_double, C.c_double]),
                       ("v2c_ens_times", [P(C.c_double)]),
                       ("v2c_ens_autotune", [C.c_int, P(C.c_double)])]:
        fn = getattr(lib, name)
        fn.argtypes = [C.c_void_p] + args
        fn.restype = C.c_int
    lib.v2c_ens_get_branches.argtypes = [C.c_void_p]
    lib.v2c_ens_get_branches.restype = C.c_int
    lib._ens_bound = True


class Ensemble:
    """Batched ensemble of independent contexts (v2c_ens_* in the C-ABI)."""

    def __init__(self, contexts):
        self.ctxs = list(contexts)
        self.lib = self.ctxs[0].lib
        _bind_ens(self.lib)
        arr = (C.c_void_p * len(self.ctxs))(*[c.h for c in self.ctxs])
        h = C.c_void_p()
        rc = self.lib.v2c_ens_create(C.byref(h), arr, len(self.ctxs))
        if rc != 0:
            raise V2CError(f"v2c_ens_create: {ERRORS.get(rc, rc)}")
        self.h = h

    def _check(self, rc, what):
        if rc != 0:
            raise V2CError(f"{what}: {ERRORS.get(rc, rc)}: {self.lib.v2c_ens_last_error(self.h).decode()}")

    def step(self, graph=True):
        self._check(self.lib.v2c_ens_step(self.h, int(graph)), "ens_step")

    def order(self):
        """Bring every model's cell ordering up to date in one batched pass."""
        self._check(self.lib.v2c_ens_order(self.h), "ens_order")

    def step_looped(self):
        self._check(self.lib.v2c_ens_step_looped(self.h), "ens_step_looped")

    def size_class(self, k):
        return self.lib.v2c_ens_size_class(self.h, k)

    def set_shape(self, cls, cells_per_group, markers_per_tile, nodes_per_tile):
        s = TileShape(cells_per_group, markers_per_tile, nodes_per_tile)
        self._check(self.lib.v2c_ens_set_shape(self.h, cls, C.byref(s)), "ens_set_shape")

    def get_shape(self, cls):
        s = TileShape()
        self._check(self.lib.v2c_ens_get_shape(self.h, cls, C.byref(s)), "ens_get_shape")
        return (s.cells_per_group, s.markers_per_tile, s.nodes_per_tile)

    def set_branches(self, b):
        """Run the step as b concurrent graph branches of model groups."""
        self._check(self.lib.v2c_ens_set_branches(self.h, b), "ens_set_branches")

    def branches(self):
        return self.lib.v2c_ens_get_branches(self.h)

    def set_blocks_per_cu(self, b):
        self._check(self.lib.v2c_ens_set_blocks_per_cu(self.h, b), "ens_set_blocks_per_cu")

    def rms(self):
        out = np.zeros(2 * len(self.ctxs))
        self._check(self.lib.v2c_ens_rms(self.h, _dptr(out)), "ens_rms")
        return out.reshape(-1, 2)

    def set_local(self, warm=0, tol=1e-10):
        """Add the per-marker local update to every step (warm < 0 removes it)."""
        self._check(self.lib.v2c_ens_set_local(self.h, int(warm), tol), "ens_set_local")

    def local_stats(self, k):
        st = LUStats()
        self._check(self.lib.v2c_ens_local_stats(self.h, k, C.byref(st)), "ens_local_stats")
        return st.as_dict()

    def set_adaptive(self, params=None, oldornew=0, dt_min=1e-3, dt_max=3.15576e7):
        """Adaptive steps from every model's calc_timestep (params None: off)."""
        p = C.byref(params) if params is not None else None
        self._check(self.lib.v2c_ens_set_adaptive(self.h, p, oldornew, dt_min, dt_max), "ens_set_adaptive")

    def times(self):
        """(timesum, timestep) of every model."""
        out = np.zeros(2 * len(self.ctxs))
        self._check(self.lib.v2c_ens_times(self.h, _dptr(out)), "ens_times")
        return out.reshape(-1, 2)

    def autotune(self, reps=3):
        ms = C.c_double(0)
        self._check(self.lib.v2c_ens_autotune(self.h, reps, C.byref(ms)), "ens_autotune")
        return ms.value

    def close(self):
        if self.h:
            self.lib.v2c_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _opt(a):
    return None if a is None else _dptr(a)


