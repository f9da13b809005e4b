""".prn model files: the C-ABI reader and writer (v2c_prn_read / v2c_prn_write)
against the reference's own save_model and load_model (load_save.c:164-347,
353-527), compiled in place, and against a committed file the reference wrote.

Bar: byte-identical files and bit-identical values.  The reader and writer
are host code, so these run without a GPU; the device exchange (a .prn loaded
into a context and written back from it) is in test_gpu_prn.py."""
import os
import tempfile

import numpy as np
import pytest

from oracle import refrun
from velo2cycles_amd import prn, synth
from velo2cycles_amd.abi import V2CError

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
needs_ref = pytest.mark.skipif(not refrun.available(), reason="reference sources not present")


def _model(seed=71, bonds=True, nx=9, ny=7, per=2):
    cfg = synth.config_grid(nx, ny, 8e3, 6e3, per, refine=nx > 8 and ny > 6, seed=seed)
    return cfg, synth.prn_model(cfg, seed, bonds=bonds)


def _expected_size(d):
    nn, nr, nb = d["xnumx"] * d["ynumy"], d["rocknum"], d["bondnum"]
    header = 4 + 5 * 8 + 3 * 8 + 4 * 8 + 4 * 8 + 4 + 8 + 4
    return (header + nr * (4 + 27 * 8) + nn * 200 + 8 * (d["xnumx"] + d["ynumy"])
            + max(nb - 1, 0) * 56 + 40 + d["marknum"] * 153)


def _same(a, b):
    for k in ("xnumx", "ynumy", "mnumx", "mnumy", "marknum", "xsize", "ysize", "pinit", "timestep",
              "timesum", "GXKOEF", "GYKOEF", "rocknum", "bondnum", "cycle"):
        assert a[k] == b[k], k
    for k in ("pkf", "timesave", "markim", "bondm", "gx", "gy", "t"):
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    # row 0 of the boundary equations is not stored (load_save.c:473)
    assert np.array_equal(np.asarray(a["bondv"])[1:], np.asarray(b["bondv"])[1:])
    assert np.array_equal(np.asarray(a["bondn"])[1:], np.asarray(b["bondn"])[1:])
    for grp in ("rock", "node", "mark"):
        for k in a[grp]:
            assert np.array_equal(a[grp][k], b[grp][k], equal_nan=True), (grp, k)


def test_round_trip_and_size():
    _, d = _model()
    d["mark"]["p"][3] = np.nan
    d["node"]["pr"][5] = -np.inf
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "m.prn")
        prn.write(p, d)
        assert os.path.getsize(p) == _expected_size(d)
        _same(prn.read(p), d)


@pytest.mark.parametrize("nx,ny,per,bonds", [(9, 7, 2, True), (2, 2, 1, False), (41, 23, 3, True)])
def test_round_trip_shapes(nx, ny, per, bonds):
    """Grids of one cell, no boundary rows (bondnum 1) and a multi-chunk node
    and marker section."""
    _, d = _model(seed=72 + nx, bonds=bonds, nx=nx, ny=ny, per=per)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "m.prn")
        prn.write(p, d)
        assert os.path.getsize(p) == _expected_size(d)
        _same(prn.read(p), d)


def test_state_mapping_round_trip():
    """to_state then from_state gives back the same file, byte for byte."""
    cfg, d = _model(seed=75)
    with tempfile.TemporaryDirectory() as td:
        a, b = os.path.join(td, "a.prn"), os.path.join(td, "b.prn")
        prn.write(a, d)
        r = prn.read(a)
        grid, mats, markers, nodal = prn.to_state(r)
        assert mats.ntypes == d["rocknum"] and list(mats.im[:mats.ntypes]) == list(d["markim"])
        assert list(mats.kt[:mats.ntypes]) == list(d["rock"]["kt"])
        assert markers.n == cfg.n and grid.nodenum == cfg.grid.nodenum
        for k in ("x", "y", "phi", "C0"):
            assert np.array_equal(markers[k], d["mark"][k])
        assert np.array_equal(nodal["RK_sliprs_final"], d["node"]["RK_sliprs_final"])
        prn.write(b, prn.from_state(r, markers, nodal))
        assert open(a, "rb").read() == open(b, "rb").read()


def test_errors():
    _, d = _model(seed=76)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "m.prn")
        with pytest.raises(V2CError, match="cannot open"):
            prn.read(os.path.join(td, "absent.prn"))
        prn.write(p, d)
        raw = bytearray(open(p, "rb").read())
        bad = os.path.join(td, "bad.prn")
        # a long of 4 bytes: the reference stops with the same complaint
        # (load_save.c:196-197)
        raw2 = bytearray(raw)
        raw2[1] = 4
        open(bad, "wb").write(bytes(raw2))
        with pytest.raises(V2CError, match="sizes"):
            prn.read(bad)
        for cut in (3, 60, len(raw) // 2, len(raw) - 1):
            open(bad, "wb").write(bytes(raw[:cut]))
            with pytest.raises(V2CError, match="truncated"):
                prn.read(bad)
        raw3 = bytearray(raw)
        raw3[4:12] = np.int64(-5).tobytes()          # xnumx
        open(bad, "wb").write(bytes(raw3))
        with pytest.raises(V2CError, match="implausible"):
            prn.read(bad)


def test_reference_file_golden():
    """A file the reference's save_model wrote: read, values as the fixture
    lists them, and written back byte for byte."""
    path = os.path.join(GOLDEN, "prn_small.prn")
    z = np.load(os.path.join(GOLDEN, "prn_small.npz"))
    d = prn.read(path)
    assert d["marknum"] == len(z["markx"]) and d["xnumx"] == len(z["gx"])
    assert np.array_equal(d["gx"], z["gx"]) and np.array_equal(d["gy"], z["gy"])
    assert np.array_equal(d["mark"]["x"], z["markx"]) and np.array_equal(d["mark"]["C0"], z["markC0"])
    assert np.array_equal(d["node"]["RK_phidts_final"], z["RK_phidts_final"])
    assert np.array_equal(d["t"], z["markt"]) and d["cycle"] == int(z["cycle"])
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "m.prn")
        prn.write(p, d)
        assert open(p, "rb").read() == open(path, "rb").read()


@needs_ref
@pytest.mark.parametrize("seed,bonds", [(81, True), (82, False)])
def test_writer_matches_reference_save_model(seed, bonds):
    cfg, d = _model(seed=seed, bonds=bonds)
    with tempfile.TemporaryDirectory() as td:
        ref, ours = os.path.join(td, "r.prn"), os.path.join(td, "o.prn")
        assert len(ref) < 50      # the reference's file name buffer (headevps.c:259)
        refrun.run(cfg, [f"save:{d['cycle']}:{ref}"], extra=refrun.prn_extra(d))
        prn.write(ours, d)
        assert open(ours, "rb").read() == open(ref, "rb").read()


@needs_ref
def test_reference_load_model_reads_our_file():
    cfg, d = _model(seed=83)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "o.prn")
        prn.write(p, d)
        # the reference starts from other data of the same shape
        _, other = _model(seed=84)
        _, _, raw = refrun.run(cfg, [f"load:{p}"], extra=refrun.prn_extra(other))
    nn, nm, nr, nb = cfg.grid.nodenum, d["marknum"], d["rocknum"], d["bondnum"]
    assert raw["xnumx"][0] == d["xnumx"] and raw["marknum"][0] == nm and raw["bondnum"][0] == nb
    assert raw["mnumx"][0] == d["mnumx"] and raw["rocknum"][0] == nr and raw["pinit"][0] == d["pinit"]
    assert np.array_equal(raw["pkf"][:4], d["pkf"]) and np.array_equal(raw["timesave"], d["timesave"])
    assert np.array_equal(raw["markim"][:nr], d["markim"])
    for k, v in d["rock"].items():
        assert np.array_equal(raw["mark" + k][:nr], v), k
    for k, v in d["node"].items():
        if k != "RK_sliprs_final_a":
            assert np.array_equal(raw[k][:nn], v), k
    for k, v in d["mark"].items():
        assert np.array_equal(raw[refrun.PRN_MARK_GLOBAL[k]][:nm], v), k
    assert np.array_equal(raw["markt"][:nm], d["t"])
    assert np.array_equal(raw["bondm"][:4 * nn], d["bondm"])
    assert np.array_equal(raw["bondn"].reshape(-1, 3)[1:nb], d["bondn"][1:])
    assert np.array_equal(raw["bondv"].reshape(-1, 4)[1:nb], d["bondv"][1:])
    assert np.array_equal(raw["gx"], d["gx"]) and np.array_equal(raw["gy"], d["gy"])
