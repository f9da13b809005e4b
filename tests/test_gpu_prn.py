"""A .prn model the reference wrote (tests/golden/prn_small.prn, from its
save_model, load_save.c:353-527) exchanged with the device: loaded into a
context and written back byte for byte, and stepped from (marker2nodes,
nodes2markers, moveandrotatemarkers) in parity with the oracle on the same
state.  Bar as tests/util.py: positions, cells and types bit-exact, FP64
within 1e-12 scaled."""
import os
import tempfile

import numpy as np
import pytest

from velo2cycles_amd import prn, synth
from velo2cycles_amd.abi import NODAL_FIELDS, Context, MarkerArrays, default_params
from tests.util import compare, run_gpu, run_oracle

pytestmark = pytest.mark.gpu
PATH = os.path.join(os.path.dirname(__file__), "golden", "prn_small.prn")


def _config(d):
    grid, mats, markers, nodal = prn.to_state(d)
    full = {k: np.zeros(grid.nodenum) for k in NODAL_FIELDS}
    full.update(nodal)
    return synth.Config("prn_small", grid, default_params(grid.xsize, grid.xnumx), mats, markers, full,
                        d["timestep"], d["timesum"])


def test_reference_file_through_the_device(gpu):
    d = prn.read(PATH)
    cfg = _config(d)
    print("case prn device round trip", cfg.n, flush=True)
    ctx = Context(cfg.grid, cfg.params, cfg.mats, capacity=cfg.n, lib=gpu)
    ctx.set_time(cfg.timesum, cfg.timestep)
    ctx.set_all_nodal(cfg.nodal)
    ctx.upload_markers(cfg.markers)
    back = MarkerArrays(cfg.n)
    ctx.download_markers(back)
    nodal = {k: ctx.get_nodal(k) for k in prn.PRN_NODE if k in NODAL_FIELDS}
    ctx.close()
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "back.prn")
        prn.write(out, prn.from_state(d, back, nodal))
        assert open(out, "rb").read() == open(PATH, "rb").read()


def test_step_from_reference_file(gpu):
    cfg = _config(prn.read(PATH))
    ops = ["m2n", "n2m", "move"]
    print("case prn step", cfg.n, flush=True)
    gm, gn, _, ctx = run_gpu(gpu, cfg, ops)
    om, on, _ = run_oracle(cfg, ops)
    bad = compare(gm, om, gn, on)
    assert not bad, bad
    ctx.close()
