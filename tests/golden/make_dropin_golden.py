"""Write the drop-in fixtures tests/golden/dropin_*.npz by running the
reference's own routines (oracle/_ref/ref_driver, compiled in place from
/root/reference/src by oracle/Makefile), single-threaded.

    python tests/golden/make_dropin_golden.py

Each fixture holds the exact driver input file (`in_bin`: the record stream
that tests/native/dropin_harness reads as well), the op list, and the
reference's outputs after those ops: every marker and nodal field, the
routines' return values (`drv_ret`; [3] the host output stand-in's sum),
`errcheck` and calc_timestep's scalars (its sticky-air statistics among them).
The drop-in tests run the same input through the reference-side binding and
libv2c on the GPU and compare with these outputs; neither the reference nor
this generator is needed there.
"""
from __future__ import annotations

import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from velo2cycles_amd import synth  # noqa: E402
from velo2cycles_amd.abi import MARKER_FIELDS, NODAL_FIELDS  # noqa: E402
from oracle import refrun  # noqa: E402

DT = synth.YEAR
STEP = ["m2n", f"before:{DT}:0", f"after:{DT}:0", "n2m", "move"]
SCALARS = ("maxvel", "maxdphi", "maxnvp", "dtL", "dtphi", "dtd", "dtnvp",
           # calc_timestep's sticky-air statistics
           "nr_1", "nr_0", "max_sigin_1", "max_sigin_0", "min_sigin_1", "min_sigin_0", "sum_sigin_1",
           "sum2_sigin_1", "sum_sigin_0", "sum2_sigin_0", "max_visc_1", "max_visc_0", "min_visc_1",
           "min_visc_0", "sum_visc_1", "sum2_visc_1", "sum_visc_0", "sum2_visc_0", "sigin_P1", "sigin_P2",
           "visco_P1", "visco_P2")


def bond_extra(bc):
    bm, bn, bv = bc
    return {"bondm": ("l", bm), "bondn": ("l", bn.ravel()), "bondv": ("d", bv.ravel()),
            "bondnum": ("l", np.array([len(bn)], np.int64))}


def step_cases():
    for tag, timesum, flags in (("t0", 0.0, {}), ("t1", DT, {}), ("tempmod", DT, dict(tempmod=1)),
                                ("euler_pmode", DT, dict(markmod=1, pmode=1))):
        cfg = synth.config_edge()
        cfg.timesum = timesum
        for k, v in flags.items():
            setattr(cfg.params, k, v)
        yield f"dropin_step_{tag}", cfg, STEP, bond_extra(synth.boundary_equations(cfg, 5))


def post_cases():
    from tests.util import load_post_fixture
    for name in ("post_rsf_p0", "post_rsf_p1", "post_dp"):
        cfg, _, maxtmstep = load_post_fixture(os.path.join(HERE, name + ".npz"))
        for oldornew in (0, 1):
            ops = [f"stress:{DT}:0", f"c2b:{DT}", f"c2bold:{DT}", "b2c", f"tstep:{oldornew}", "copy:0"]
            yield (f"dropin_{name}_ts{oldornew}", cfg, ops,
                   {"maxtmstep": ("d", np.array([maxtmstep]))})


def bc_cases():
    from tests.test_oracle_golden import load_bc_fixture
    for name in ("bc_walls_chains", "bc_compressible"):
        cfg, bc, _ = load_bc_fixture(os.path.join(HERE, name + ".npz"))
        yield f"dropin_{name}", cfg, ["bcvxvy", "bcp", "bctk"], bond_extra(bc)


def loop_cases():
    """The reference loop around the host Stokes solve (i2evps.c:89-113 with
    solve_momentum_continuity, momentum_continuity_iterate.c:44-129): the
    binding's routines with the host stand-ins (tests/native/host_stand_in.h)
    where the reference runs host code, so the drop-in test exercises the
    binding's residency hand-overs.  One Picard iteration; two (rk 0, 1);
    rkmod 3 with four stages (dt / 2 for the first two)."""
    def iteration(dt, rk):
        return [f"before:{dt}:{rk}", "b2c", "hostsolve", f"stress:{dt}:{rk}", f"c2b:{dt}", f"after:{dt}:{rk}",
                "tstep:1"]
    head = ["m2n", f"c2bold:{DT}"]
    tail = ["hostout", "n2m", "move"]
    specs = (("rk1", (41, 21, 40e3, 20e3, 3, False), {}, iteration(DT, 0) + ["copy:1"]),
             ("rk2", (61, 31, 60e3, 30e3, 3, True), {}, iteration(DT, 0) + iteration(DT, 1) + ["copy:2"]),
             ("rkmod3", (41, 21, 40e3, 20e3, 3, True), dict(rkmod=3),
              iteration(DT / 2, 0) + iteration(DT / 2, 1) + iteration(DT, 2) + iteration(DT, 3) + ["copy:4"]))
    for tag, (nx, ny, xs, ys, per, ref), flags, body in specs:
        cfg = synth.config_grid(nx, ny, xs, ys, per, refine=ref, fault_frac=0.02)
        cfg.timesum = DT
        # viscosity cap 3e26 rather than the table's 1e26: calc_timestep
        # branches on nu_m2c > 1e26 (calc_RSF.c:390, 417), and a node whose
        # markers all sit at a 1e26 cap is then decided by the last bit of
        # marker2nodes' weighted average, which no reordered sum reproduces
        for t in range(256):
            cfg.mats.n1[t] = 3e26
        for k, v in flags.items():
            setattr(cfg.params, k, v)
        extra = bond_extra(synth.boundary_equations(cfg, 5))
        extra["maxtmstep"] = ("d", np.array([DT]))
        yield f"dropin_loop_{tag}", cfg, head + body + tail, extra


def write_case(name, cfg, ops, extra):
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        refrun.write_input(fin, cfg, extra=extra)
        in_bin = np.fromfile(fin, dtype=np.uint8)
        raw = refrun.run_file(refrun.DRIVER, fin, ops)
    mk, nd = refrun.split_output(raw, cfg.n, cfg.grid.nodenum)
    d = {"in_bin": in_bin, "ops": np.array(json.dumps(list(ops))), "n": np.array(cfg.n),
         "nodenum": np.array(cfg.grid.nodenum), "drv_ret": raw["drv_ret"][:4],
         "errcheck": raw["errcheck"]}
    for f in MARKER_FIELDS:
        d["out_m_" + f] = mk[f]
    d["out_m_t"] = mk["t"]
    for k in NODAL_FIELDS:
        d["out_n_" + k] = nd[k]
    for k in SCALARS:
        d["s_" + k] = raw[k]
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **d)
    print(f"{path}: {cfg.n} markers, {cfg.grid.nodenum} nodes, ops {ops}")


def main():
    assert refrun.available(), "needs the reference sources and oracle/_ref/ref_driver"
    for gen in (step_cases, post_cases, bc_cases, loop_cases):
        for name, cfg, ops, extra in gen():
            write_case(name, cfg, ops, extra)


if __name__ == "__main__":
    main()
