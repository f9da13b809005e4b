"""Pins of the oracle that do not need the reference: brute force, closed forms
and invariants, each chosen so that a plausible slip (wrong index, sign, clamp
or operand order) fails."""
import math

import numpy as np
import pytest

from velo2cycles_amd import synth
from velo2cycles_amd.abi import MarkerArrays
from oracle.pyoracle import Oracle

DT = synth.YEAR


def brute_search(g, x):
    """Largest k <= n-2 with g[k] <= x (0 if none); NaN -> n-2."""
    n = len(g)
    if x != x:
        return n - 2
    k = 0
    for i in range(n - 1):
        if g[i] <= x:
            k = i
    return k


def test_cell_search_brute_force():
    cfg = synth.config_grid(37, 23, 50e3, 30e3, 1, refine=True)
    o = Oracle(cfg)
    gx, gy = cfg.grid.gx, cfg.grid.gy
    rng = np.random.default_rng(1)
    xs = list(rng.uniform(-1e3, 51e3, 4000)) + list(gx) + [-0.0, 0.0, 50e3, 1e300, -1e300, float("nan")]
    ys = list(rng.uniform(-1e3, 31e3, 4000)) + list(gy) + [float("nan")]
    for x in xs:
        assert o.nxsearch(x) == brute_search(gx, x), x
    for y in ys:
        assert o.nysearch(y) == brute_search(gy, y), y


def _uniform_cloud(cfg, **fields):
    """Every marker of one Newtonian type with the given constant fields."""
    m = cfg.markers
    m.t[:] = 8
    for k, v in fields.items():
        m[k][:] = v
    return cfg


def test_partition_of_unity_m2n():
    """A constant marker field interpolates to exactly that constant (up to
    rounding) wherever a node received weight (interpolation.c:726-821)."""
    cfg = synth.config_grid(31, 21, 30e3, 20e3, 4, refine=True)
    _uniform_cloud(cfg, xy=1.234e6, xx=-7.5e5, exy=3e-15, exx=-2e-15, e=0.25, C=4e6, phi=9.0,
                   p=2.5e8, k=900.0, vx=1e-9, vy=-2e-9)
    cfg.timesum = 0.0
    cfg.params.band_y = -1e9     # keep friction_marker from resetting phi
    o = Oracle(cfg)
    o.marker2nodes()
    n = o.nodal
    for k, want in [("sxye", 1.234e6), ("sxxe", -7.5e5), ("exye", 3e-15), ("exxe", -2e-15),
                    ("pe", 0.25), ("cohesion", 4e6), ("phis", 50.0), ("sppe", 2.5e8),
                    ("tk", 900.0), ("mvx", 1e-9), ("mvy", -2e-9), ("nu_m2c", 1e19), ("nd_m2c", 1e19),
                    ("bm", 50e9), ("ro", 2800.0), ("gg", 3e10), ("gd", 3e10)]:
        v = n[k]
        got = v[v != 0]
        assert got.size > 0, k
        np.testing.assert_allclose(got, want, rtol=2e-15, err_msg=k)


def test_marker_at_cell_centres_hits_centre_nodes_only():
    """cx == cy == 0.5 gives zero basic/vx/vy weight and full centre weight
    (strict inequalities of interpolation.c:550-624)."""
    cfg = synth.config_grid(11, 9, 10e3, 8e3, 1, jitter=0.0)
    o = Oracle(cfg)
    o.marker2nodes()
    n = o.nodal
    for k in ("nu_m2c", "ro", "mvx", "mvy", "mrx", "mry"):
        assert not np.any(n[k]), k
    assert np.count_nonzero(n["bm"]) == 10 * 8


def test_bilinear_exactness_n2m():
    """A nodal field linear in x,y gathers exactly (to rounding) at interior markers."""
    cfg = synth.config_grid(21, 15, 20e3, 14e3, 3, refine=True)
    g = cfg.grid
    X, Y = np.meshgrid(g.gx, g.gy, indexing="ij")
    lin = (3.0 + 2e-4 * X - 5e-4 * Y).ravel()
    cfg.nodal["exy"] = lin.copy()
    # centre nodes sit at the cell midpoints
    XC = np.concatenate([[g.gx[0]], (g.gx[:-1] + g.gx[1:]) / 2])
    YC = np.concatenate([[g.gy[0]], (g.gy[:-1] + g.gy[1:]) / 2])
    XC, YC = np.meshgrid(XC, YC, indexing="ij")
    cfg.nodal["exx"] = (1.0 - 1e-4 * XC + 7e-4 * YC).ravel()
    o = Oracle(cfg)
    o.nodes2markers()
    x, y = cfg.markers["x"], cfg.markers["y"]
    inner = (x > g.gx[2]) & (x < g.gx[-3]) & (y > g.gy[2]) & (y < g.gy[-3])
    np.testing.assert_allclose(o.markers["exy"][inner], (3.0 + 2e-4 * x - 5e-4 * y)[inner], rtol=0, atol=1e-13)
    np.testing.assert_allclose(o.markers["exx"][inner], (1.0 - 1e-4 * x + 7e-4 * y)[inner], rtol=0, atol=1e-13)


def test_uniform_flow_advection_and_zero_spin():
    """Uniform velocity, zero spin: RK4 moves every marker by v*dt and leaves
    the stress untouched (moveandrotatemarkers.c:26-68)."""
    cfg = synth.config_grid(21, 15, 20e3, 14e3, 3, refine=True)
    for k in ("vx", "vy", "esp", "exx", "exy"):
        cfg.nodal[k][:] = 0.0
    cfg.nodal["vx"][:] = 1.5e-9
    cfg.nodal["vy"][:] = -0.5e-9
    o = Oracle(cfg)
    x0, y0 = cfg.markers["x"].copy(), cfg.markers["y"].copy()
    xx0, xy0 = cfg.markers["xx"].copy(), cfg.markers["xy"].copy()
    o.moveandrotatemarkers()
    g = cfg.grid
    inside = (x0 >= 0) & (x0 <= g.xsize) & (y0 >= 0) & (y0 <= g.ysize)
    np.testing.assert_allclose(o.markers["x"][inside], (x0 + DT * 1.5e-9)[inside], rtol=0, atol=1e-9)
    np.testing.assert_allclose(o.markers["y"][inside], (y0 + DT * -0.5e-9)[inside], rtol=0, atol=1e-9)
    assert np.array_equal(o.markers["xx"], xx0) and np.array_equal(o.markers["xy"], xy0)


def test_rotation_preserves_stress_invariant():
    cfg = synth.config_grid(21, 15, 20e3, 14e3, 3, refine=True)
    cfg.nodal["esp"][:] = 3e-9     # theta = esp*dt ~ 0.1 rad
    o = Oracle(cfg)
    s0 = np.hypot(cfg.markers["xx"], cfg.markers["xy"])
    o.moveandrotatemarkers()
    s1 = np.hypot(o.markers["xx"], o.markers["xy"])
    np.testing.assert_allclose(s1, s0, rtol=1e-14)
    # the rotation angle is 2*theta: compare against a direct rotation
    th = 3e-9 * DT
    a0 = np.arctan2(cfg.markers["xy"], cfg.markers["xx"])
    a1 = np.arctan2(o.markers["xy"], o.markers["xx"])
    d = np.angle(np.exp(1j * (a1 - a0)))
    np.testing.assert_allclose(d, -2 * th, atol=1e-12)


def _rsf_node_cfg():
    cfg = synth.config_grid(9, 7, 8e3, 6e3, 1)
    cfg.params.inertyn = 1                  # no radiation damping
    n = cfg.grid.nodenum
    rng = np.random.default_rng(4)
    for k, v in (("as", 0.012), ("bs", 0.015), ("Ls", 0.02), ("V0s", 4e-9), ("mus_s", 0.6)):
        cfg.nodal[k] = np.full(n, v)
    cfg.nodal["phis"] = rng.uniform(8, 12, n)
    cfg.nodal["nu_m2c"] = np.full(n, 1e21)
    cfg.nodal["RK_sliprs_final"] = 10 ** rng.uniform(-11, -8, n)
    cfg.nodal["sxxs"] = np.full(n, 1e6)
    cfg.nodal["sxy"] = np.full(n, 1e6)
    cfg.timesum = DT
    return cfg


def test_rsf_yield_and_sliprate_are_inverse():
    """syield(V) (calc_RSF.c:129) and V(sigin) (calc_RSF.c:192) are an
    asinh/sinh pair: setting sigin = syield recovers V."""
    cfg = _rsf_node_cfg()
    o = Oracle(cfg)
    o.compute_before_solve(DT, 0)
    n = o.nodal
    interior = n["syields"] != 0
    V = n["sliprs"].copy()
    n["sxxs"][:] = n["syields"]
    n["sxy"][:] = 0.0
    n["exxpl"][:] = 1.0           # cap V <= EII*W does not bind
    n["exy"][:] = 0.0
    o.compute_after_solve(DT, 0)
    np.testing.assert_allclose(n["RK_sliprs_solved"][interior], V[interior], rtol=1e-12)


def test_aging_law_steady_state_and_branch_continuity():
    """Large V*dt/L: phi' -> ln(V0/V) (steady state of dphi/dt = V0/L e^-phi - V/L);
    the small-step branch (V*dt/L <= 1e-6) joins the exact one continuously."""
    cfg = _rsf_node_cfg()
    V = 1e-6
    cfg.nodal["RK_sliprs_final"][:] = V
    o = Oracle(cfg)
    o.compute_before_solve(1e7, 0)          # V dt / L = 500
    phi = o.nodal["phis_it"][o.nodal["syields"] != 0]
    np.testing.assert_allclose(phi, math.log(4e-9 / V), rtol=1e-12)

    def state(v, dt, phi0, L=0.02, V0=4e-9):
        if v * dt / L <= 1e-6:
            return math.log(math.exp(phi0) * (1 - v * dt / L) + V0 * dt / L)
        return math.log(V0 / v + (math.exp(phi0) - V0 / v) * math.exp(-v * dt / L))
    for phi0 in (5.0, 10.0, 12.0):
        lo = state(1e-9, 0.02 * 1e-6 / 1e-9, phi0)
        hi = state(1e-9 * (1 + 1e-12), 0.02 * 1e-6 / 1e-9, phi0)
        assert abs(lo - hi) <= 1e-9 * abs(lo)


def test_ductile_branches_and_newton_residual():
    cfg = synth.config_edge()
    o = Oracle(cfg)
    M = cfg.mats
    # Newtonian and sticky-air types return the table viscosity
    assert o.ductilevisc(8, 1e7, 0, 900, 1e3)[0] == M.nu[8]
    assert o.ductilevisc(1, 1e7, 0, 900, 1e3)[0] == M.nu[1]
    # zero stress: diffusion creep only, 0.5*k1*exp(E/RT) clamped
    t, T, pb = 6, 1500.0, 2e4
    k1 = M.nu[t] / M.ss[t] ** (M.mm[t] - 1)
    e1 = math.exp(min((M.dh[t] + M.dv[t] * pb) / (8.314 * T), 150.0))
    want = min(max(0.5 * k1 * e1, M.n0[t]), M.n1[t])
    assert o.ductilevisc(t, 0.0, 0.0, T, pb)[0] == pytest.approx(want, rel=1e-15)
    # build-defined NEWTON mode solves the reference residual to 1e-10 and
    # agrees with the reference's quasi-Newton within its 1e-6 stop rule
    for t, T in ((2, 700.0), (3, 900.0), (5, 1300.0), (6, 1550.0), (7, 1100.0)):
        ref, it = o.ductilevisc(t, 2e7, -1e7, T, 3e3)
        new, itn, res = o.ductilevisc_newton(t, 2e7, -1e7, T, 3e3)
        assert res <= 1e-10
        assert new == pytest.approx(ref, rel=2e-5)
