"""The FP64 reference MLP that checks the MFMA kernels is itself pinned by
central finite differences (CPU)."""
import numpy as np

from velo2cycles_amd.warmstart import IN, NPARAM, init_params, reference_grad


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = init_params(3).astype(np.float64) * 1.3
    Xs = rng.normal(size=(50, IN))
    Ts = rng.normal(size=(50, 2))
    Ms = (rng.uniform(size=(50, 2)) > 0.3).astype(np.float64)
    g, loss, _ = reference_grad(p, Xs, Ts, Ms, emulate_bf16=False)
    idx = rng.choice(NPARAM, 60, replace=False)
    h = 1e-6
    for i in idx:
        pp = p.copy(); pp[i] += h
        pm = p.copy(); pm[i] -= h
        fd = (reference_grad(pp, Xs, Ts, Ms, False)[1] - reference_grad(pm, Xs, Ts, Ms, False)[1]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
