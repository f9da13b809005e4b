"""Learned warm start of the per-marker local update: sample generation from
plain FP64 Newton solves, training on the GPU (libv2c's MFMA kernels),
evaluation of iterations and wall time, plain against warm.

The network (9 -> 64 -> 64 -> 2, tanh) predicts log10(root / default start)
for the ductile Newton unknown and for the fault slip rate; Newton to the
same residual remains the arbiter, so a poor guess costs iterations, never
accuracy.  Inputs: log10 strain-rate invariant, T, P, log10 stress invariant,
log10 theta, a-b, and for fault markers log10 of the slip-rate default start,
log10 L and log10 of the previous slip rate.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import synth
from .abi import Context, _dptr

IN, H, OUT = 9, 64, 2
OFF = {"W1": 0, "b1": H * IN}
OFF["W2"] = OFF["b1"] + H
OFF["b2"] = OFF["W2"] + H * H
OFF["W3"] = OFF["b2"] + H
OFF["b3"] = OFF["W3"] + OUT * H
NPARAM = OFF["b3"] + OUT
NNORM = 2 * IN + 4
ROW = IN + 2                     # sample row: IN features, 2 labels


def _bind(lib):
    P = C.POINTER
    fp = P(C.c_float)
    for name, args in [("v2c_mlp_set", [fp, fp]), ("v2c_mlp_get", [fp, fp]),
                       ("v2c_mlp_predict", [C.c_double]),
                       ("v2c_mlp_predict_outputs", [C.c_double, C.c_int]),
                       ("v2c_mlp_samples", [C.c_double, C.c_double, fp, fp]),
                       ("v2c_mlp_train", [fp, fp, fp, C.c_long, C.c_int, C.c_int, C.c_float, C.c_uint, fp]),
                       ("v2c_mlp_grad", [fp, fp, fp, C.c_long, fp, fp]),
                       ("v2c_mlp_forward", [fp, C.c_long, fp])]:
        f = getattr(lib, name)
        f.argtypes = [C.c_void_p] + args
        f.restype = C.c_int


def _fptr(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def init_params(seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, dtype=np.float32)
    p[OFF["W1"]:OFF["b1"]] = rng.normal(0, 1 / np.sqrt(IN), H * IN)
    p[OFF["W2"]:OFF["b2"]] = rng.normal(0, 1 / np.sqrt(H), H * H)
    p[OFF["W3"]:OFF["b3"]] = rng.normal(0, 1 / np.sqrt(H), OUT * H)
    return p


class WarmStart:
    """The MLP on one context: parameters, normalisation, training, use."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        _bind(ctx.lib)
        self.norm = None
        self.params = None

    # --- data ------------------------------------------------------------
    def samples(self, dt: float, tol: float = 1e-10):
        """Features/labels from one plain local update of the resident markers."""
        n = self.ctx.lib.v2c_marker_count(self.ctx.h)
        X = np.zeros((n, ROW), dtype=np.float32)
        M = np.zeros((n, 2), dtype=np.float32)
        self.ctx._check(self.ctx.lib.v2c_mlp_samples(self.ctx.h, dt, tol, _fptr(X), _fptr(M)), "mlp_samples")
        keep = M.sum(1) > 0
        return X[keep], M[keep]

    def fit_norm(self, X, M):
        mu = X[:, :IN].mean(0)
        sd = X[:, :IN].std(0)
        sd[sd == 0] = 1.0
        om, osd = np.zeros(2), np.ones(2)
        for o in range(2):
            y = X[M[:, o] > 0, IN + o]
            if y.size:
                om[o], osd[o] = y.mean(), max(y.std(), 1e-6)
        self.norm = np.concatenate([mu, sd, om, osd]).astype(np.float32)
        return self.norm

    def standardise(self, X):
        n = self.norm
        Xs = ((X[:, :IN] - n[:IN]) / n[IN:2 * IN]).astype(np.float32)
        Ts = ((X[:, IN:IN + 2] - n[2 * IN:2 * IN + 2]) / n[2 * IN + 2:]).astype(np.float32)
        return np.ascontiguousarray(Xs), np.ascontiguousarray(Ts)

    # --- parameters ------------------------------------------------------
    def load(self, params, norm=None):
        if norm is not None:
            self.norm = np.ascontiguousarray(norm, dtype=np.float32)
        self.params = np.ascontiguousarray(params, dtype=np.float32)
        self.ctx._check(self.ctx.lib.v2c_mlp_set(self.ctx.h, _fptr(self.params), _fptr(self.norm)), "mlp_set")

    def fetch(self):
        p = np.zeros(NPARAM, dtype=np.float32)
        nrm = np.zeros(NNORM, dtype=np.float32)
        self.ctx._check(self.ctx.lib.v2c_mlp_get(self.ctx.h, _fptr(p), _fptr(nrm)), "mlp_get")
        self.params = p
        return p

    # --- training --------------------------------------------------------
    def train(self, X, M, epochs=8, batch=4096, lr=2e-3, seed=1):
        Xs, Ts = self.standardise(X)
        Ms = np.ascontiguousarray(M, dtype=np.float32)
        loss = np.zeros(epochs, dtype=np.float32)
        t0 = time.perf_counter()
        self.ctx._check(self.ctx.lib.v2c_mlp_train(self.ctx.h, _fptr(Xs), _fptr(Ts), _fptr(Ms), len(Xs), epochs,
                                                   batch, lr, seed, _fptr(loss)), "mlp_train")
        self.train_seconds = time.perf_counter() - t0
        self.fetch()
        return loss

    def grad(self, Xs, Ts, Ms):
        g = np.zeros(NPARAM, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self.ctx._check(self.ctx.lib.v2c_mlp_grad(self.ctx.h, _fptr(Xs), _fptr(Ts), _fptr(Ms), len(Xs), _fptr(g),
                                                  _fptr(loss)), "mlp_grad")
        return g, float(loss[0])

    def forward(self, Xs):
        Y = np.zeros((len(Xs), 2), dtype=np.float32)
        self.ctx._check(self.ctx.lib.v2c_mlp_forward(self.ctx.h, _fptr(Xs), len(Xs), _fptr(Y)), "mlp_forward")
        return Y

    def predict(self, dt, outputs=3):
        """Guesses for the resident markers: outputs bit 0 ductile, bit 1 slip rate."""
        self.ctx._check(self.ctx.lib.v2c_mlp_predict_outputs(self.ctx.h, dt, outputs), "mlp_predict")


# ----------------------------------------------------------------------------- FP64 reference
def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def reference_grad(params, Xs, Ts, Ms, emulate_bf16=True):
    """FP64 forward/backward of the masked MSE (the CPU check of the MFMA
    kernels).  With emulate_bf16 the matmul operands are rounded to bf16 as
    the GPU rounds them."""
    q = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    p = params.astype(np.float64)
    o = OFF
    W1 = p[o["W1"]:o["b1"]].reshape(H, IN); b1 = p[o["b1"]:o["W2"]]
    W2 = p[o["W2"]:o["b2"]].reshape(H, H); b2 = p[o["b2"]:o["W3"]]
    W3 = p[o["W3"]:o["b3"]].reshape(OUT, H); b3 = p[o["b3"]:NPARAM]
    X = q(Xs)
    Z1 = X @ q(W1).T + b1
    H1 = np.tanh(Z1)
    Z2 = q(H1) @ q(W2).T + b2
    H2 = np.tanh(Z2)
    Y = q(H2) @ q(W3).T + b3
    cnt = Ms.sum()
    E = Y - Ts
    loss = float((Ms * E * E).sum() / cnt)
    dY = 2 * Ms * E / cnt
    g = np.zeros(NPARAM)
    g[o["W3"]:o["b3"]] = (q(dY).T @ q(H2)).ravel()
    g[o["b3"]:NPARAM] = dY.sum(0)
    dZ2 = (q(dY) @ q(W3)) * (1 - H2 * H2)
    g[o["W2"]:o["b2"]] = (q(dZ2).T @ q(H1)).ravel()
    g[o["b2"]:o["W3"]] = dZ2.sum(0)
    dZ1 = (q(dZ2) @ q(W2)) * (1 - H1 * H1)
    g[o["W1"]:o["b1"]] = (q(dZ1).T @ X).ravel()
    g[o["b1"]:o["W2"]] = dZ1.sum(0)
    return g, loss, Y


# ----------------------------------------------------------------------------- pipeline
def upload_cloud(ctx, cfg):
    ctx.set_time(cfg.timesum, cfg.timestep)
    ctx.upload_markers(cfg.markers)
    r = cfg.rsf
    ctx.rsf_upload(r["fault"], r["V"], r["phi"], r["a"], r["b"], r["L"], r["V0"], r["mu0"])


def generate(ctx, seeds, dt, maker=synth.config_small, max_rows=2_000_000, keep_other=1.0):
    """Training rows from plain solves on the clouds of `seeds`.  keep_other < 1
    keeps every row with a slip-rate label but only that fraction (seeded) of
    the others, so more clouds fit in max_rows when fault markers are rare."""
    ws = WarmStart(ctx)
    Xs, Ms = [], []
    for s in seeds:
        cfg = maker(seed=s)
        cfg.timestep = dt
        upload_cloud(ctx, cfg)
        X, M = ws.samples(dt)
        if keep_other < 1.0:
            rng = np.random.default_rng(s)
            keep = (M[:, 1] > 0) | (rng.uniform(size=len(M)) < keep_other)
            X, M = X[keep], M[keep]
        Xs.append(X)
        Ms.append(M)
        if sum(len(x) for x in Xs) >= max_rows:
            break
    X = np.concatenate(Xs)[:max_rows]
    M = np.concatenate(Ms)[:max_rows]
    return X, M


# mode -> (guesses predicted beforehand: bit 0 ductile, bit 1 slip rate; warm
# argument of the local update)
MODES = {"plain": (0, 0), "warm": (3, 1), "warm_rsf": (0, 2)}


def evaluate(ctx, cfg, ws: WarmStart, dt, tol=1e-10, repeats=3, modes=("plain", "warm")):
    """Plain vs warm local update on the same cloud: iterations, residual
    evaluations, fallbacks and kernel times (the median over `repeats`, and
    the best as ms_total_best).  Modes:
    "plain" (default starts), "warm" (network guesses for both solves,
    predicted for every marker by v2c_mlp_predict before the solves),
    "warm_rsf" (the slip-rate guess evaluated by the network inside the
    slip-rate solve, which runs over the fault markers only).  Times: the
    ductile solve (ms_ductile), the slip-rate solve (ms_rsf, network included
    for warm_rsf; it runs beside the ductile solve), the local update from
    fork to join (ms_wall), the separate inference (ms_mlp), and
    ms_total = ms_wall + ms_mlp."""
    runs = {mode: [] for mode in modes}
    stats = {}
    # repeats interleaved across the modes, so that drifts of the machine's
    # clock or load fall on every mode alike
    for r in range(repeats):
        for mode in modes:
            outputs, warm = MODES[mode]
            upload_cloud(ctx, cfg)
            ctx.enable_timing(True)
            ctx.reset_timing()
            if outputs:
                ws.predict(dt, outputs)
            st = ctx.local_update(dt, warm=warm, tol=tol)
            ms_d, _ = ctx.kernel_time("local")
            ms_r, _ = ctx.kernel_time("local_rsf")
            ms_mlp, _ = ctx.kernel_time("mlp")
            ms_wall, _ = ctx.kernel_time("local_wall")
            ctx.enable_timing(False)
            runs[mode].append((ms_wall + ms_mlp, ms_d, ms_r, ms_mlp, ms_wall))
            stats[mode] = st
            if r == repeats - 1:
                stats[mode + "/out"] = ctx.local_download(cfg.n)
    res = {}
    for mode in modes:
        t = np.array(runs[mode])
        med = np.median(t, axis=0)
        res[mode] = dict(stats=stats[mode], ms_total=float(med[0]), ms_ductile=float(med[1]),
                         ms_rsf=float(med[2]), ms_mlp=float(med[3]), ms_wall=float(med[4]),
                         ms_total_best=float(t[:, 0].min()), repeats=len(t), out=stats[mode + "/out"])
    return res
