"""Learned warm start for the CDMFT bath fit.

A small MLP (n_in -> H -> H -> n_out, tanh hidden layers) maps model parameters such as
(U, mu, t') to the bath parameters the plain CDMFT loop converges to. It is trained on
the device (qed_mlp_train: fused FP32 forward/backward kernels and Adam) on solutions of
the plain loop, evaluated with BF16 matrix-core inference (qed_mlp_forward_bf16), and
replaces the starting point x0 of the CDMFT iteration (pyqcm/cdmft.py:300-302). The plain
loop then runs from that point unchanged: it is both the fallback and the arbiter of the
prediction, since convergence is still decided by its own tests.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


class MLP:
    """Device-resident MLP (include/qed.h qed_mlp_*). Inputs and targets are normalized
    with the statistics of the training set (kept here)."""

    def __init__(self, n_in=3, n_hidden=128, n_out=16, seed=0):
        self.n_in, self.n_hidden, self.n_out = n_in, n_hidden, n_out
        h = ctypes.c_void_p()
        _lib.call("qed_mlp_create", n_in, n_hidden, n_out, ctypes.c_uint64(seed), ctypes.byref(h))
        self._h = h
        self.x_mean = np.zeros(n_in)
        self.x_std = np.ones(n_in)
        self.y_mean = np.zeros(n_out)
        self.y_std = np.ones(n_out)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.call("qed_mlp_destroy", h)
            except Exception:
                pass
            self._h = None

    @property
    def n_params(self):
        c = ctypes.c_int()
        _lib.call("qed_mlp_get_params", self._h, None, ctypes.byref(c))
        return c.value

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        c = ctypes.c_int()
        _lib.call("qed_mlp_get_params", self._h, _lib.fp(p), ctypes.byref(c))
        return p

    def set_params(self, p):
        p = _f32(p)
        _lib.call("qed_mlp_set_params", self._h, _lib.fp(p), len(p))

    # raw (normalized-space) calls
    def forward_raw(self, x, bf16=False):
        x = _f32(np.atleast_2d(x))
        y = np.zeros((len(x), self.n_out), np.float32)
        fn = "qed_mlp_forward_bf16" if bf16 else "qed_mlp_forward"
        _lib.call(fn, self._h, len(x), _lib.fp(x), _lib.fp(y))
        return y

    def loss_grad_raw(self, x, t):
        x, t = _f32(np.atleast_2d(x)), _f32(np.atleast_2d(t))
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.n_params, np.float32)
        _lib.call("qed_mlp_loss_grad", self._h, len(x), _lib.fp(x), _lib.fp(t), _lib.fp(loss), _lib.fp(g))
        return float(loss[0]), g

    def train_raw(self, x, t, epochs, batch=64, lr=1e-3, seed=0):
        x, t = _f32(np.atleast_2d(x)), _f32(np.atleast_2d(t))
        hist = np.zeros(epochs, np.float32)
        _lib.call("qed_mlp_train", self._h, len(x), _lib.fp(x), _lib.fp(t), int(epochs), int(batch),
                  ctypes.c_float(lr), ctypes.c_uint64(seed), _lib.fp(hist))
        return hist

    # physical-units interface
    def fit(self, X, Y, epochs=2000, batch=64, lr=1e-3, seed=0):
        X, Y = np.asarray(X, float), np.asarray(Y, float)
        self.x_mean, self.x_std = X.mean(0), X.std(0) + 1e-12
        self.y_mean, self.y_std = Y.mean(0), Y.std(0) + 1e-12
        return self.train_raw((X - self.x_mean) / self.x_std, (Y - self.y_mean) / self.y_std,
                              epochs, batch, lr, seed)

    def predict(self, X, bf16=True):
        X = np.atleast_2d(np.asarray(X, float))
        y = self.forward_raw((X - self.x_mean) / self.x_std, bf16=bf16).astype(float)
        return y * self.y_std + self.y_mean


def plaquette_features(values):
    """(U, mu, t') of a parameter dict: the inputs of the 2x2+8-bath warm start."""
    return np.array([values["U"], values["mu"], values.get("tp", 0.0)], float)


# ------------------------------------------------------------------ 2x2 + 8-bath problem

PLAQUETTE_COLD_START = np.array([-1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0] + [0.5] * 8)


def plaquette_points(n, seed):
    """Seeded model points: U ~ U[0.5, 12], mu ~ U[0, U], t' ~ U[-0.5, 0.5] (t = 1)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        U = rng.uniform(0.5, 12.0)
        out.append({"t": 1.0, "tp": rng.uniform(-0.5, 0.5), "U": U, "mu": rng.uniform(0.0, U)})
    return out


def plaquette_problem(name="plaq_ws"):
    from . import cdmft as C
    M = C.plaquette_8b_model(name)
    return C.CDMFTProblem(M, C.PLAQUETTE_VARIA, C.plaquette_sectors)


def solve_points(problem, points, x0, batch=64, verbose=False, alpha=None):
    """Plain CDMFT (batched engine, `batch` problems active at a time) from the starting
    points x0 (one per point or one for all) and, optionally, Broyden inverse-Jacobian
    starts `alpha` (one per point); returns (x (n x nvar, canonical gauge), converged,
    niter, solves, inverse Jacobians reached, in the canonical coordinates of x))."""
    from . import cdmft as C
    x0 = np.atleast_2d(x0)
    if len(x0) == 1:
        x0 = np.repeat(x0, len(points), 0)
    B = C.BatchedCDMFT(problem, points, list(x0), solver="batched", verbose=verbose, capacity=batch,
                       alpha_list=alpha)
    good = B.converged & ~B.failed
    if verbose:
        print("%d points: %d converged" % (len(points), int(good.sum())), flush=True)
    Xr = np.array(B.x)
    maps = [C.canonical_map(x) for x in Xr]
    ij = np.array([Pm @ a @ Pm.T for Pm, a in zip(maps, B.alpha)])
    return C.canonical_bath(Xr), good, np.array(B.niter), np.array(B.n_solves), ij


DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def load_training_set(pattern="plaquette_warmstart"):
    """All committed training files (data/<pattern>*.npz): features and converged baths of
    the points the plain loop converged."""
    X, Y = [], []
    for f in sorted(os.listdir(DATA_DIR)):
        if f.startswith(pattern) and f.endswith(".npz"):
            d = np.load(os.path.join(DATA_DIR, f))
            ok = d["converged"].astype(bool)
            X.append(d["features"][ok])
            Y.append(d["bath"][ok])
    if not X:
        raise FileNotFoundError("no training data in " + DATA_DIR)
    return np.concatenate(X), np.concatenate(Y)


def load_jacobians(pattern="plaquette_warmstart"):
    """Features and Broyden inverse Jacobians (canonical coordinates) of the converged
    points of the training files that hold them; (None, None) if none does."""
    F, J = [], []
    for f in sorted(os.listdir(DATA_DIR)):
        if f.startswith(pattern) and f.endswith(".npz"):
            d = np.load(os.path.join(DATA_DIR, f))
            if "inverse_jacobian" not in d.files:
                continue
            ok = d["converged"].astype(bool)
            F.append(d["features"][ok])
            J.append(d["inverse_jacobian"][ok])
    if not F:
        return None, None
    return np.concatenate(F), np.concatenate(J)


def nearest_jacobians(feats, F, J, scale):
    """For each feature row, the inverse Jacobian of the nearest training point (features
    divided by `scale`)."""
    d = ((feats[:, None, :] - F[None, :, :]) / scale) ** 2
    return [J[k] for k in d.sum(-1).argmin(1)]


def generate_training_set(n, seed, path, batch=64, chunk=128, verbose=True):
    """Training data for the 2x2+8-bath warm start: n seeded points solved by the plain
    loop from PLAQUETTE_COLD_START, `chunk` points at a time; the .npz (features U, mu, t';
    converged baths) is rewritten after every chunk with the points solved so far."""
    P = plaquette_problem()
    pts = plaquette_points(n, seed)
    feats = np.array([plaquette_features(v) for v in pts])
    parts = []
    for c0 in range(0, n, chunk):
        parts.append(solve_points(P, pts[c0:c0 + chunk], PLAQUETTE_COLD_START, batch, verbose))
        X, conv, nit, ns, ij = (np.concatenate(a) for a in zip(*parts))
        k = len(conv)
        np.savez(path, features=feats[:k], bath=X, converged=conv, iterations=nit, solves=ns, seed=seed,
                 inverse_jacobian=ij)
        if verbose:
            print("points 0..%d solved, %d converged, saved to %s" % (k, int(conv.sum()), path), flush=True)
    return feats, X, conv


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="warm-start training data for the 2x2+8-bath CDMFT")
    ap.add_argument("--points", type=int, default=384)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--out", default=os.path.join(DATA_DIR, "plaquette_warmstart.npz"))
    a = ap.parse_args()
    from . import _lib as L
    L.call("qed_init", 0)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    generate_training_set(a.points, a.seed, a.out)
    print("wrote", a.out)
