"""Learned rank predictor for the randomized sparsification (rank mode 1).

The predictor maps 32 features of an interface block (log sizes, level, tolerance, neighbour
counts, column-norm statistics and the R-diagonal decay of a 16-column Gaussian sketch, all
computed on the GPU during factorize) to the fraction rank / min(k, w) that geqp3 would keep
at the sparsification tolerance.  The planner sizes the interface's sketch from it.

Everything here runs the library's HIP kernels (spaqr_mlp_* in include/spaqr_gpu.h):
forward, backward and Adam.  Training data are (features, rank) pairs logged from the plain
geqp3 factorizations of the 2D Poisson and 3D variable-coefficient diffusion problems
(configurations 1 and 2) at several tolerances.

    python -m pspaqr_public_amd.rank_model --out pspaqr_public_amd/rank_mlp_weights.npy
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from ._lib import check, lib

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_FILE = os.path.join(HERE, "rank_mlp_weights.npy")
NF, NH = 32, 64


def n_params() -> int:
    return lib().spaqr_mlp_params()


def init_weights(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights, zero biases (flat layout of include/spaqr_gpu.h)."""
    rng = np.random.RandomState(seed)

    def glorot(fo, fi):
        a = np.sqrt(6.0 / (fi + fo))
        return rng.uniform(-a, a, size=(fo, fi))

    W = np.concatenate([glorot(NH, NF).ravel(), np.zeros(NH), glorot(NH, NH).ravel(), np.zeros(NH),
                        glorot(1, NH).ravel(), np.zeros(1)]).astype(np.float32)
    assert W.size == n_params()
    return W


def predict(X: np.ndarray, W: np.ndarray) -> np.ndarray:
    X = np.ascontiguousarray(X, dtype=np.float32)
    W = np.ascontiguousarray(W, dtype=np.float32)
    out = np.zeros(X.shape[0], dtype=np.float32)
    check(lib().spaqr_mlp_predict(X.ctypes.data, X.shape[0], W.ctypes.data, out.ctypes.data), None, "mlp_predict")
    return out


def grad(X: np.ndarray, y: np.ndarray, W: np.ndarray):
    X = np.ascontiguousarray(X, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    W = np.ascontiguousarray(W, dtype=np.float32)
    g = np.zeros(W.size, dtype=np.float32)
    L = np.zeros(1, dtype=np.float32)
    check(lib().spaqr_mlp_grad(X.ctypes.data, y.ctypes.data, X.shape[0], W.ctypes.data, g.ctypes.data, L.ctypes.data),
          None, "mlp_grad")
    return float(L[0]), g


def train(X: np.ndarray, y: np.ndarray, W0: np.ndarray, epochs: int = 3000, lr: float = 3e-3):
    """Full-batch Adam on the GPU; returns (weights, loss history)."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    W = np.array(W0, dtype=np.float32, copy=True)
    hist = np.zeros(max(epochs, 1), dtype=np.float32)
    check(lib().spaqr_mlp_train(X.ctypes.data, y.ctypes.data, X.shape[0], W.ctypes.data, int(epochs), float(lr),
                                hist.ctypes.data), None, "mlp_train")
    return W, hist[:epochs]


def load_weights(path: str = WEIGHTS_FILE) -> np.ndarray:
    if not os.path.exists(path):
        raise FileNotFoundError("rank predictor weights not found: %s (train them with "
                                "python -m pspaqr_public_amd.rank_model)" % path)
    W = np.load(path)
    if W.dtype != np.float32 or W.size != n_params():
        raise ValueError("rank predictor weights have the wrong shape or type: %s" % path)
    return W


def collect(cases, log=print):
    """Factorize each (config, side, tol) with the plain geqp3 path (FP64) and return the
    logged (features, rank fraction, meta)."""
    from .problems import make_problem
    from .ptree import build_tree

    Xs, ys, metas = [], [], []
    for cfg, side, tol in cases:
        P = make_problem(cfg, side)
        T = build_tree(P, 64, tol=tol)
        T.set_rank_log(True)
        T.factorize()
        X, meta = T.rank_log()
        kw = np.minimum(meta[:, 1], meta[:, 2]).astype(np.float64)
        y = meta[:, 0] / np.maximum(kw, 1)
        log("[rank_model] %s tol %g: %d interfaces, mean rank fraction %.3f" % (P.name, tol, len(y), y.mean()
                                                                                 if len(y) else 0))
        Xs.append(X)
        ys.append(y.astype(np.float32))
        metas.append(meta)
        del T
    return np.concatenate(Xs), np.concatenate(ys), np.concatenate(metas)


TRAIN_CASES = [(1, None, 1e-1), (1, None, 1e-2), (1, None, 1e-3), (2, None, 1e-1), (2, None, 1e-2), (2, None, 1e-3)]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--out", default=WEIGHTS_FILE)
    ap.add_argument("--epochs", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)
    t0 = time.perf_counter()
    X, y, meta = collect(TRAIN_CASES)
    t1 = time.perf_counter()
    W0 = init_weights(args.seed)
    W, hist = train(X, y, W0, args.epochs, args.lr)
    t2 = time.perf_counter()
    p = predict(X, W)
    kw = np.minimum(meta[:, 1], meta[:, 2])
    err = np.abs(p - y) * kw
    summary = dict(samples=int(len(y)), epochs=args.epochs, lr=args.lr, loss_first=float(hist[0]),
                   loss_last=float(hist[-1]), mean_abs_rank_err=float(err.mean()),
                   p90_abs_rank_err=float(np.percentile(err, 90)), collect_s=t1 - t0, train_s=t2 - t1)
    np.save(args.out, W)
    print(json.dumps(summary))
    return 0


if __name__ == "__main__":
    sys.exit(main())
