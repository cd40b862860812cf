"""Training data, weights and evaluation of the learned form's network.

The network predicts y = log(alpha_SCS,i(i w_k) / alpha_i(i w_k)) per atom and
grid point from the device descriptors (see mbd_amd/csrc/net.hip); its targets
come only from the plain FP64 path (``MbdEngine.scs("fp64")``).  Training runs
on the GPU with the library's own kernels (``engine.net_train``).

Weights are stored as ``net_weights.npz`` beside this file: the float32 blob,
the layout it was made for, a format version and its SHA-256.

    python -m mbd_amd.learned --out mbd_amd/net_weights.npz     # on a GPU
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

from . import engine, generators

VERSION = 1
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "net_weights.npz")


def training_systems(seed: int = 11):
    """Config-4-style chain molecules plus small periodic cells and sheets."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(48):
        n = int(round(np.exp(rng.uniform(np.log(8), np.log(128)))))
        geo = generators.chain_molecule(rng, max(n - 2, 1))
        out.append((geo, generators.Settings()))
    for k in range(8):
        n = 256
        box = 38.0 * (n / 4096) ** (1.0 / 3.0)
        out.append(generators.random_periodic(n, box, seed=100 + k))
    for k in range(4):
        out.append(generators.hex_sheet(8, 8, seed=200 + k, rc_scs=40.0))
    return out


def sample(eng: "engine.MbdEngine", geo, st):
    """(X, y) for one system: descriptors and FP64 log screening ratios, or
    None when the FP64 solve is unphysical (a non-positive alpha_SCS, or a
    non-SPD relay matrix) and the system cannot serve as training data."""
    eng.set_geometry(geo, st)
    X, aw = eng.net_descriptors()
    aeff, cpol, _, _, _, diag = eng.scs("fp64")
    alpha = np.concatenate([aeff[None, :], cpol.T], axis=0).reshape(-1)  # s = k N + i
    if diag["info_scs"] != 0 or not np.all(alpha > 0):
        return None
    return X, np.log(alpha / aw)


def init_blob(X, seed: int = 3):
    n_in, H, nb = engine.net_layout()
    rng = np.random.default_rng(seed)
    mean = X.mean(axis=0)
    std = X.std(axis=0) + 1e-6
    params = []
    dims = [n_in, H, H, H, 1]
    for l in range(4):
        fi, fo = dims[l], dims[l + 1]
        W = rng.normal(0.0, np.sqrt(1.0 / fi), (fo, fi))
        params.append(W.T.reshape(-1))  # column-major fo x fi
        params.append(np.zeros(fo))
    blob = np.concatenate([mean, std] + params).astype(np.float32)
    assert blob.size == nb
    return blob


def save(path, blob, meta):
    n_in, H, nb = engine.net_layout()
    digest = hashlib.sha256(blob.tobytes()).hexdigest()
    np.savez(path, blob=blob, version=VERSION, n_in=n_in, hidden=H,
             sha256=digest, meta=json.dumps(meta))
    return digest


def load(path=DEFAULT_WEIGHTS):
    """The weight blob, after checking its layout, version and hash."""
    z = np.load(path, allow_pickle=False)
    blob = z["blob"].astype(np.float32)
    n_in, H, nb = engine.net_layout()
    if int(z["version"]) != VERSION or int(z["n_in"]) != n_in or int(z["hidden"]) != H:
        raise ValueError(f"{path}: weights made for another network layout")
    if hashlib.sha256(blob.tobytes()).hexdigest() != str(z["sha256"]):
        raise ValueError(f"{path}: weight hash mismatch")
    return blob


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-3)
    args = ap.parse_args()
    t0 = time.time()
    eng = engine.MbdEngine(0)
    Xs, ys = [], []
    dropped = 0
    for geo, st in training_systems():
        xy = sample(eng, geo, st)
        if xy is None:
            dropped += 1
            continue
        Xs.append(xy[0])
        ys.append(xy[1])
    X = np.concatenate(Xs).astype(np.float32)
    y = np.concatenate(ys).astype(np.float32)
    print(f"learned: {X.shape[0]} samples from {len(Xs)} systems, {dropped} unphysical "
          f"systems dropped ({time.time() - t0:.1f} s)", flush=True)
    blob0 = init_blob(X)
    blob, hist = engine.net_train(blob0, X, y, epochs=args.epochs, batch=args.batch, lr=args.lr)
    rmse = float(np.sqrt(np.mean((engine.net_forward(blob, X) - y) ** 2)))
    print(f"learned: loss {hist[0]:.3e} -> {hist[-1]:.3e}, train rmse {rmse:.3e} "
          f"({time.time() - t0:.1f} s)", flush=True)
    digest = save(args.out, blob, {"samples": int(X.shape[0]), "epochs": args.epochs,
                                   "batch": args.batch, "lr": args.lr,
                                   "loss_first": float(hist[0]), "loss_last": float(hist[-1]),
                                   "train_rmse": rmse})
    print(f"learned: wrote {args.out} sha256={digest}", flush=True)
    eng.close()


if __name__ == "__main__":
    sys.exit(main())
