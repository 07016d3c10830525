"""Learned planner: training data from the exact count pass, GPU training,
tile-level error.

Labels are log-gaps to upper bounds computed from the features (output nnz
vs min(products, area); distinct columns vs summed fiber lengths), so the
network's job is the collision/packing factor and a tile unlike the training
set is predicted near its bound rather than near zero.

The exact planner (count-then-fill) records, for every left tile, six cheap
features and two labels (exact output nnz, distinct output columns) when
called with ``collect_stats=True``. ``train`` fits the MLP on the GPU with the
library's own forward/backward kernels (fcc_mlp_train); ``tile_mape`` scores
predictions against exact labels. The exact planner stays the fallback: a
misprediction only costs a rerun or a global-hash spill, never a wrong result.
"""
from __future__ import annotations

import os

import numpy as np

from . import workloads as W

PARAMS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mlp_params.npy")


def collect(ctx, wl, tile_hints=(0,), device_tensors=None):
    """Exact contractions of ``wl`` with statistics on; returns (features, labels)."""
    if device_tensors is None:
        a, b = wl.tensors()
        b = a if wl.self_contraction else b
    else:
        a, b = device_tensors
    fs, ls = [], []
    for th in tile_hints:
        r = ctx.contract(a, list(wl.a_modes), b, list(wl.b_modes), tile_hint=th, collect_stats=True)
        r.free()
        f, l = ctx.tile_stats()
        fs.append(f)
        ls.append(l)
    return np.concatenate(fs), np.concatenate(ls)


def training_workloads():
    """Configurations 1-2 at full size and at smaller scales (several row tiles)."""
    return [(W.c1(), (0, 8, 64)), (W.c2(), (0, 8, 32)), (W.c1(0.5), (0, 32)), (W.c2(0.5), (0, 16)),
            (W.c2(0.25), (0, 4))]


def train(ctx, features, labels, epochs=60, batch=256, lr=2e-3, seed=7):
    loss = ctx.train_mlp(features, labels, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return ctx.mlp_params(), loss


def predict(ctx, features):
    """Predicted (nnz, distinct columns) per tile with the context's MLP (GPU)."""
    p = ctx.mlp_params()
    dummy = np.zeros((features.shape[0], 2), dtype=np.float32)
    pred, _, _ = ctx.mlp_forward_backward(features, dummy)
    nout = 2
    om = p[-2 * nout:-nout]
    os_ = p[-nout:]
    gap = np.minimum(pred * os_ + om, 0.0)
    ub = np.stack([np.minimum(features[:, 3], features[:, 5]), features[:, 2]], 1)
    return np.maximum(np.expm1(np.minimum(gap + ub, 80.0)), 0.0)


def tile_mape(ctx, features, labels):
    """Mean |pred - actual| / max(actual, 1) of tile output nnz (and of distinct columns)."""
    pr = predict(ctx, features)
    ub = np.stack([np.minimum(features[:, 3], features[:, 5]), features[:, 2]], 1)
    act = np.expm1(labels + ub)
    e_nnz = np.mean(np.abs(pr[:, 0] - act[:, 0]) / np.maximum(act[:, 0], 1.0))
    e_col = np.mean(np.abs(pr[:, 1] - act[:, 1]) / np.maximum(act[:, 1], 1.0))
    return float(e_nnz), float(e_col)


def main(argv=None):
    """Train on configurations 1-2 and save the parameters (default
    fastcc_amd/mlp_params.npy; run on a GPU host)."""
    import argparse
    import json
    from . import Context
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=PARAMS_PATH)
    args = ap.parse_args(argv)
    ctx = Context(0)
    fs, ls = [], []
    for wl, hints in training_workloads():
        f, l = collect(ctx, wl, hints)
        print(json.dumps({"workload": wl.name, "tile_hints": list(hints), "tiles": int(f.shape[0])}), flush=True)
        fs.append(f)
        ls.append(l)
    F, L = np.concatenate(fs), np.concatenate(ls)
    rng = np.random.default_rng(0)
    idx = rng.permutation(F.shape[0])
    nval = max(1, F.shape[0] // 10)
    va, tr = idx[:nval], idx[nval:]
    params, loss = train(ctx, F[tr], L[tr])
    m_tr = tile_mape(ctx, F[tr], L[tr])
    m_va = tile_mape(ctx, F[va], L[va])
    np.save(args.out, params)
    print(json.dumps({"train_tiles": int(tr.size), "val_tiles": int(va.size), "loss": loss,
                      "mape_nnz_train": m_tr[0], "mape_cols_train": m_tr[1],
                      "mape_nnz_val": m_va[0], "mape_cols_val": m_va[1], "saved": args.out}), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
