"""Learned tile-shape tuner.

The reference fixes its tile shape at compile time (THREADBUNCH / SEG_H /
STEP, spmv_cuda/common.h:39-48) and picks the csr-vector width by nnz/m
(main.cu:238-287).  Here a small MLP picks, per matrix, one of CLASSES -- the
segmented sum at a tile shape (items per lane x waves per block x tiles per
wave), the same over a column-panel copy of A, or csr-vector at a width --
from 8 row-length statistics that the GPU reduces from row_ptr/col.  The network predicts log2(time / best time) for
every class; the tuner takes the argmin.  It is trained on timing sweeps this
module generates (the exhaustive autotuner, bhs_*_autotune, which also measures
the tuner's regret), by the library's own HIP forward/backward/Adam kernels.

Command line (GPU):
  python -m benchmark_spmv_using_csr_amd.tuner sweep --seeds 0:1000 --out sweep_a.npz
  python -m benchmark_spmv_using_csr_amd.tuner train --data sweep_a.npz sweep_b.npz --out tuner_net.json
  python -m benchmark_spmv_using_csr_amd.tuner eval  --data ... --net tuner_net.json
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

from . import _lib, matgen
from ._lib import check, spmv_lib

PKG = os.path.dirname(os.path.abspath(__file__))
DATA_DIR = os.path.join(PKG, "data")
DEFAULT_NET = os.path.join(DATA_DIR, "tuner_net.json")

# (variant, items per lane, waves per block, tiles per wave, csr-vector width or
# column panels): the segmented sum, csr-vector, and the segmented sum over a
# column-panel copy of A
CLASSES = [(0, S, wpb, T, 0) for S in (4, 8, 16) for wpb in (1, 4) for T in (1, 4, 8, 16)] + \
          [(1, 8, 4, 1, tpv) for tpv in (2, 4, 8, 16, 32, 64)] + \
          [(2, S, 1, T, npan) for npan in (2, 4) for S in (4, 8) for T in (2, 4, 8, 16)]
K = len(CLASSES)
# The method without the tuner: one fixed tile for every matrix.
DEFAULT_CLASS = CLASSES.index((0, 8, 4, 4, 0))
N_FEATURES = 8
HIDDEN = 64

# Workload families of the training set (matgen.family_matrix).
FAMILIES = ("uniform rows", "power law", "banded", "3x3 blocks")


def class_name(c):
    v, S, wpb, T, tpv = c
    if v == 1:
        return f"csr-vector tpv={tpv}"
    if v == 2:
        return f"segsum {tpv} column panels S={S} wpb={wpb} T={T}"
    return f"segsum S={S} wpb={wpb} T={T}"


def transform(raw: np.ndarray) -> np.ndarray:
    """The library's feature transform (bhs_tuner_transform) on rows of raw features."""
    raw = np.atleast_2d(np.asarray(raw, np.float64))
    out = np.empty((raw.shape[0], N_FEATURES), np.float32)
    L = spmv_lib()
    for i in range(raw.shape[0]):
        r = np.ascontiguousarray(raw[i])
        check(L.bhs_tuner_transform(r.ctypes.data_as(C.c_void_p),
                                    out[i].ctypes.data_as(C.c_void_p)), "tuner_transform")
    return out


def targets(us: np.ndarray):
    """log2(time / best) per class and a mask of the classes that ran."""
    us = np.atleast_2d(np.asarray(us, np.float64))
    mask = us > 0
    best = np.where(mask, us, np.inf).min(axis=1, keepdims=True)
    t = np.where(mask, np.log2(np.where(mask, us, 1.0) / best), 0.0)
    return t.astype(np.float32), mask.astype(np.float32)


class Mlp:
    """Device-resident MLP (bhs_mlp_*)."""

    def __init__(self, hidden=HIDDEN, n_out=K, device=0):
        self.hidden, self.n_out = hidden, n_out
        self.handle = C.c_void_p()
        check(spmv_lib().bhs_mlp_create(C.byref(self.handle), device, N_FEATURES, hidden, n_out),
              "mlp_create")
        self.n_params = spmv_lib().bhs_mlp_param_count(N_FEATURES, hidden, n_out)

    def set(self, params, mean, std):
        p = np.ascontiguousarray(params, np.float32)
        mu = np.ascontiguousarray(mean, np.float32)
        sd = np.ascontiguousarray(std, np.float32)
        check(spmv_lib().bhs_mlp_set_params(self.handle, p.ctypes.data_as(C.c_void_p), len(p),
                                            mu.ctypes.data_as(C.c_void_p),
                                            sd.ctypes.data_as(C.c_void_p)), "mlp_set_params")

    def params(self):
        p = np.empty(self.n_params, np.float32)
        check(spmv_lib().bhs_mlp_get_params(self.handle, p.ctypes.data_as(C.c_void_p), len(p)),
              "mlp_get_params")
        return p

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        out = np.empty((X.shape[0], self.n_out), np.float32)
        check(spmv_lib().bhs_mlp_forward(self.handle, X.ctypes.data_as(C.c_void_p), X.shape[0],
                                         out.ctypes.data_as(C.c_void_p)), "mlp_forward")
        return out

    def loss_grad(self, X, T, M):
        X, T, M = (np.ascontiguousarray(a, np.float32) for a in (X, T, M))
        g = np.empty(self.n_params, np.float32)
        loss = C.c_double(0)
        check(spmv_lib().bhs_mlp_loss_grad(self.handle, X.ctypes.data_as(C.c_void_p),
                                           T.ctypes.data_as(C.c_void_p),
                                           M.ctypes.data_as(C.c_void_p), X.shape[0],
                                           C.byref(loss), g.ctypes.data_as(C.c_void_p)),
              "mlp_loss_grad")
        return loss.value, g

    def train(self, X, T, M, epochs=3000, lr=3e-3, weight_decay=1e-4):
        X, T, M = (np.ascontiguousarray(a, np.float32) for a in (X, T, M))
        hist = np.empty(epochs, np.float64)
        check(spmv_lib().bhs_mlp_train(self.handle, X.ctypes.data_as(C.c_void_p),
                                       T.ctypes.data_as(C.c_void_p), M.ctypes.data_as(C.c_void_p),
                                       X.shape[0], epochs, lr, weight_decay,
                                       hist.ctypes.data_as(C.c_void_p)), "mlp_train")
        return hist

    def close(self):
        if self.handle:
            spmv_lib().bhs_mlp_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def loss_grad_f64(params, norm, X, T, M, hidden=HIDDEN, n_out=K, device=0):
    """The library's forward/backward kernels in double (gradient checks)."""
    params, norm, X, T, M = (np.ascontiguousarray(a, np.float64) for a in (params, norm, X, T, M))
    g = np.empty(len(params), np.float64)
    loss = C.c_double(0)
    check(spmv_lib().bhs_mlp_loss_grad_f64(device, N_FEATURES, hidden, n_out,
                                           params.ctypes.data_as(C.c_void_p),
                                           norm.ctypes.data_as(C.c_void_p),
                                           X.ctypes.data_as(C.c_void_p),
                                           T.ctypes.data_as(C.c_void_p),
                                           M.ctypes.data_as(C.c_void_p), X.shape[0],
                                           C.byref(loss), g.ctypes.data_as(C.c_void_p)),
          "mlp_loss_grad_f64")
    return loss.value, g


def init_params(hidden=HIDDEN, n_out=K, seed=0):
    """He-uniform weights, zero biases (numpy, seeded)."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((N_FEATURES, hidden), (hidden, hidden), (hidden, n_out)):
        lim = np.sqrt(6.0 / fan_in)
        parts.append(rng.uniform(-lim, lim, fan_out * fan_in))
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


def save_net(path, params, mean, std, meta=None):
    d = {"n_features": N_FEATURES, "hidden": HIDDEN, "classes": CLASSES,
         "default_class": DEFAULT_CLASS, "mean": [float(v) for v in mean],
         "std": [float(v) for v in std], "params": [float(v) for v in params], "meta": meta or {}}
    with open(path, "w") as f:
        json.dump(d, f)


def load_net(path=DEFAULT_NET):
    with open(path) as f:
        d = json.load(f)
    d["classes"] = [tuple(c) for c in d["classes"]]
    return d


def make_net(d, device=0) -> Mlp:
    net = Mlp(d["hidden"], len(d["classes"]), device)
    net.set(np.array(d["params"], np.float32), np.array(d["mean"], np.float32),
            np.array(d["std"], np.float32))
    return net


# ---------------------------------------------------------------------------
# dataset: seeded matrices, GPU row features, exhaustive sweep
# ---------------------------------------------------------------------------

def dataset_matrix(seed):
    """Matrix `seed` of the tuner set: m log-uniform in [1e3, 4e6], family
    seed mod 4, fp64 when (seed // 4) is even (both precisions in every family)."""
    A = matgen.family_matrix(seed, 1e3, 4e6)
    dtype = np.float64 if (seed // 4) % 2 == 0 else np.float32
    return A, dtype


def sweep(seeds, reps=7, flush_bytes=1 << 30, log=print):
    from .spmv import bhsparse_spmv
    rows = []
    t0 = time.time()
    for i, seed in enumerate(seeds):
        A, dtype = dataset_matrix(seed)
        x = matgen.vector(seed + 7, A.n)
        op = bhsparse_spmv(dtype)
        check(op.init_platform(0), "init_platform")
        y = np.zeros(A.m, dtype)
        check(op.prepare_mem(A.m, A.n, A.nnz, A.row_ptr, A.col, A.val.astype(dtype),
                             x.astype(dtype), y), "prepare_mem")
        raw = op.row_features()
        us, best = op.autotune(CLASSES, reps=reps, flush_bytes=flush_bytes)
        op.close()
        rows.append((seed, seed % 4, int(dtype == np.float64), A.m, A.nnz, raw, us))
        if log:
            log(f"[sweep {i + 1}/{len(seeds)} {time.time() - t0:.0f}s] seed={seed} fam={seed % 4} "
                f"{np.dtype(dtype).name} m={A.m} nnz={A.nnz} best={class_name(CLASSES[best])} "
                f"{us[best]:.1f}us default={us[DEFAULT_CLASS]:.1f}us")
        del A, x, y
    return {
        "seed": np.array([r[0] for r in rows], np.int64),
        "family": np.array([r[1] for r in rows], np.int64),
        "fp64": np.array([r[2] for r in rows], np.int64),
        "m": np.array([r[3] for r in rows], np.int64),
        "nnz": np.array([r[4] for r in rows], np.int64),
        "raw": np.array([r[5] for r in rows], np.float64),
        "us": np.array([r[6] for r in rows], np.float64),
        "classes": np.array(CLASSES, np.int64),
    }


def load_data(paths):
    parts = [dict(np.load(p, allow_pickle=False)) for p in paths]
    out = {k: np.concatenate([p[k] for p in parts]) for k in parts[0] if k != "classes"}
    out["classes"] = parts[0]["classes"]
    order = np.argsort(out["seed"])
    return {k: (v[order] if k != "classes" else v) for k, v in out.items()}


def split(data, n_train=1500):
    tr = data["seed"] < n_train
    return tr, ~tr


def train_net(data, epochs=4000, lr=3e-3, weight_decay=1e-4, seed=0, device=0):
    tr, _ = split(data)
    X = transform(data["raw"][tr])
    T, M = targets(data["us"][tr])
    mean = X.mean(axis=0)
    std = X.std(axis=0)
    std = np.where(std > 1e-6, std, 1.0).astype(np.float32)
    net = Mlp(HIDDEN, K, device)
    net.set(init_params(HIDDEN, K, seed), mean, std)
    hist = net.train(X, T, M, epochs=epochs, lr=lr, weight_decay=weight_decay)
    return net, mean, std, hist


def evaluate(data, scores, mask_sel):
    """Regret of argmin(scores) vs the exhaustive best, and speedup vs the
    default class, over the selected rows."""
    us = data["us"][mask_sel]
    sc = np.where(us > 0, scores, np.inf)
    pick = sc.argmin(axis=1)
    best = np.where(us > 0, us, np.inf).min(axis=1)
    t_pick = us[np.arange(len(us)), pick]
    t_def = us[:, DEFAULT_CLASS]
    regret = t_pick / best - 1
    return {
        "n": int(len(us)),
        "mean_regret": float(regret.mean()),
        "median_regret": float(np.median(regret)),
        "p90_regret": float(np.percentile(regret, 90)),
        "geomean_time_ratio_to_best": float(np.exp(np.log(t_pick / best).mean())),
        "total_time_ratio_to_best": float(t_pick.sum() / best.sum()),
        "geomean_speedup_vs_default": float(np.exp(np.log(t_def / t_pick).mean())),
        "total_speedup_vs_default": float(t_def.sum() / t_pick.sum()),
        "best_possible_geomean_speedup_vs_default": float(np.exp(np.log(t_def / best).mean())),
        "frac_exact_best": float((t_pick == best).mean()),
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("sweep")
    s.add_argument("--seeds", default="0:2000")
    s.add_argument("--reps", type=int, default=7)
    s.add_argument("--family", type=int, default=-1,
                   help="sweep only the seeds of this family (seed mod 4)")
    s.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--data", nargs="+", required=True)
    t.add_argument("--epochs", type=int, default=4000)
    t.add_argument("--lr", type=float, default=3e-3)
    t.add_argument("--out", required=True)
    e = sub.add_parser("eval")
    e.add_argument("--data", nargs="+", required=True)
    e.add_argument("--net", default=DEFAULT_NET)
    a = ap.parse_args(argv)
    if a.cmd == "sweep":
        lo, hi = (int(v) for v in a.seeds.split(":"))
        seeds = [s_ for s_ in range(lo, hi) if a.family < 0 or s_ % 4 == a.family]
        d = sweep(seeds, reps=a.reps, log=lambda m: print(m, flush=True))
        np.savez(a.out, **d)
        print(f"wrote {a.out}", flush=True)
    elif a.cmd == "train":
        data = load_data(a.data)
        net, mean, std, hist = train_net(data, epochs=a.epochs, lr=a.lr)
        tr, te = split(data)
        X = transform(data["raw"])
        sc = net.forward(X)
        meta = {"epochs": a.epochs, "lr": a.lr, "final_loss": float(hist[-1]),
                "train": evaluate(data, sc[tr], tr), "held_out": evaluate(data, sc[te], te)}
        save_net(a.out, net.params(), mean, std, meta)
        print(json.dumps(meta), flush=True)
    else:
        data = load_data(a.data)
        d = load_net(a.net)
        net = make_net(d)
        tr, te = split(data)
        sc = net.forward(transform(data["raw"]))
        print(json.dumps({"train": evaluate(data, sc[tr], tr),
                          "held_out": evaluate(data, sc[te], te)}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
