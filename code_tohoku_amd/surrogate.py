"""The KI-cooling surrogate's weights: training on the GPU and the saved copy.

The network (3 -> 64 -> 64 -> 1, SiLU; code_tohoku_amd/csrc/mcf_mlp.h) is
trained by the library's own kernels (``mcf_mlp_train``: IMC samples drawn on
the device, MFMA forward pass, FP32 backward pass, Adam).  The trained
parameters are stored as MLP_NPARAM little-endian float32 values in
``data/ki_cool_mlp.f32`` with a JSON report beside them.

    python -m code_tohoku_amd.surrogate [--samples 10000000] [--epochs 24]

retrains on the current GPU and rewrites both files.
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from . import ics, mcf

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "data", "ki_cool_mlp.f32")
REPORT = os.path.join(HERE, "data", "ki_cool_mlp.json")
KB = 8.63359


def load_weights(path: str = WEIGHTS) -> np.ndarray:
    w = np.fromfile(path, dtype="<f4")
    if w.shape != (mcf.MLP_NPARAM,) or not np.all(np.isfinite(w)):
        raise ValueError(f"{path}: not a vector of {mcf.MLP_NPARAM} finite float32 values")
    return w


def stiff_batch(n: int, seed: int):
    """Backward-Euler solves of the stiff branch (dt/t_cool in [0.2, 1e3]) on
    log-uniform n in [0.1, 1e4] cm^-3 and T in [10, 1e4] K."""
    rng = np.random.default_rng(seed)
    N = 10 ** rng.uniform(-1, 4, n)
    T = 10 ** rng.uniform(1, 4, n) * 1e-3
    r = 10 ** rng.uniform(np.log10(0.2), 3, n)
    rho, gi = 1.27 * N, np.full(n, 2.0 / 3.0)
    P0 = KB * N * T
    TK = T * 1e3
    laml = np.minimum(1e7 * np.exp(-1.184e5 / (TK + 1e3)), 5e3)
    lamc = 1.4e-2 * np.sqrt(TK) * np.exp(-92.0 / TK)
    L = (N * N * (laml + lamc) - N) * 39.4656
    dt = r * P0 / (gi * np.abs(L))
    return P0, N.copy(), dt, rho, gi


def evaluate(dom: "mcf.Domain", n: int = 1 << 20, seed: int = 12345) -> dict:
    """Fallback fraction, agreement with IMC and kernel times on stiff cells."""
    args = stiff_batch(n, seed)
    dom.cool_solve(*args, 0)
    Pi, _, ms_imc = dom.cool_solve(*args, 0)
    dom.cool_solve(*args, 1)
    Ps, fl, ms_sur = dom.cool_solve(*args, 1)
    acc = fl == 0
    rel = np.abs(Ps[acc] - Pi[acc]) / Pi[acc]
    return {"cells": n, "fallback_fraction": float(1 - acc.mean()), "imc_ms": ms_imc, "surrogate_ms": ms_sur,
            "speedup": ms_imc / ms_sur, "rel_diff_vs_imc_median": float(np.median(rel)),
            "rel_diff_vs_imc_p99": float(np.quantile(rel, 0.99))}


def train(samples: int = 10_000_000, epochs: int = 24, batch: int = 8192, lr: float = 3e-3,
          lr_min: float = 1e-5, seed: int = 1) -> tuple[np.ndarray, dict]:
    p = ics.converging_flow_input(16)
    dom = mcf.Domain(p, nsplit=(1, 1, 1))
    try:
        t0 = time.perf_counter()
        el, vl = dom.mlp_train(nsamples=samples, nval=1 << 18, seed=seed, batch=batch, epochs=epochs, lr=lr,
                               lr_min=lr_min)
        dt = time.perf_counter() - t0
        w = dom.mlp_get_weights()
        rep = {"samples": samples, "epochs": epochs, "batch": batch, "lr": lr, "lr_min": lr_min, "seed": seed,
               "train_seconds": dt, "train_loss": [float(v) for v in el], "heldout_loss": [float(v) for v in vl],
               "stiff_eval": evaluate(dom)}
        return w, rep
    finally:
        dom.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=10_000_000)
    ap.add_argument("--epochs", type=int, default=24)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=WEIGHTS)
    a = ap.parse_args()
    w, rep = train(a.samples, a.epochs, a.batch, seed=a.seed)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    w.astype("<f4").tofile(a.out)
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(rep, f, indent=1)
    print(json.dumps(rep))


if __name__ == "__main__":
    main()
