"""Shared test helpers: golden fixtures and field comparisons."""
import json
import os

import numpy as np

from code_tohoku_amd.restart import FIELD_NAMES, Input3D, empty_state

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_case(name):
    """Golden case written by tests/golden/make_golden.py: (params, nsplit,
    arrays) with arrays holding index 1..N+1 of each axis."""
    z = np.load(os.path.join(GOLDEN, f"ref_{name}.npz"))
    meta = json.loads(str(z["meta"]))
    p = Input3D(**meta["params"])
    return p, tuple(meta["nsplit"]), {k: z[k] for k in z.files if k != "meta"}


def case_cooling(name):
    """The Fcool a golden case was written with: "ki" (the Koyama-Inutsuka
    substitution) or "chem" (the reference's chemistry-network form)."""
    z = np.load(os.path.join(GOLDEN, f"ref_{name}.npz"))
    return json.loads(str(z["meta"])).get("cooling", "ki")


def expand(crop, p):
    """Index-1..N+1 block -> full state on -1..N+2 (other ghosts zero)."""
    nx, ny, nz = p.n
    st = empty_state(p.n)
    st[:, 2:nz + 3, 2:ny + 3, 2:nx + 3] = crop
    return st


def interior(st, p):
    nx, ny, nz = p.n
    return st[:, 2:nz + 2, 2:ny + 2, 2:nx + 2]


def faces(st, p):
    """Normal face-B values on the far faces (index N+1)."""
    nx, ny, nz = p.n
    return (st[5, 2:nz + 2, 2:ny + 2, nx + 2], st[6, 2:nz + 2, ny + 2, 2:nx + 2], st[7, nz + 2, 2:ny + 2, 2:nx + 2])


def rel_linf(a, b):
    """Per-field relative L-infinity error max|a-b| / max|b| over the interior."""
    out = {}
    for f in range(a.shape[0]):
        den = np.abs(b[f]).max()
        out[FIELD_NAMES[f]] = float(np.abs(a[f] - b[f]).max() / (den if den > 0 else 1.0))
    return out


def worst(errs):
    return max(errs.values())
