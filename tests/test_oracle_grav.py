"""Self-gravity restatement (oracle/mcf_grav.c) against the reference build.

The fixtures come from the reference MPI build run with ifgrv = 2 (GRAVTY of
3DMHD-CfCA/selfG_MPIfull8.f90) with its multigrid arrays zero-filled on
entry (oracle/patch_ref.py, option "grav"; the unedited code reads some of
them before writing them).  The inputs are regenerated here from the seed
the fixture records and checked against its SHA-256.

What the fixtures pin, measured by planting each mistake in a copy of the
restatement (CPU): double-precision real-FFT twiddles instead of CMPLX's
single precision move Phi by 4.6e-10 (8 ranks) and 3.5e-10 (16 ranks); the
plain kappa on the Nyquist column by 2.3e-8; the whole-domain levels wrapped
with period n-1 instead of n by 6.2e-4 and the residual kept inside its
level by 1.6e-3 (16 ranks: only the (4,2,2) split has whole-domain levels
below the coarsest split level).  The restatement itself agrees to 2.4e-15.
The sum over ranks of the boundary spectra is the MPI library's
(MPI_ALLREDUCE), so agreement is to rounding, not bit for bit.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from code_tohoku_amd import ics  # noqa: E402
from code_tohoku_amd.restart import Input3D  # noqa: E402
import oracle  # noqa: E402

GOLD = os.path.join(ROOT, "tests", "golden")
TOL = 1e-13


def load(name):
    z = np.load(os.path.join(GOLD, f"ref_{name}.npz"))
    meta = json.loads(str(z["meta"]))
    p = Input3D(**meta["params"])
    st = ics.random_state(p, seed=meta["seed"])
    sha = hashlib.sha256(np.ascontiguousarray(st, dtype="<f8").tobytes()).hexdigest()
    assert sha == meta["input_sha256"], "ics.random_state no longer gives the fixture's input"
    return z, meta, p, st


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("name", ["grav_32_8", "grav_32_16", "grav_32_8p"])
def test_oracle_gravity_matches_reference(name):
    z, meta, p, st = load(name)
    n = p.n[0]
    for k in meta["steps"]:
        o = oracle.Oracle(p, st, nsplit=tuple(meta["nsplit"]))
        dts, _ = o.run(k)
        np.testing.assert_allclose(dts, z[f"dts{k}"], rtol=1e-13, atol=0)
        phi = o.phi()[2:n + 2, 2:n + 2, 2:n + 2]
        e = rel(phi, z[f"phi{k}"])
        assert e <= TOL, (name, k, e)
        if f"state{k}" in z.files:
            so = o.state()[:, 2:n + 2, 2:n + 2, 2:n + 2]
            for t, f in enumerate(meta["state_fields"]):
                e = rel(so[f], z[f"state{k}"][t])
                assert e <= TOL, (name, k, f, e)


def test_gravity_split_matters():
    """The reference's potential depends on the block split (Gauss-Seidel
    per rank, boundary planes 1..3P per x block): the 16-rank fixture is not
    the 8-rank solution of the same input."""
    z, meta, p, st = load("grav_32_16")
    o = oracle.Oracle(p, st, nsplit=(2, 2, 2))
    o.run(1)
    n = p.n[0]
    assert rel(o.phi()[2:n + 2, 2:n + 2, 2:n + 2], z["phi1"]) > 1e-6


def _rlft3(data, speq, isign):
    import ctypes as C
    L = oracle.lib()
    pd = C.POINTER(C.c_double)
    f = L.mcfg_rlft3
    f.argtypes = [pd, pd, C.c_int, C.c_int, C.c_int]
    nn2, nn1 = data.shape
    f(data.ctypes.data_as(pd), speq.ctypes.data_as(pd), nn1, nn2, isign)


@pytest.mark.parametrize("nn1,nn2", [(16, 16), (32, 8), (8, 32)])
def test_rlft3_is_the_dft(nn1, nn2):
    """Forward rlft3 gives sum x exp(+2 pi i (p y/nn1 + q z/nn2)) (the
    conjugate of numpy's fft2 for real data) up to the single-precision
    unpacking twiddles; the inverse returns nn1*nn2/2 times the input."""
    rng = np.random.default_rng(1)
    x = rng.standard_normal((nn2, nn1))
    d = x.copy()
    speq = np.zeros(2 * nn2)
    _rlft3(d, speq, 1)
    F = np.conj(np.fft.fft2(x))  # [q, p]
    got = d.reshape(nn2, nn1 // 2, 2)
    got = got[..., 0] + 1j * got[..., 1]
    scale = np.abs(F).max()
    assert np.abs(got - F[:, :nn1 // 2]).max() <= 1e-6 * scale
    sp = speq[0::2] + 1j * speq[1::2]
    assert np.abs(sp - F[:, nn1 // 2]).max() <= 1e-6 * scale
    _rlft3(d, speq, -1)
    np.testing.assert_allclose(d / (nn1 * nn2 / 2), x, rtol=0, atol=1e-6 * np.abs(x).max())
