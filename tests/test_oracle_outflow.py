"""The outflow (free) x boundary, BC = 1, in the oracle.

The 3D reference has only periodic (3) and inflow (4) x boundaries; BC = 1 is
the zero-gradient rule of the reference's 1D code
(1DMHDKIcool/1DMHDwCC3KIcool.f90:352-372) applied to every ghost array, with
the face field keeping the divergence-free copy of the inflow rule.  No
reference output exists for it in 3D, so it is pinned by what the
mathematics fixes:

* a uniform state, at rest or moving through both x ends, stays uniform;
* a mirror-symmetric blast stays mirror-symmetric about the midplane (a
  ghost copied from the wrong cell on one side breaks this);
* a planar state stays planar and does not depend on the transverse size
  (64x8x8 and 64x16x16 give the same rows bit for bit).
"""
import numpy as np
import pytest

import oracle
from code_tohoku_amd import ics
from code_tohoku_amd.restart import Input3D, empty_state

KB = ics.KB


def box(nx=32, ny=8, nz=8, ql=0.5):
    h = ql / (nx // 2)
    p = Input3D(np1x=nx // 2, np2x=nx // 2, np1y=ny // 2, np2y=ny // 2, np1z=nz // 2, np2z=nz // 2,
                ql1x=ql, ql2x=ql, ql1y=h * (ny // 2), ql2y=h * (ny // 2), ql1z=h * (nz // 2), ql2z=h * (nz // 2),
                vinitx1=0.0, vinitx2=0.0, binitx1=0.0, binitx2=0.0, binity1=0.0, binity2=0.0,
                cfl=0.4, ifthrm=0)
    p.bcx1 = p.bcx2 = 1
    return p


def uniform_state(p, rho=2.0, pres=200.0, v=(0.0, 0.0, 0.0), b=(10.0, 5.0, 0.0)):
    st = empty_state(p.n)
    st[0] = rho
    st[1], st[2], st[3] = v
    st[4] = pres
    st[5], st[6], st[7] = b
    st[8] = rho
    floors = (0.0, 1e-20, 1e-20, 0.1949628 * 0.09, 1e-20, 1e-20, 1e-20, 1e-20)
    for s in range(1, 8):
        st[8 + s] = floors[s]
    return st


def interior(st, p):
    nx, ny, nz = p.n
    return st[:, 2:nz + 2, 2:ny + 2, 2:nx + 2]


@pytest.mark.parametrize("v", [(0.0, 0.0, 0.0), (3.0, 1.0, -2.0), (-4.0, 0.5, 0.0)])
def test_uniform_state_stays_uniform(v):
    p = box()
    o = oracle.Oracle(p, uniform_state(p, v=v), nsplit=(2, 2, 2))
    dts, _ = o.run(10)
    assert len(dts) == 10
    out = interior(o.state(), p)
    ref = interior(uniform_state(p, v=v), p)
    for f in range(8):
        scale = max(np.abs(ref[f]).max(), 1.0)
        assert np.abs(out[f] - ref[f]).max() <= 1e-12 * scale, f


def test_blast_stays_mirror_symmetric():
    """rho, p, B_x uniform except a pressure bump at the midplane; v = 0:
    rho, p, B_x, B_y even and v_x odd about x = 1/2 after 60 steps, with
    the waves through both outflow ends."""
    p = box(nx=48)
    st = uniform_state(p, b=(10.0, 0.0, 0.0))
    (xe, dxe), _, _ = ics.grid(p)
    xc = xe - 0.5 * dxe
    st[4] = (200.0 * (1.0 + 20.0 * np.exp(-((xc - 0.5) / 0.06) ** 2)))[None, None, :]
    o = oracle.Oracle(p, st, nsplit=(2, 2, 2))
    o.run(60)
    out = interior(o.state(), p)
    assert np.abs(out[1]).max() > 1.0, "the blast did not move the gas"
    for f, sign in ((0, 1), (4, 1), (5, 1), (1, -1)):
        a = out[f]
        b = sign * a[:, :, ::-1]
        assert np.abs(a - b).max() <= 1e-11 * np.abs(a).max(), f


def test_planar_state_independent_of_transverse_size():
    """Brio-Wu along x with outflow ends, 80 steps: every row of a run is the
    same and the rows of the 64x8x8 and 64x16x16 runs agree.  Both to
    round-off, not bit for bit: the y and z sweeps form Lagrangian widths
    from absolute cell positions, whose rounding differs from row to row."""
    rows = []
    for ny in (8, 16):
        p = ics.brio_wu_input(64, ny=ny, periodic_x=False)
        p.bcx1 = p.bcx2 = 1
        o = oracle.Oracle(p, ics.brio_wu_state(p), nsplit=(2, 2, 2))
        dts, _ = o.run(80)
        out = interior(o.state(), p)
        for f in range(8):
            scale = max(np.abs(out[f]).max(), 1e-300)
            assert np.abs(out[f] - out[f][:1, :1, :]).max() <= 1e-12 * scale, (ny, f)
        rows.append((dts, out[:, 0, 0, :]))
    np.testing.assert_allclose(rows[0][0], rows[1][0], rtol=1e-12)
    for f in range(8):
        scale = max(np.abs(rows[1][1][f]).max(), 1e-300)
        assert np.abs(rows[0][1][f] - rows[1][1][f]).max() <= 1e-11 * scale, f
    # the shocks reached the outflow ends: the edge cells moved
    assert abs(rows[0][1][0][0] - 2.0) > 1e-3 or abs(rows[0][1][0][-1] - 0.25) > 1e-3
