"""The chemistry parts of the oracle: SHIELD column densities, the chemistry
Fcool and the chemistry network (M3:2576-2774, 2825-2895, 2924-3141).

The fixtures chem_rad_24x8x8, chem_rad_48x8x8 and chem_norad_24x8x8 (the
reference build that keeps its own chemistry Fcool) pin these bit for bit
(test_oracle_golden.py).  The tests here show that each part is exercised by
them and pin what the mathematics fixes:

* every column seen from the left plus the one seen from the right is the
  row total, for all five columns and any block split (a block total added
  to the wrong side, or a block left out, breaks it);
* on a uniform row the column from the left at cell i is (i - 1/2) dx n;
* the columns do not depend on the block split beyond round-off;
* the network conserves H, He and C nuclei (its reactions only move nuclei
  between species), so a wrong sign or species in any reaction breaks it.
"""
import numpy as np
import pytest

import oracle
from code_tohoku_amd import ics
from code_tohoku_amd.restart import Input3D
from helpers import expand, interior, load_case, rel_linf, worst


def chem_box(nx=32, ny=8, nz=8, periodic=False):
    p = Input3D(np1x=nx // 2, np2x=nx // 2, np1y=ny // 2, np2y=ny // 2, np1z=nz // 2, np2z=nz // 2,
                ql1x=10.0, ql2x=10.0, ql1y=2.5, ql2y=2.5, ql1z=2.5, ql2z=2.5)
    if periodic:
        p.bcx1 = p.bcx2 = 3
    p.ifchem, p.ifrad = 2, 2
    return p


@pytest.mark.parametrize("change", ["ki_fcool", "no_shield", "no_network"])
def test_chem_fixture_exercises_each_part(change):
    """Running the chem_rad fixture with the KI Fcool, without SHIELD or
    without the network must not reproduce it."""
    p, nsplit, arr = load_case("chem_rad_24x8x8")
    cooling = "ki" if change == "ki_fcool" else "chem"
    if change == "no_shield":
        p.ifrad = 0
    if change == "no_network":
        p.ifchem = 0
    o = oracle.Oracle(p, expand(arr["input"], p), nsplit=nsplit, cooling=cooling)
    o.run(6)
    errs = rel_linf(interior(o.state(), p), interior(expand(arr["out6"], p), p))
    assert worst(errs) > 1e-8, errs


@pytest.mark.parametrize("nsplit", [(1, 2, 2), (2, 2, 2), (4, 2, 2)])
def test_columns_left_plus_right_is_row_total(nsplit):
    p = chem_box()
    st = ics.random_state(p, seed=21)
    o = oracle.Oracle(p, st, nsplit=nsplit, cooling="chem")
    o.run(2)
    col = o.columns()
    nx, ny, nz = p.n
    inner = col[:, :, 2:nz + 2, 2:ny + 2, 2:nx + 2]
    tot = inner[:, 0] + inner[:, 1]
    for q, name in enumerate(oracle.COLUMNS):
        t = tot[q]
        scale = np.abs(inner[q]).max()
        assert scale > 0, name
        dev = np.abs(t - t[:, :, :1]).max()
        assert dev <= 1e-13 * scale, (name, dev, scale)
        # and the total is the row sum of the per-cell columns (2 * dxh * v)
        assert np.all(inner[q, 0, :, :, 0] < inner[q, 0, :, :, -1]) or name == "tCII"


def test_uniform_row_columns_closed_form():
    """Uniform gas: Ntot from the left at cell i is (i - 1/2) dx ndtot and
    from the right (N - i + 1/2) dx ndtot, for every block split."""
    for nsplit in ((1, 1, 1), (2, 2, 2), (4, 2, 2)):
        p = chem_box(periodic=True)
        st = ics.random_state(p, seed=3, vamp=0.0, bamp=0.0, pt_range=(5.0, 5.0))
        nx, ny, nz = p.n
        # uniform composition and density everywhere
        for f in [0, 4] + list(range(8, 16)):
            st[f] = st[f][2, 2, 2]
        o = oracle.Oracle(p, st, nsplit=nsplit, cooling="chem")
        col = o.columns()
        (xe, dxe), _, _ = ics.grid(p)
        dx = dxe[2]
        s = st[8:16, 2, 2, 2]
        ndtot = s[1] + s[0] + 2.0 * s[2] + s[3] + s[4]
        i = np.arange(1, nx + 1)
        left = (i - 0.5) * dx * ndtot
        right = (nx - i + 0.5) * dx * ndtot
        got = col[0, :, 2:nz + 2, 2:ny + 2, 2:nx + 2]
        np.testing.assert_allclose(got[0], np.broadcast_to(left, got[0].shape), rtol=1e-13, atol=0)
        np.testing.assert_allclose(got[1], np.broadcast_to(right, got[1].shape), rtol=1e-13, atol=0)
        nh2 = col[1, 0, 2:nz + 2, 2:ny + 2, 2:nx + 2]
        np.testing.assert_allclose(nh2, np.broadcast_to((i - 0.5) * dx * s[2], nh2.shape), rtol=1e-13, atol=0)


def test_columns_independent_of_block_split():
    p = chem_box()
    st = ics.random_state(p, seed=23)
    cols = []
    for nsplit in ((1, 2, 2), (4, 2, 2)):
        o = oracle.Oracle(p, st, nsplit=nsplit, cooling="chem")
        cols.append(o.columns())
    nx, ny, nz = p.n
    a = cols[0][:, :, 2:nz + 2, 2:ny + 2, 2:nx + 2]
    b = cols[1][:, :, 2:nz + 2, 2:ny + 2, 2:nx + 2]
    for q in range(5):
        scale = np.abs(a[q]).max()
        assert np.abs(a[q] - b[q]).max() <= 1e-12 * scale, oracle.COLUMNS[q]


def test_network_conserves_nuclei():
    """SOURCE with only the network (ifthrm = 0, ifrad = 0): per cell, H
    (H + p + 2 H2), He (He + He+) and C (C + C+ + CO) nuclei are conserved
    to round-off and the density is recomputed from the species."""
    p = chem_box()
    p.ifthrm, p.ifrad = 0, 0
    st = ics.random_state(p, seed=29)
    nx, ny, nz = p.n
    # make every species well above its floor so no clamp applies
    sl = (slice(None), slice(2, nz + 2), slice(2, ny + 2), slice(2, nx + 2))
    st[8 + 5] = np.maximum(st[8 + 5], 1e-6)
    st[8 + 6] = np.maximum(st[8 + 6], 1e-6)
    st[8 + 4] = np.maximum(st[8 + 4], 1e-4)
    o = oracle.Oracle(p, st, nsplit=(2, 2, 2), cooling="chem")
    before = o.state()[sl[0], sl[1], sl[2], sl[3]]
    o.source(0.05)
    after = o.state()[sl[0], sl[1], sl[2], sl[3]]
    sb, sa = before[8:], after[8:]
    assert np.abs(sa - sb).max() > 1e-6, "the network changed nothing"
    for name, idx, w in (("H", [0, 1, 2], [1, 1, 2]), ("He", [3, 4], [1, 1]), ("C", [5, 6, 7], [1, 1, 1])):
        nb = sum(wi * sb[i] for i, wi in zip(idx, w))
        na = sum(wi * sa[i] for i, wi in zip(idx, w))
        rel = np.abs(na - nb).max() / np.abs(nb).max()
        assert rel <= 1e-13, (name, rel)
    rho = sa[1] + sa[0] + 2.0 * sa[2] + 4.0 * sa[3] + 4.0 * sa[4]
    np.testing.assert_allclose(after[0], rho, rtol=1e-15, atol=0)
    # pressure and velocities untouched
    np.testing.assert_array_equal(after[1:5], before[1:5])
