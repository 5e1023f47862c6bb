"""Known-answer and invariant checks of the oracle's point routines.

Each check is fixed by mathematics or by a data file of the reference, not
by restating the routine: the golden whole-step fixtures
(test_oracle_golden.py) pin the composition, these pin the pieces so that a
plausible slip (a dropped term, a swapped operand, a wrong sign or index)
fails a named test.
"""
import os

import numpy as np
import pytest
from scipy.optimize import brentq

import oracle

KB = 8.63359
DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "data")


# ---------------------------------------------------------------- RIEMAN
@pytest.mark.parametrize("rho,v,p,bt2,gam", [(1.0, 0.3, 2.0, 0.5, 5 / 3), (3.7, -1.2, 40.0, 9.0, 1.4),
                                              (0.25, 0.0, 16.0, 0.0, 5 / 3)])
def test_rieman_identical_states(rho, v, p, bt2, gam):
    """Identical left/right states: no wave, P* = p + B_t^2/2 and V* = v."""
    pa, va = oracle.rieman((rho, v, p, bt2), (rho, v, p, bt2), gam)
    assert pa == pytest.approx(p + 0.5 * bt2, rel=2e-16 * 8)
    assert va == pytest.approx(v, rel=1e-15, abs=1e-15)


def test_rieman_mirror_states_zero_velocity():
    """Mirror-symmetric collision: V* = 0 exactly, P* above both pressures."""
    pa, va = oracle.rieman((1.3, 5.0, 30.0, 4.0), (1.3, -5.0, 30.0, 4.0), 5 / 3)
    assert va == 0.0
    assert pa > 32.0


def test_rieman_two_shock_iteration():
    """For B_t = 0 the reference's mass flux qM = sqrt(rho (pp + sqrt(pp^2 -
    8 (P*-P)((g+1)P* + (g-1)P))))/2 with pp = (g+3)P* + (g-3)P reduces to the
    textbook two-shock flux W = sqrt(rho ((g+1) P* + (g-1) P) / 2), since the
    discriminant is ((g-1)P* + (g+1)P)^2.  For a mirror-symmetric collision
    the update is then P* <- P + W(P*) v; four of them, from P* = P, must give
    the oracle's P*, which sits within 1% of the converged Rankine-Hugoniot
    pressure."""
    rho, v, p, g = 1.0, 3.0, 10.0, 5 / 3
    pa, va = oracle.rieman((rho, v, p, 0.0), (rho, -v, p, 0.0), g)
    W = lambda ps: np.sqrt(rho * ((g + 1) * ps + (g - 1) * p) / 2)  # noqa: E731
    ps = p
    for _ in range(4):
        ps = p + W(ps) * v
    assert pa == pytest.approx(ps, rel=1e-13)
    conv = brentq(lambda x: x - p - W(x) * v, p, 10 * p)
    assert abs(pa - conv) / conv < 1e-2
    assert va == 0.0


# ---------------------------------------------------------------- CHAREQ
@pytest.mark.parametrize("seed", range(5))
def test_chareq_characteristic_relations(seed):
    """The Alfven-characteristic solution carries the invariants of both
    sides: sqrt(rhoL) V* - B* = sqrt(rhoL) vL - BL and
    sqrt(rhoR) V* + B* = sqrt(rhoR) vR + BR."""
    r = np.random.default_rng(seed)
    vl, bl, rl, vr, br, rr = r.normal(), r.normal(), r.uniform(0.2, 5), r.normal(), r.normal(), r.uniform(0.2, 5)
    bay, vay = oracle.chareq((vl, bl, rl), (vr, br, rr))
    assert np.sqrt(rl) * vay - bay == pytest.approx(np.sqrt(rl) * vl - bl, abs=1e-13)
    assert np.sqrt(rr) * vay + bay == pytest.approx(np.sqrt(rr) * vr + br, abs=1e-13)


def test_chareq_uniform():
    bay, vay = oracle.chareq((0.7, -1.1, 2.0), (0.7, -1.1, 2.0))
    assert vay == pytest.approx(0.7, rel=1e-15) and bay == pytest.approx(-1.1, rel=1e-15)


# ---------------------------------------------------------------- VLIMIT
def test_vlimit_linear_profile_both_branches():
    """Linear data on a uniform grid: the van Albada slope (T < 3) and the
    minmod slope (T >= 3) both return the exact gradient."""
    n = 8
    dx = np.full(n + 4, 0.25)
    x = (np.arange(n + 4) - 1.5) * 0.25
    q = 3.0 + 2.0 * x
    for T in (1.0, 3.0, 7.0):
        g = oracle.vlimit1d(q, np.full(n + 4, T), dx)
        np.testing.assert_allclose(g, 2.0, rtol=1e-9)


def test_vlimit_extremum_and_switch():
    n = 6
    dx = np.full(n + 4, 1.0)
    q = np.array([0, 1, 2, 3, 2, 1, 0, 1, 2, 3], dtype=float)
    for T in (1.0, 5.0):
        g = oracle.vlimit1d(q, np.full(n + 4, T), dx)
        assert g[2] == 0.0 and g[5] == 0.0      # extrema at positions 2 and 5 (q index 3, 6)
    # non-uniform slopes: minmod takes the smaller one-sided gradient
    q = np.array([0, 0, 1, 3, 3.5, 3.5, 3.5, 3.5, 3.5, 3.5])
    gm = oracle.vlimit1d(q, np.full(10, 9.0), dx)
    assert gm[1] == pytest.approx(1.0)          # position 1 (q index 2): deltas 1 and 2
    ga = oracle.vlimit1d(q, np.full(10, 1.0), dx)
    # van Albada: (2 dp dm + eps)/(dp^2+dm^2+eps) * (q+ - q-)/(2 dx) = 4/5 * 3/2
    assert ga[1] == pytest.approx(0.8 * 1.5, rel=1e-9)


# ------------------------------------------------------------ KI cooling
@pytest.mark.parametrize("fname", ["KIcool.dat", "INPUT_KIcool.dat"])
def test_fcool_ki_equilibrium_curve(fname):
    """The reference's thermal-equilibrium curve (KIcool/nibun.f90 bisection
    of n Lambda(T) = Gamma, T = p/n) is a zero of the KI net cooling."""
    d = np.loadtxt(os.path.join(DATA, fname), delimiter=None if fname != "KIcoolbr.dat" else ",")
    n, pkb = d[:, 0], d[:, 1]
    for ni, pi in zip(n, pkb):
        T3 = pi / ni / 1.0e3
        c = oracle.fcool_ki(1.27 * ni, T3)
        assert abs(c) / (ni * 39.4656) < 1e-6, (ni, pi, c)


def test_fcool_ki_limits():
    """Heating -N AbsL dominates in cold thin gas; the Lyman-alpha term is
    capped at 5e3 above ~14 600 K."""
    assert oracle.fcool_ki(1.27 * 1e-3, 0.01) == pytest.approx(-1e-3 * 39.4656, rel=1e-6)
    hot1 = oracle.fcool_ki(1.27, 20.0)
    hot2 = oracle.fcool_ki(1.27, 40.0)
    lamc = lambda T: 1.4e-2 * np.sqrt(T * 1e3) * np.exp(-92.0 / (T * 1e3))  # noqa: E731
    assert hot1 == pytest.approx((5e3 + lamc(20.0) - 1.0) * 39.4656, rel=1e-12)
    assert hot2 - hot1 == pytest.approx((lamc(40.0) - lamc(20.0)) * 39.4656, rel=1e-9)


# ------------------------------------------------------------------- IMC
@pytest.mark.parametrize("P,n,dt,rho", [(60.0, 1.0, 0.05, 1.27), (30.0, 20.0, 0.01, 25.4), (200.0, 0.5, 0.2, 0.6)])
def test_imc_solves_backward_euler(P, n, dt, rho):
    """IMC's bisection result lies within its stopping tolerance of the root
    of Pm + gammi1 Lambda(Pm/(kb n)) dt - P found by an independent solver."""
    g1 = 2.0 / 3.0
    f = lambda x: x + oracle.fcool_ki(rho, x / (KB * n)) * dt * g1 - P  # noqa: E731
    root = brentq(f, 1e-10, 1e10, xtol=1e-14, rtol=1e-15)
    pm = oracle.imc(P, n, dt, rho, g1)
    assert abs(pm - root) < 5e-5 * max(1.0, root)
    assert 1e-10 <= pm <= 1e10
