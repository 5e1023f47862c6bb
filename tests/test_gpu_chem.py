"""HIP chemistry path (SHIELD, chemistry Fcool, network) against the
reference fixtures and the oracle, through the C ABI.

Tolerances as in test_gpu_parity.py: FP64 relative L-infinity error per
field, 1e-12 for one step, 1e-11 over several; the device exp/log/pow differ
from glibc's by at most an ulp, and SHIELD's prefix sums run as wave scans
whose additions are grouped differently from the reference's sequential
sums (round-off only).
"""
import numpy as np
import pytest

import oracle
from code_tohoku_amd import ics, mcf
from code_tohoku_amd.restart import Input3D
from helpers import case_cooling, expand, interior, load_case, rel_linf, worst

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-12
MULTI_TOL = 1e-11
COOLING = {"ki": mcf.COOL_KI, "chem": mcf.COOL_CHEM}


@pytest.mark.parametrize("name,steps", [("chem_rad_24x8x8", 1), ("chem_rad_24x8x8", 6), ("chem_rad_48x8x8", 6),
                                        ("chem_norad_24x8x8", 6)])
def test_chem_reference_fixture(name, steps):
    p, nsplit, arr = load_case(name)
    print("case", name, steps, flush=True)
    g = mcf.Domain(p, nsplit=nsplit, cooling=COOLING[case_cooling(name)])
    g.upload(expand(arr["input"], p))
    dts = g.run(steps)
    assert len(dts) == steps
    np.testing.assert_allclose(dts, arr[f"dts{steps}"], rtol=1e-13, atol=0)
    errs = rel_linf(interior(g.download(), p), interior(expand(arr[f"out{steps}"], p), p))
    print(name, steps, "worst", worst(errs))
    assert worst(errs) <= (STEP_TOL if steps == 1 else MULTI_TOL), errs


def _col_errs(a, b, p):
    nx, ny, nz = p.n
    out = {}
    for q, name in enumerate(mcf.COLUMNS):
        for d in range(2):
            x = a[q, d, 2:nz + 2, 2:ny + 2, 2:nx + 2]
            y = b[q, d, 2:nz + 2, 2:ny + 2, 2:nx + 2]
            out[f"{name}{d + 1}"] = float(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300))
    return out


@pytest.mark.parametrize("nsplit", [(2, 2, 2), (4, 2, 2)])
def test_shield_columns_vs_oracle(nsplit):
    """The column densities after INITIA's 20 SHIELD sweeps and after two
    steps against the oracle's."""
    p, _, arr = load_case("chem_rad_48x8x8")
    st = expand(arr["input"], p)
    print("shield", nsplit, flush=True)
    o = oracle.Oracle(p, st, nsplit=nsplit, cooling="chem")
    g = mcf.Domain(p, nsplit=nsplit, cooling=mcf.COOL_CHEM)
    g.upload(st)
    errs = _col_errs(g.columns(), o.columns(), p)
    print("init columns worst", max(errs.values()))
    assert max(errs.values()) <= STEP_TOL, errs
    dts_o, _ = o.run(2)
    dts_g = g.run(2)
    np.testing.assert_allclose(dts_g, dts_o, rtol=1e-13, atol=0)
    errs = _col_errs(g.columns(), o.columns(), p)
    print("2-step columns worst", max(errs.values()))
    assert max(errs.values()) <= MULTI_TOL, errs


@pytest.mark.parametrize("ifchem,ifrad,cooling", [(2, 2, "ki"), (2, 0, "ki"), (0, 2, "chem"), (0, 0, "chem")])
def test_chem_combinations_vs_oracle(ifchem, ifrad, cooling):
    """Each switch on its own: the network with the KI Fcool, SHIELD feeding
    only the chemistry Fcool, and the chemistry Fcool with zero columns; 48^3
    converging flow (4 x blocks), 3 steps."""
    p = ics.converging_flow_input(48)
    base = Input3D()
    for k in ("Hini", "pini", "H2ini", "Heini", "Hepini", "Cini", "COini", "Cpini", "pinit_over_kb"):
        setattr(p, k, getattr(base, k))
    p.ifchem, p.ifrad = ifchem, ifrad
    st = ics.converging_flow_state(p)
    print("combo", ifchem, ifrad, cooling, flush=True)
    o = oracle.Oracle(p, st, nsplit=(4, 2, 2), cooling=cooling)
    dts_o, _ = o.run(3)
    g = mcf.Domain(p, nsplit=(4, 2, 2), cooling=COOLING[cooling])
    g.upload(st)
    dts_g = g.run(3)
    np.testing.assert_allclose(dts_g, dts_o, rtol=1e-13, atol=0)
    errs = rel_linf(interior(g.download(), p), interior(o.state(), p))
    print("combo worst", worst(errs))
    assert worst(errs) <= MULTI_TOL, errs


def test_chem_mixed_precision():
    """MCF_PREC_MIXED with the chemistry Fcool: the sweeps and CT in mixed
    precision, SHIELD, the cooling and the network in FP64; 6 steps of the
    chem_rad fixture on its dt sequence, against the reference at the
    mixed-precision bar used for the other configurations (1e-5)."""
    p, nsplit, arr = load_case("chem_rad_24x8x8")
    g = mcf.Domain(p, nsplit=nsplit, precision=mcf.PREC_MIXED, cooling=mcf.COOL_CHEM)
    g.upload(expand(arr["input"], p))
    for dt in arr["dts6"]:
        g.step(float(dt))
    errs = rel_linf(interior(g.download(), p), interior(expand(arr["out6"], p), p))
    print("chem mixed worst", worst(errs))
    assert worst(errs) <= 1e-5, errs


def test_columns_need_chemistry():
    """mcf_download_columns on a KI run without network or SHIELD: no
    columns are carried, and the call says so."""
    p = ics.converging_flow_input(16, nsplit=(1, 1, 1))
    g = mcf.Domain(p, nsplit=(1, 1, 1))
    with pytest.raises(mcf.MCFError, match="no column densities"):
        g.columns()


def _chem_256(ifthrm, ifchem, ifrad):
    p = ics.converging_flow_input(256)
    base = Input3D()
    for k in ("Hini", "pini", "H2ini", "Heini", "Hepini", "Cini", "COini", "Cpini", "pinit_over_kb"):
        setattr(p, k, getattr(base, k))
    p.ifthrm, p.ifchem, p.ifrad = ifthrm, ifchem, ifrad
    return p, ics.converging_flow_state(p)


def test_network_conserves_nuclei_256():
    """At the full 256^3 size: SOURCE with only the network (no thermal
    update, no SHIELD) conserves H (H + p + 2 H2), He (He + He+) and C
    (C + C+ + CO) nuclei in every cell to round-off, recomputes the density
    from the species and leaves velocities and pressure untouched."""
    p, st = _chem_256(0, 2, 0)
    print("network 256^3", flush=True)
    g = mcf.Domain(p, nsplit=(4, 4, 4), cooling=mcf.COOL_CHEM)
    g.upload(st)
    before = interior(g.download(), p)
    g.source(0.05)
    after = interior(g.download(), p)
    g.close()
    sb, sa = before[8:], after[8:]
    assert np.abs(sa - sb).max() > 0.0
    for name, idx, w in (("H", [0, 1, 2], [1, 1, 2]), ("He", [3, 4], [1, 1]), ("C", [5, 6, 7], [1, 1, 1])):
        nb = sum(wi * sb[i] for i, wi in zip(idx, w))
        na = sum(wi * sa[i] for i, wi in zip(idx, w))
        rel = float(np.abs(na - nb).max() / np.abs(nb).max())
        print(name, "nuclei rel change", rel)
        assert rel <= 1e-13, (name, rel)
    rho = sa[1] + sa[0] + 2.0 * sa[2] + 4.0 * sa[3] + 4.0 * sa[4]
    np.testing.assert_allclose(after[0], rho, rtol=1e-15, atol=0)
    np.testing.assert_array_equal(after[1:5], before[1:5])


def test_shield_columns_row_totals_256():
    """At the full 256^3 size with 4 x blocks: after INITIA's 20 SHIELD
    sweeps every column seen from the left plus the one seen from the right
    is the row total, for all five columns."""
    p, st = _chem_256(2, 2, 2)
    print("shield 256^3", flush=True)
    g = mcf.Domain(p, nsplit=(4, 4, 4), cooling=mcf.COOL_CHEM)
    g.upload(st)
    col = g.columns()
    g.close()
    n = 256
    inner = col[:, :, 2:n + 2, 2:n + 2, 2:n + 2]
    for q, name in enumerate(mcf.COLUMNS):
        tot = inner[q, 0] + inner[q, 1]
        dev = float(np.abs(tot - tot[:, :, :1]).max() / np.abs(inner[q]).max())
        print(name, "row-total deviation", dev)
        assert dev <= 1e-13, (name, dev)


def test_chem_backward_euler_cells_vs_oracle():
    """The chemistry Fcool in SOURCE's backward-Euler branch (IMC with the
    chemistry cooling, k_src_imc<CHEM>): the 24^3 converging flow at 1e3
    times the density and 3e3 times the pressure (T = 18000 K), where
    Stblty's dt floor (2.5e-4) exceeds 0.2 t_cool in every cell; two steps
    against the oracle at the multi-step bar (measured: 62530 backward-Euler
    cells, worst field 1.8e-13).  The bisection's midpoints come from the
    device's pow/log10, which may differ from glibc's in the last bit; the
    branch decisions have not differed."""
    p = ics.converging_flow_input(24)
    base = Input3D()
    for k in ("Hini", "pini", "H2ini", "Heini", "Hepini", "Cini", "COini", "Cpini", "pinit_over_kb"):
        setattr(p, k, getattr(base, k))
    p.ifchem, p.ifrad = 2, 2
    st = ics.converging_flow_state(p)
    st[0] *= 1e3
    st[4] *= 3e3
    st[8:16] *= 1e3
    print("chem stiff", flush=True)
    o = oracle.Oracle(p, st, nsplit=(2, 2, 2), cooling="chem")
    dts_o, _ = o.run(2)
    g = mcf.Domain(p, nsplit=(2, 2, 2), cooling=mcf.COOL_CHEM)
    g.upload(st)
    g.cool_stats(reset=True)
    dts_g = g.run(2)
    nst, _ = g.cool_stats()
    print("backward-Euler cells", nst)
    assert nst > 10000
    print("dts", dts_g, dts_o)
    errs = rel_linf(interior(g.download(), p), interior(o.state(), p))
    print("chem stiff errors", errs)
    np.testing.assert_allclose(dts_g, dts_o, rtol=1e-12, atol=0)
    assert worst(errs) <= MULTI_TOL, errs
