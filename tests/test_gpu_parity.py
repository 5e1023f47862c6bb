"""HIP path (through the C ABI) against the oracle and the reference fixtures.

Tolerances: FP64 relative L-infinity error per field, max|a-b|/max|b|.
The GPU evaluates exp/log10/pow with the device math library while the
reference and the oracle use glibc; the two differ by at most an ulp, so
any step that involves the KI cooling (SOURCE, the Stblty limit of dt) may
differ at round-off level.  Everything else (divisions and square roots are
IEEE correctly rounded on both sides, no fused multiply-add) is expected to
match bit for bit.  The bar is 1e-12 per step; multi-step runs use 1e-11.
"""
import numpy as np
import pytest

import oracle
from code_tohoku_amd import ics, mcf
from code_tohoku_amd.restart import Input3D
from helpers import expand, faces, interior, load_case, rel_linf, worst

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-12
MULTI_TOL = 1e-11


def small(nx=32, ny=16, nz=16, periodic=False):
    p = Input3D(np1x=nx // 2, np2x=nx // 2, np1y=ny // 2, np2y=ny // 2, np1z=nz // 2, np2z=nz // 2,
                ql1x=10.0, ql2x=10.0, ql1y=5.0, ql2y=5.0, ql1z=5.0, ql2z=5.0)
    if periodic:
        p.bcx1 = p.bcx2 = 3
    return p


@pytest.mark.parametrize("name,steps", [("inflow_24x8x8", 1), ("periodic_24x8x8", 1), ("inflow_24x8x8", 6),
                                        ("periodic_24x8x8", 6), ("inflow_48x8x8", 6), ("converging_24", 8)])
def test_reference_fixture(name, steps):
    p, nsplit, arr = load_case(name)
    g = mcf.Domain(p, nsplit=nsplit)
    g.upload(expand(arr["input"], p))
    dts = g.run(steps)
    assert len(dts) == steps
    np.testing.assert_allclose(dts, arr[f"dts{steps}"], rtol=1e-13, atol=0)
    got = g.download()
    ref = expand(arr[f"out{steps}"], p)
    errs = rel_linf(interior(got, p), interior(ref, p))
    tol = STEP_TOL if steps == 1 else MULTI_TOL
    print(name, steps, "worst", worst(errs))
    assert worst(errs) <= tol, errs
    for a, b in zip(faces(got, p), faces(ref, p)):
        assert np.abs(a - b).max() <= tol * np.abs(b).max()


@pytest.mark.parametrize("precision", [mcf.PREC_FP64, mcf.PREC_MIXED])
def test_outflow_brio_wu_vs_oracle(precision):
    """Outflow x ends (BC = 1): the Brio-Wu tube at 64x8x8 for 60 steps,
    with the fast waves through both ends, against the oracle (FP64 dt and
    fields at 1e-12; mixed precision on the oracle's dt sequence at 1e-5)."""
    p = ics.brio_wu_input(64, ny=8, periodic_x=False)
    p.bcx1 = p.bcx2 = 1
    st = ics.brio_wu_state(p)
    o = oracle.Oracle(p, st, nsplit=(2, 2, 2))
    dts_o, _ = o.run(60)
    g = mcf.Domain(p, nsplit=(2, 2, 2), precision=precision)
    g.upload(st)
    if precision == mcf.PREC_FP64:
        dts_g = g.run(60)
        np.testing.assert_allclose(dts_g, dts_o, rtol=1e-13, atol=0)
        tol = STEP_TOL
    else:
        for dt in dts_o:
            g.step(dt)
        tol = 1e-5
    errs = rel_linf(interior(g.download(), p)[:8], interior(o.state(), p)[:8])
    print("outflow Brio-Wu worst", worst(errs))
    assert worst(errs) <= tol, errs


def test_outflow_config1_64cubed_planar():
    """Config 1 (64^3 Brio-Wu, outflow x, periodic y/z, 200 steps) on the
    GPU: the dt sequence and every row agree with the 64x8x8 run, which the
    oracle pins (test_outflow_brio_wu_vs_oracle)."""
    out = {}
    for ny in (8, 64):
        p = ics.brio_wu_input(64, ny=ny, periodic_x=False)
        p.bcx1 = p.bcx2 = 1
        g = mcf.Domain(p, nsplit=(2, 2, 2))
        g.upload(ics.brio_wu_state(p))
        dts = g.run(200, tsave=1e30)
        assert len(dts) == 200
        out[ny] = (dts, interior(g.download(), p)[:8])
        g.close()
    np.testing.assert_allclose(out[64][0], out[8][0], rtol=1e-12)
    big, small = out[64][1], out[8][1]
    for f in range(8):
        scale = max(np.abs(small[f]).max(), 1e-300)
        assert np.abs(big[f] - small[f][:1, :1, :]).max() <= 1e-10 * scale, f


def test_dtf_inflow_fixture():
    """Inflow density read from the DTF field shifted by -v t (the reference
    build with the DTF read enabled): 8 steps against the reference."""
    p, nsplit, arr = load_case("converging_dtf_24")
    g = mcf.Domain(p, nsplit=nsplit)
    g.upload(expand(arr["input"], p))
    g.set_inflow_density(arr["dtf"])
    dts = g.run(8)
    np.testing.assert_allclose(dts, arr["dts8"], rtol=1e-13, atol=0)
    errs = rel_linf(interior(g.download(), p), interior(expand(arr["out8"], p), p))
    print("dtf fixture worst", worst(errs))
    assert worst(errs) <= MULTI_TOL, errs


def test_dtf_inflow_wrapped_shift_vs_oracle():
    """The DTF shift at late times: started at t = 7.3 the inflowing gas has
    moved 76 pc, so the ghost density wraps the 20 pc field several times;
    two steps at 32^3 (block split 2) against the oracle."""
    p = ics.converging_flow_input(32)
    st = ics.converging_flow_state(p)
    dtf = ics.inflow_density_field(st, p)
    o = oracle.Oracle(p, st, nsplit=(2, 2, 2), time=7.3)
    o.set_dtf(dtf)
    g = mcf.Domain(p, nsplit=(2, 2, 2))
    g.upload(st)
    g.set_time(7.3)
    g.set_inflow_density(dtf)
    dts_o, _ = o.run(2)
    dts_g = g.run(2, tsave=7.5)
    np.testing.assert_allclose(dts_g, dts_o, rtol=1e-13)
    errs = rel_linf(interior(g.download(), p), interior(o.state(), p))
    assert worst(errs) <= MULTI_TOL, errs


def test_brio_wu_200_steps_bitexact():
    """Config 1 geometry, 200 steps without cooling: the dt is CFL-bound and
    no transcendental enters the state, so the GPU must reproduce the
    reference bit for bit."""
    p, nsplit, arr = load_case("briowu_64x8x8")
    g = mcf.Domain(p, nsplit=nsplit)
    g.upload(expand(arr["input"], p))
    dts = g.run(200)
    np.testing.assert_array_equal(dts, arr["dts200"])
    errs = rel_linf(interior(g.download(), p), interior(expand(arr["out200"], p), p))
    assert worst(errs) == 0.0, errs


@pytest.mark.parametrize("periodic", [False, True])
def test_stages_vs_oracle(periodic):
    """Each entry point of the ABI against the same routine of the oracle,
    from identical inputs (one stage at a time, re-synchronised)."""
    p = small(periodic=periodic)
    st = ics.random_state(p, seed=21 if periodic else 22)
    o = oracle.Oracle(p, st, nsplit=(2, 2, 2))
    g = mcf.Domain(p, nsplit=(2, 2, 2))
    g.upload(st)
    dt, tag = o.timestep(0.25)
    gdt, gtag = g.timestep(0.25)
    assert gtag == tag
    assert abs(gdt - dt) <= 1e-14 * dt
    stages = [("source", lambda d: d.source(0.5 * dt))] + \
             [(f"sweep{a}", (lambda a_: (lambda d: d.sweep(a_, dt)))(a)) for a in (0, 1, 2)] + \
             [("ct", lambda d: d.ct(dt)), ("source2", lambda d: d.source(0.5 * dt)), ("dissip", lambda d: d.dissip())]
    for name, fn in stages:
        fn(o)
        fn(g)
        ref = o.state()
        got = g.download()
        errs = rel_linf(interior(got, p), interior(ref, p))
        assert worst(errs) <= STEP_TOL, (name, errs)
        gb, ob = g.bcc(), o.bcc()
        nx, ny, nz = p.n
        sl = (slice(None), slice(2, nz + 2), slice(2, ny + 2), slice(2, nx + 2))
        assert np.abs(gb[sl] - ob[sl]).max() <= STEP_TOL * np.abs(ob[sl]).max(), name
        # resynchronise so that every stage starts from identical inputs
        g.upload(ref)


@pytest.mark.parametrize("periodic", [False, True])
def test_source_ct_ghosts_vs_oracle(periodic):
    """SOURCE and CT fill their ghosts the way the reference's BC_MPI /
    BC_MPI_OT passes do (x rules or periodic x, then y, then z), so after each
    stage the whole arrays, ghost layers and corners included, equal the
    oracle's: a direct check of the fill kernels (x rules, the joint y/z
    pass, the corrector's ghost-layer copy)."""
    p = small(24, 16, 16, periodic=periodic)
    st = ics.random_state(p, seed=23 if periodic else 24)
    o = oracle.Oracle(p, st, nsplit=(2, 2, 2))
    g = mcf.Domain(p, nsplit=(2, 2, 2))
    g.upload(st)
    dt, _ = o.timestep(0.25)
    g.timestep(0.25)
    for name, fn in (("source", lambda d: d.source(0.5 * dt)), ("ct", lambda d: d.ct(dt))):
        fn(o)
        fn(g)
        ref = o.state()
        got = g.download()
        for f in range(ref.shape[0]):
            scale = max(np.abs(ref[f]).max(), 1e-300)
            assert np.abs(got[f] - ref[f]).max() <= STEP_TOL * scale, (name, f)
        g.upload(ref)


@pytest.mark.parametrize("periodic", [False, True])
def test_source_many_backward_euler_cells_vs_oracle(periodic):
    """SOURCE on a ragged grid (72 x 44 x 136: partial x tiles and row
    groups, three z segments) at 1e4 times the density and pressure, where
    most cells take the backward-Euler branch, so every block of the
    explicit kernel queues cells through its own range and one atomic: the
    whole arrays after the call against the oracle."""
    p = Input3D(np1x=36, np2x=36, np1y=22, np2y=22, np1z=68, np2z=68,
                ql1x=10.0, ql2x=10.0, ql1y=5.0, ql2y=5.0, ql1z=5.0, ql2z=5.0)
    if periodic:
        p.bcx1 = p.bcx2 = 3
    st = ics.random_state(p, seed=51)
    for f in (0, 4) + tuple(range(8, 16)):
        st[f] *= 1.0e4
    o = oracle.Oracle(p, st, nsplit=(2, 2, 2))
    g = mcf.Domain(p, nsplit=(2, 2, 2))
    g.upload(st)
    print("source backward-Euler", periodic, flush=True)
    dt, _ = o.timestep(0.25)
    g.timestep(0.25)
    g.cool_stats(reset=True)
    o.source(0.5 * dt)
    g.source(0.5 * dt)
    nst, _ = g.cool_stats()
    ncell = 72 * 44 * 136
    assert nst > ncell, (nst, ncell)  # over half the cells in each of the two stages
    ref = o.state()
    got = g.download()
    for f in range(ref.shape[0]):
        scale = max(np.abs(ref[f]).max(), 1e-300)
        assert np.abs(got[f] - ref[f]).max() <= STEP_TOL * scale, f


def test_multistep_vs_oracle_64():
    p = small(64, 32, 32)
    st = ics.random_state(p, seed=5)
    o = oracle.Oracle(p, st, nsplit=(2, 2, 2))
    g = mcf.Domain(p, nsplit=(2, 2, 2))
    g.upload(st)
    dts_o, _ = o.run(3)
    dts_g = g.run(3)
    np.testing.assert_allclose(dts_g, dts_o, rtol=1e-13)
    errs = rel_linf(interior(g.download(), p), interior(o.state(), p))
    assert worst(errs) <= MULTI_TOL, errs


@pytest.mark.parametrize("precision", [mcf.PREC_FP64, mcf.PREC_MIXED])
def test_multistep_vs_oracle_long_columns(precision):
    """y longer than one column segment of the y/z sweep kernels (256 cells)
    and not a multiple of it, so the march runs in two segments whose halo
    cells each evaluate in their own block's view; z shorter than one
    segment: the oracle's result (FP64) and agreement to FP32 round-off
    (mixed)."""
    p = small(24, 264, 40)
    st = ics.random_state(p, seed=6)
    o = oracle.Oracle(p, st, nsplit=(2, 2, 2))
    g = mcf.Domain(p, nsplit=(2, 2, 2), precision=precision)
    g.upload(st)
    dts_o, _ = o.run(2)
    if precision == mcf.PREC_FP64:
        dts_g = g.run(2)
        np.testing.assert_allclose(dts_g, dts_o, rtol=1e-13)
        errs = rel_linf(interior(g.download(), p), interior(o.state(), p))
        assert worst(errs) <= MULTI_TOL, errs
    else:
        for dt in dts_o:
            g.step(dt)
        got = interior(g.download(), p)
        ref = interior(o.state(), p)
        for f in range(8):
            assert np.abs(got[f] - ref[f]).max() <= 1e-5 * np.abs(ref[f]).max(), f


def test_one_wave_columns_vs_oracle():
    """FP64 on a grid whose y/z launches reach the device's CU count (here
    256 x 264 x 128: 2 x 4 x 32 = 256 blocks for the y sweep), so the y and z
    sweeps run as the one-wave column march, the y axis in two segments:
    one step against the oracle."""
    p = small(256, 264, 128, periodic=True)
    st = ics.random_state(p, seed=9)
    o = oracle.Oracle(p, st, nsplit=(4, 2, 2))
    g = mcf.Domain(p, nsplit=(4, 2, 2))
    g.upload(st)
    print("one-wave columns 256x264x128", flush=True)
    dts_o, _ = o.run(1)
    dts_g = g.run(1)
    np.testing.assert_allclose(dts_g, dts_o, rtol=1e-13)
    errs = rel_linf(interior(g.download(), p), interior(o.state(), p))
    assert worst(errs) <= STEP_TOL, errs


def test_axis_longer_than_sweep_col_uses_split_column_kernels():
    """y longer than the fused column sweep takes (2048 cells): the y sweep
    runs as k_rhsmoc_col + k_remap_col; two FP64 steps against the oracle."""
    p = Input3D(np1x=8, np2x=8, np1y=1028, np2y=1028, np1z=4, np2z=4,
                ql1x=5.0, ql2x=5.0, ql1y=642.5, ql2y=642.5, ql1z=2.5, ql2z=2.5)
    st = ics.random_state(p, seed=17)
    o = oracle.Oracle(p, st, nsplit=(1, 8, 2))
    g = mcf.Domain(p, nsplit=(1, 8, 2))
    g.upload(st)
    dts_o, _ = o.run(2)
    dts_g = g.run(2)
    np.testing.assert_allclose(dts_g, dts_o, rtol=1e-13)
    errs = rel_linf(interior(g.download(), p), interior(o.state(), p))
    assert worst(errs) <= MULTI_TOL, errs


@pytest.mark.parametrize("precision", [mcf.PREC_FP64, mcf.PREC_MIXED])
def test_multistep_vs_oracle_long_z(precision):
    """z longer than one column segment (256 cells) and than SOURCE's
    z segments (64 planes mixed, 32 FP64), not a multiple of either: the
    oracle's result (FP64) and agreement to FP32 round-off (mixed)."""
    p = small(24, 16, 264)
    st = ics.random_state(p, seed=14)
    o = oracle.Oracle(p, st, nsplit=(2, 2, 2))
    g = mcf.Domain(p, nsplit=(2, 2, 2), precision=precision)
    g.upload(st)
    dts_o, _ = o.run(2)
    if precision == mcf.PREC_FP64:
        dts_g = g.run(2)
        np.testing.assert_allclose(dts_g, dts_o, rtol=1e-13)
        errs = rel_linf(interior(g.download(), p), interior(o.state(), p))
        assert worst(errs) <= MULTI_TOL, errs
    else:
        for dt in dts_o:
            g.step(dt)
        got = interior(g.download(), p)
        ref = interior(o.state(), p)
        for f in range(8):
            assert np.abs(got[f] - ref[f]).max() <= 1e-5 * np.abs(ref[f]).max(), f


@pytest.mark.parametrize("periodic", [False, True])
def test_multistep_vs_oracle_long_rows(periodic):
    """x longer than one wave of the fused x sweep (56 cells), not a multiple
    of it, with inflow and with periodic x boundaries."""
    p = small(72, 16, 16, periodic=periodic)
    st = ics.random_state(p, seed=31 if periodic else 32)
    o = oracle.Oracle(p, st, nsplit=(3, 2, 2))
    g = mcf.Domain(p, nsplit=(3, 2, 2))
    g.upload(st)
    dts_o, _ = o.run(2)
    dts_g = g.run(2)
    np.testing.assert_allclose(dts_g, dts_o, rtol=1e-13)
    errs = rel_linf(interior(g.download(), p), interior(o.state(), p))
    assert worst(errs) <= MULTI_TOL, errs


def test_nonuniform_inflow_x_uses_stage_kernels():
    """x inflow with two spacings (16 cells over 10 pc, 8 over 10 pc): the x
    sweep runs as the three stage kernels, so REMAP's x ghosts come from the
    mode-2 fill rules (momentum from the filled density, energy from
    density, Bcc_x and Blg); FP64 result against the oracle."""
    p = Input3D(np1x=16, np2x=8, np1y=8, np2y=8, np1z=8, np2z=8,
                ql1x=10.0, ql2x=10.0, ql1y=5.0, ql2y=5.0, ql1z=5.0, ql2z=5.0)
    st = ics.random_state(p, seed=41)
    o = oracle.Oracle(p, st, nsplit=(2, 2, 2))
    g = mcf.Domain(p, nsplit=(2, 2, 2))
    g.upload(st)
    dts_o, _ = o.run(2)
    dts_g = g.run(2)
    np.testing.assert_allclose(dts_g, dts_o, rtol=1e-13)
    errs = rel_linf(interior(g.download(), p), interior(o.state(), p))
    assert worst(errs) <= MULTI_TOL, errs


def test_mixed_x_boundaries_rejected():
    """An inflow end opposite a periodic end is refused at set-up."""
    p = small(24, 16, 16)
    p.bcx2 = 3
    with pytest.raises(mcf.MCFError, match="both inflow"):
        mcf.Domain(p, nsplit=(2, 2, 2))


def test_nonuniform_periodic_x_uses_stage_kernels():
    """Periodic x with two spacings (16 cells over 10 pc, 8 over 10 pc) runs
    the x sweep as the three stage kernels; FP64 result against the oracle."""
    p = Input3D(np1x=16, np2x=8, np1y=8, np2y=8, np1z=8, np2z=8,
                ql1x=10.0, ql2x=10.0, ql1y=5.0, ql2y=5.0, ql1z=5.0, ql2z=5.0)
    p.bcx1 = p.bcx2 = 3
    st = ics.random_state(p, seed=33)
    o = oracle.Oracle(p, st, nsplit=(2, 2, 2))
    g = mcf.Domain(p, nsplit=(2, 2, 2))
    g.upload(st)
    dts_o, _ = o.run(2)
    dts_g = g.run(2)
    np.testing.assert_allclose(dts_g, dts_o, rtol=1e-13)
    errs = rel_linf(interior(g.download(), p), interior(o.state(), p))
    assert worst(errs) <= MULTI_TOL, errs


@pytest.mark.parametrize("precision", [mcf.PREC_FP64, mcf.PREC_MIXED])
def test_short_axis_uses_stage_kernels(precision):
    """An axis shorter than 8 cells runs the y/z sweep as three stage
    kernels (RHS, MOC, REMAP with fills between them) instead of the fused
    column march; both must give the oracle's result (FP64) and agree with
    each other to FP32 round-off (mixed)."""
    p = small(24, 16, 6)
    st = ics.random_state(p, seed=12)
    o = oracle.Oracle(p, st, nsplit=(2, 2, 2))
    g = mcf.Domain(p, nsplit=(2, 2, 2), precision=precision)
    g.upload(st)
    dts_o, _ = o.run(2)
    if precision == mcf.PREC_FP64:
        dts_g = g.run(2)
        np.testing.assert_allclose(dts_g, dts_o, rtol=1e-13)
        errs = rel_linf(interior(g.download(), p), interior(o.state(), p))
        assert worst(errs) <= MULTI_TOL, errs
    else:
        for dt in dts_o:
            g.step(dt)
        got = interior(g.download(), p)
        ref = interior(o.state(), p)
        for f in range(8):
            assert np.abs(got[f] - ref[f]).max() <= 1e-5 * np.abs(ref[f]).max(), f


@pytest.mark.parametrize("n", [64, 128])
def test_divergence_free(n):
    """CT keeps the discrete div B of the face field at round-off."""
    p = small(n, n, n, periodic=True)
    st = ics.random_state(p, seed=9, bamp=4.0)
    g = mcf.Domain(p, nsplit=(2, 2, 2))
    g.upload(st)
    g.run(4)
    out = g.download()
    bscale = np.abs(out[5:8]).max() / min(p.ql1x / p.np1x, p.ql1y / p.np1y)
    div = np.abs(ics.face_divergence(out, p)).max()
    assert div <= 1e-12 * bscale, (div, bscale)


def test_uniform_state_preserved():
    """A uniform periodic state at rest with a uniform field stays uniform and
    at rest (v = 0 exactly); only p moves, identically in every cell."""
    p = small(64, 64, 64, periodic=True)
    st = np.zeros((16,) + tuple(np.array(p.n)[::-1] + 4))
    cm = ics.bbrv_cm(p)
    rho = 1.3
    st[0] = rho
    for s in range(8):
        st[8 + s] = cm[s] * rho
    ntot = st[9] + st[8] + st[10] + st[11] + st[12]
    st[4] = ics.KB * ntot * 6.0
    st[5], st[6], st[7] = 3.0, 3.0, 1.0
    g = mcf.Domain(p, nsplit=(2, 2, 2))
    g.upload(st)
    g.run(3)
    out = interior(g.download(), p)
    for f in range(16):
        assert out[f].max() == out[f].min(), f
    assert np.all(out[1:4] == 0.0)
