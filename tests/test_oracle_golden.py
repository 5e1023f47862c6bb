"""The CPU restatement (oracle/) against the reference's own outputs.

The fixtures were written by the reference MPI build (oracle/build_ref.sh,
tests/golden/make_golden.py) from the same inputs; the oracle must reproduce
them bit for bit, including the per-step dt sequence.  This pins every
routine of the step that the fixtures exercise: SOURCE with KI cooling
(explicit and implicit branches), RHS/RIEMAN, MOC/CHAREQ with the
rank-boundary ghost semantics, REMAP, the CT family, DISSIP and the dt
reduction, in both x-inflow and periodic geometries and for the 8- and
16-rank block splits, on seeded random states and on the converging-flow
initial state (Kolmogorov density, tanh inflow profile, B at 45 deg).
"""
import numpy as np
import pytest

import oracle
from helpers import case_cooling, expand, faces, interior, load_case, rel_linf, worst

CASES = [("inflow_24x8x8", 1), ("inflow_24x8x8", 6), ("periodic_24x8x8", 1), ("periodic_24x8x8", 6),
         ("inflow_48x8x8", 6), ("briowu_64x8x8", 200), ("converging_24", 8),
         ("chem_rad_24x8x8", 1), ("chem_rad_24x8x8", 6), ("chem_rad_48x8x8", 6), ("chem_norad_24x8x8", 6)]


@pytest.mark.parametrize("name,steps", CASES)
def test_oracle_matches_reference(name, steps):
    p, nsplit, arr = load_case(name)
    o = oracle.Oracle(p, expand(arr["input"], p), nsplit=nsplit, cooling=case_cooling(name))
    dts, _ = o.run(steps)
    assert len(dts) == steps
    np.testing.assert_array_equal(dts, arr[f"dts{steps}"])
    got = o.state()
    ref = expand(arr[f"out{steps}"], p)
    errs = rel_linf(interior(got, p), interior(ref, p))
    assert worst(errs) == 0.0, errs
    for a, b in zip(faces(got, p), faces(ref, p)):
        np.testing.assert_array_equal(a, b)
    assert o.time[0] == float(arr[f"time{steps}"])


def test_block_split_matters():
    """The reference's result depends on its MPI block split (MOC reads the
    pre-sweep pressure in rank-local ghosts): running the oracle with the
    wrong split must NOT reproduce the fixture, so the fixture does test the
    block-edge emulation."""
    p, nsplit, arr = load_case("periodic_24x8x8")
    o = oracle.Oracle(p, expand(arr["input"], p), nsplit=(1, 1, 1))
    o.run(6)
    errs = rel_linf(interior(o.state(), p), interior(expand(arr["out6"], p), p))
    assert worst(errs) > 1e-12


@pytest.mark.parametrize("name,steps", [("inflow_48x8x8", 6), ("periodic_24x8x8", 6)])
def test_oracle_thread_count_independent(name, steps):
    """The oracle splits its pencil and plane loops over OpenMP threads: one
    thread and four threads must give the fixture bit for bit (each cell is
    computed by one thread with the serial expressions; the dt minima are
    exact)."""
    p, nsplit, arr = load_case(name)
    outs = []
    for th in (1, 4):
        o = oracle.Oracle(p, expand(arr["input"], p), nsplit=nsplit, threads=th)
        assert o.threads == th
        dts, _ = o.run(steps)
        np.testing.assert_array_equal(dts, arr[f"dts{steps}"])
        outs.append(o.state())
    np.testing.assert_array_equal(outs[0], outs[1])


def test_oracle_dtf_inflow_matches_reference():
    """With the DTF read enabled (oracle/patch_ref.py, option dtf) the
    inflow ghost density is the DTF field shifted by -v t (M3:818-826,
    854-861): the oracle given the same field reproduces the reference bit
    for bit, and the uniform inflow density does not."""
    p, nsplit, arr = load_case("converging_dtf_24")
    for use, exact in ((True, True), (False, False)):
        o = oracle.Oracle(p, expand(arr["input"], p), nsplit=nsplit)
        if use:
            o.set_dtf(arr["dtf"])
        dts, _ = o.run(8)
        errs = rel_linf(interior(o.state(), p), interior(expand(arr["out8"], p), p))
        if exact:
            np.testing.assert_array_equal(dts, arr["dts8"])
            assert worst(errs) == 0.0, errs
        else:
            assert worst(errs) > 1e-6
