#!/usr/bin/env python3
"""Benchmark: MHD zone-updates/s of the full timestep at 256^3 on one MI355X,
mixed precision + cooling surrogate against the plain FP64 path.

Workload: the thermally bistable converging flow at 256^3 -- 20 pc box, WNM
n = 1 cm^-3 at 6000 K, +-10 km/s inflow along x, seeded Kolmogorov density
perturbation (0.1 rms), B = 5 muG at 45 deg, KI cooling and conduction on,
the reference's 4x4x4 block split.  A step is one EVOLVE iteration of the
reference (3DMHD-CfCA/MCF3DSGincfca3.f90:562-657): the dt reduction,
SOURCE(dt/2), three sweeps, CT, SOURCE(dt/2) and DISSIP; the state stays
resident in HBM.

The headline ``value`` is the mixed-precision path (FP32 reconstruction,
Riemann and EMF; FP64 state, pressure recovery and face B) with the MLP
cooling surrogate; the plain FP64 path is timed on the same state in the
same run (``fp64``).  The timed steps carry no profiling events; ten more
steps with every launch bracketed by HIP events give the per-kernel and
per-stage tables.  ``roofline`` is the stage (dt, SOURCE call, x/y/z sweep,
CT) of the kernel that takes the most time per step: its compulsory HBM bytes (each
persistent array it must read or write, once per interior cell: 30 doubles
per cell for a sweep, 14 for CT, 8 for a SOURCE call, 13 for dt) over its
measured time; two child ``rocprofv3 --pmc`` runs (FETCH_SIZE, then
WRITE_SIZE, calibrated by a plain copy kernel) give the HBM bytes its
kernels moved.  ``stiff`` times the same problem at 1e4 times the density,
where SOURCE's backward-Euler branch (IMC or the surrogate) does the work.
The CPU restatement (oracle/, OpenMP on the host's cores) is timed on a
bounded sample of the same problem as the CPU baseline.

``chemistry`` times the same problem with the reference's chemistry network,
SHIELD column densities and chemistry-network Fcool switched on.

``gravity`` times the same problem with self-gravity (ifgrv = 2): the
multigrid Poisson solve once per step and two velocity kicks, on the 4x4x4
split and on the 8x8x8 split of the reference's 512-rank runs; with the CPU
baseline on, the oracle's solve of the same density on 8x8x8 is timed on
the host beside it.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-pmc] [--no-stiff] [--no-chem]
                       [--no-grav]
"""
import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E
COPY_WORDS = 1 << 27            # calibration copy: 1 GiB read + 1 GiB written


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_state(n):
    from code_tohoku_amd import ics
    p = ics.converging_flow_input(n)
    return p, ics.converging_flow_state(p)


def make_domain(p, st, nsplit, mixed, surrogate, weights):
    from code_tohoku_amd import mcf
    d = mcf.Domain(p, nsplit=nsplit, precision=mcf.PREC_MIXED if mixed else mcf.PREC_FP64,
                   cooling=mcf.COOL_SURROGATE if surrogate else mcf.COOL_KI)
    d.upload(st)
    if surrogate:
        d.mlp_set_weights(weights)
    return d


def surrogate_weights():
    from code_tohoku_amd import surrogate
    try:
        return surrogate.load_weights(), "data/ki_cool_mlp.f32"
    except (OSError, ValueError):
        log("[bench] no saved surrogate weights: training on the GPU (1e7 IMC samples)")
        w, _ = surrogate.train(samples=10_000_000, epochs=16)
        return w, "trained in this run"


def timed(d, steps, warmup, profile_steps=10):
    """Warm-up, then the timed steps (no profiling events in them), then
    profile_steps more steps with every launch bracketed by HIP events for
    the per-kernel and per-stage tables: returns (seconds, dts, {kernel:
    (launches, ms, touched bytes)}, {stage: (runs, ms, compulsory bytes)},
    profiled steps)."""
    d.run(warmup, tsave=1e30)
    d.sync()
    t0 = time.perf_counter()
    dts = d.run(steps, tsave=1e30)
    d.sync()
    el = time.perf_counter() - t0
    assert len(dts) == steps
    rep, stg, np_ = {}, {}, 0
    if profile_steps > 0:
        d.profile(True)
        np_ = len(d.run(profile_steps, tsave=1e30))
        d.sync()
        rep = d.profile_report()
        stg = d.profile_stages()
        d.profile(False)
    return el, dts, rep, stg, np_


def report_kernels(tag, rep, stg, steps):
    tot = sum(v[1] for v in rep.values())
    log(f"[bench] {tag}: kernel time {tot / steps:.3f} ms/step")
    for k, (runs, ms, by) in sorted(stg.items(), key=lambda kv: -kv[1][1]):
        if runs:
            log(f"  stage {k:10s} runs/step {runs / steps:4.1f}  ms/run {ms / runs:8.4f}  ms/step {ms / steps:8.4f}"
                f"  {by / (ms * 1e-3) / 1e9 if ms > 0 else 0:8.1f} GB/s compulsory")
    for k, (cnt, ms, by) in sorted(rep.items(), key=lambda kv: -kv[1][1])[:16]:
        log(f"  {k:18s} launches {cnt:6d}  avg {1e3 * ms / cnt:9.2f} us  total/step {ms / steps:8.3f} ms"
            f"  {by / (ms * 1e-3) / 1e9:8.1f} GB/s touched")


def kernel_stage(name):
    """The step stage a kernel belongs to (None for the ghost fills, which
    several stages launch).  Sweep kernels carry their axis as the first
    template argument (k_rhsmoc_col<1> is the y sweep's)."""
    if name in ("k_sweep_x", "k_rhsmoc_x", "k_remap_x") or name.endswith("<0>"):
        return "sweep_x"
    if name.endswith("<1>"):
        return "sweep_y"
    if name.endswith("<2>"):
        return "sweep_z"
    if name in ("k_cc1", "k_ct_march", "k_ct_march_mixed", "k_cc3", "k_cc4"):
        return "ct"
    if name in ("k_src_col", "k_shell_copy", "k_src_stiff", "k_src_imc", "k_shield", "k_chem"):
        return "source"
    if name in ("k_dtmin", "k_dtfinal"):
        return "dt"
    if name in ("k_dissip_div", "k_dissip_upd"):
        return "dissip"
    if name == "k_grav_kick":
        return "grav_kick"
    if name.startswith("k_grav_") or name.startswith("k_pb_"):
        return "grav_solve"
    return None


def stage_rows(stg, rep, steps, traffic=None):
    """Per-stage table: compulsory bytes (each persistent array the stage
    must read or write, once per interior cell) over the stage's measured
    time, and, where the counter runs measured them, the HBM bytes its own
    kernels moved per run (ghost fills excluded)."""
    rows = {}
    for name, (runs, ms, by) in stg.items():
        if runs == 0 or by == 0:
            continue
        ach = by / (ms * 1e-3) / 1e9
        kern = {k: round(1e3 * v[1] / v[0], 2) for k, v in rep.items() if kernel_stage(k) == name}
        row = {"runs_per_step": runs / steps, "ms_per_run": ms / runs, "ms_per_step": ms / steps,
               "compulsory_bytes_per_run": by / runs, "achieved_GBs": round(ach, 1),
               "frac": round(ach / HBM_PEAK_GBS, 4), "kernels_avg_us": kern}
        if traffic is not None and traffic[0] is not None:
            per, _ = traffic
            tb = [per[k]["bytes_per_step"] for k in kern if k in per]
            if tb:
                row["traffic_bytes_per_run"] = sum(tb) / (runs / steps)
        rows[name] = row
    return rows


def roofline(rows, traffic, rep):
    """The roofline of the dominant kernel's stage: the stage of the kernel
    that takes the most time per step (a sweep's stage is that one kernel).
    The stage with the most time per step is named beside it."""
    kname = max((k for k in rep if kernel_stage(k) in rows), key=lambda k: rep[k][1])
    name = kernel_stage(kname)
    r = rows[name]
    slowest = max(rows.items(), key=lambda kv: kv[1]["ms_per_step"])[0]
    out = {"bound": "hbm", "stage": name, "dominant_kernel": kname, "slowest_stage": slowest,
           "kernels_avg_us": r["kernels_avg_us"], "achieved": r["achieved_GBs"],
           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": r["frac"],
           "traffic": r.get("traffic_bytes_per_run"),
           "alg_bytes_per_run": r["compulsory_bytes_per_run"], "ms_per_run": r["ms_per_run"],
           "accounting": "compulsory bytes of the stage: a sweep reads rho, v(3), p, 8 species, Bcc(3), face B_n and "
                         "writes rho, v(3), p, 8 species (30 doubles per cell); CT 8 + 6; SOURCE 7 + 1 per call; "
                         "dt 13"}
    if traffic is not None and traffic[0] is None:
        out["traffic_note"] = traffic[1]
    elif traffic is not None:
        out["traffic_detail"] = traffic[1]
    return out


def kernel_table(rep, traffic, steps, top=10):
    """Per-kernel time, bytes of the arrays each kernel touches (a
    diagnostic: it counts the sweep's own intermediates) and, where the
    counter runs measured it, HBM traffic per launch."""
    per = traffic[0] if traffic is not None else None
    rows = []
    for name, (cnt, ms, by) in sorted(rep.items(), key=lambda kv: -kv[1][1])[:top]:
        row = {"kernel": name, "stage": kernel_stage(name), "launches_per_step": cnt / steps,
               "avg_launch_us": round(1e3 * ms / cnt, 2), "ms_per_step": round(ms / steps, 4),
               "touched_bytes_per_launch": by / cnt, "touched_GBs": round(by / (ms * 1e-3) / 1e9, 1)}
        if per is not None and name in per:
            row["traffic_bytes_per_launch"] = per[name]["bytes"]
            row["traffic_GBs"] = round(per[name]["bytes"] / (ms / cnt * 1e-3) / 1e9, 1)
        rows.append(row)
    return rows


def cpu_baseline(n: int, steps: int):
    """The oracle (C restatement, OpenMP over all the threads OpenMP is given)
    on `steps` FP64 steps of the same problem."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from code_tohoku_amd import ics

    p = ics.converging_flow_input(n)
    st = ics.converging_flow_state(p)
    o = oracle.Oracle(p, st, nsplit=(max(1, n // 64),) * 3)
    del st
    t0 = time.perf_counter()
    dts, _ = o.run(steps)
    el = time.perf_counter() - t0
    return {"value": n ** 3 * len(dts) / el, "unit": "zone-updates/s", "cores": o.threads, "kind": "port",
            "sample": f"CPU restatement (oracle/, C + OpenMP, {o.threads} threads), {len(dts)} FP64 steps of the "
                      f"same converging-flow problem at {n}^3, {el:.1f} s"}


def cpu_grav(n, nsplit):
    """The oracle's GRAVTY(dt,2) (one multigrid Poisson solve with the FFT
    boundary potential) on the host, for the gravity leg: the same 256^3
    converging-flow density on the same block split."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    p, st = build_state(n)
    p.ifgrv = 2
    o = oracle.Oracle(p, st, nsplit=nsplit)
    del st
    t0 = time.perf_counter()
    o.grav_solve()
    el = time.perf_counter() - t0
    return {"solve_ms": round(1e3 * el, 1), "threads": o.threads, "kind": "port",
            "sample": f"CPU restatement (oracle/mcf_grav.c, C + OpenMP, {o.threads} threads), one GRAVTY(dt,2) "
                      f"solve at {n}^3 on the {'x'.join(map(str, nsplit))} split"}


def stiff_state(n):
    """The converging flow at 1e4 times the density and pressure (n = 1e4
    cm^-3, T = 6000 K): the cooling time is below Stblty's dt floor, so SOURCE
    takes its backward-Euler branch in most cells."""
    p, st = build_state(n)
    st[0] *= 1.0e4
    st[4] *= 1.0e4
    st[8:16] *= 1.0e4
    return p, st


def stiff_compare(n, steps, warmup, w):
    """ms/step of the stiff problem: FP64 + IMC, mixed + IMC, mixed +
    surrogate, with the backward-Euler cell counts of the timed steps."""
    p, st = stiff_state(n)
    nsplit = (max(1, n // 64),) * 3
    out = {"workload": f"converging_flow_{n}^3 at 1e4 x density (n = 1e4 cm^-3, T = 6000 K)", "steps": steps}
    for tag, mixed, sur in (("fp64_imc", False, False), ("mixed_imc", True, False), ("mixed_surrogate", True, True)):
        d = make_domain(p, st, nsplit, mixed, sur, w)
        d.run(warmup, tsave=1e30)
        d.sync()
        d.cool_stats(reset=True)
        t0 = time.perf_counter()
        dts = d.run(steps, tsave=1e30)
        d.sync()
        el = time.perf_counter() - t0
        nst, nfb = d.cool_stats()
        d.close()
        out[tag] = {"ms_per_step": 1e3 * el / steps, "zone_updates_per_s": n ** 3 * steps / el,
                    "backward_euler_cells": nst, "imc_fallbacks": nfb,
                    "fallback_fraction": (nfb / nst) if (sur and nst) else None, "dt_mean": float(np.mean(dts))}
        log(f"[bench] stiff {tag}: {out[tag]}")
    out["surrogate_speedup_vs_mixed_imc"] = out["mixed_imc"]["ms_per_step"] / out["mixed_surrogate"]["ms_per_step"]
    out["mixed_surrogate_speedup_vs_fp64"] = out["fp64_imc"]["ms_per_step"] / out["mixed_surrogate"]["ms_per_step"]
    return out


def chem_compare(n, steps, warmup):
    """The same converging flow with the reference's chemistry switched on:
    the chemistry network (ifchem = 2), SHIELD column densities (ifrad = 2)
    and the chemistry-network Fcool, on the FP64 and the mixed path (whose
    SOURCE runs in FP64 here).  ms/step and the per-kernel times of SHIELD,
    the network and SOURCE's thermal kernels."""
    from code_tohoku_amd import mcf
    p, st = build_state(n)
    p.ifchem, p.ifrad = 2, 2
    nsplit = (max(1, n // 64),) * 3
    out = {"workload": f"converging_flow_{n}^3 with ifchem = 2, ifrad = 2 and the chemistry-network Fcool",
           "steps": steps,
           "accounting": "compulsory bytes per cell: SHIELD reads 8 species, p and tCII(2) and writes 10 column "
                         "arrays (21 doubles); the network reads 8 species, p and 8 columns and writes 8 species "
                         "and rho (26 doubles)"}
    cells = float(n) ** 3
    for tag, prec in (("fp64", mcf.PREC_FP64), ("mixed", mcf.PREC_MIXED)):
        d = mcf.Domain(p, nsplit=nsplit, precision=prec, cooling=mcf.COOL_CHEM)
        t0 = time.perf_counter()
        d.upload(st)
        d.sync()
        up = time.perf_counter() - t0
        el, dts, rep, stg, nprof = timed(d, steps, warmup, profile_steps=5)
        d.close()
        kern = {k: {"avg_us": round(1e3 * v[1] / v[0], 2), "launches_per_step": v[0] / nprof}
                for k, v in rep.items() if k in ("k_shield", "k_chem", "k_src_col", "k_src_imc", "k_dtmin")}
        for k, nb in (("k_shield", 21.0), ("k_chem", 26.0)):
            if k in kern:
                gbs = 8.0 * nb * cells / (kern[k]["avg_us"] * 1e-6) / 1e9
                kern[k]["compulsory_GBs"] = round(gbs, 1)
                kern[k]["frac"] = round(gbs / HBM_PEAK_GBS, 4)
        src = stg.get("source")
        out[tag] = {"ms_per_step": 1e3 * el / steps, "zone_updates_per_s": n ** 3 * steps / el,
                    "upload_with_20_shield_s": round(up, 3),
                    "source_ms_per_step": (src[1] / nprof) if src else None, "kernels": kern,
                    "dt_mean": float(np.mean(dts))}
        log(f"[bench] chemistry {tag}: {out[tag]}")
    return out


def grav_compare(n, steps, warmup):
    """The same converging flow with self-gravity (ifgrv = 2): GRAVTY(dt,3)
    before the first SOURCE, GRAVTY(dt,2) and (dt,3) after the second
    (M3:607, 653).  The multigrid follows the reference's block split, so
    its cost depends on it: the 4x4x4 split of the headline and the 8x8x8
    split of the reference's 512-rank runs, mixed precision (the MHD part)
    with KI cooling, and FP64 on 8x8x8."""
    from code_tohoku_amd import mcf
    p, st = build_state(n)
    p.ifgrv = 2
    out = {"workload": f"converging_flow_{n}^3 with self-gravity (ifgrv = 2), KI cooling", "steps": steps}
    for tag, nsplit, prec in (("mixed_4x4x4", (4, 4, 4), mcf.PREC_MIXED), ("mixed_8x8x8", (8, 8, 8), mcf.PREC_MIXED),
                              ("fp64_8x8x8", (8, 8, 8), mcf.PREC_FP64)):
        d = mcf.Domain(p, nsplit=nsplit, precision=prec)
        d.upload(st)
        el, dts, rep, stg, nprof = timed(d, steps, warmup, profile_steps=3)
        d.close()
        gs, gk = stg.get("grav_solve"), stg.get("grav_kick")
        glaunch = sum(v[0] for k, v in rep.items() if kernel_stage(k) == "grav_solve") / nprof
        top = sorted(((k, v) for k, v in rep.items() if kernel_stage(k) in ("grav_solve", "grav_kick")),
                     key=lambda kv: -kv[1][1])[:12]
        out[tag] = {"ms_per_step": 1e3 * el / steps, "zone_updates_per_s": n ** 3 * steps / el,
                    "grav_solve_ms_per_step": (gs[1] / nprof) if gs else None,
                    "grav_kick_ms_per_step": (gk[1] / nprof) if gk else None,
                    "grav_solve_launches_per_step": glaunch,
                    "kernels": {k: {"avg_us": round(1e3 * v[1] / v[0], 2), "launches_per_step": v[0] / nprof,
                                    "ms_per_step": round(v[1] / nprof, 3)} for k, v in top},
                    "dt_mean": float(np.mean(dts))}
        # the finest level's Gauss-Seidel sweep: compulsory bytes (old phi
        # and rhs read, new phi written: 24 B per node, n^3 nodes) against
        # HBM peak, and the time per hyperplane of the block's march (3C-5
        # planes for C cells per block and axis), the latency that bounds it
        fin = rep.get(f"k_grav_relax<{int(round(np.log2(n)))}>")
        if fin:
            avg = fin[1] / fin[0] * 1e-3
            cmax = n // min(nsplit)
            out[tag]["relax_finest"] = {"avg_us": round(avg * 1e6, 2), "compulsory_bytes": 24 * n ** 3,
                                        "achieved_GBps": round(24 * n ** 3 / avg / 1e9, 1),
                                        "frac_of_hbm_peak": round(24 * n ** 3 / avg / 1e9 / HBM_PEAK_GBS, 4),
                                        "planes_per_sweep": 3 * cmax - 5,
                                        "us_per_plane": round(avg * 1e6 / (3 * cmax - 5), 3)}
        log(f"[bench] gravity {tag}: {out[tag]['ms_per_step']:.3f} ms/step, solve "
            f"{out[tag]['grav_solve_ms_per_step']:.3f} ms in {glaunch:.0f} launches")
    return out


# ---------------------------------------------------------------------------
# HBM traffic of the dominant kernel: child rocprofv3 --pmc runs
# ---------------------------------------------------------------------------
PMC_STEPS = 2


def pmc_child(args):
    """Run under rocprofv3 --pmc: upload, PMC_STEPS steps of the headline
    path, then the calibration copy."""
    p, st = build_state(args.n)
    w, _ = surrogate_weights()
    d = make_domain(p, st, (max(1, args.n // 64),) * 3, True, True, w)
    del st
    d.run(PMC_STEPS, tsave=1e30)
    d.sync()
    d.copy_probe(COPY_WORDS, 1)
    d.close()


def short_name(kernel_name):
    """rocprofv3's kernel name -> the library's profile name: namespace and
    argument list dropped, a leading integer template argument (the sweep
    axis) kept: "void mcf::k_rhsmoc_col<1, float>(...)" -> "k_rhsmoc_col<1>"."""
    base = kernel_name.split("(")[0].replace("void ", "").replace("mcf::", "").strip()
    if "<" in base:
        head, args_ = base.split("<", 1)
        first = args_.rstrip(">").split(",")[0].strip()
        base = f"{head}<{first}>" if first.isdigit() else head
    return base


def pmc_traffic(args):
    """{kernel: HBM bytes} of the headline path's steps from FETCH_SIZE and
    WRITE_SIZE (one rocprofv3 run each, calibrated by k_copy), or a reason.
    Only dispatches from the first dt reduction on are counted (the upload's
    CC(4) is not part of a step)."""
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None, "rocprofv3 not found"
    res = {}
    tmp = tempfile.mkdtemp(prefix="bench_pmc_")
    try:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            out = os.path.join(tmp, counter)
            cmd = ["timeout", "-s", "KILL", "240", prof, "--pmc", counter, "--output-format", "csv", "-d", out,
                   "-o", "pmc", "--", sys.executable, os.path.abspath(__file__), "--pmc-child", "--n", str(args.n)]
            log(f"[bench] PMC pass {counter}")
            r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, cwd=tmp)
            files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
            if r.returncode != 0 and not files:
                return None, f"rocprofv3 pass {counter} failed (rc {r.returncode})"
            disp = {}
            for f in files:
                for row in csv.DictReader(open(f)):
                    if row.get("Counter_Name") != counter:
                        continue
                    did = int(row["Dispatch_Id"])
                    name, v = disp.get(did, (short_name(row["Kernel_Name"]), 0.0))
                    disp[did] = (name, v + float(row["Counter_Value"]))
            first = min((i for i, (nm, _) in disp.items() if nm == "k_dtmin"), default=None)
            per = {}
            for did, (nm, v) in disp.items():
                if nm != "k_copy" and (first is None or did < first):
                    continue
                per.setdefault(nm, []).append(v)
            res[counter] = per
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    try:
        copy_f = res["FETCH_SIZE"]["k_copy"][-1]
        copy_w = res["WRITE_SIZE"]["k_copy"][-1]
    except (KeyError, IndexError):
        return None, "counters missing for the calibration copy"
    nbytes = 8.0 * COPY_WORDS
    sf, sw = nbytes / (copy_f * 1024.0), nbytes / (copy_w * 1024.0)
    out = {}
    for k, kf in res["FETCH_SIZE"].items():
        kw = res["WRITE_SIZE"].get(k)
        if not kw or k == "k_copy":
            continue
        fetch = sum(kf) / len(kf) * 1024.0 * sf
        write = sum(kw) / len(kw) * 1024.0 * sw
        out[k] = {"bytes": fetch + write, "fetch_bytes": fetch, "write_bytes": write, "launches": len(kf),
                  "bytes_per_step": (fetch + write) * len(kf) / PMC_STEPS}
    return out, {"fetch_scale": sf, "write_scale": sw, "steps": PMC_STEPS,
                 "calibration": f"k_copy, {int(nbytes)} B read and {int(nbytes)} B written"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pmc", action="store_true")
    ap.add_argument("--no-stiff", action="store_true")
    ap.add_argument("--no-chem", action="store_true")
    ap.add_argument("--no-grav", action="store_true")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: the library runs the whole domain on a single MI355X "
              f"and has no multi-GPU path (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    if args.pmc_child:
        pmc_child(args)
        return

    n = args.n
    nsplit = (max(1, n // 64),) * 3
    # counter runs first, as child processes, before this process uses the GPU
    pmc = (None, "not measured (--no-pmc)") if args.no_pmc else pmc_traffic(args)

    log(f"[bench] building the converging-flow state {n}^3 ...")
    p, st = build_state(n)
    w, wsrc = surrogate_weights()

    # headline: mixed precision + surrogate
    d = make_domain(p, st, nsplit, True, True, w)
    log(f"[bench] mixed+surrogate: warmup {args.warmup}, timed {args.steps} steps")
    d.cool_stats(reset=True)
    el, dts, rep, stg, nprof = timed(d, args.steps, args.warmup)
    nstiff, nfall = d.cool_stats()
    report_kernels("mixed+surrogate", rep, stg, nprof)
    from code_tohoku_amd import surrogate
    solve = surrogate.evaluate(d, n=1 << 20)
    log(f"[bench] surrogate on 2^20 stiff cells: {solve}")
    cms = d.copy_probe(1 << 28, 10)
    d.close()
    # the plain FP64 path on the same state
    f = make_domain(p, st, nsplit, False, False, None)
    del st
    log(f"[bench] fp64: warmup {args.warmup}, timed {args.steps} steps")
    el64, dts64, rep64, stg64, nprof64 = timed(f, args.steps, args.warmup)
    report_kernels("fp64", rep64, stg64, nprof64)
    f.close()

    ms_step = 1e3 * el / args.steps
    value = n ** 3 * args.steps / el
    value64 = n ** 3 * args.steps / el64
    log(f"[bench] mixed+surrogate {ms_step:.3f} ms/step {value:.4e} zone-updates/s; "
        f"fp64 {1e3 * el64 / args.steps:.3f} ms/step {value64:.4e}; ratio {value / value64:.3f}")

    rows = stage_rows(stg, rep, nprof, pmc)
    rows64 = stage_rows(stg64, rep64, nprof64)
    step_bytes = sum(r["compulsory_bytes_per_run"] * r["runs_per_step"] for r in rows.values())
    step_gbs = step_bytes / (ms_step * 1e-3) / 1e9
    out = {"metric": "MHD zone-updates/s at 256^3 (mixed+surrogate vs FP64) + % HBM roofline",
           "value": value, "unit": "zone-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+f64 (mixed; state f64, surrogate bf16/f32)", "data": "synthetic",
           "config": {"workload": f"converging_flow_{n}^3_ki_cooling", "grid": [n, n, n],
                      "block_split": list(nsplit), "precision": "mixed", "cooling": "surrogate",
                      "surrogate_weights": wsrc},
           "roofline": roofline(rows, pmc, rep),
           "stages": rows,
           "kernels": kernel_table(rep, pmc, nprof),
           "speedup_vs_fp64": value / value64,
           "fp64": {"value": value64, "ms_per_step": 1e3 * el64 / args.steps, "roofline": roofline(rows64, None, rep64),
                    "stages": rows64, "kernels": kernel_table(rep64, None, nprof64),
                    "dt_mean": float(sum(dts64) / len(dts64))},
           "surrogate": {"backward_euler_cells": nstiff, "imc_fallbacks": nfall,
                         "fallback_fraction": (nfall / nstiff) if nstiff else None,
                         "note": "cells of the run's steps (warm-up, timed, profiled) that took SOURCE's "
                                 "backward-Euler branch; none at n = 1 (Stblty keeps dt below 0.2 t_cool)",
                         "stiff_solve": solve},
           "step_roofline": {"compulsory_bytes_per_step": step_bytes, "achieved": round(step_gbs, 1),
                             "unit": "GB/s", "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
           "dt_mean": float(sum(dts) / len(dts)),
           "copy_probe": {"bytes": 16 * (1 << 28), "ms": cms, "GBs": round(16 * (1 << 28) / (cms * 1e-3) / 1e9, 1)}}
    if not args.no_stiff:
        out["stiff"] = stiff_compare(n, 10, 2, w)
    if not args.no_chem:
        out["chemistry"] = chem_compare(n, 10, 2)
    if not args.no_grav:
        out["gravity"] = grav_compare(n, 5, 1)
    if not args.no_cpu:
        log(f"[bench] CPU baseline: oracle {args.cpu_steps} steps at {args.cpu_n}^3")
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_steps)
        if "gravity" in out:
            log(f"[bench] CPU baseline: oracle gravity solve at {n}^3")
            cg = cpu_grav(n, (8, 8, 8))
            gs = out["gravity"]["mixed_8x8x8"]["grav_solve_ms_per_step"]
            cg["gpu_solve_ms"] = gs
            cg["gpu_speedup"] = round(cg["solve_ms"] / gs, 1) if gs else None
            out["gravity"]["cpu_baseline"] = cg
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
