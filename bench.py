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
steps with every launch bracketed by HIP events give the per-kernel table
and the roofline of the dominant kernel; two child ``rocprofv3 --pmc`` runs
(FETCH_SIZE, then WRITE_SIZE, calibrated by a plain copy kernel) give its
HBM traffic per launch.  The CPU restatement (oracle/) is timed on a bounded
sample of the same problem as the CPU baseline.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-pmc]
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

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E
ALG_BYTES_PER_ZONE = 920.0      # whole FP64 step, 5 active species (compulsory traffic)
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
    the per-kernel table: returns (seconds, dts, {kernel: (launches, ms,
    bytes)}, profiled steps)."""
    d.run(warmup, tsave=1e30)
    d.sync()
    t0 = time.perf_counter()
    dts = d.run(steps, tsave=1e30)
    d.sync()
    el = time.perf_counter() - t0
    assert len(dts) == steps
    rep, np_ = {}, 0
    if profile_steps > 0:
        d.profile(True)
        np_ = len(d.run(profile_steps, tsave=1e30))
        d.sync()
        rep = d.profile_report()
        d.profile(False)
    return el, dts, rep, np_


def report_kernels(tag, rep, steps):
    tot = sum(v[1] for v in rep.values())
    log(f"[bench] {tag}: kernel time {tot / steps:.3f} ms/step")
    for k, (cnt, ms, by) in sorted(rep.items(), key=lambda kv: -kv[1][1])[:14]:
        log(f"  {k:18s} launches {cnt:6d}  avg {1e3 * ms / cnt:9.2f} us  total/step {ms / steps:8.3f} ms"
            f"  {by / (ms * 1e-3) / 1e9:8.1f} GB/s")


def cpu_baseline(n: int, steps: int):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from code_tohoku_amd import ics

    p = ics.converging_flow_input(n)
    st = ics.converging_flow_state(p)
    o = oracle.Oracle(p, st, nsplit=(max(1, n // 64),) * 3)
    t0 = time.perf_counter()
    dts, _ = o.run(steps)
    el = time.perf_counter() - t0
    return {"value": n ** 3 * len(dts) / el, "unit": "zone-updates/s", "cores": 1, "kind": "port",
            "sample": f"CPU restatement (oracle/, C, 1 thread), {len(dts)} FP64 steps of the same converging-flow "
                      f"problem at {n}^3, {el:.1f} s"}


# ---------------------------------------------------------------------------
# HBM traffic of the dominant kernel: child rocprofv3 --pmc runs
# ---------------------------------------------------------------------------
def pmc_child(args):
    """Run under rocprofv3 --pmc: one warm-up and one step of the headline
    path, then the calibration copy."""
    p, st = build_state(args.n)
    w, _ = surrogate_weights()
    d = make_domain(p, st, (max(1, args.n // 64),) * 3, True, True, w)
    del st
    d.run(2, tsave=1e30)
    d.sync()
    d.copy_probe(COPY_WORDS, 1)
    d.close()


def pmc_traffic(args):
    """{kernel: HBM bytes per launch} of the headline path from FETCH_SIZE and
    WRITE_SIZE (one rocprofv3 run each, calibrated by k_copy), or a reason."""
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
            per = {}
            for f in files:
                for row in csv.DictReader(open(f)):
                    if row.get("Counter_Name") != counter:
                        continue
                    base = row["Kernel_Name"].split("(")[0].replace("void ", "").replace("mcf::", "").split("<")[0]
                    d = per.setdefault(base, {})
                    d[row["Dispatch_Id"]] = d.get(row["Dispatch_Id"], 0.0) + float(row["Counter_Value"])
            res[counter] = per
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    try:
        copy_f = list(res["FETCH_SIZE"]["k_copy"].values())[-1]
        copy_w = list(res["WRITE_SIZE"]["k_copy"].values())[-1]
    except (KeyError, IndexError):
        return None, "counters missing for the calibration copy"
    nbytes = 8.0 * COPY_WORDS
    sf, sw = nbytes / (copy_f * 1024.0), nbytes / (copy_w * 1024.0)
    out = {}
    for k, kf in res["FETCH_SIZE"].items():
        kw = res["WRITE_SIZE"].get(k)
        if not kw or k == "k_copy":
            continue
        fetch = sum(kf.values()) / len(kf) * 1024.0 * sf
        write = sum(kw.values()) / len(kw) * 1024.0 * sw
        out[k] = {"bytes": fetch + write, "fetch_bytes": fetch, "write_bytes": write, "launches": len(kf)}
    return out, {"fetch_scale": sf, "write_scale": sw,
                 "calibration": f"k_copy, {int(nbytes)} B read and {int(nbytes)} B written"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-n", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pmc", action="store_true")
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
    el, dts, rep, nprof = timed(d, args.steps, args.warmup)
    nstiff, nfall = d.cool_stats()
    report_kernels("mixed+surrogate", rep, nprof)
    from code_tohoku_amd import surrogate
    solve = surrogate.evaluate(d, n=1 << 20)
    log(f"[bench] surrogate on 2^20 stiff cells: {solve}")
    cms = d.copy_probe(1 << 28, 10)
    d.close()
    # the plain FP64 path on the same state
    f = make_domain(p, st, nsplit, False, False, None)
    del st
    log(f"[bench] fp64: warmup {args.warmup}, timed {args.steps} steps")
    el64, dts64, rep64, nprof64 = timed(f, args.steps, args.warmup)
    report_kernels("fp64", rep64, nprof64)
    f.close()

    ms_step = 1e3 * el / args.steps
    value = n ** 3 * args.steps / el
    value64 = n ** 3 * args.steps / el64
    log(f"[bench] mixed+surrogate {ms_step:.3f} ms/step {value:.4e} zone-updates/s; "
        f"fp64 {1e3 * el64 / args.steps:.3f} ms/step {value64:.4e}; ratio {value / value64:.3f}")

    def roof(rep, traffic):
        name, (cnt, ms, by) = max(rep.items(), key=lambda kv: kv[1][1])
        achieved = by / (ms * 1e-3) / 1e9
        r = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "avg_launch_us": round(1e3 * ms / cnt, 2),
             "bytes_per_launch": by / cnt}
        if traffic is not None:
            per, info = traffic
            if per is not None and name in per:
                t = per[name]
                r["traffic"] = t["bytes"]
                r["traffic_GBs"] = round(t["bytes"] / (ms / cnt * 1e-3) / 1e9, 1)
                r["traffic_detail"] = {"fetch_bytes": t["fetch_bytes"], "write_bytes": t["write_bytes"], **info}
            elif per is None:
                r["traffic_note"] = info
        return r

    def kernel_table(rep, traffic, steps, top=8):
        """Per-kernel achieved bandwidth (algorithmic bytes) and, where the
        counter runs measured it, HBM traffic per launch."""
        per = traffic[0] if traffic is not None else None
        rows = []
        for name, (cnt, ms, by) in sorted(rep.items(), key=lambda kv: -kv[1][1])[:top]:
            row = {"kernel": name, "launches_per_step": cnt / steps, "avg_launch_us": round(1e3 * ms / cnt, 2),
                   "ms_per_step": round(ms / steps, 4), "alg_bytes_per_launch": by / cnt,
                   "achieved_GBs": round(by / (ms * 1e-3) / 1e9, 1),
                   "frac": round(by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
            if per is not None and name in per:
                row["traffic_bytes_per_launch"] = per[name]["bytes"]
                row["traffic_GBs"] = round(per[name]["bytes"] / (ms / cnt * 1e-3) / 1e9, 1)
            rows.append(row)
        return rows

    step_gbs = ALG_BYTES_PER_ZONE * n ** 3 / (ms_step * 1e-3) / 1e9
    out = {"metric": "MHD zone-updates/s at 256^3 (mixed+surrogate vs FP64) + % HBM roofline",
           "value": value, "unit": "zone-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+f64 (mixed; state f64, surrogate bf16/f32)", "data": "synthetic",
           "config": {"workload": f"converging_flow_{n}^3_ki_cooling", "grid": [n, n, n],
                      "block_split": list(nsplit), "precision": "mixed", "cooling": "surrogate",
                      "surrogate_weights": wsrc},
           "roofline": roof(rep, pmc),
           "kernels": kernel_table(rep, pmc, nprof),
           "speedup_vs_fp64": value / value64,
           "fp64": {"value": value64, "ms_per_step": 1e3 * el64 / args.steps, "roofline": roof(rep64, None),
                    "kernels": kernel_table(rep64, None, nprof64),
                    "dt_mean": float(sum(dts64) / len(dts64))},
           "surrogate": {"backward_euler_cells": nstiff, "imc_fallbacks": nfall,
                         "fallback_fraction": (nfall / nstiff) if nstiff else None,
                         "note": "cells of the run's steps (warm-up, timed, profiled) that took SOURCE's backward-Euler branch",
                         "stiff_solve": solve},
           "step_roofline": {"alg_bytes_per_zone": ALG_BYTES_PER_ZONE, "achieved": round(step_gbs, 1),
                             "unit": "GB/s", "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
           "dt_mean": float(sum(dts) / len(dts)),
           "copy_probe": {"bytes": 16 * (1 << 28), "ms": cms, "GBs": round(16 * (1 << 28) / (cms * 1e-3) / 1e9, 1)}}
    if not args.no_cpu:
        log(f"[bench] CPU baseline: oracle {args.cpu_steps} steps at {args.cpu_n}^3")
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
