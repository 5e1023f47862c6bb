#!/usr/bin/env python3
"""Benchmark: MHD zone-updates/s of the full timestep on one MI355X.

Workload (config 2): 256^3 thermally bistable converging flow in FP64 --
20 pc box, WNM n = 1 cm^-3 at 6000 K, +-10 km/s inflow along x, seeded
Kolmogorov density perturbation (0.1 rms), B = 5 muG at 45 deg, KI cooling
and conduction on.  A step is one EVOLVE iteration of the reference
(3DMHD-CfCA/MCF3DSGincfca3.f90:562-657): the dt reduction, SOURCE(dt/2),
three sweeps, CT, SOURCE(dt/2) and DISSIP; the state stays resident in HBM.

Prints one JSON line.  Per-kernel HIP-event timings over the timed region
give the roofline of the dominant kernel; the CPU restatement (oracle/) is
timed on a bounded sample of the same problem as the CPU baseline.
usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E
ALG_BYTES_PER_ZONE = 920.0      # whole FP64 step, 5 active species (compulsory traffic)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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
            "sample": f"CPU restatement (oracle/, C, 1 thread), {len(dts)} steps of the same converging-flow "
                      f"problem at {n}^3, {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-n", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build runs the whole domain on a single MI355X "
              f"and has no multi-GPU path (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    from code_tohoku_amd import ics, mcf

    n = args.n
    nsplit = (max(1, n // 64),) * 3
    log(f"[bench] building config-2 state {n}^3 ...")
    p = ics.converging_flow_input(n)
    st = ics.converging_flow_state(p)
    prec = mcf.PREC_MIXED if args.precision == "mixed" else mcf.PREC_FP64
    g = mcf.Domain(p, nsplit=nsplit, precision=prec)
    g.upload(st)
    del st
    log(f"[bench] warmup {args.warmup} steps")
    g.run(args.warmup, tsave=1e30)
    g.sync()
    g.profile(True)
    log(f"[bench] timed {args.steps} steps")
    t0 = time.perf_counter()
    dts = g.run(args.steps, tsave=1e30)
    g.sync()
    el = time.perf_counter() - t0
    rep = g.profile_report()
    g.profile(False)
    assert len(dts) == args.steps
    ms_step = 1e3 * el / args.steps
    value = n ** 3 * args.steps / el
    tot_ms = sum(v[1] for v in rep.values())
    log(f"[bench] {ms_step:.3f} ms/step, {value:.4e} zone-updates/s; kernel time {tot_ms / args.steps:.3f} ms/step")
    for k, (cnt, ms, by) in sorted(rep.items(), key=lambda kv: -kv[1][1]):
        log(f"  {k:18s} launches {cnt:6d}  avg {1e3 * ms / cnt:9.2f} us  total/step {ms / args.steps:8.3f} ms"
            f"  {by / (ms * 1e-3) / 1e9:8.1f} GB/s")
    dom = max(rep.items(), key=lambda kv: kv[1][1])
    name, (cnt, ms, by) = dom
    achieved = by / (ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "avg_launch_us": round(1e3 * ms / cnt, 2),
            "bytes_per_launch": by / cnt}
    step_gbs = ALG_BYTES_PER_ZONE * n ** 3 / (ms_step * 1e-3) / 1e9
    out = {"metric": "MHD zone-updates/s at 256^3 (mixed+surrogate vs FP64) + % HBM roofline",
           "value": value, "unit": "zone-updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"converging_flow_{n}^3_fp64_ki_cooling", "grid": [n, n, n],
                      "block_split": list(nsplit), "precision": "fp64", "path": "plain"},
           "roofline": roof,
           "step_roofline": {"alg_bytes_per_zone": ALG_BYTES_PER_ZONE, "achieved": round(step_gbs, 1),
                             "unit": "GB/s", "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
           "dt_mean": float(sum(dts) / len(dts))}
    nw = 1 << 28
    cms = g.copy_probe(nw, 10)
    out["copy_probe"] = {"bytes": 16 * nw, "ms": cms, "GBs": round(16 * nw / (cms * 1e-3) / 1e9, 1)}
    log(f"[bench] copy kernel: {16 * nw / 2**30:.1f} GiB moved in {cms:.3f} ms = {out['copy_probe']['GBs']} GB/s")
    g.close()
    if not args.no_cpu:
        log(f"[bench] CPU baseline: oracle {args.cpu_steps} steps at {args.cpu_n}^3")
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
