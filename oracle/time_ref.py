#!/usr/bin/env python3
"""Time the reference MPI build on the converging-flow problem (CPU baseline).

TEST INFRASTRUCTURE, run in this container only (the reference never
travels): the KI-cooling build of oracle/build_ref.sh, 8 MPI ranks (block
split 2x2x2), the seeded converging-flow state injected through the restart
files, and the per-step time taken as (wall(13 steps) - wall(3 steps)) / 10
so that start-up, restart reading and the final SAVEU cancel.

usage: python oracle/time_ref.py [N ...]      (default 128 256)
Prints one JSON object per grid size.
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)
from code_tohoku_amd import ics  # noqa: E402
import ref_runner  # noqa: E402


def time_reference(n: int, nranks: int = 8, short: int = 3, long: int = 13) -> dict:
    p = ics.converging_flow_input(n)
    st = ics.converging_flow_state(p)
    walls = {}
    for k in (short, long):
        t0 = time.perf_counter()
        r = ref_runner.run_reference(p, st, k, nranks=nranks, tag=f"timing_{n}_{k}")
        walls[k] = time.perf_counter() - t0
        assert len(r["dts"]) == k, (n, k, len(r["dts"]))
    per_step = (walls[long] - walls[short]) / (long - short)
    return {"grid": n, "ranks": nranks, "cores": os.cpu_count(), "wall_s": walls, "s_per_step": per_step,
            "zone_updates_per_s": n ** 3 / per_step,
            "method": f"(wall {long} steps - wall {short} steps) / {long - short}"}


if __name__ == "__main__":
    sizes = [int(a) for a in sys.argv[1:]] or [128, 256]
    for n in sizes:
        print(json.dumps(time_reference(n)), flush=True)
