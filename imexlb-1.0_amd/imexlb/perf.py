"""Throughput of the time step on every workload and storage form.

    python -m imexlb.perf [--decks cyl400x100,sphere512x256x256,cyl8192x2048]
                          [--forms f64,f32,f16_f32,bf16_f32,bf16sr_f32] [--seconds 2.0]
                          [--no-graphs]

For each (workload, form): set up the lattice on device 0, warm up, time
``--timed-steps`` steps launch by launch with HIP events (the average
duration of the fused step kernel), then advance for about ``--seconds`` of
GPU time the ordinary way (graph replay of step pairs) for MLUPS over the
wall clock.  One JSON line per case: MLUPS, the kernel's average duration,
and the algorithmic HBM rate of that kernel,
2*Q*s bytes per node update (s = bytes per stored population), against the
MI355X peak of 8000 GB/s.  Lattices that fit in the 256 MB last-level cache
(the 400x100 cylinder) are launch-bound, not HBM-bound; their GB/s is
reported for completeness.

--no-graphs launches every step on its own (IMEXLB_GRAPHS=0) instead of
replaying captured step pairs: the form used under a profiler's kernel
trace, where every launch then appears as its own dispatch.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import _lib
from .inputs import deck as get_deck, nodes as count_nodes
from .lattice import setup

HBM_PEAK_GBS = 8000.0
FORMS = {"f64": (_lib.F64, 8), "f32": (_lib.F32, 4), "f16_f32": (_lib.F16_F32, 2), "bf16_f32": (_lib.BF16_F32, 2),
         "bf16sr_f32": (_lib.BF16SR_F32, 2)}


def measure(deck_name, form, seconds=2.0, timed_steps=200, seed=2021):
    prec, s = FORMS[form]
    dk = get_deck(deck_name)
    n = count_nodes(dk)
    Q = 9 if dk.dim == 2 else 27
    r_vec = np.random.default_rng(seed).random(100)
    lat, _ = setup(dk, r_vec, device=0, precision=prec)
    lat.advance(10)
    lat.synchronize()
    # per-launch timing of a short run
    lat.profile(True)
    lat.advance(timed_steps)
    launches, ms = lat.profile_read()
    lat.profile(False)
    # size the wall-clock run so it takes about `seconds`
    t0 = time.perf_counter()
    lat.advance(20)
    lat.synchronize()
    per = (time.perf_counter() - t0) / 20
    steps = int(max(20, min(200000, seconds / max(per, 1e-7))))
    steps += steps % 2  # whole fused pairs
    t0 = time.perf_counter()
    lat.advance(steps)
    lat.synchronize()
    wall = time.perf_counter() - t0
    um = lat.monitor()
    lat.close()
    k_ms = ms[0] / max(launches[0], 1)
    gbps = n * 2 * Q * s / (k_ms * 1e-3) / 1e9
    return {"workload": deck_name, "form": form, "nodes": n, "steps": steps,
            "MLUPS": round(n * steps / wall / 1e6, 1),
            "ms_per_step": round(wall / steps * 1e3, 5),
            "timed_steps": timed_steps, "fused_kernel_ms": round(k_ms, 5),
            "fill_kernel_ms": round(ms[1] / max(launches[1], 1), 5),
            "algorithmic_bytes_per_update": 2 * Q * s, "achieved_GBps": round(gbps, 1),
            "frac_of_peak": round(gbps / HBM_PEAK_GBS, 4), "umax": um}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--decks", default="cyl400x100,sphere512x256x256,cyl8192x2048")
    ap.add_argument("--forms", default="f64,f32,f16_f32,bf16_f32,bf16sr_f32")
    ap.add_argument("--seconds", type=float, default=2.0)
    ap.add_argument("--timed-steps", type=int, default=200)
    ap.add_argument("--no-graphs", action="store_true", help="launch every step on its own (IMEXLB_GRAPHS=0)")
    a = ap.parse_args(argv)
    if a.no_graphs:
        os.environ["IMEXLB_GRAPHS"] = "0"  # read by every lattice context at creation
    for name in a.decks.split(","):
        for form in a.forms.split(","):
            print(f"[perf] {name} {form}", file=sys.stderr, flush=True)
            print(json.dumps(measure(name, form, a.seconds, a.timed_steps)), flush=True)


if __name__ == "__main__":
    main()
