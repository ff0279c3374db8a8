"""Accuracy of the reduced-precision storage forms against the FP64 run.

    python -m imexlb.accuracy sphere512x256x256 [--steps 5000]
    python -m imexlb.accuracy cyl8192x2048 [--steps 50000]

Every form starts from the same state (InitUP with a seeded perturbation)
and runs the same number of steps; drag and lift are logged every `interv`
steps (the deck's Monitor interval).  Reported per form: MLUPS, relative L2
error of the final velocity field against FP64, relative error of the final
drag and of the mean drag over the second half of the run, the Strouhal
number St = f D / U of the lift signal (its dominant frequency) when the run
is long enough to show shedding, and the mass check: the mean density
deviation per node (sum of f over fluid nodes, FP64, imexlb_mass) and its
difference from the FP64 run.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import _lib
from .inputs import deck as get_deck
from .inputs import derive, nodes
from .lattice import setup

FORMS = {"f64": _lib.F64, "f32": _lib.F32, "f16": _lib.F16_F32, "bf16": _lib.BF16_F32}


def run_form(deck, prec, steps, every, r_vec):
    lat, d = setup(deck, r_vec, device=0, precision=prec)
    forces = []
    t_run = 0.0
    done = 0
    while done < steps:
        n = min(every, steps - done)
        t0 = time.perf_counter()
        lat.advance(n)
        lat.synchronize()
        t_run += time.perf_counter() - t0
        done += n
        forces.append(lat.forces())
    _, u = lat.download_fields()
    mass = lat.mass()
    lat.close()
    return np.array(forces), u, t_run, mass


def strouhal(lift, every, D, U):
    """Dominant frequency of the lift signal (second half), as f D / U."""
    x = np.asarray(lift, dtype=float)
    x = x[len(x) // 2:]
    if len(x) < 16:
        return None
    x = x - x.mean()
    spec = np.abs(np.fft.rfft(x * np.hanning(len(x))))
    freqs = np.fft.rfftfreq(len(x), d=every)  # cycles per lattice step
    k = int(np.argmax(spec[1:])) + 1
    if spec[k] <= 1e-12 * (np.abs(x).max() + 1e-300):
        return None
    return float(freqs[k] * D / U)


def study(name, steps, forms, seed=2021):
    deck = get_deck(name)
    d = derive(deck)
    r_vec = np.random.default_rng(seed).random(100)
    every = max(d.interv, 1)
    D = 2.0 * d.radius
    out = {"deck": name, "steps": steps, "log_every": every, "forms": {}}
    ref = None
    for form in forms:
        print(f"[accuracy] {name} {form}: {steps} steps", file=sys.stderr, flush=True)
        F, u, t, mass = run_form(deck, FORMS[form], steps, every, r_vec)
        half = F[len(F) // 2:]
        res = {"MLUPS": round(nodes(deck) * steps / t / 1e6, 1), "drag_final": float(F[-1, 0]),
               "density_deviation_per_node": mass / nodes(deck),
               "drag_mean_2nd_half": float(half[:, 0].mean()), "lift_final": float(F[-1, 1]),
               "strouhal": strouhal(F[:, 1], every, D, d.uu)}
        if form == "f64":
            ref = (F, u, mass)
        elif ref is not None:
            F64, u64, m64 = ref
            res["density_deviation_per_node_diff"] = float(abs(mass - m64) / nodes(deck))
            res["u_rel_L2_err"] = float(np.linalg.norm(u - u64) / np.linalg.norm(u64))
            res["drag_final_rel_err"] = float(abs(F[-1, 0] - F64[-1, 0]) / abs(F64[-1, 0]))
            m64 = F64[len(F64) // 2:, 0].mean()
            res["drag_mean_rel_err"] = float(abs(half[:, 0].mean() - m64) / abs(m64))
        out["forms"][form] = res
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("deck")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--forms", default="f64,f32,f16,bf16")
    a = ap.parse_args(argv)
    steps = a.steps
    if steps is None:
        steps = derive(get_deck(a.deck)).max_step + 1
    print(json.dumps(study(a.deck, steps, a.forms.split(","))), flush=True)


if __name__ == "__main__":
    main()
