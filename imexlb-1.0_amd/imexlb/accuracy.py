"""Accuracy of the reduced-precision storage forms against the FP64 run.

    python -m imexlb.accuracy sphere512x256x256 [--steps 5000]
    python -m imexlb.accuracy cyl8192x2048 [--steps 50000]

Every form starts from the same state (InitUP with a seeded perturbation)
and runs the same number of steps; drag and lift are logged every 10 steps
(--log-every; at most the deck's Monitor interval).  Reported per form:
MLUPS, relative L2 error of the final velocity field against FP64, relative
error of the final
drag and of the mean drag over the second half of the run, the Strouhal
number St = D / (U T) of the lift signal with its standard error (T from the
interpolated zero crossings of the second half, see strouhal()) when that
half holds at least three shedding periods, and the mass check: the mean density
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

FORMS = {"f64": _lib.F64, "f32": _lib.F32, "f16": _lib.F16_F32, "bf16": _lib.BF16_F32, "bf16sr": _lib.BF16SR_F32}


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


def upward_crossings(x, every, hysteresis=0.25):
    """Times (in lattice steps, sample i at i*every) where x - mean(x) passes
    upward through zero, by linear interpolation between the two samples
    that bracket it.  A crossing counts once the signal has come from below
    -h and goes on to reach +h (h = hysteresis times its standard
    deviation), so noise around zero does not count as extra cycles; its
    time is the last upward zero crossing on that way up."""
    x = np.asarray(x, dtype=float)
    x = x - x.mean()
    h = hysteresis * x.std()
    out, low, last = [], False, None
    for i in range(len(x) - 1):
        if x[i] < -h:
            low, last = True, None
        if x[i] < 0.0 <= x[i + 1]:
            last = i + (-x[i]) / (x[i + 1] - x[i])
        if low and last is not None and x[i + 1] > h:
            out.append(last * every)
            low, last = False, None
    return np.array(out)


def strouhal(lift, every, D, U, min_periods=3):
    """Shedding frequency of the lift signal's second half as St = D / (U T).

    T is the mean of the periods between consecutive upward zero crossings
    (crossings interpolated between samples, upward_crossings), so a signal
    sampled every `every` steps resolves periods far finer than the sample
    spacing.  Returns (St, St_err, periods) with St_err the standard error
    of that mean carried to St (the scatter of the individual periods over
    the square root of their number: cycle-to-cycle variation of the wake,
    and for a strictly periodic signal the interpolation error only), or
    (None, None, n) when the second half holds fewer than min_periods whole
    periods."""
    x = np.asarray(lift, dtype=float)
    x = x[len(x) // 2:]
    if len(x) < 8 or not np.isfinite(x).all():
        return None, None, 0
    periods = np.diff(upward_crossings(x, every))
    n = len(periods)
    if n < min_periods:
        return None, None, n
    T = periods.mean()
    sT = periods.std(ddof=1) / np.sqrt(n)
    st = D / (U * T)
    return float(st), float(st * sT / T), int(n)


def study(name, steps, forms, seed=2021, log_every=10):
    deck = get_deck(name)
    d = derive(deck)
    r_vec = np.random.default_rng(seed).random(100)
    every = max(min(d.interv, log_every), 1)
    D = 2.0 * d.radius
    out = {"deck": name, "steps": steps, "log_every": every, "forms": {}}
    ref = None
    for form in forms:
        print(f"[accuracy] {name} {form}: {steps} steps", file=sys.stderr, flush=True)
        F, u, t, mass = run_form(deck, FORMS[form], steps, every, r_vec)
        half = F[len(F) // 2:]
        res = {"MLUPS": round(nodes(deck) * steps / t / 1e6, 1), "drag_final": float(F[-1, 0]),
               "density_deviation_per_node": mass / nodes(deck),
               "drag_mean_2nd_half": float(half[:, 0].mean()), "lift_final": float(F[-1, 1])}
        st, st_err, periods = strouhal(F[:, 1], every, D, d.uu)
        res.update({"strouhal": st, "strouhal_err": st_err, "shedding_periods": periods})
        if form == "f64":
            ref = (F, u, mass)
        elif ref is not None:
            F64, u64, m64 = ref
            res["density_deviation_per_node_diff"] = float(abs(mass - m64) / nodes(deck))
            res["u_rel_L2_err"] = float(np.linalg.norm(u - u64) / np.linalg.norm(u64))
            res["drag_final_rel_err"] = float(abs(F[-1, 0] - F64[-1, 0]) / abs(F64[-1, 0]))
            m64 = F64[len(F64) // 2:, 0].mean()
            res["drag_mean_rel_err"] = float(abs(half[:, 0].mean() - m64) / abs(m64))
            st64 = out["forms"]["f64"]["strouhal"]
            if st is not None and st64 is not None:
                res["strouhal_diff"] = st - st64
                res["strouhal_diff_err"] = float(np.hypot(st_err, out["forms"]["f64"]["strouhal_err"]))
        out["forms"][form] = res
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("deck")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--forms", default="f64,f32,f16,bf16")
    ap.add_argument("--log-every", type=int, default=10, help="steps between drag/lift samples (at most interv)")
    a = ap.parse_args(argv)
    steps = a.steps
    if steps is None:
        steps = derive(get_deck(a.deck)).max_step + 1
    print(json.dumps(study(a.deck, steps, a.forms.split(","), log_every=a.log_every)), flush=True)


if __name__ == "__main__":
    main()
