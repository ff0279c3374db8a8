"""Benchmark: lattice updates per second of the MI355X time step.

Workload (one GPU): the D3Q27 sphere, 512x256x256 (513x257x257 = 33,883,137
nodes), Re 200, in the reference's own boundary semantics, FP64 storage and
arithmetic (results equal to the reference's).  A step is one trip of the
reference's loop body (Collision, BoundaryCondition, Propagation,
PostProcessing) over the whole lattice; inputs are resident in HBM before
the timed region.  The CPU baseline is the oracle (the C restatement of the
reference) run on the host cores over a few steps of the same lattice.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-variants]
                    [--warm-start full|off] [--warm-start-args "..."]
    python bench.py --copy-calibration

The metric's second half, wall time to drag within 1 % with the learned
warm start against the plain (cold) start, comes from the warm-start leg
(``warm_start`` in the JSON line): the whole pipeline of
imexlb.surrogate on the 512 x 128 channel -- data from the batched solver,
training of the 4-level network by the library's kernels, BF16 inference,
then cold, warm and mean-field starts of the held-out cases -- with its
wall-time ratios (inference and start included) and every phase's time.

``roofline.traffic`` (HBM bytes per launch of the step kernel) is read from
the counter summary in profiles/: FETCH_SIZE and WRITE_SIZE of the step
kernel, each scaled by what the counter reads for a plain 256 MiB device copy
(``--copy-calibration`` runs that copy).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "imexlb-1.0_amd"))

import numpy as np  # noqa: E402

import imexlb as I  # noqa: E402

HBM_PEAK_GBS = 8000.0
DECK = "sphere512x256x256"
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_step_sphere512x256x256_f64_hbm_traffic.md")
CAL_BYTES = 256 * 1024 * 1024


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def gpu_leg(deck, r_vec, precision, steps, warmup):
    lat, d = I.setup(deck, r_vec, device=0, precision=precision)
    info = lat.info()
    lat.advance(warmup)
    lat.synchronize()
    lat.profile(True)
    t0 = time.perf_counter()
    lat.advance(steps)
    lat.synchronize()
    t1 = time.perf_counter()
    n, ms = lat.profile_read()
    lat.profile(False)
    um = lat.monitor()
    lat.close()
    return {"seconds": t1 - t0, "fused_launches": n[0], "fused_ms": ms[0], "fill_launches": n[1],
            "fill_ms": ms[1], "other_ms": ms[2], "info": info, "umax": um}


def cpu_leg(deck, r_vec, steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    d = I.derive(deck)
    o = O.Oracle(deck.dim, *deck.n)
    o.set_params(deck.ome, deck.rho0, d.uu)
    o.set_geometry(d.pos, d.radius)
    o.init_up(r_vec)
    o.init_pdf()
    o.run(0, 0, 0)  # first step touches every page of the arrays: untimed
    t0 = time.perf_counter()
    o.run(1, steps, 0)
    t1 = time.perf_counter()
    o.close()
    return t1 - t0


def warm_start_leg(extra):
    """The learned warm-start pipeline (imexlb.surrogate) at the sizes its
    options give; the wall-time ratios warm / cold to drag within 1 % of the
    converged value (median over the held-out cases that converge in both
    and match the cold run's drag and lift), with the network's inference
    and the starts included."""
    import imexlb.surrogate as S
    try:
        a = S.build_parser().parse_args(extra)
        r = S.pipeline(a, log=log)
    except Exception as exc:  # the MLUPS legs stand; the line records the failure
        return {"error": f"{type(exc).__name__}: {exc}"}
    return {
        "wall_ratio": r["median_wall_ratio_warm_over_cold"],
        "wall_ratio_single_case": r["median_wall_ratio_single_case"],
        "steps_ratio": r["median_ratio_warm_over_cold"],
        "steps_ratio_mean_field": r["median_ratio_meanfield_over_cold"],
        "target": "wall_ratio <= 1/3",
        "frac_cases_wall_within_one_third": r["frac_wall_within_one_third"],
        "median_steps": {"cold": r["median_steps_cold"], "warm": r["median_steps_warm"]},
        "cases": {"train": r["train_cases"], "test": r["test_cases"], "scored": r["cases_scored"],
                  "drag_match": r["drag_match"], "lift_match": r["lift_match"]},
        "heldout_rel_rms": {"network_bf16": r["heldout_rel_rms"], "mean_field": r["mean_field_rel_rms"]},
        "re_range": r["re_range"],
        "re_note": "Re capped at 300 (the plan names 40-400): BGK diverges on this grid above u D / nu of about 16",
        "config": {"workload": "D2Q9 channel 512x128, cylinders and ellipses (axis ratio 0.5-2, 0-90 deg), "
                               "FP16 storage", "network": f"4-level encoder-decoder {r['chans']} channels, "
                               f"{r['params']} parameters, FP32 training, BF16 inference",
                   "epochs": r["epochs"], "eval_steps": r["eval_steps"], "seed": r["seed"]},
        "seconds": {"data": r["data_s"], "train": r["train_s"], "pipeline": r["pipeline_s"],
                    "start_batch_cold": r["start_seconds"]["batch_cold"],
                    "start_batch_warm": r["start_seconds"]["batch_warm"]},
    }


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo) and its logical CPU count."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{name} ({os.cpu_count()} logical CPUs visible)"


def traffic_per_launch():
    """HBM bytes per step-kernel launch from the committed counter summary
    (its ``traffic_bytes_per_launch = N`` line), or None."""
    try:
        with open(TRAFFIC_SUMMARY) as fh:
            for line in fh:
                m = line.strip().strip("`").split("=")
                if len(m) == 2 and m[0].strip() == "traffic_bytes_per_launch":
                    return int(m[1].strip().replace(",", ""))
    except OSError:
        pass
    return None


def copy_calibration():
    """A plain device-to-device copy of 256 MiB, five times: the kernel whose
    FETCH_SIZE / WRITE_SIZE readings scale the step kernel's."""
    import torch
    a = torch.ones(CAL_BYTES // 8, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": CAL_BYTES, "copies": 5}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=16)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--warm-start", choices=("full", "off"), default="full",
                    help="run the learned warm-start pipeline (full: its default sizes)")
    ap.add_argument("--warm-start-args", default="",
                    help="options passed to imexlb.surrogate (e.g. '--train 1000 --epochs 20')")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the lattice is not decomposed across GPUs).", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return

    deck = I.deck(DECK)
    nodes = I.nodes(deck)
    Q = 27
    r_vec = np.random.default_rng(2021).random(100)

    log(f"[bench] {DECK} FP64: setup, {args.warmup} warm-up + {args.steps} timed steps")
    g = gpu_leg(deck, r_vec, I.F64, args.steps, args.warmup)
    sec = g["seconds"]
    mlups = nodes * args.steps / sec / 1e6
    avg_ms = g["fused_ms"] / max(g["fused_launches"], 1)
    bytes_per_launch = nodes * 2 * Q * 8
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    log(f"[bench] FP64 {mlups:.0f} MLUPS, fused kernel {avg_ms:.3f} ms -> {achieved:.0f} GB/s")

    variants = {}
    if not args.no_variants:
        for name, prec, s in (("f32", I.F32, 4), ("f16_f32", I.F16_F32, 2), ("bf16_f32", I.BF16_F32, 2),
                              ("bf16sr_f32", I.BF16SR_F32, 2)):
            log(f"[bench] {DECK} {name}: {args.steps} timed steps")
            try:
                v = gpu_leg(deck, r_vec, prec, args.steps, args.warmup)
            except Exception as exc:  # the FP64 line stands; the form records its failure
                variants[name] = {"error": f"{type(exc).__name__}: {exc}"}
                continue
            vm = nodes * args.steps / v["seconds"] / 1e6
            va = v["fused_ms"] / max(v["fused_launches"], 1)
            variants[name] = {"MLUPS": round(vm, 1), "fused_kernel_ms": round(va, 4),
                              "achieved_GBps": round(nodes * 2 * Q * s / (va * 1e-3) / 1e9, 1),
                              "speedup_vs_f64": round(vm / mlups, 3)}

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"[bench] CPU oracle: {args.cpu_steps} steps of the same lattice on {threads} threads")
        try:
            cs = cpu_leg(deck, r_vec, args.cpu_steps)
            cpu = {"value": round(nodes * args.cpu_steps / cs / 1e6, 2), "unit": "MLUPS", "cores": threads,
                   "kind": "port",
                   "sample": f"{args.cpu_steps} full steps of the 513x257x257 sphere lattice (oracle, OpenMP)",
                   "cpu_model": cpu_model()}
        except Exception as exc:  # the GPU line stands; the baseline records its failure
            cpu = {"error": f"{type(exc).__name__}: {exc}", "kind": "port"}

    warm = None
    if args.warm_start != "off":
        log(f"[bench] learned warm start: imexlb.surrogate {args.warm_start_args}")
        warm = warm_start_leg(args.warm_start_args.split())

    out = {
        "metric": "MLUPS (D3Q27 sphere, FP64) & % HBM roofline",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: InitUP perturbation from a seeded generator; geometry and boundaries as the reference",
        "config": {"workload": "D3Q27 BGK sphere 512x256x256 (513x257x257 nodes), radius 20, Re 200, FP64",
                   "nodes": nodes, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_per_launch()},
        "cpu_baseline": cpu,
        "kernel": {"name": "k_step<3,double,double,17> (fused step writing its boundary copies)",
                   "avg_ms": round(avg_ms, 4),
                   "launches": g["fused_launches"], "algorithmic_bytes_per_launch": bytes_per_launch,
                   "fill_avg_ms": round(g["fill_ms"] / max(g["fill_launches"], 1), 4)},
        "variants": variants,
        "warm_start": warm,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
