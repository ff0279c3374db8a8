"""Batched ensemble of independent D2Q9 cylinder cases (C ABI imexlb_batch_*).

    python -m imexlb.ensemble [--cases 96] [--steps 20000] [--seed 7]

generates the cases (Nx = 4*U{64..256}, Ny = Nx/4, radius = Ny/10, Re
log-uniform in [40, 400], inlet velocity 0.05, body at x = 0.2 Nx), tunes the
step-kernel shape, and reports the aggregate MLUPS of the batch against the
same cases run one after another as single lattices.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time

import numpy as np

from . import _lib
from ._lib import check, lib
from .inputs import Deck, derive

_DP = C.POINTER(C.c_double)


class Batch:
    """Handle on an imexlb_batch: cases given as Decks (2D)."""

    def __init__(self, decks, device=0, precision=_lib.F64, shapes=None):
        """shapes: None (circles) or one (aspect, angle_deg) per case."""
        self.decks = list(decks)
        self.ncase = len(self.decks)
        nxny = np.zeros(2 * self.ncase, dtype=np.int32)
        params = np.zeros(5 * self.ncase)
        for i, dk in enumerate(self.decks):
            if dk.dim != 2:
                raise ValueError("batched cases are D2Q9")
            d = derive(dk)
            nxny[2 * i:2 * i + 2] = dk.n[:2]
            params[5 * i:5 * i + 5] = (d.pos, d.radius, dk.ome, dk.rho0, d.uu)
        h = C.c_void_p()
        if shapes is None:
            check(lib().imexlb_batch_create(C.byref(h), self.ncase, nxny.ctypes.data_as(C.POINTER(C.c_int32)),
                                            params.ctypes.data_as(_DP), device, precision), "imexlb_batch_create")
        else:
            sh = np.ascontiguousarray(np.asarray(shapes, dtype=np.float64).reshape(self.ncase, 2))
            check(lib().imexlb_batch_create_shapes(C.byref(h), self.ncase,
                                                   nxny.ctypes.data_as(C.POINTER(C.c_int32)),
                                                   params.ctypes.data_as(_DP), sh.ctypes.data_as(_DP), device,
                                                   precision), "imexlb_batch_create_shapes")
        self._h = h
        self.nodes = [(dk.n[0] + 1) * (dk.n[1] + 1) for dk in self.decks]

    def close(self):
        if self._h:
            lib().imexlb_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self, r_vec):
        r = np.ascontiguousarray(r_vec, dtype=np.float64)
        check(lib().imexlb_batch_init(self._h, r.ctypes.data_as(_DP)), "imexlb_batch_init")

    def init_fields(self, ps, us):
        """Warm start every case from its own p, u (interior layout, as
        download_fields returns them)."""
        p = np.ascontiguousarray(np.concatenate([np.ravel(x) for x in ps]), dtype=np.float64)
        u = np.ascontiguousarray(np.concatenate([np.ravel(x) for x in us]), dtype=np.float64)
        if p.size != sum(self.nodes) or u.size != 2 * sum(self.nodes):
            raise ValueError("fields do not match the cases")
        check(lib().imexlb_batch_init_fields(self._h, p.ctypes.data_as(_DP), u.ctypes.data_as(_DP)),
              "imexlb_batch_init_fields")

    def init_fields_f32(self, fields, p_scale, u_scale, r_vec=None):
        """Warm start every case from a device FP32 tensor (ncase, 3, H, W)
        of p, u_x, u_y (see Lattice.init_fields_f32)."""
        if fields.dim() != 4 or fields.shape[0] != self.ncase or fields.shape[1] != 3 or not fields.is_contiguous():
            raise ValueError("fields: contiguous (ncase, 3, H, W)")
        rv = None if r_vec is None else np.ascontiguousarray(r_vec, dtype=np.float64)
        check(lib().imexlb_batch_init_fields_f32(self._h, C.c_void_p(fields.data_ptr()), fields.shape[2],
                                                 fields.shape[3], p_scale, u_scale,
                                                 None if rv is None else rv.ctypes.data_as(_DP)),
              "imexlb_batch_init_fields_f32")

    def body_mask_f32(self, out):
        """Body masks of every case into channel 0 of the device FP32 tensor
        out (ncase, C, H, W) (1 = body node)."""
        if out.dim() != 4 or out.shape[0] != self.ncase or not out.is_contiguous():
            raise ValueError("out: contiguous (ncase, C, H, W)")
        check(lib().imexlb_batch_body_mask_f32(self._h, C.c_void_p(out.data_ptr()), out.shape[2], out.shape[3],
                                               out[0].numel()), "imexlb_batch_body_mask_f32")

    def advance(self, n):
        check(lib().imexlb_batch_advance(self._h, int(n)), "imexlb_batch_advance")

    def set_config(self, block, cells_per_thread, persistent):
        check(lib().imexlb_batch_set_config(self._h, block, cells_per_thread, 1 if persistent else 0),
              "imexlb_batch_set_config")

    def tune(self, trial_steps=20):
        best = (C.c_int32 * 3)()
        ms = C.c_double()
        check(lib().imexlb_batch_tune(self._h, trial_steps, best, C.byref(ms)), "imexlb_batch_tune")
        return {"block": best[0], "cells_per_thread": best[1], "persistent": best[2], "ms_per_step": ms.value}

    def forces(self):
        F = np.zeros(3 * self.ncase)
        check(lib().imexlb_batch_forces(self._h, F.ctypes.data_as(_DP)), "imexlb_batch_forces")
        return F.reshape(self.ncase, 3)[:, :2]

    def monitor(self):
        """Monitor of every case: (um, diverged) with um[i] the case's
        maximum |(u_x, u_y)| (NaN for a case with a non-finite velocity)
        and diverged the indices of those cases."""
        um = np.zeros(self.ncase)
        rc = lib().imexlb_batch_monitor(self._h, um.ctypes.data_as(_DP))
        if rc not in (_lib.OK, _lib.E_NAN):
            check(rc, "imexlb_batch_monitor")
        diverged = np.nonzero(np.isnan(um))[0]
        assert (rc == _lib.E_NAN) == (diverged.size > 0)
        return um, diverged

    def download_fields(self, i):
        n = self.nodes[i]
        p = np.zeros(n)
        u = np.zeros(2 * n)
        check(lib().imexlb_batch_download_fields(self._h, i, p.ctypes.data_as(_DP), u.ctypes.data_as(_DP)),
              "imexlb_batch_download_fields")
        return p, u

    def synchronize(self):
        check(lib().imexlb_batch_synchronize(self._h), "imexlb_batch_synchronize")


def cylinder_cases(ncase, seed=7, uu=0.05, xpos=0.2):
    """Seeded cylinder decks: Nx = 4*U{64..256}, Ny = Nx/4, radius = Ny/10,
    Re log-uniform in [40, 400]; the relaxation follows from Re."""
    rng = np.random.default_rng(seed)
    decks = []
    for _ in range(ncase):
        nx = 4 * int(rng.integers(64, 257))
        ny = nx // 4
        radius = ny / 10.0
        re = float(np.exp(rng.uniform(np.log(40.0), np.log(400.0))))
        D = 2.0 * radius
        nu = uu * D / re
        ome = 1.0 / (3.0 * nu + 0.5)
        # DataInput derives uu = Re nu / charlength with charlength = L nx, dt = 1
        Lc = D / nx
        decks.append(Deck(2, (nx, ny, 0), Lc, xpos, re, ome, 1.0, 1.0, 1.0, 0.01))
    return decks


def main(argv=None):
    from .lattice import setup
    ap = argparse.ArgumentParser()
    ap.add_argument("--cases", type=int, default=96)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--seq-steps", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--precision", default="f16")
    ap.add_argument("--config", default=None,
                    help="block,cells_per_thread,persistent instead of the autotuner (e.g. 128,2,0)")
    ap.add_argument("--no-sequential", action="store_true", help="skip the 96 single-lattice runs")
    a = ap.parse_args(argv)
    prec = {"f64": _lib.F64, "f32": _lib.F32, "f16": _lib.F16_F32}[a.precision]
    decks = cylinder_cases(a.cases, a.seed)
    nodes = sum((d.n[0] + 1) * (d.n[1] + 1) for d in decks)
    r_vec = np.random.default_rng(a.seed).random(100)
    print(f"[ensemble] {a.cases} cases, {nodes} nodes, {a.precision}", file=sys.stderr, flush=True)
    b = Batch(decks, precision=prec)
    b.init(r_vec)
    if a.config:
        blk, cpt, per = (int(v) for v in a.config.split(","))
        b.set_config(blk, cpt, bool(per))
        best = {"block": blk, "cells_per_thread": cpt, "persistent": per, "ms_per_step": None}
    else:
        best = b.tune(20)
    print(f"[ensemble] config {best}", file=sys.stderr, flush=True)
    b.init(r_vec)
    b.synchronize()
    t0 = time.perf_counter()
    b.advance(a.steps)
    b.synchronize()
    t_batch = time.perf_counter() - t0
    F = b.forces()
    _, diverged = b.monitor()
    b.close()
    out = {"cases": a.cases, "nodes": nodes, "precision": a.precision, "tuned": best,
           "batch_steps": a.steps, "batch_s": round(t_batch, 3),
           "batch_MLUPS": round(nodes * a.steps / t_batch / 1e6, 1),
           "nonfinite_cases": int(diverged.size), "diverged": diverged.tolist(),
           "drag_range": [float(np.nanmin(F[:, 0])), float(np.nanmax(F[:, 0]))]}
    if a.no_sequential:
        print(json.dumps(out), flush=True)
        return
    # the same cases as single lattices, one after another
    lats = [setup(d, r_vec, device=0, precision=prec)[0] for d in decks]
    for lat in lats:
        lat.advance(2)
        lat.synchronize()
    t0 = time.perf_counter()
    for lat in lats:
        lat.advance(a.seq_steps)
    for lat in lats:
        lat.synchronize()
    t_seq = time.perf_counter() - t0
    for lat in lats:
        lat.close()
    out.update({"sequential_steps": a.seq_steps, "sequential_s": round(t_seq, 3),
                "sequential_MLUPS": round(nodes * a.seq_steps / t_seq / 1e6, 1)})
    out["speedup"] = round(out["batch_MLUPS"] / out["sequential_MLUPS"], 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
