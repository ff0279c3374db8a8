"""The benchmark configurations at their own sizes, on the GPU.

* The 96-case ensemble (seeded cylinder decks, Nx = 4 U{64..256},
  Ny = Nx/4, radius Ny/10, Re log-uniform 40-400): the batched kernel
  (k_batch_step, k_batch_fill) against an oracle run of every case, bit for
  bit in FP64 (each case follows lbm.f90:52-201 at one rank); FP16 storage
  against FP64.
* The 8192 x 2048 cylinder at Re 1000: FP64 bit for bit against the oracle;
  FP32 / FP16 / BF16 storage against FP64.
* The 512 x 256 x 256 sphere: the storage forms' drag and velocity against
  FP64.

Error bounds of the reduced-precision forms are about three times the
errors measured on MI355X with these seeds and step counts (written beside
each bound).
"""
import numpy as np
import pytest

import imexlb as I
from cases import assert_state_equal, get_deck, oracle_setup
from imexlb.ensemble import Batch, cylinder_cases

pytestmark = pytest.mark.gpu

rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _oracle_fields(o):
    p = o.interior(o.p, 1)[..., 0].ravel()
    u = o.interior(o.u, 2)
    return p, np.concatenate([u[..., 0].ravel(), u[..., 1].ravel()])


def test_ensemble96_matches_oracle(r_vec):
    """Every case of the 96-case ensemble, advanced 50 steps by the batched
    kernel in two shapes (one tile per block of 256 threads; persistent
    blocks of 128 threads with 2 nodes each), equals the oracle run of the
    same deck: p, u and Monitor bit for bit, the force within 1e-12 of its
    magnitude (the summation orders differ)."""
    decks = cylinder_cases(96, seed=7)
    nodes = sum((dk.n[0] + 1) * (dk.n[1] + 1) for dk in decks)
    assert 10_000_000 < nodes < 12_500_000
    steps = 50
    runs = []
    for shape in ((256, 1, 0), (128, 2, 1)):
        print(f"[configs] ensemble96 FP64 shape {shape}", flush=True)
        b = Batch(decks, precision=I.F64)
        b.set_config(*shape)
        b.init(r_vec)
        b.advance(steps)
        um, diverged = b.monitor()
        assert diverged.size == 0
        F = b.forces()
        fields = [b.download_fields(i) for i in range(len(decks))]
        b.close()
        runs.append((um, F, fields))
    print("[configs] ensemble96: oracle runs", flush=True)
    for i, dk in enumerate(decks):
        o, _ = oracle_setup(dk, r_vec)
        o.run(0, steps - 1, 0)
        po, uo = _oracle_fields(o)
        Fo, umo = o.forces(), o.monitor()
        scale = np.abs(Fo).max()
        for um, F, fields in runs:
            p, u = fields[i]
            assert np.array_equal(p, po) and np.array_equal(u, uo), i
            assert um[i] == umo, i
            assert scale > 0 and np.all(np.abs(F[i] - Fo) <= 1e-12 * scale), (i, F[i], Fo)
        o.close()


def test_ensemble96_fp16_close_to_fp64(r_vec):
    """FP16 storage against FP64 on the 96-case ensemble after 1,000 steps.
    Three cases (the coarsest grids at the highest Re, tau down to 0.507,
    where BGK is unstable) have diverged by then, in both forms alike
    (Batch.monitor flags them); on the other 93 the relative L2 error of u
    per case (measured: at most 1.4e-3, median 6.0e-4) and the relative drag
    error (at most 2.2e-2 where the early drag is small, median 1.2e-3) stay
    within bounds three times those."""
    decks = cylinder_cases(96, seed=7)
    out = {}
    for prec in (I.F64, I.F16_F32):
        print(f"[configs] ensemble96 precision {prec}: 1000 steps", flush=True)
        b = Batch(decks, precision=prec)
        b.init(r_vec)
        b.advance(1000)
        _, diverged = b.monitor()
        out[prec] = ([b.download_fields(i)[1] for i in range(len(decks))], b.forces(), diverged)
        b.close()
    (u64, F64, d64), (u16, F16, d16) = out[I.F64], out[I.F16_F32]
    assert d64.tolist() == d16.tolist() == [5, 77, 81]
    ok = np.setdiff1d(np.arange(len(decks)), d64)
    eu = np.array([rel(u16[i], u64[i]) for i in ok])
    ed = np.abs(F16[ok, 0] - F64[ok, 0]) / np.abs(F64[ok, 0])
    print(f"[configs] ensemble96 FP16: u error max {eu.max():.2e} median {np.median(eu):.2e}; "
          f"drag error max {ed.max():.2e} median {np.median(ed):.2e}", flush=True)
    assert np.isfinite(eu).all() and np.isfinite(ed).all()
    assert eu.max() < 4.5e-3 and np.median(eu) < 2e-3 and ed.max() < 7e-2 and np.median(ed) < 3.5e-3


def test_cylinder_8192x2048_bitwise(r_vec):
    """The 8192 x 2048 cylinder (16.8 M nodes, Re 1000), FP64: 4 steps equal
    the oracle bit for bit (f, p, u, Monitor); the force within 1e-12."""
    deck = get_deck("cyl8192x2048")
    print("[configs] cyl8192x2048 FP64 vs oracle", flush=True)
    lat, d = I.setup(deck, r_vec, device=0)
    o, _ = oracle_setup(deck, r_vec)
    lat.advance(4)
    o.run(0, 3, 0)
    assert lat.monitor() == o.monitor()
    F, Fo = lat.forces(), o.forces()
    assert np.all(np.abs(F - Fo) <= 1e-12 * np.abs(Fo).max()), (F, Fo)
    assert_state_equal(lat, o)
    lat.close()
    o.close()


def test_sphere_above_2_31_slots_bitwise(r_vec):
    """A D3Q27 sphere lattice of 640 x 352 x 352 (81.0 M padded cells, 2.19 G
    population slots: past 2^31, so the boundary tables hold int64 slots and
    the step takes its K_IDX64 form), FP64, the sphere deck's flow
    parameters: 2 steps equal the oracle bit for bit (f, p, u, Monitor).
    About 55 GB of device memory and 75 GB of host memory."""
    deck = I.parse_deck("640 352 352\n0.078125 0.25 200\n1.9083969465648855 1 1\n0.024997 0.0005001\n", 3)
    cells = (640 + 3) * (352 + 3) * (352 + 3)
    assert 27 * cells > 2 ** 31
    print("[configs] sphere 640x352x352 FP64 (int64 slots) vs oracle", flush=True)
    lat, d = I.setup(deck, r_vec, device=0)
    lat.advance(2)
    o, _ = oracle_setup(deck, r_vec)
    o.run(0, 1, 0)
    assert lat.monitor() == o.monitor()
    assert_state_equal(lat, o)
    lat.close()
    o.close()


_FP64_RUNS = {}  # (workload, steps) -> the FP64 run every form is compared with


def _run_form(name, prec, steps, r_vec):
    key = (name, steps)
    if prec == I.F64 and key in _FP64_RUNS:
        return _FP64_RUNS[key]
    lat, _ = I.setup(get_deck(name), r_vec, device=0, precision=prec)
    lat.advance(steps)
    _, u = lat.download_fields()
    F = lat.forces()
    m = lat.mass()
    lat.close()
    if prec == I.F64:
        _FP64_RUNS.clear()  # hold one workload's reference at a time
        _FP64_RUNS[key] = (u, F, m)
    return u, F, m


# measured on MI355X after 2,000 steps (u error, drag error): FP32 1.14e-4, 2.7e-5;
# FP16 storage 7.9e-4, 5.5e-4; BF16 storage 1.76e-2, 5.7e-3; BF16 with stochastic
# rounding 7.7e-3, 3.4e-2 (the Re 1000 flow at tau 0.519 amplifies the rounding of
# every form, FP32 arithmetic included; the dither's noise shows in the drag)
CYL_BOUNDS = {I.F32: (3.5e-4, 1e-4), I.F16_F32: (2.5e-3, 1.7e-3), I.BF16_F32: (5.3e-2, 1.7e-2),
              I.BF16SR_F32: (2.3e-2, 1.0e-1)}


@pytest.mark.parametrize("prec", [I.F32, I.F16_F32, I.BF16_F32, I.BF16SR_F32])
def test_cylinder_8192x2048_forms(prec, r_vec):
    """FP32 arithmetic with FP32 / FP16 / BF16 storage against FP64 on the
    8192 x 2048 cylinder after 2,000 steps: relative L2 error of u and
    relative error of the drag within the stated bounds."""
    steps = 2000
    print(f"[configs] cyl8192x2048 precision {prec}: {steps} steps", flush=True)
    u64, F64, _ = _run_form("cyl8192x2048", I.F64, steps, r_vec)
    u, F, _ = _run_form("cyl8192x2048", prec, steps, r_vec)
    eu, ed = rel(u, u64), abs(F[0] - F64[0]) / abs(F64[0])
    print(f"[configs] cyl8192x2048 precision {prec}: u error {eu:.2e}, drag error {ed:.2e}", flush=True)
    bu, bd = CYL_BOUNDS[prec]
    assert eu < bu and ed < bd


# measured on MI355X after 500 steps (u error, drag error): FP32 2.1e-5, 1.3e-4;
# FP16 storage 3.0e-4, 4.9e-4; BF16 storage 3.8e-3, 2.5e-2; BF16 with stochastic
# rounding 3.0e-3, 8.7e-3
SPHERE_BOUNDS = {I.F32: (6.3e-5, 4e-4), I.F16_F32: (9e-4, 1.5e-3), I.BF16_F32: (1.2e-2, 7.5e-2),
                 I.BF16SR_F32: (9.0e-3, 2.6e-2)}


@pytest.mark.parametrize("prec", [I.F32, I.F16_F32, I.BF16_F32, I.BF16SR_F32])
def test_sphere_forms(prec, r_vec):
    """The storage forms on the 512 x 256 x 256 sphere after 500 steps
    against FP64: relative L2 error of u and relative drag error.  FP32 and
    FP16 storage keep the drag within the 0.5 % of the plan; the BF16 forms
    are held to their own measured bounds: round-to-nearest does not meet
    0.5 % (2.5 % here, about 9 % after 5,000 steps: its 8-bit significand
    rounds the slowly varying populations with a bias); stochastic rounding
    has no bias but its noise shows in the instantaneous drag (0.9 % here,
    0.12 % after 5,000 steps)."""
    steps = 500
    print(f"[configs] sphere precision {prec}: {steps} steps", flush=True)
    u64, F64, m64 = _run_form("sphere512x256x256", I.F64, steps, r_vec)
    u, F, m = _run_form("sphere512x256x256", prec, steps, r_vec)
    eu, ed = rel(u, u64), abs(F[0] - F64[0]) / abs(F64[0])
    print(f"[configs] sphere precision {prec}: u error {eu:.2e}, drag error {ed:.2e}, "
          f"mass {m:.6e} vs {m64:.6e}", flush=True)
    bu, bd = SPHERE_BOUNDS[prec]
    assert eu < bu and ed < bd
    if prec in (I.F32, I.F16_F32):
        assert ed < 5e-3
