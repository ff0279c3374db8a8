"""Batched ensemble: every case of a batch evolves exactly like the same
deck run as a single lattice (bit for bit in FP64, fields and forces), for
every step-kernel shape the autotuner can pick."""
import numpy as np
import pytest

import imexlb as I
from imexlb.ensemble import Batch, cylinder_cases

def _small_cases():
    """Five cylinder cases of uneven size and Reynolds number, small enough to
    compare with single lattices: Nx = 48..112, Ny = Nx/4, diameter Nx/10,
    inlet velocity 0.05, Re 20..80 (the relaxation follows from Re)."""
    out = []
    for k in range(5):
        nx = 48 + 16 * k
        ny = nx // 4
        Lc = 0.1
        re = 20.0 + 15.0 * k
        nu = 0.05 * (Lc * nx) / re
        ome = 1.0 / (3.0 * nu + 0.5)
        out.append(I.Deck(2, (nx, ny, 0), Lc, 0.25, re, ome, 1.0, 1.0, 1.0, 0.01))
    return out


def test_small_cases_are_stable_decks():
    for dk in _small_cases():
        d = I.derive(dk)
        assert abs(d.uu - 0.05) < 1e-12 and 2 * d.radius < dk.n[1] / 2


@pytest.mark.gpu
@pytest.mark.parametrize("block,cpt,persist", [(64, 1, 0), (256, 2, 1), (512, 4, 0), (128, 4, 1), (256, 1, 0)])
def test_batch_matches_single_lattices(block, cpt, persist, r_vec):
    decks = _small_cases()
    print(f"[batch] block={block} cpt={cpt} persist={persist}", flush=True)
    b = Batch(decks, precision=I.F64)
    b.set_config(block, cpt, persist)
    b.init(r_vec)
    b.advance(37)
    Fb = b.forces()
    for i, dk in enumerate(decks):
        lat, _ = I.setup(dk, r_vec, device=0)
        lat.advance(37)
        p, u = lat.download_fields()
        pb, ub = b.download_fields(i)
        assert np.isfinite(p).all() and np.isfinite(u).all(), i
        assert np.array_equal(p, pb) and np.array_equal(u, ub), i
        assert np.array_equal(lat.forces(), Fb[i]), i
        lat.close()
    b.close()


@pytest.mark.gpu
def test_batch_tune_and_continue(r_vec):
    """Tuning picks a valid shape; advancing after it (odd and even step
    counts, through the graph and the single-step path) stays exact."""
    decks = _small_cases()
    b = Batch(decks, precision=I.F64)
    b.init(r_vec)
    b.synchronize()  # valid before any step
    best = b.tune(4)
    assert best["block"] in (64, 128, 256, 512) and best["cells_per_thread"] in (1, 2, 4)
    assert best["ms_per_step"] > 0
    b.init(r_vec)
    for n in (3, 8, 1):
        b.advance(n)
    lat, _ = I.setup(decks[2], r_vec, device=0)
    lat.advance(12)
    p, u = lat.download_fields()
    pb, ub = b.download_fields(2)
    assert np.isfinite(p).all() and np.isfinite(u).all()
    assert np.array_equal(p, pb) and np.array_equal(u, ub)
    lat.close()
    b.close()


@pytest.mark.gpu
def test_batch_fp16_close_to_fp64(r_vec):
    decks = _small_cases()
    b64 = Batch(decks, precision=I.F64)
    b16 = Batch(decks, precision=I.F16_F32)
    for b in (b64, b16):
        b.init(r_vec)
        b.advance(200)
    for i in range(len(decks)):
        _, u64 = b64.download_fields(i)
        _, u16 = b16.download_fields(i)
        assert np.isfinite(u64).all(), i
        err = np.linalg.norm(u16 - u64) / np.linalg.norm(u64)
        assert err < 1e-2, (i, err)
    b64.close()
    b16.close()


@pytest.mark.gpu
def test_batch_ellipses_match_single_lattices(r_vec):
    """Cases with elliptic bodies (imexlb_batch_create_shapes) evolve like
    the same shapes run as single lattices, bit for bit."""
    decks = _small_cases()[:3]
    shapes = [(1.8, 20.0), (0.6, 70.0), (1.0, 0.0)]
    b = Batch(decks, precision=I.F64, shapes=shapes)
    b.init(r_vec)
    b.advance(25)
    for i, (dk, (aspect, angle)) in enumerate(zip(decks, shapes)):
        d = I.derive(dk)
        lat = I.Lattice(2, dk.n[0], dk.n[1], 0, device=0)
        lat.set_params(dk.ome, dk.rho0, d.uu)
        lat.set_geometry(d.pos, d.radius, aspect, angle)
        lat.init_up(r_vec)
        lat.init_pdf()
        lat.advance(25)
        p, u = lat.download_fields()
        pb, ub = b.download_fields(i)
        assert np.isfinite(p).all() and np.array_equal(p, pb) and np.array_equal(u, ub), i
        lat.close()
    b.close()


@pytest.mark.gpu
def test_batch_warm_start_matches_single_lattices(r_vec):
    """imexlb_batch_init_fields (warm start from given p, u) equals the
    single-lattice warm-start hook (imexlb_upload_fields + InitPDF)."""
    decks = _small_cases()[:3]
    rng = np.random.default_rng(5)
    ps, us = [], []
    for dk in decks:
        n = (dk.n[0] + 1) * (dk.n[1] + 1)
        ps.append(rng.normal(0.0, 1e-3, n))
        us.append(np.concatenate([0.05 + rng.normal(0.0, 5e-3, n), rng.normal(0.0, 5e-3, n)]))
    b = Batch(decks, precision=I.F64)
    b.init_fields(ps, us)
    b.advance(21)
    Fb = b.forces()
    for i, dk in enumerate(decks):
        lat, _ = I.setup(dk, r_vec, device=0)
        lat.upload_fields(ps[i], us[i])
        lat.init_pdf()
        lat.advance(21)
        p, u = lat.download_fields()
        pb, ub = b.download_fields(i)
        assert np.isfinite(p).all() and np.array_equal(p, pb) and np.array_equal(u, ub), i
        assert np.array_equal(lat.forces(), Fb[i]), i
        lat.close()
    b.close()
