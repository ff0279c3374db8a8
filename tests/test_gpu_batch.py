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
@pytest.mark.parametrize("nsteps", [1, 2, 37])
def test_batch_fill_path_matches_pushed_copies(nsteps, r_vec, monkeypatch):
    """The step kernel writes each case's boundary copies from its push
    table; with IMEXLB_PUSH=0 the batch runs the boundary fill before every
    step instead.  Both give the same fields and forces, bit for bit, also
    for a first call of one or two steps (the step after the start still
    takes every case's fill)."""
    decks = _small_cases()
    runs = []
    for push in ("1", "0"):
        monkeypatch.setenv("IMEXLB_PUSH", push)
        b = Batch(decks, precision=I.F64)
        b.init(r_vec)
        b.advance(nsteps)
        b.advance(3)
        runs.append(([b.download_fields(i) for i in range(len(decks))], b.forces()))
        b.close()
    for (p0, u0), (p1, u1) in zip(runs[0][0], runs[1][0]):
        assert np.array_equal(p0, p1) and np.array_equal(u0, u1)
    assert np.array_equal(runs[0][1], runs[1][1])


@pytest.mark.gpu
def test_batch_64bit_slot_tables(r_vec, monkeypatch):
    """The batch's boundary fill with int64 slot tables (IMEXLB_SLOTS64=1,
    the form an ensemble above 2^31 population slots takes) gives the same
    fields and forces as its int32 tables, bit for bit."""
    decks = _small_cases()
    runs = []
    for s64 in ("0", "1"):
        monkeypatch.setenv("IMEXLB_SLOTS64", s64)
        b = Batch(decks, precision=I.F64)
        b.set_config(128, 2, 0)
        b.init(r_vec)
        b.advance(41)
        runs.append((b, b.forces()))
    (b0, F0), (b1, F1) = runs
    assert np.array_equal(F0, F1)
    for i in range(len(decks)):
        p0, u0 = b0.download_fields(i)
        p1, u1 = b1.download_fields(i)
        assert np.isfinite(p0).all()
        assert np.array_equal(p0, p1) and np.array_equal(u0, u1), i
    b0.close()
    b1.close()


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


def _fluid(dk, shape=(1.0, 0.0)):
    """Fluid flags (ny+1, nx+1) of a deck's nodes from the geometry table."""
    d = I.derive(dk)
    nx, ny = dk.n[0], dk.n[1]
    lat = I.Lattice(2, nx, ny, 0, device=-1)
    lat.set_params(dk.ome, dk.rho0, d.uu)
    lat.set_geometry(d.pos, d.radius, *shape)
    geo = lat.geometry_geo().reshape(ny + 7, nx + 7)[3:ny + 4, 3:nx + 4]
    lat.close()
    return (geo != 0).astype(np.float64)


def _stack_fields(dk, f, ps, us, r_vec, shape=(1.0, 0.0)):
    """p, u (interior layout) that an FP32 stack f (3, H, W) gives a deck's
    nodes: node (x, y) reads f[:, min(y, H-1), min(x, W-1)]; p = ps f0,
    u_x = us f1 r, u_y = us f2, r = 1 + ((r_vec[id mod 100] - 0.5) 2) 0.2
    with the reference id of node (x, y) (initialization.f90:273), r = 1
    without r_vec; zero on body nodes."""
    nx, ny = dk.n[0], dk.n[1]
    H, W = f.shape[1:]
    y, x = np.mgrid[0:ny + 1, 0:nx + 1]
    g = f[:, np.minimum(y, H - 1), np.minimum(x, W - 1)].astype(np.float64)
    r = 1.0 if r_vec is None else 1.0 + ((np.asarray(r_vec)[((x + 3) + (nx + 7) * (y + 3)) % 100] - 0.5) * 2.0) * 0.2
    fl = _fluid(dk, shape)
    return (ps * g[0] * fl).ravel(), np.concatenate([(us * g[1] * r * fl).ravel(), (us * g[2] * fl).ravel()])


@pytest.mark.gpu
def test_init_fields_f32_single_lattice(r_vec):
    """imexlb_init_fields_f32 sets p, u as the stack formula states (bit
    for bit), and a stack of InitUP's normalised inflow profile reproduces
    InitUP (up to the profile's FP32 rounding), which pins the perturbation's
    node indexing."""
    import torch
    dk = _small_cases()[1]
    nx, ny = dk.n[0], dk.n[1]
    d = I.derive(dk)
    rng = np.random.default_rng(31)
    f = rng.normal(0.0, 1.0, (3, ny, nx - 5)).astype(np.float32)  # last row and 6 columns repeat
    fd = torch.tensor(f, device="cuda")
    lat, _ = I.setup(dk, r_vec, device=0)
    for rv in (r_vec, None):
        lat.init_fields_f32(fd, 2.5e-3, 0.05, rv)
        p, u = lat.download_fields()
        pe, ue = _stack_fields(dk, f, 2.5e-3, 0.05, rv)
        assert np.array_equal(p, pe) and np.array_equal(u, ue)
    y = np.arange(ny + 1, dtype=np.float64)
    prof = np.zeros((3, ny + 1, nx + 1), np.float32)
    prof[1] = (4.0 * y * (ny - y) / float(ny * ny))[:, None]
    lat.init_fields_f32(torch.tensor(prof, device="cuda"), 1.0, d.uu, r_vec)
    p, u = lat.download_fields()
    lat.init_up(r_vec)
    p0, u0 = lat.download_fields()
    assert not p.any() and not p0.any()
    assert np.allclose(u, u0, rtol=1e-6, atol=0.0)
    lat.init_pdf()
    lat.advance(5)
    assert np.isfinite(lat.download_fields()[1]).all()
    lat.close()


@pytest.mark.gpu
def test_batch_init_fields_f32_matches_host_fields(r_vec):
    """A batch warm-started from a device FP32 stack
    (imexlb_batch_init_fields_f32) equals the same batch warm-started from
    the host arrays the stack formula gives (imexlb_batch_init_fields), bit
    for bit, at the start and after 17 steps."""
    import torch
    decks = _small_cases()[:3]
    shapes = [(1.0, 0.0), (1.6, 35.0), (0.7, 80.0)]
    H = min(dk.n[1] for dk in decks)
    W = min(dk.n[0] for dk in decks)
    rng = np.random.default_rng(32)
    f = np.zeros((3, 3, H, W), np.float32)
    f[:, 0] = rng.normal(0.0, 0.2, (3, H, W))
    f[:, 1] = 1.0 + rng.normal(0.0, 0.1, (3, H, W))
    f[:, 2] = rng.normal(0.0, 0.1, (3, H, W))
    host = [_stack_fields(dk, f[i], 2.5e-3, 0.05, r_vec, shapes[i]) for i, dk in enumerate(decks)]
    a = Batch(decks, precision=I.F64, shapes=shapes)
    a.init_fields([h[0] for h in host], [h[1] for h in host])
    b = Batch(decks, precision=I.F64, shapes=shapes)
    b.init_fields_f32(torch.tensor(f, device="cuda"), 2.5e-3, 0.05, r_vec)
    for steps in (0, 17):
        a.advance(steps)
        b.advance(steps)
        for i in range(3):
            pa, ua = a.download_fields(i)
            pb, ub = b.download_fields(i)
            assert np.isfinite(pa).all() and np.array_equal(pa, pb) and np.array_equal(ua, ub), (steps, i)
        if steps:  # forces need a completed step
            assert np.array_equal(a.forces(), b.forces())
    a.close()
    b.close()


@pytest.mark.gpu
def test_body_masks_match_geometry():
    """imexlb_body_mask_f32 and imexlb_batch_body_mask_f32 write 1 exactly on
    the geometry's body nodes, leave other channels alone, and reject masks
    larger than a case."""
    import torch
    decks = _small_cases()[:3]
    shapes = [(1.0, 0.0), (1.6, 35.0), (0.7, 80.0)]
    H = min(dk.n[1] for dk in decks) + 1
    W = min(dk.n[0] for dk in decks) + 1
    b = Batch(decks, precision=I.F64, shapes=shapes)
    out = torch.full((3, 2, H, W), 7.0, device="cuda")
    b.body_mask_f32(out)
    b.synchronize()
    got = out.cpu().numpy()
    for i, dk in enumerate(decks):
        assert np.array_equal(got[i, 0], 1.0 - _fluid(dk, shapes[i])[:H, :W]), i
        assert (got[i, 1] == 7.0).all()
    with pytest.raises(I.ImexlbError):
        b.body_mask_f32(torch.zeros((3, 1, H + 1, W), device="cuda"))
    b.close()
    dk = decks[1]
    d = I.derive(dk)
    lat = I.Lattice(2, dk.n[0], dk.n[1], 0, device=0)
    lat.set_params(dk.ome, dk.rho0, d.uu)
    lat.set_geometry(d.pos, d.radius, *shapes[1])
    m = torch.zeros((dk.n[1] + 1, dk.n[0] + 1), device="cuda")
    lat.body_mask_f32(m)
    lat.synchronize()
    assert np.array_equal(m.cpu().numpy(), 1.0 - _fluid(dk, shapes[1]))
    lat.close()


@pytest.mark.gpu
def test_batch_without_graphs_matches(r_vec, monkeypatch):
    """IMEXLB_GRAPHS=0 (every step launched, no captured pairs) gives the same
    batch, bit for bit, as graph replay."""
    decks = _small_cases()[:3]
    a = Batch(decks, precision=I.F64)
    monkeypatch.setenv("IMEXLB_GRAPHS", "0")
    b = Batch(decks, precision=I.F64)
    for x in (a, b):
        x.init(r_vec)
        x.advance(37)
    assert np.array_equal(a.forces(), b.forces())
    for i in range(3):
        pa, ua = a.download_fields(i)
        pb, ub = b.download_fields(i)
        assert np.array_equal(pa, pb) and np.array_equal(ua, ub), i
    a.close()
    b.close()


@pytest.mark.gpu
def test_device_start_and_check_errors(r_vec):
    """Argument and order errors of the device-memory entry points: a field
    stack for a 3D lattice, a mask larger than the lattice, the mass before
    any populations exist, and a stack on a lattice without geometry."""
    import torch
    dk = _small_cases()[0]
    d = I.derive(dk)
    nx, ny = dk.n[0], dk.n[1]
    lat = I.Lattice(2, nx, ny, 0, device=0)
    lat.set_params(dk.ome, dk.rho0, d.uu)
    stack = torch.zeros((3, ny + 1, nx + 1), device="cuda")
    with pytest.raises(I.ImexlbError) as e:
        lat.init_fields_f32(stack, 1.0, 1.0)  # no geometry yet
    assert e.value.code == -4
    lat.set_geometry(d.pos, d.radius)
    with pytest.raises(I.ImexlbError) as e:
        lat.mass()  # no populations yet
    assert e.value.code == -4
    with pytest.raises(I.ImexlbError) as e:
        lat.body_mask_f32(torch.zeros((ny + 2, nx + 1), device="cuda"))
    assert e.value.code == -1
    lat.close()
    lat3 = I.Lattice(3, 16, 8, 8, device=0)
    lat3.set_params(1.6, 1.0, 0.02)
    lat3.set_geometry(8.0, 2.0)
    with pytest.raises(I.ImexlbError) as e:
        lat3.init_fields_f32(torch.zeros((3, 9, 17), device="cuda"), 1.0, 1.0)
    assert e.value.code == -1
    lat3.close()


@pytest.mark.gpu
def test_batch_monitor_flags_diverged_cases(r_vec):
    """imexlb_batch_monitor: per-case Umax equal to the single lattices'
    Monitor for stable cases, NaN and IMEXLB_E_NAN (reported, not raised, by
    Batch.monitor) for a case run far past BGK's stability limit."""
    decks = _small_cases()[:3]
    bad = decks[1]
    # the same grid at Re 1e6 and ome 1.99: inlet velocity far above 1
    decks[1] = I.Deck(2, bad.n, bad.Lc, bad.Xpos, 1.0e6, 1.99, 1.0, 1.0, 1.0, 0.01)
    assert I.derive(decks[1]).uu > 10
    b = Batch(decks, precision=I.F64)
    b.init(r_vec)
    um0, div0 = b.monitor()
    assert div0.size == 0 and np.isfinite(um0).all()
    b.advance(30)
    um, diverged = b.monitor()
    assert diverged.tolist() == [1] and np.isnan(um[1])
    for i in (0, 2):
        lat, _ = I.setup(decks[i], r_vec, device=0)
        lat.advance(30)
        assert lat.monitor() == um[i], i
        lat.close()
    b.close()
