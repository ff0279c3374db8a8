"""Parity of the HIP path (through the C ABI) with the oracle.

FP64 form: bit-for-bit equality of f, p, u and of every Monitor value after
N steps (IEEE values; the only admitted difference is the sign of a zero).
Low-precision forms: error against the FP64 HIP run, bounded as stated per
test.
"""
import numpy as np
import pytest

import imexlb as I
import oracle as O
from cases import assert_state_equal, get_deck, interior_mask, oracle_setup

pytestmark = pytest.mark.gpu


def _run_pair(name, r_vec, steps, fused=True, precision=I.F64):
    deck = get_deck(name)
    print(f"[parity] deck={name} steps={steps} fused={fused} precision={precision}", flush=True)
    lat, d = I.setup(deck, r_vec, device=0, precision=precision)
    lat.set_fused(fused)
    o, _ = oracle_setup(deck, r_vec)
    return lat, o, d


def _advance_with_monitor(lat, o, d, steps):
    """main.f90:48-63 for iter = 0..steps-1 on both sides; returns Monitor series."""
    interv = max(d.interv, 1)
    um_dev, um_or = [], []
    it = 0
    while it < steps:
        if it % interv == 0:
            um_dev.append(lat.monitor())
        n = min(interv - it % interv, steps - it)
        lat.advance(n)
        it += n
    um_or = list(o.run(0, steps - 1, interv))
    return np.array(um_dev), np.array(um_or)


@pytest.mark.parametrize("name,steps", [("tiny2d", 70), ("tiny3d", 44), ("inletbody2d", 145),
                                        ("wallsphere3d", 52), ("ysphere3d", 40), ("outletsphere3d", 40),
                                        ("ragged2d", 90), ("ragged3d", 40), ("channel3", 60)])
@pytest.mark.parametrize("fused", [True, False])
def test_full_run_bitwise_small(name, steps, fused, r_vec):
    lat, o, d = _run_pair(name, r_vec, steps, fused)
    um_dev, um_or = _advance_with_monitor(lat, o, d, steps)
    assert np.array_equal(um_dev, um_or), (um_dev, um_or)
    assert_state_equal(lat, o)
    lat.close()
    o.close()


def test_cylinder_400x100_bitwise(r_vec):
    """The 400x100 cylinder at Re 100: 1000 steps, Monitor every 100."""
    lat, o, d = _run_pair("cyl400x100", r_vec, 1000)
    um_dev, um_or = _advance_with_monitor(lat, o, d, 1000)
    assert len(um_dev) == 10
    assert np.array_equal(um_dev, um_or), (um_dev, um_or)
    assert_state_equal(lat, o)
    lat.close()
    o.close()


def test_reference_inputs_bitwise(r_vec):
    """The reference's own example decks input.in (200 steps) and input_3d.in (6 steps)."""
    for name, steps in (("input.in", 200), ("input_3d.in", 6)):
        lat, o, d = _run_pair(name, r_vec, steps)
        lat.advance(steps)
        o.run(0, steps - 1, 0)
        assert lat.monitor() == o.monitor()
        assert_state_equal(lat, o)
        lat.close()
        o.close()


def test_sphere_512x256x256_bitwise(r_vec):
    """The benchmark lattice itself, D3Q27 sphere 513x257x257 nodes: 3 steps,
    then f, p, u on all 33.9 M interior nodes and Monitor equal the oracle's
    bit for bit, and the drag/lift vector agrees to 1e-12 (summation order)."""
    lat, o, d = _run_pair("sphere512x256x256", r_vec, 3)
    lat.advance(3)
    o.run(0, 2, 0)
    assert lat.monitor() == o.monitor()
    F, Fo = lat.forces(), o.forces()
    assert np.all(np.abs(F - Fo) <= 1e-12 * np.abs(Fo).max()), (F, Fo)
    assert_state_equal(lat, o)
    lat.close()
    o.close()


@pytest.mark.parametrize("remap", ["0", "1", "3", "4"])
def test_block_order_does_not_change_results(remap, r_vec, monkeypatch):
    """The block order over the XCDs is a per-lattice choice (an eighth of
    the blocks per XCD for D3Q27 FP64 by default, runs of R blocks for
    R >= 2); forced to each, 2D and 3D runs equal the oracle."""
    monkeypatch.setenv("IMEXLB_XCD_REMAP", remap)
    for name, steps in (("inletbody2d", 40), ("wallsphere3d", 20)):
        lat, o, d = _run_pair(name, r_vec, steps)
        lat.advance(steps)
        o.run(0, steps - 1, 0)
        assert lat.monitor() == o.monitor()
        assert_state_equal(lat, o)
        lat.close()
        o.close()


@pytest.mark.parametrize("small_cells", ["0", str(1 << 30)])
def test_small_lattice_mode_does_not_change_results(small_cells, r_vec, monkeypatch):
    """2D lattices up to IMEXLB_SMALL_CELLS padded cells step with the
    small-lattice kernel form (one-wave blocks, each tile's boundary copies
    spread over its lanes through LDS, cached stores); forced off and on,
    the decks with walls, a body cutting the inlet, a 3-row channel whose
    tile holds more copies than the kernel loads up front, and the 400x100
    cylinder equal the oracle."""
    monkeypatch.setenv("IMEXLB_SMALL_CELLS", small_cells)
    for name, steps in (("tiny2d", 61), ("inletbody2d", 70), ("ragged2d", 45), ("channel3", 50),
                        ("cyl400x100", 200)):
        lat, o, d = _run_pair(name, r_vec, steps)
        lat.advance(steps)
        o.run(0, steps - 1, 0)
        assert lat.monitor() == o.monitor()
        assert_state_equal(lat, o)
        lat.close()
        o.close()


@pytest.mark.parametrize("nx,ny,precision", [(2048, 512, I.F64), (4096, 1024, I.F16_F32)])
def test_midsize_lattices_small_form_bitwise(nx, ny, precision, r_vec, monkeypatch):
    """2D lattices whose population buffers fit in the 256 MB Infinity Cache
    (1.06 M cells in FP64, 4.2 M in FP16 storage) step with the small-lattice
    form by default; forced off (IMEXLB_SMALL_CELLS=0), the other form gives
    the same populations and fields, bit for bit."""
    dk = I.parse_deck(f"{nx} {ny}\n0.05 0.2 100\n1.8867924528301887 1 1\n0.01 0.0025001\n", 2)
    runs = []
    for small in (None, "0"):
        print(f"[midsize] {nx}x{ny} precision={precision} IMEXLB_SMALL_CELLS={small}", flush=True)
        if small is None:
            monkeypatch.delenv("IMEXLB_SMALL_CELLS", raising=False)
        else:
            monkeypatch.setenv("IMEXLB_SMALL_CELLS", small)
        lat, _ = I.setup(dk, r_vec, device=0, precision=precision)
        lat.advance(21)
        runs.append((lat.download_pdf(I.PDF_F), lat.download_fields()))
        lat.close()
    (f0, (p0, u0)), (f1, (p1, u1)) = runs
    assert np.isfinite(f0).all()
    assert np.array_equal(f0, f1)
    assert np.array_equal(p0, p1) and np.array_equal(u0, u1)


@pytest.mark.parametrize("push", ["1", "0"])
def test_64bit_slot_tables_match_oracle(push, r_vec, monkeypatch):
    """Lattices above 2^31 population slots keep their boundary tables in
    int64 slots (the step's K_IDX64 form, the fill's int64 tables);
    IMEXLB_SLOTS64=1 selects them for small decks too.  Fused with pushed
    copies or with the separate fill, fused and per phase: equal to the
    oracle."""
    monkeypatch.setenv("IMEXLB_SLOTS64", "1")
    monkeypatch.setenv("IMEXLB_PUSH", push)
    for name, steps, fused in (("inletbody2d", 70, True), ("wallsphere3d", 36, True), ("tiny2d", 31, False),
                               ("cyl400x100", 100, True)):
        lat, o, d = _run_pair(name, r_vec, steps, fused)
        lat.advance(steps)
        o.run(0, steps - 1, 0)
        assert lat.monitor() == o.monitor()
        assert_state_equal(lat, o)
        lat.close()
        o.close()


def test_separate_fill_path_matches_oracle(r_vec, monkeypatch):
    """With IMEXLB_PUSH=0 the boundary copies run as their own fill kernel
    before every pull (instead of being written by the previous step);
    results stay equal to the oracle, also under graph replay and across a
    body that cuts the inlet (previous-step body values)."""
    monkeypatch.setenv("IMEXLB_PUSH", "0")
    for name, steps in (("inletbody2d", 70), ("wallsphere3d", 36), ("cyl400x100", 100)):
        lat, o, d = _run_pair(name, r_vec, steps)
        lat.advance(steps)
        o.run(0, steps - 1, 0)
        assert lat.monitor() == o.monitor()
        assert_state_equal(lat, o)
        lat.close()
        o.close()


@pytest.mark.parametrize("aspect,angle", [(1.8, 35.0), (0.5, 80.0)])
def test_ellipse_bitwise(aspect, angle, r_vec):
    """An elliptic body (the warm-start study's second shape): 60 steps equal
    the oracle bit for bit."""
    nx, ny, pos, radius, ome, uu = 64, 32, 20.0, 5.0, 1.7, 0.05
    print(f"[parity] ellipse aspect={aspect} angle={angle}", flush=True)
    lat = I.Lattice(2, nx, ny, 0, device=0)
    lat.set_params(ome, 1.0, uu)
    lat.set_geometry(pos, radius, aspect, angle)
    lat.init_up(r_vec)
    lat.init_pdf()
    o = O.Oracle(2, nx, ny, 0)
    o.set_params(ome, 1.0, uu)
    o.set_geometry(pos, radius, aspect, angle)
    o.init_up(r_vec)
    o.init_pdf()
    lat.advance(60)
    o.run(0, 59, 0)
    assert lat.monitor() == o.monitor()
    assert_state_equal(lat, o)
    lat.close()
    o.close()


def test_phase_order_errors(r_vec):
    """The phase state machine refuses calls out of the reference's order
    (main.f90:44-61: SetGeometry, InitUP, InitPDF, then Collision,
    BoundaryCondition, Propagation, PostProcessing) with IMEXLB_E_STATE, a
    negative step count with IMEXLB_E_ARG, and a refused call changes
    nothing: the run carried on after them equals the oracle bit for bit."""
    deck = get_deck("tiny2d")
    d = I.derive(deck)
    print("[parity] phase-order errors on tiny2d", flush=True)
    lat = I.Lattice(2, *deck.n[:2], device=0)
    lat.set_params(deck.ome, deck.rho0, d.uu)

    def refused(call, code=I.E_STATE):
        with pytest.raises(I.ImexlbError) as e:
            call()
        assert e.value.code == code, (e.value.code, code)

    refused(lambda: lat.init_up(r_vec))      # before SetGeometry
    lat.set_geometry(d.pos, d.radius)
    refused(lat.monitor)                     # no fields yet
    refused(lat.init_pdf)                    # before InitUP
    refused(lat.collision)
    lat.init_up(r_vec)
    refused(lat.collision)                   # before InitPDF
    lat.init_pdf()
    lat.set_fused(False)
    refused(lat.boundary_condition)          # before Collision
    refused(lat.propagation)
    refused(lat.post_processing)
    lat.collision()
    refused(lat.collision)                   # twice
    refused(lat.propagation)                 # before BoundaryCondition
    lat.boundary_condition()
    refused(lat.post_processing)             # before Propagation
    lat.propagation()
    refused(lambda: lat.set_fused(True))     # between Propagation and PostProcessing
    refused(lambda: lat.advance(-1), I.E_ARG)
    lat.post_processing()
    lat.set_fused(True)
    lat.advance(20)
    lat.collision()
    lat.boundary_condition()
    lat.propagation()
    refused(lambda: lat.set_fused(False))    # fused step pending
    lat.post_processing()
    o, _ = oracle_setup(deck, r_vec)
    o.run(0, 21, 0)
    assert lat.monitor() == o.monitor()
    assert_state_equal(lat, o)
    lat.close()
    o.close()


def _phase_steps(lat, n):
    """n steps through the reference's four phase calls (main.f90:57-61)."""
    for _ in range(n):
        lat.collision()
        lat.boundary_condition()
        lat.propagation()
        lat.post_processing()


def test_deferred_phase_steps_match_oracle(r_vec):
    """Through the four phase calls a fused lattice defers its steps and runs
    them as graph replays when the state is next needed.  Interleaved with
    Monitor, downloads, a forces call, advance() and a parameter change in
    the middle of the run (the pending steps keep the old parameters; the
    next Collision uses the new ones, as the reference's does), and over
    many 16-step chunks through phase calls alone: equal to the oracle."""
    deck = get_deck("tiny2d")
    d = I.derive(deck)
    print("[parity] deferred phase steps on tiny2d", flush=True)
    lat, _ = I.setup(deck, r_vec, device=0)
    o, _ = oracle_setup(deck, r_vec)

    def check():
        assert lat.monitor() == o.monitor()
        assert_state_equal(lat, o)

    _phase_steps(lat, 37)
    o.run(0, 36, 0)
    check()
    _phase_steps(lat, 5)
    o.run(37, 41, 0)
    F, Fo = lat.forces(), o.forces()
    assert np.all(np.abs(F - Fo) <= 1e-12 * np.abs(Fo).max()), (F, Fo)
    _phase_steps(lat, 3)
    lat.advance(12)  # joins the 3 pending steps
    o.run(42, 56, 0)
    check()
    _phase_steps(lat, 9)
    lat.set_params(1.7, deck.rho0, d.uu)  # the 9 pending steps run first, with the old ome
    o.run(57, 65, 0)
    o.set_params(1.7, deck.rho0, d.uu)
    _phase_steps(lat, 11)
    o.run(66, 76, 0)
    check()
    _phase_steps(lat, 1100)  # 68 chunks of 16 run on their own, the last 12 on Monitor
    o.run(77, 1176, 0)
    check()
    lat.close()
    o.close()


def test_graph_replay_matches_oracle(r_vec):
    """advance() replays pairs of fused steps as HIP graphs of 64, 8 and 1
    pairs, largest first; after every call, with eager Monitor and download
    calls in between, the state equals the oracle's bit for bit."""
    lat, o, d = _run_pair("cyl400x100", r_vec, 0)
    total = 0
    for n in (1, 6, 7, 40, 3, 64, 301, 128):
        lat.advance(n)
        o.run(total, total + n - 1, 0)
        total += n
        assert lat.monitor() == o.monitor()
        assert_state_equal(lat, o)
    lat.close()
    o.close()


def test_warm_start_hook(r_vec):
    """Warm start (InitUP/InitPDF replaced by a given field, lbm.f90:24-48):
    a developed p, u uploaded into a fresh lattice, InitPDF, 200 steps, equals
    the oracle started from the same field, bit for bit; both layouts."""
    deck = get_deck("cyl400x100")
    d = I.derive(deck)
    src, _ = oracle_setup(deck, r_vec)
    src.run(0, 299, 0)  # a developed field
    p_ref, u_ref = src.p.copy(), src.u.copy()
    for layout in (I.LAYOUT_REF, I.LAYOUT_INTERIOR):
        lat = I.Lattice(2, *deck.n[:2], device=0)
        lat.set_params(deck.ome, deck.rho0, d.uu)
        lat.set_geometry(d.pos, d.radius)
        if layout == I.LAYOUT_REF:
            lat.upload_fields(p_ref, u_ref, layout)
        else:
            p_in = src.interior(p_ref, 1)[..., 0].ravel()
            u_in = src.interior(u_ref, 2)
            lat.upload_fields(p_in, np.concatenate([u_in[..., 0].ravel(), u_in[..., 1].ravel()]), layout)
        lat.init_pdf()
        o = __import__("oracle").Oracle(2, *deck.n[:2])
        o.set_params(deck.ome, deck.rho0, d.uu)
        o.set_geometry(d.pos, d.radius)
        o.p[:] = p_ref
        o.u[:] = u_ref
        o.init_pdf()
        lat.advance(200)
        o.run(0, 199, 0)
        assert_state_equal(lat, o)
        lat.close()
        o.close()
    src.close()


def test_phase_by_phase(r_vec):
    """State after each of Collision, BoundaryCondition, Propagation and
    PostProcessing (unfused), and after the fused step, equals the oracle's."""
    deck = get_deck("tiny2d")
    lat, d = I.setup(deck, r_vec, device=0)
    lat.set_fused(False)
    o, _ = oracle_setup(deck, r_vec)
    fid = o.list("fluid_id")
    slot, src, _ = lat.bc_table()
    for step in range(3):
        lat.collision(); o.collision()
        fb = lat.download_pdf(I.PDF_FB, I.LAYOUT_REF).reshape(o.A, o.nq)
        assert np.array_equal(fb[fid], o.fb.reshape(o.A, o.nq)[fid])
        lat.boundary_condition(); o.boundary_condition()
        fbf = lat.download_pdf(I.PDF_FB, I.LAYOUT_REF)
        assert np.array_equal(fbf[slot], o.fb[slot])
        lat.propagation(); o.propagation()
        assert_state_equal(lat, o, what=("f",))
        lat.post_processing(); o.post_processing()
        assert_state_equal(lat, o)
    lat.close()
    o.close()


def test_fused_and_unfused_agree_midstep(r_vec):
    """Downloads between phases in fused mode see the reference's state."""
    deck = get_deck("tiny3d")
    lat, d = I.setup(deck, r_vec, device=0)
    o, _ = oracle_setup(deck, r_vec)
    for _ in range(2):
        lat.advance(1)
        o.run(0, 0, 0)
    lat.collision(); o.collision()
    assert_state_equal(lat, o)  # f, p, u of this step are still readable after Collision
    lat.boundary_condition(); o.boundary_condition()
    lat.propagation(); o.propagation()
    lat.post_processing(); o.post_processing()
    assert_state_equal(lat, o)
    lat.close()
    o.close()


def test_state_errors(r_vec):
    deck = get_deck("tiny2d")
    lat, d = I.setup(deck, r_vec, device=0)
    with pytest.raises(I.ImexlbError):
        lat.propagation()  # before Collision and BoundaryCondition
    lat.collision()
    with pytest.raises(I.ImexlbError):
        lat.post_processing()
    lat.close()




@pytest.mark.parametrize("name,steps", [("cyl400x100", 500), ("tiny3d", 40), ("inletbody2d", 100)])
@pytest.mark.parametrize("fused", [True, False])
def test_forces_match_oracle(name, steps, fused, r_vec):
    """Drag and lift by momentum exchange, FP64, within 1e-12 of the
    oracle's (the summation order differs)."""
    lat, o, d = _run_pair(name, r_vec, steps, fused)
    lat.advance(steps)
    o.run(0, steps - 1, 0)
    Fd, Fo = lat.forces(), o.forces()
    scale = np.abs(Fo).max()
    assert scale > 0
    assert np.all(np.abs(Fd - Fo) <= 1e-12 * scale), (Fd, Fo)
    lat.close()
    o.close()


@pytest.mark.parametrize("name,steps,fused", [("tiny2d", 23, True), ("tiny3d", 9, True), ("wallsphere3d", 7, False)])
def test_mass_matches_oracle(name, steps, fused, r_vec):
    """imexlb_mass equals the sum over the oracle's fluid interior nodes of
    sum_q f_q (FP64; within 1e-12 of the sum of magnitudes, since the
    summation orders differ), and the device's own f gives the same sum."""
    lat, o, d = _run_pair(name, r_vec, steps, fused)
    lat.advance(steps)
    o.run(0, steps - 1, 0)
    m = interior_mask(o) & (np.asarray(o.geo) != 0)
    fo = np.asarray(o.f).reshape(o.A, o.nq)[m]
    ref, scale = fo.sum(), np.abs(fo).sum()
    got = lat.mass()
    assert scale > 0 and abs(got - ref) <= 1e-12 * scale, (got, ref)
    fd = lat.download_pdf(I.PDF_F, I.LAYOUT_REF).reshape(o.A, o.nq)[m]
    assert abs(got - fd.sum()) <= 1e-12 * scale
    assert lat.mass() == got  # the order depends on the lattice only
    lat.close()
    o.close()


# measured on MI355X (relative L2 error of u after 2,000 steps): FP32 1.2e-6,
# FP16 storage 6.1e-4, BF16 storage 1.2e-2, BF16 with stochastic rounding
# 6.2e-3; the bounds are three times those
@pytest.mark.parametrize("precision,tol", [(I.F32, 3.6e-6), (I.F16_F32, 1.8e-3), (I.BF16_F32, 3.7e-2),
                                           (I.BF16SR_F32, 1.9e-2)])
def test_precision_forms_error(precision, tol, r_vec):
    """FP32 arithmetic with FP32 / FP16 / BF16 / stochastically rounded BF16
    storage against the FP64 run of the 400x100 cylinder after 2000 steps:
    relative L2 error of u below tol."""
    deck = get_deck("cyl400x100")
    ref, _ = I.setup(deck, r_vec, device=0)
    ref.advance(2000)
    _, u64 = ref.download_fields()
    lat, _ = I.setup(deck, r_vec, device=0, precision=precision)
    lat.advance(2000)
    _, u = lat.download_fields()
    err = np.linalg.norm(u - u64) / np.linalg.norm(u64)
    print(f"precision {precision}: rel L2 error of u = {err:.3e}")
    assert np.isfinite(err) and err < tol
    ref.close()
    lat.close()


def test_bf16_stochastic_rounding_repeats_and_dithers(r_vec):
    """BF16 storage with stochastic rounding: two runs of the 400x100
    cylinder give the same populations bit for bit (the dither is a hash of
    the cell and its pressure), and they differ from round-to-nearest BF16."""
    deck = get_deck("cyl400x100")
    runs = []
    for prec in (I.BF16SR_F32, I.BF16SR_F32, I.BF16_F32):
        lat, _ = I.setup(deck, r_vec, device=0, precision=prec)
        lat.advance(200)
        runs.append(lat.download_pdf(I.PDF_FB))
        lat.close()
    assert np.isfinite(runs[0]).all()
    assert np.array_equal(runs[0], runs[1])
    assert not np.array_equal(runs[0], runs[2])


@pytest.mark.parametrize("fail_at", range(11, 21))
def test_failed_set_geometry_leaves_no_geometry(fail_at, r_vec, monkeypatch):
    """A SetGeometry whose device work fails part-way (IMEXLB_FAIL_ALLOC
    makes the context's fail_at-th allocation fail; a 2D context makes 10
    at creation, SetGeometry the next 10) leaves the context without a
    geometry: InitUP, advance, Monitor, the list queries and the downloads
    return IMEXLB_E_STATE instead of running on half-built tables.  A second
    SetGeometry then succeeds and the run equals the oracle bit for bit."""
    deck = get_deck("tiny2d")
    d = I.derive(deck)
    print(f"[parity] injected allocation failure at {fail_at}", flush=True)
    monkeypatch.setenv("IMEXLB_FAIL_ALLOC", str(fail_at))
    lat = I.Lattice(2, *deck.n[:2], device=0)
    monkeypatch.delenv("IMEXLB_FAIL_ALLOC")
    lat.set_params(deck.ome, deck.rho0, d.uu)
    with pytest.raises(I.ImexlbError) as err:
        lat.set_geometry(d.pos, d.radius)
    assert err.value.code == I.E_OOM
    for call in (lambda: lat.init_up(r_vec), lambda: lat.init_pdf(), lambda: lat.advance(2), lat.monitor,
                 lambda: lat.geometry_list("b_user"), lat.download_fields, lat.forces, lat.mass):
        with pytest.raises(I.ImexlbError) as err:
            call()
        assert err.value.code == I.E_STATE
    lat.set_geometry(d.pos, d.radius)  # later allocations succeed
    lat.init_up(r_vec)
    lat.init_pdf()
    lat.advance(20)
    o, _ = oracle_setup(deck, r_vec)
    o.run(0, 19, 0)
    assert lat.monitor() == o.monitor()
    assert_state_equal(lat, o)
    lat.close()
    o.close()


def test_monitor_reports_nonfinite(r_vec):
    """A NaN or an infinity in u at a fluid node makes Monitor return
    IMEXLB_E_NAN (a diverged lattice is not reported as Umax = 0); one at
    the first fluid node, which the reference's Monitor skips
    (lbm.f90:212), is not seen; after InitPDF and a few steps the NaN has
    spread and Monitor reports it again."""
    deck = get_deck("tiny2d")
    lat, d = I.setup(deck, r_vec, device=0)
    lat.advance(5)
    p, u = lat.download_fields(I.LAYOUT_REF)
    fid = lat.geometry_list("fluid_id")
    base = lat.monitor()
    for bad in (np.nan, np.inf, -np.inf):
        u2 = u.copy()
        u2[2 * fid[len(fid) // 2] + 1] = bad  # u_y of a node mid-list
        lat.upload_fields(p, u2, I.LAYOUT_REF)
        with pytest.raises(I.ImexlbError) as err:
            lat.monitor()
        assert err.value.code == I.E_NAN and "1 fluid node" in str(err.value)
    u2 = u.copy()
    u2[2 * fid[0]] = np.nan  # the entry Monitor leaves out
    lat.upload_fields(p, u2, I.LAYOUT_REF)
    assert lat.monitor() == base
    u2[2 * fid[len(fid) // 2]] = np.nan
    lat.upload_fields(p, u2, I.LAYOUT_REF)
    lat.init_pdf()
    lat.advance(4)
    with pytest.raises(I.ImexlbError) as err:
        lat.monitor()
    assert err.value.code == I.E_NAN
    lat.close()
