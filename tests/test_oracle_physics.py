"""Pins the oracle's arithmetic (InitPDF, Collision, Propagation, the moment
pass of PostProcessing, Monitor) by what the mathematics fixes:

* moments of the equilibrium: sum f = 3p, sum f e = rho0 u and
  sum f e e = p I + rho0 u u (the quadratic terms of lbm.f90:70 are needed
  for the last one);
* BGK collision conserves p and u and leaves an equilibrium in place;
* a single population streams to x + e_q;
* a shear wave decays at the rate nu k^2 with nu = (1/ome - 1/2)/3 in a
  periodic box (2D and 3D);
* Monitor is the largest |(u_x, u_y)| over fluid entries 1..F-1.
"""
import numpy as np
import pytest

import oracle as O


def _box(dim, n, ome=1.2, rho0=1.0):
    """Oracle lattice with no body (centre far outside the box)."""
    o = O.Oracle(dim, *n)
    o.set_params(ome, rho0, 0.0)
    o.set_geometry(-1.0e6, 1.0)
    assert len(o.list("fluid_id")) == np.prod([k + 1 for k in n[:dim]])
    return o


def _random_state(o, rng, umax=0.08):
    fid = o.list("fluid_id")
    p = o.p
    u = o.u.reshape(o.A, o.dim)
    p[fid] = rng.uniform(-0.01, 0.01, fid.size)
    u[fid] = rng.uniform(-umax, umax, (fid.size, o.dim))
    return fid


@pytest.mark.parametrize("dim", [2, 3])
def test_equilibrium_moments(dim):
    rng = np.random.default_rng(3)
    o = _box(dim, (9, 7, 5), rho0=1.3)
    fid = _random_state(o, rng)
    o.init_pdf()
    f = o.f.reshape(o.A, o.nq)[fid]
    p = o.p[fid]
    u = o.u.reshape(o.A, dim)[fid]
    e = o.e.astype(float)
    assert np.allclose(f.sum(1), 3 * p, atol=1e-15, rtol=0)
    assert np.allclose(f @ e, 1.3 * u, atol=1e-15, rtol=0)
    second = np.einsum("nq,qa,qb->nab", f, e, e)
    expect = p[:, None, None] * np.eye(dim) + 1.3 * np.einsum("na,nb->nab", u, u)
    assert np.allclose(second, expect, atol=1e-15, rtol=0)
    o.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_collision_conserves_and_fixes_equilibrium(dim):
    rng = np.random.default_rng(4)
    o = _box(dim, (8, 6, 4), ome=1.7)
    fid = _random_state(o, rng)
    o.init_pdf()
    # equilibrium is a fixed point of the relaxation
    o.collision()
    assert np.allclose(o.fb.reshape(o.A, o.nq)[fid], o.f.reshape(o.A, o.nq)[fid], atol=1e-16, rtol=1e-14)
    # perturb f away from equilibrium, take its moments, relax: moments kept
    f = o.f.reshape(o.A, o.nq)
    f[fid] += rng.uniform(-1e-3, 1e-3, (fid.size, o.nq))
    o.moments()
    o.collision()
    fb = o.fb.reshape(o.A, o.nq)[fid]
    e = o.e.astype(float)
    assert np.allclose(fb.sum(1), f[fid].sum(1), atol=1e-15, rtol=0)
    assert np.allclose(fb @ e, f[fid] @ e, atol=1e-15, rtol=0)
    assert not np.allclose(fb, f[fid], atol=1e-6)
    o.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_streaming_moves_population_by_e(dim):
    o = _box(dim, (10, 10, 10))
    c0 = (5, 4, 6)[:dim]
    for q in range(o.nq):
        fb = o.fb
        fb[:] = 0.0
        fb[o.id(*c0) * o.nq + q] = 1.0
        o.propagation()
        f = o.f.reshape(o.A, o.nq)
        hit = np.argwhere(f != 0.0)
        target = tuple(c + int(o.e[q, k]) for k, c in enumerate(c0))
        assert hit.tolist() == [[o.id(*target), q]]
    o.close()


def _shear_decay(dim, ome, steps):
    n = (7, 63, 0) if dim == 2 else (3, 3, 31)
    o = _box(dim, n, ome=ome)
    fid = o.list("fluid_id")
    ids = fid
    coord = ((ids // o.dyy) % (n[1] + 7) - 3) if dim == 2 else (ids // o.dzz - 3)
    period = (n[1] + 1) if dim == 2 else (n[2] + 1)
    k = 2 * np.pi / period
    U = 1e-4
    u = o.u.reshape(o.A, dim)
    u[fid] = 0.0
    u[fid, 0] = U * np.sin(k * coord)
    o.p[fid] = 0.0
    o.init_pdf()
    amps = []
    for s in range(steps + 1):
        amps.append(2.0 / coord.size * np.sum(u[fid, 0] * np.sin(k * coord)))
        if s == steps:
            break
        o.collision()
        o.pass_f()
        o.propagation()
        o.moments()
    amps = np.array(amps)
    t = np.arange(len(amps))
    # fit log amplitude after the first steps (the initial f lacks the
    # non-equilibrium part of a developed shear flow)
    slope = np.polyfit(t[20:], np.log(amps[20:]), 1)[0]
    o.close()
    return -slope / k ** 2, (1.0 / ome - 0.5) / 3.0


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("ome", [0.8, 1.5])
def test_shear_wave_viscosity(dim, ome):
    nu_meas, nu_theory = _shear_decay(dim, ome, 300)
    assert nu_meas == pytest.approx(nu_theory, rel=0.01)


def test_monitor_is_max_speed_skipping_first_fluid_entry():
    rng = np.random.default_rng(5)
    o = _box(2, (12, 9))
    fid = _random_state(o, rng)
    u = o.u.reshape(o.A, 2)
    u[fid[0]] = (10.0, 10.0)  # the first fluid entry is not part of the maximum
    speeds = np.sqrt(u[fid[1:], 0] ** 2 + u[fid[1:], 1] ** 2)
    assert o.monitor() == speeds.max()
    o.close()


def test_forces_known_answers():
    """Momentum exchange on hand-set populations: one link to a plain solid
    cell gives e_q * fb(x,q); a link to a body-list cell adds the value the
    User phase mirrored into it; a lattice at rest feels no force."""
    o = O.Oracle(2, 30, 16)
    o.set_params(1.5, 1.0, 0.0)
    o.set_geometry(12.0, 3.0)
    o.boundary_condition()
    assert np.array_equal(o.forces(), [0.0, 0.0])
    fb = o.fb.reshape(o.A, o.nq)
    geo = o.geo
    users = set(o.list("b_user").tolist())
    fid = o.list("fluid_id")
    plain = body = None
    for c in fid:
        for q in range(9):
            nb = c + o.e[q, 0] + o.dyy * o.e[q, 1]
            if geo[nb] == 0:
                if nb in users and body is None:
                    body = (c, q, nb)
                elif nb not in users and plain is None:
                    plain = (c, q, nb)
    c, q, nb = plain
    fb[:] = 0.0
    fb[c, q] = 0.25
    assert np.array_equal(o.forces(), 0.25 * o.e[q])
    c, q, nb = body
    fb[:] = 0.0
    fb[c, q] = 0.5
    fb[nb, 8 - q] = 0.125  # what the User phase left in the body cell
    assert np.array_equal(o.forces(), (0.5 + 0.125) * o.e[q])
    o.close()


def test_forces_drag_positive_in_stream(r_vec):
    """A cylinder in the channel flow is pushed downstream."""
    import imexlb as I
    deck = I.deck("cyl400x100")
    d = I.derive(deck)
    o = O.Oracle(2, *deck.n)
    o.set_params(deck.ome, deck.rho0, d.uu)
    o.set_geometry(d.pos, d.radius)
    o.init_up(r_vec)
    o.init_pdf()
    o.run(0, 199, 0)
    F = o.forces()
    assert F[0] > 0 and abs(F[1]) < F[0]
    o.close()
