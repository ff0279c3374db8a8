"""Pins the oracle's PostProcessing (lbm.f90:159-201; lbm_3d.f90:161-203)
by closed forms that do not restate its floating-point expressions.

* Moment pass: on a textbook equilibrium built here (w_q, e_q from the
  direction picture, not from the oracle), p = sum f / 3 and
  u = sum f e / rho0 come back as the p, u it was built from, with
  rho0 = 1.3 (a dropped division by rho0 would return 1.3 u); a single
  population f_q = 1 gives p = 1/3 and u = e_q / rho0.
* Inlet (lbm.f90:184-192): at every cell bl(idn)+1, i.e. x = 0 of every row
  the x = -1 ring holds, ghost rows included, u_x equals the exact rational
  uu rho0 4 h (n - h) / n^2 (h = y in 2D, z in 3D) to within two roundings;
  it is exactly uu rho0 at h = n/2 for n a power of two, exactly 0 at
  h = 0 and h = n, negative at h = -1; u_y = 0 there; in 3D u_z keeps its
  moment value.
* Outlet (lbm.f90:194-199): p = 0 at every cell br(idn)-1 (x = nx), ghost
  rows included, while x = nx-1 keeps its moment and x = nx+1 is untouched.

Arrays are filled with sentinels first, so a write to the wrong column,
row or component shows.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

SENT_P, SENT_U = 7.0, 9.0


def _weights(dim):
    """Lattice weights and velocities from the direction pictures
    initialization.f90:106-111 (D2Q9) and the D3Q27 tensor product, in the
    reference numbering, independently of the oracle's tables."""
    e, w = [], []
    if dim == 2:
        for q in range(9):
            ex, ey = q // 3 - 1, q % 3 - 1
            e.append((ex, ey))
            w.append({0: 4 / 9, 1: 1 / 9, 2: 1 / 36}[abs(ex) + abs(ey)])
    else:
        for q in range(27):
            ex, ey, ez = (q % 9) // 3 - 1, q % 3 - 1, q // 9 - 1
            e.append((ex, ey, ez))
            w.append({0: 8 / 27, 1: 2 / 27, 2: 1 / 54, 3: 1 / 216}[abs(ex) + abs(ey) + abs(ez)])
    return np.array(e, dtype=float), np.array(w)


def _lattice(dim, n, rho0=1.3, uu=0.04):
    o = O.Oracle(dim, *n)
    o.set_params(1.5, rho0, uu)
    # a small body well inside the box, so solid nodes exist too
    o.set_geometry(n[0] / 2.0, 2.0)
    return o


def _coords(o, ids):
    ids = np.asarray(ids)
    x = ids % o.dyy - 3
    y = (ids // o.dyy) % (o.n[1] + 7) - 3
    z = ids // o.dzz - 3 if o.dim == 3 else np.zeros_like(ids)
    return x, y, z


def _fill_sentinels(o):
    o.p[:] = SENT_P
    o.u[:] = SENT_U


@pytest.mark.parametrize("dim,n", [(2, (20, 16, 0)), (3, (12, 10, 8))])
def test_moments_of_a_textbook_equilibrium_with_rho0(dim, n):
    rho0 = 1.3
    o = _lattice(dim, n, rho0=rho0)
    e, w = _weights(dim)
    assert np.array_equal(e, o.e.astype(float)) and np.allclose(w, o.t, rtol=1e-15, atol=0)
    rng = np.random.default_rng(11)
    fid = o.list("fluid_id")
    p = rng.uniform(-0.01, 0.01, fid.size)
    u = rng.uniform(-0.05, 0.05, (fid.size, dim))
    eu = u @ e.T
    uu2 = (u * u).sum(1)
    feq = w[None, :] * (3 * p[:, None] + rho0 * (3 * eu + 4.5 * eu ** 2 - 1.5 * uu2[:, None]))
    _fill_sentinels(o)
    f = o.f.reshape(o.A, o.nq)
    f[:] = 0.0
    f[fid] = feq
    o.moments()
    assert np.allclose(o.p[fid], p, rtol=0, atol=1e-16)
    assert np.allclose(o.u.reshape(o.A, dim)[fid], u, rtol=0, atol=1e-16)
    # solid nodes are not in the loop: their sentinels stay
    solid = np.nonzero((o.geo == 0) & (_coords(o, np.arange(o.A))[0] >= 0))[0]
    inner = solid[np.all([(c >= 0) & (c <= m) for c, m in zip(_coords(o, solid), o.n)][:dim], axis=0)]
    assert inner.size > 0 and np.all(o.p[inner] == SENT_P)
    o.close()


@pytest.mark.parametrize("dim,n", [(2, (20, 16, 0)), (3, (12, 10, 8))])
def test_moments_of_one_population(dim, n):
    """f_q = 1 alone: p = 1/3 and u = e_q / rho0, for every q."""
    rho0 = 1.3
    o = _lattice(dim, n, rho0=rho0)
    e, _ = _weights(dim)
    fid = o.list("fluid_id")
    node = fid[len(fid) // 2]
    f = o.f.reshape(o.A, o.nq)
    for q in range(o.nq):
        f[:] = 0.0
        f[node, q] = 1.0
        o.moments()
        assert o.p[node] == 1.0 / 3.0
        assert np.array_equal(o.u.reshape(o.A, dim)[node], e[q] / rho0), q
    o.close()


def _exact_inlet(uu, rho0, h, n):
    """uu rho0 4 h (n - h) / n^2 with uu rho0 rounded once (as the first
    product of the reference's expression) and the rest exact."""
    return Fraction(uu * rho0) * 4 * h * (n - h) / (n * n)


@pytest.mark.parametrize("dim,n", [(2, (20, 16, 0)), (2, (19, 23, 0)), (3, (12, 10, 8)), (3, (9, 6, 11))])
def test_inlet_profile_and_outlet_pressure(dim, n):
    rho0, uu = 1.3, 0.04
    o = _lattice(dim, n, rho0=rho0, uu=uu)
    rng = np.random.default_rng(5)
    fid = o.list("fluid_id")
    f = o.f.reshape(o.A, o.nq)
    f[:] = 0.0
    f[fid] = rng.uniform(0.0, 0.2, (fid.size, o.nq))
    _fill_sentinels(o)
    o.moments()
    p_mom, u_mom = o.p.copy(), o.u.reshape(o.A, dim).copy()
    _fill_sentinels(o)
    o.post_processing()
    p, u = o.p, o.u.reshape(o.A, dim)
    nh = n[1] if dim == 2 else n[2]

    bl = o.list("bl")
    assert bl.size > 0
    x, y, z = _coords(o, bl)
    assert np.all(x == -1)
    h_of = y if dim == 2 else z
    # the list reaches the ghost rows on both sides of the channel
    assert h_of.min() == -1 and h_of.max() >= nh + 1
    for cell, h in zip(bl + 1, h_of):
        got = u[cell, 0]
        ex = _exact_inlet(uu, rho0, int(h), nh)
        # two roundings after the first product, each within half an ulp
        assert abs(Fraction(got) - ex) <= abs(ex) * Fraction(1, 2 ** 51), (h, got, float(ex))
        if h in (0, nh):
            assert got == 0.0
        if h < 0 or h > nh:
            assert got < 0.0
        assert u[cell, 1] == 0.0
        if dim == 3:
            # u_z is not reset by the inlet (lbm_3d.f90:188-193)
            assert u[cell, 2] == (u_mom[cell, 2] if cell in set(fid) else SENT_U)
    # the ring x = -1 itself and the column x = 1 are not written by the inlet
    assert np.all(u[bl, 0] == SENT_U)
    col1 = bl + 2
    fset = set(fid.tolist())
    for cell in col1:
        assert u[cell, 0] == (u_mom[cell, 0] if cell in fset else SENT_U)

    br = o.list("br")
    x, y, z = _coords(o, br)
    assert np.all(x == o.n[0] + 1)
    assert np.all(p[br - 1] == 0.0)                  # x = nx, every row of the ring
    assert np.all(p[br] == SENT_P)                   # x = nx + 1 untouched
    inner = [c for c in br - 2 if c in fset]         # x = nx - 1 keeps its moment
    assert inner and all(p[c] == p_mom[c] for c in inner)
    assert any(p_mom[c] != 0.0 for c in inner)
    o.close()


def test_inlet_centre_is_exactly_uu_rho0():
    """ny a power of two: every factor after uu*rho0 is a power of two, so
    u(0, ny/2) == uu*rho0 exactly (2D); likewise u(0, y, nz/2) in 3D."""
    for dim, n in ((2, (20, 16, 0)), (3, (12, 10, 8))):
        rho0, uu = 1.3, 0.04
        o = _lattice(dim, n, rho0=rho0, uu=uu)
        o.post_processing()
        nh = n[1] if dim == 2 else n[2]
        bl = o.list("bl")
        x, y, z = _coords(o, bl)
        h_of = y if dim == 2 else z
        u = o.u.reshape(o.A, dim)
        centre = bl[h_of == nh // 2] + 1
        assert centre.size > 0 and np.all(u[centre, 0] == uu * rho0)
        o.close()
