"""Pins the oracle's BoundaryCondition + Propagation (lbm.f90:96-155, 80-92;
PassF cart_mpi.f90:240-277) with hand-derived known answers.

Every fluid slot of fb gets a unique value, ghost cells get garbage and solid
cells 0; after BoundaryCondition and Propagation the populations that fluid
nodes next to a wall, the inlet, the outlet and the body pulled in must be the
values the reference's copy rules name, derived by hand below.
"""
import numpy as np
import pytest

import imexlb as I
import oracle as O


def _setup(dim, n, pos, radius):
    o = O.Oracle(dim, *n)
    o.set_params(1.5, 1.0, 0.01)
    o.set_geometry(pos, radius)
    fb = o.fb.reshape(o.A, o.nq)
    fb[:] = -7.0e9  # garbage in every ghost cell
    ids = np.arange(o.A)
    x = ids % o.dyy - 3
    y = (ids // o.dyy) % (n[1] + 7) - 3
    inside = (x >= 0) & (x <= n[0]) & (y >= 0) & (y <= n[1])
    if dim == 3:
        z = ids // o.dzz - 3
        inside &= (z >= 0) & (z <= n[2])
    fb[inside & (o.geo == 0)] = 0.0
    fid = o.list("fluid_id")
    fb[fid] = (fid[:, None] * o.nq + np.arange(o.nq)[None, :] + 1).astype(float)
    o.boundary_condition()
    o.propagation()
    return o


def val(o, cell, q):
    """unique value that the post-collision slot (cell, q) was given"""
    return float(o.id(*cell) * o.nq + q + 1)


def f_at(o, cell, q):
    return o.f[o.id(*cell) * o.nq + q]


def q2(ex, ey):
    return (ex + 1) * 3 + (ey + 1)


def test_2d_known_answers():
    nx, ny = 30, 16
    o = _setup(2, (nx, ny), pos=12.0, radius=3.0)
    # Down wall: (0,0) pulls q=(-1,+1) from ghost (1,-1); the Down phase copies
    # (1,-1)+2e_q = (-1,1) with the opposite direction, and (-1,1) holds the
    # periodic image (nx,1) because the inlet phase runs later.
    assert f_at(o, (0, 0), q2(-1, 1)) == val(o, (nx, 1), q2(1, -1))
    # Down wall away from the corners: (x,0) pulls q=(ex,+1) from (x-ex,-1),
    # mirrored from (x+ex, 1) with the opposite direction.
    for ex in (-1, 0, 1):
        assert f_at(o, (7, 0), q2(ex, 1)) == val(o, (7 + ex, 1), q2(-ex, -1))
    # Up wall: (x,ny) pulls q=(ex,-1) from (x-ex,ny+1) = mirror of (x+ex, ny-1).
    for ex in (-1, 0, 1):
        assert f_at(o, (20, ny), q2(ex, -1)) == val(o, (20 + ex, ny - 1), q2(-ex, 1))
    # Inlet: (0,y) pulls q=(+1,ey) from (-1,y-ey), a zero-gradient copy of (0,y-ey).
    for ey in (-1, 0, 1):
        assert f_at(o, (0, 8), q2(1, ey)) == val(o, (0, 8 - ey), q2(1, ey))
    # Outlet: (nx,y) pulls q=(-1,ey) from (nx+1,y-ey), which copies
    # (nx+1,y-ey)+e_q = (nx,y): the node keeps its own post-collision value.
    for ey in (-1, 0, 1):
        assert f_at(o, (nx, 8), q2(-1, ey)) == val(o, (nx, 8), q2(-1, ey))
    # Bulk: plain pull.
    assert f_at(o, (25, 3), q2(1, 1)) == val(o, (24, 2), q2(1, 1))
    o.close()


def test_2d_body_mirror_and_zero():
    nx, ny = 30, 16
    o = _setup(2, (nx, ny), pos=12.0, radius=3.0)
    bu = set(o.list("b_user").tolist())
    geo = o.geo
    n_mirror = n_zero = 0
    for c in o.list("fluid_id"):
        x, y = c % o.dyy - 3, c // o.dyy - 3
        for q in range(9):
            ex, ey = o.e[q]
            s = o.id(x - ex, y - ey)
            if not (0 <= x - ex <= nx and 0 <= y - ey <= ny) or geo[s] != 0:
                continue
            got = f_at(o, (x, y), q)
            if s in bu:  # User phase: fb(s,q) = fb(s + 2 e_q, opposite q)
                assert got == val(o, (x - ex + 2 * ex, y - ey + 2 * ey), 8 - q)
                n_mirror += 1
            else:  # never-written solid cell: rest state 0
                assert got == 0.0
                n_zero += 1
    assert n_mirror > 0 and n_zero > 0
    o.close()


def q3(ex, ey, ez):
    return (ez + 1) * 9 + (ex + 1) * 3 + (ey + 1)


def test_3d_known_answers():
    nx, ny, nz = 12, 10, 8
    o = _setup(3, (nx, ny, nz), pos=5.0, radius=2.0)
    # y is periodic in 3D (no y-wall phase): (x,0,z) pulls q=(ex,+1,ez) from
    # (x-ex,-1,z-ez), the periodic image (x-ex, ny, z-ez).
    for ex, ez in [(0, 0), (1, 0), (-1, 1), (1, -1)]:
        assert f_at(o, (6, 0, 4), q3(ex, 1, ez)) == val(o, (6 - ex, ny, 4 - ez), q3(ex, 1, ez))
    # z walls mirror: (x,y,0) pulls q=(ex,ey,+1) from (x-ex,y-ey,-1), mirrored
    # from (x+ex, y+ey, 1) with the opposite direction.
    for ex, ey in [(0, 0), (1, -1), (-1, 1)]:
        assert f_at(o, (9, 5, 0), q3(ex, ey, 1)) == val(o, (9 + ex, 5 + ey, 1), q3(-ex, -ey, -1))
    # inlet zero gradient and outlet self-copy
    assert f_at(o, (0, 5, 4), q3(1, 1, -1)) == val(o, (0, 4, 5), q3(1, 1, -1))
    assert f_at(o, (nx, 5, 4), q3(-1, 0, 1)) == val(o, (nx, 5, 4), q3(-1, 0, 1))
    o.close()


def test_relevant_ghost_reads_resolve_to_interior():
    """After BoundaryCondition no garbage ghost value reaches a fluid node."""
    for dim, n, pos, r in [(2, (30, 16), 12.0, 3.0), (3, (12, 10, 8), 5.0, 2.0)]:
        o = _setup(dim, n, pos, r)
        fid = o.list("fluid_id")
        f = o.f.reshape(o.A, o.nq)[fid]
        assert np.all(f > -1.0), "a stale ghost value was streamed into the fluid"
        o.close()


@pytest.mark.parametrize("name", ["cyl400x100", "input.in"])
def test_monitor_series_is_deterministic(name, r_vec):
    deck = I.deck(name)
    d = I.derive(deck)
    runs = []
    for _ in range(2):
        o = O.Oracle(deck.dim, *deck.n)
        o.set_params(deck.ome, deck.rho0, d.uu)
        o.set_geometry(d.pos, d.radius)
        o.init_up(r_vec)
        o.init_pdf()
        runs.append(o.run(0, 2 * d.interv, d.interv))
        o.close()
    assert np.array_equal(runs[0], runs[1]) and len(runs[0]) == 3
