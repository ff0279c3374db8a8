"""Pins the oracle's set-up functions (DataInput, InitLattice, SetGeometry,
PassInt, InitUP) against facts fixed independently of its text:

* derived values of five input decks computed in IEEE double, with the
  integer truncation of max_step (19,999 for the 400x100 cylinder, not 20,000);
* the direction picture at initialization.f90:106-111 and the lattice
  isotropy identities;
* fluid / solid / body-list / link counts of five decks from an independent
  analysis restatement of the reference's scan;
* the periodic image rule of PassInt and the disc/sphere geometry.
"""
import numpy as np
import pytest

import imexlb as I
import oracle as O

# (nu, uu, t_intv, max_step, interv) per deck, IEEE double
KNOWN_DERIVED = {
    "input.in": (0.004, 0.025, 64000.0, 32000, 640),
    "input_3d.in": (0.004, 0.075, 64000.0, 6400, 640),
    "cyl400x100": (0.01, 0.05, 40000.00000000001, 19999, 100),
    "sphere512x256x256": (0.008, 0.04, 200000.0000000002, 4999, 100),
    "cyl8192x2048": (0.0064, 0.05, 2560000.000000001, 49999, 100),
}

# nodes, array_size, fluid, solid, b_user size, mirror links, zero links
KNOWN_COUNTS = {
    "cyl400x100": (40501, 43549, 40184, 317, 19, 64, 136),
    "input.in": (66177, 70065, 65980, 197, 17, 61, 99),
    "input_3d.in": (8536833, 9458775, 8534724, 2109, 197, 2213, 5109),
}
KNOWN_COUNTS_LARGE = {
    "cyl8192x2048": (16787457, 16848945, 16774604, 12853, 92, 301, 939),
    "sphere512x256x256": (33883137, 35898711, 33849736, 33401, 1257, 13085, 33021),
}


@pytest.mark.parametrize("name", sorted(KNOWN_DERIVED))
def test_data_input_known_values(name):
    deck = I.deck(name)
    d = O.data_input(deck.n[0], deck.Lc, deck.Xpos, deck.Re, deck.ome, deck.dt, deck.rho0, deck.max_t,
                     deck.inter)
    nu, uu, t_intv, max_step, interv = KNOWN_DERIVED[name]
    assert d.nu == pytest.approx(nu, rel=1e-13)
    assert d.uu == pytest.approx(uu, rel=1e-13)
    assert d.t_intv == pytest.approx(t_intv, rel=1e-15)
    assert d.max_step == max_step and d.interv == interv
    assert d.radius == d.charlength / 2 and d.pos == deck.Xpos * deck.n[0]
    # the product's own DataInput agrees exactly
    p = I.derive(deck)
    assert (p.nu, p.uu, p.t_intv, p.max_step, p.interv, p.radius, p.pos) == (
        d.nu, d.uu, d.t_intv, d.max_step, d.interv, d.radius, d.pos)


def test_lattice_picture_2d():
    e, t = O.init_lattice(2)
    # initialization.f90:106-111:   2 5 8 / 1 4 7 / 0 3 6  (x to the right, y up)
    picture = [[2, 5, 8], [1, 4, 7], [0, 3, 6]]
    for row, ey in zip(picture, (1, 0, -1)):
        for q, ex in zip(row, (-1, 0, 1)):
            assert tuple(e[q]) == (ex, ey)
    assert t[4] == 4.0 / 9.0 and t[1] == 1.0 / 9.0 and t[0] == 1.0 / 36.0


@pytest.mark.parametrize("dim", [2, 3])
def test_lattice_isotropy(dim):
    e, t = O.init_lattice(dim)
    nq = len(t)
    assert abs(t.sum() - 1.0) < 1e-15
    assert np.allclose(t @ e, 0.0, atol=1e-16)
    second = np.einsum("q,qa,qb->ab", t, e, e)
    assert np.allclose(second, np.eye(dim) / 3.0, atol=1e-15)
    fourth = np.einsum("q,qa,qb,qc,qd->abcd", t, e, e, e, e)
    d = np.eye(dim)
    iso = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d)) / 9
    assert np.allclose(fourth, iso, atol=1e-15)
    for q in range(nq):  # opposite direction is nq-1-q
        assert np.array_equal(e[nq - 1 - q], -e[q])
    if dim == 3:  # initialization_3d.f90:114: q = (k+1)*9 + (i+1)*3 + (j+1)
        for q in range(27):
            i, j, k = e[q]
            assert q == (k + 1) * 9 + (i + 1) * 3 + (j + 1)


def _geometry(name):
    deck = I.deck(name)
    d = I.derive(deck)
    o = O.Oracle(deck.dim, *deck.n)
    o.set_params(deck.ome, deck.rho0, d.uu)
    o.set_geometry(d.pos, d.radius)
    return o, deck, d


def _links(o):
    """Count fluid->solid pull links whose solid end is / is not in b_user."""
    fid = o.list("fluid_id")
    geo = o.geo
    bu = np.zeros(o.A, dtype=bool)
    bu[o.list("b_user")] = True
    mirror = zero = 0
    for q in range(o.nq):
        off = o.e[q, 0] + o.dyy * o.e[q, 1] + (o.dzz * o.e[q, 2] if o.dim == 3 else 0)
        s = fid - off
        s = s[geo[s] == 0]
        mirror += int(bu[s].sum())
        zero += int((~bu[s]).sum())
    return mirror, zero


def _check_counts(name, known):
    o, deck, d = _geometry(name)
    nodes, A, F, solid, nuser, mirror, zero = known
    assert I.nodes(deck) == nodes and o.A == A
    assert len(o.list("fluid_id")) == F and nodes - F == solid
    assert len(o.list("b_user")) == nuser
    assert o.ub_reads == 0
    if mirror is not None:
        assert _links(o) == (mirror, zero)
    o.close()


@pytest.mark.parametrize("name", sorted(KNOWN_COUNTS))
def test_geometry_counts(name):
    _check_counts(name, KNOWN_COUNTS[name])


@pytest.mark.slow
@pytest.mark.parametrize("name", sorted(KNOWN_COUNTS_LARGE))
def test_geometry_counts_large(name):
    known = KNOWN_COUNTS_LARGE[name]
    _check_counts(name, known[:5] + (None, None))


def test_mask_is_the_disc_and_lists_are_the_ring():
    o, deck, d = _geometry("cyl400x100")
    nx, ny, _ = deck.n
    ids = np.arange(o.A)
    x = ids % o.dyy - 3
    y = ids // o.dyy - 3
    inside = (x >= 0) & (x <= nx) & (y >= 0) & (y <= ny)
    r2 = (x - d.pos) ** 2 + (y - ny / 2) ** 2
    geo = o.geo
    assert np.all(geo[inside & (r2 > d.radius ** 2 + 1e-9)] == 1)
    assert np.all(geo[inside & (r2 < d.radius ** 2 - 1e-9)] == 0)
    # fluid_id: the fluid interior cells in increasing id (scan order)
    assert np.array_equal(o.list("fluid_id"), ids[inside & (geo == 1)])
    # wall lists: the ghost ring x=-1 / x=nx+1 / y=-1 / y=ny+1 over [-1, n+2]
    ring = (x >= -1) & (x <= nx + 2) & (y >= -1) & (y <= ny + 2) & ~inside
    assert np.array_equal(o.list("bl"), ids[ring & (x == -1)])
    assert np.array_equal(o.list("br"), ids[ring & (x == nx + 1)])
    assert np.array_equal(o.list("bd"), ids[ring & (y == -1)])
    assert np.array_equal(o.list("bu"), ids[ring & (y == ny + 1)])
    assert len(o.list("bl")) == ny + 4 and len(o.list("bu")) == nx + 4
    o.close()


def test_pass_int_is_the_periodic_image():
    o, deck, d = _geometry("input.in")
    nx, ny, _ = deck.n
    geo = o.geo.reshape(ny + 7, nx + 7)
    Lx, Ly = nx + 1, ny + 1
    for yy in range(-3, ny + 4):
        for xx in (-3, -2, -1, nx + 1, nx + 2, nx + 3):
            assert geo[yy + 3, xx + 3] == geo[(yy % Ly) + 3, (xx % Lx) + 3]
    for xx in range(-3, nx + 4):
        for yy in (-3, -1, ny + 1, ny + 3):
            assert geo[yy + 3, xx + 3] == geo[(yy % Ly) + 3, (xx % Lx) + 3]
    o.close()


def test_body_list_rule_2d():
    """Each b_user entry is, in its x-row, the first solid cell of the scan
    range whose cell at flat offset e_x(q) + dy*(2q+1) is fluid for some q
    (initialization.f90:232-253), and no other row has such a cell."""
    o, deck, d = _geometry("cyl400x100")
    nx, ny, _ = deck.n
    geo = o.geo
    ex = o.e[:, 0]
    found = []
    for yy in range(-1, ny + 2):
        for xx in range(-1, nx + 3):
            c = o.id(xx, yy)
            if geo[c] != 0:
                continue
            nb = c + ex + o.dyy * (2 * np.arange(9) + 1)
            nb = nb[(nb >= 0) & (nb < o.A)]
            if np.any(geo[nb] == 1):
                found.append(c)
                break
    assert np.array_equal(o.list("b_user"), np.array(found))
    o.close()


def test_init_up_closed_form(r_vec):
    o, deck, d = _geometry("cyl400x100")
    o.init_up(r_vec)
    nx, ny, _ = deck.n
    for (xx, yy) in [(0, 0), (0, 37), (123, 50), (400, 99), (80, 50)]:
        c = o.id(xx, yy)
        r = 1.0 + (r_vec[c % 100] - 0.5) * 2.0 * 0.2
        expect = d.uu * 4.0 * yy * (ny - yy) / ny ** 2 * o.geo[c] * r
        assert o.u[2 * c] == pytest.approx(expect, rel=1e-15, abs=0)
        assert o.u[2 * c + 1] == 0.0 and o.p[c] == 0.0
    assert o.u[2 * o.id(80, 50)] == 0.0  # cylinder centre is solid
    o.close()


# ---------------------------------------------------------------------------
# elliptic bodies (an extension for the warm-start study; no reference)

def _ellipse_oracle(aspect, angle, nx=96, ny=48, pos=40.0, radius=9.0):
    o = O.Oracle(2, nx, ny, 0)
    o.set_params(1.5, 1.0, 0.05)
    o.set_geometry(pos, radius, aspect, angle)
    return o


def _solid_mask(o):
    nx, ny = o.n[0], o.n[1]
    ids = np.array([[o.id(x, y, 0) for x in range(nx + 1)] for y in range(ny + 1)])
    return o.geo[ids] == 0


def test_ellipse_unit_aspect_is_the_circle():
    a = _ellipse_oracle(1.0, 0.0)
    b = O.Oracle(2, 96, 48, 0)
    b.set_params(1.5, 1.0, 0.05)
    b.set_geometry(40.0, 9.0)
    assert np.array_equal(_solid_mask(a), _solid_mask(b))
    a.close()
    b.close()


@pytest.mark.parametrize("aspect,angle", [(2.0, 0.0), (0.5, 30.0), (1.7, 75.0)])
def test_ellipse_area_and_symmetry(aspect, angle):
    """Solid node count ~ pi r^2 (the ellipse keeps the circle's area), and
    the mask is symmetric under the half turn about the centre (40, 24)."""
    o = _ellipse_oracle(aspect, angle)
    m = _solid_mask(o)
    assert abs(m.sum() - np.pi * 81.0) < 0.06 * np.pi * 81.0, m.sum()
    # node (x, y) <-> (80 - x, 48 - y)
    assert np.array_equal(m[:, :81], m[::-1, 80::-1])
    o.close()


def test_ellipse_quarter_turn_swaps_the_axes():
    """(aspect, 90 deg) and (1/aspect, 0 deg) describe the same ellipse; cos 90
    deg is not exactly 0 in floating point, so allow a node or two to differ."""
    a = _solid_mask(_ellipse_oracle(2.0, 90.0))
    b = _solid_mask(_ellipse_oracle(0.5, 0.0))
    assert (a != b).sum() <= 2
