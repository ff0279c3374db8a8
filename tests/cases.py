"""Shared test decks and helpers (4-record input format of input.in)."""
import numpy as np

import imexlb as I
import oracle as O

# small decks that the oracle runs in well under a second per step
TINY = {
    # cylinder of radius 6 at (12, 12) in a 49x25 channel
    "tiny2d": (2, "48 24\n0.25 0.25 20\n1.6 1 1\n0.02 0.001\n"),
    # sphere of radius 3 at (6, 6, 6) in a 25x13x13 box
    "tiny3d": (3, "24 12 12\n0.25 0.25 20\n1.6 1 1\n0.05 0.005\n"),
    # body cut by the inlet plane and both walls: disc of radius 10 centred
    # on the inlet column x = 0 of a 41x21 channel
    "inletbody2d": (2, "40 20\n0.5 0.0 20\n1.5 1 1\n0.02 0.001\n"),
    # sphere touching the z walls: radius 6 centred at (8, 6, 6) in 25x13x13
    "wallsphere3d": (3, "24 12 12\n0.5 0.333333333333333 20\n1.5 1 1\n0.02 0.001\n"),
    # sphere wider than the periodic y extent: radius 6 at (15, 4, 10) in
    # 31x9x21, so the body crosses both periodic y faces
    "ysphere3d": (3, "30 8 20\n0.4 0.5 20\n1.6 1 1\n0.02 0.001\n"),
    # sphere cut by the outlet plane: radius 3 centred at (24, 6, 6) in 25x13x13
    "outletsphere3d": (3, "24 12 12\n0.25 1.0 20\n1.6 1 1\n0.02 0.001\n"),
    # odd sizes, body off the half-integer centre line: 38x24 and 24x18x14
    "ragged2d": (2, "37 23\n0.2 0.3 20\n1.7 1 1\n0.02 0.001\n"),
    "ragged3d": (3, "23 17 13\n0.3 0.35 20\n1.55 1 1\n0.02 0.001\n"),
    # three rows of 62 nodes: row y = 1 is the source of both walls'
    # bounce-back copies, so one 64-cell tile of the device layout holds 379
    # boundary copies (more than the small-lattice step loads up front)
    "channel3": (2, "61 2\n0.02 0.5 2\n1.6 1 1\n0.02 0.001\n"),
}


def get_deck(name):
    if name in TINY:
        dim, text = TINY[name]
        return I.parse_deck(text, dim)
    return I.deck(name)


def oracle_setup(deck, r_vec):
    d = I.derive(deck)
    o = O.Oracle(deck.dim, *deck.n)
    o.set_params(deck.ome, deck.rho0, d.uu)
    o.set_geometry(d.pos, d.radius)
    o.init_up(r_vec)
    o.init_pdf()
    return o, d


def interior_mask(o):
    ids = np.arange(o.A)
    nx, ny, nz = o.n
    x = ids % o.dyy - 3
    y = (ids // o.dyy) % (ny + 7) - 3
    m = (x >= 0) & (x <= nx) & (y >= 0) & (y <= ny)
    if o.dim == 3:
        z = ids // o.dzz - 3
        m &= (z >= 0) & (z <= nz)
    return m


def assert_state_equal(lat, o, what=("f", "p", "u"), chunk=1 << 22):
    """Bit-for-bit equality (as IEEE values; +0 == -0) of f, p, u on interior
    cells, compared in chunks of cells so full-size lattices need no copies."""
    m = interior_mask(o)

    def compare(name, dev, ref, ncomp):
        dev = dev.reshape(o.A, ncomp)
        ref = ref.reshape(o.A, ncomp)
        nbad, first = 0, []
        for lo in range(0, o.A, chunk):
            hi = min(lo + chunk, o.A)
            mm = m[lo:hi]
            bad = np.nonzero(mm & ~(dev[lo:hi] == ref[lo:hi]).all(axis=1))[0]
            nbad += bad.size
            first.extend((bad[:5] + lo).tolist())
        assert nbad == 0, f"{name} differs on {nbad} interior cells, first ids {first[:5]}"

    if "f" in what:
        compare("f", lat.download_pdf(I.PDF_F, I.LAYOUT_REF), o.f, o.nq)
    if "p" in what or "u" in what:
        p_dev, u_dev = lat.download_fields(I.LAYOUT_REF)
        if "p" in what:
            compare("p", p_dev, o.p, 1)
        if "u" in what:
            compare("u", u_dev, o.u, o.dim)
