"""Host-side logic of the product, on the CPU (no compute calls without a GPU):

* libimexlb.so loads and exports every entry point include/imexlb.h declares;
* input decks parse like the reference's list-directed read;
* a host-only context (device -1) builds the reference's mask and lists
  exactly like the oracle's SetGeometry;
* the resolved boundary table names, for every slot a fluid node pulls from a
  ghost or body cell, the value the oracle's literal PassF + five copy phases
  leave there;
* errors are reported as status codes.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import imexlb as I
from imexlb import _lib
from cases import get_deck, interior_mask
import oracle as O

LISTS = ("fluid_id", "bl", "br", "bu", "bd", "bf", "bb", "b_user")


def test_library_exports_every_declared_symbol(root):
    declared = set()
    for h in ("imexlb.h", "imexlb_nn.h"):
        header = open(os.path.join(root, "include", h)).read()
        declared |= set(re.findall(r"^\s*(?:int|int64_t|const char \*)\s*(imexlb_\w+)\s*\(", header, re.M))
    assert len(declared) >= 45
    L = C.CDLL(_lib.LIB_PATH)
    for name in sorted(declared):
        assert hasattr(L, name), f"{name} declared but not exported"
    assert declared == set(_lib.SIGNATURES), "ctypes signatures out of step with the header"


def test_parse_deck_list_directed():
    text = "512\t128\n0.03125\t0.125\t100\n1.953125\t1\t1\n0.5\t0.01\n\n----\nnx ny\nL Xpos Re\n"
    d = I.parse_deck(text, 2)
    assert d.n == (512, 128, 0) and d.Re == 100.0 and d.max_t == 0.5 and d.inter == 0.01
    d3 = I.parse_deck("512 128\n128\n0.03125 0.125 300 extra\n1.953125,1,1\n0.1 0.01\n", 3)
    assert d3.n == (512, 128, 128) and d3.Re == 300.0 and d3.rho0 == 1.0
    assert I.parse_deck("8 8\n1.d0 0.5 10\n1.2 1 1\n1 0.1\n", 2).Lc == 1.0
    with pytest.raises(ValueError):
        I.parse_deck("8 8\n1 0.5 10\n", 2)


def _host_ctx(deck):
    d = I.derive(deck)
    lat = I.Lattice(deck.dim, *deck.n, device=-1)
    lat.set_params(deck.ome, deck.rho0, d.uu)
    lat.set_geometry(d.pos, d.radius)
    o = O.Oracle(deck.dim, *deck.n)
    o.set_params(deck.ome, deck.rho0, d.uu)
    o.set_geometry(d.pos, d.radius)
    return lat, o


@pytest.mark.parametrize("name", ["tiny2d", "tiny3d", "inletbody2d", "wallsphere3d", "ysphere3d", "outletsphere3d",
                                  "ragged2d", "ragged3d", "cyl400x100", "input.in"])
def test_geometry_matches_oracle(name):
    lat, o = _host_ctx(get_deck(name))
    assert np.array_equal(lat.geometry_geo(), o.geo)
    for k in LISTS:
        assert np.array_equal(lat.geometry_list(k), o.list(k)), k
    lat.close()
    o.close()


def _check_bc_table(lat, o):
    slot, src, stats = lat.bc_table()
    fb = o.fb.reshape(o.A, o.nq)
    fb[:] = -7.0e9
    m = interior_mask(o)
    fb[m & (o.geo == 0)] = 0.0
    fid = o.list("fluid_id")
    fb[fid] = (fid[:, None] * o.nq + np.arange(o.nq)[None, :] + 1).astype(float)
    # interior body cells carry the previous step's values: give them unique
    # negative values
    users = o.list("b_user")
    users = users[m[users]]
    fb[users] = -(users[:, None] * o.nq + np.arange(o.nq)[None, :] + 1).astype(float)
    o.boundary_condition()
    prev = -src - 2
    expect = np.where(src >= 0, (src + 1).astype(float), np.where(src == -1, 0.0, -(prev + 1).astype(float)))
    assert np.array_equal(o.fb[slot], expect)
    # completeness: the (fluid node, q) pulls from a ghost cell, plus every
    # direction of every interior body cell
    want = [(users[:, None] * o.nq + np.arange(o.nq)[None, :]).ravel()]
    for q in range(o.nq):
        off = o.e[q, 0] + o.dyy * o.e[q, 1] + (o.dzz * o.e[q, 2] if o.dim == 3 else 0)
        s = fid - off
        want.append(s[~m[s]] * o.nq + q)
    assert np.array_equal(np.sort(np.concatenate(want)), np.sort(slot))
    return stats


def _table_step(o, slot, src, bufs, cur):
    """One fused step as the device runs it (boundary fill from the table,
    pull, moments, inlet/outlet, collision into the other buffer), in numpy."""
    nq, dim = o.nq, o.dim
    fid = o.list("fluid_id")
    fb, prev = bufs[cur].reshape(-1), bufs[1 - cur].reshape(-1)
    fb[slot] = np.where(src >= 0, fb[np.maximum(src, 0)], np.where(src == -1, 0.0, prev[np.maximum(-src - 2, 0)]))
    offs = o.e[:, 0] + o.dyy * o.e[:, 1] + (o.dzz * o.e[:, 2] if dim == 3 else 0)
    f = np.stack([bufs[cur][fid - offs[q], q] for q in range(nq)], 1)
    return f, fid


def test_table_driven_steps_reproduce_oracle_populations(r_vec):
    """Pulling through the resolved table, buffer after buffer, gives the
    oracle's post-propagation populations for 30 steps on decks whose body
    touches the inlet plane or the walls."""
    for name in ("inletbody2d", "wallsphere3d"):
        deck = get_deck(name)
        lat, o = _host_ctx(deck)
        slot, src, _ = lat.bc_table()
        d = I.derive(deck)
        o.init_up(r_vec)
        o.init_pdf()
        bufs = [np.zeros((o.A, o.nq)), np.zeros((o.A, o.nq))]
        cur = 0
        for step in range(30):
            o.collision()
            cur = 1 - cur
            fid = o.list("fluid_id")
            bufs[cur][fid] = o.fb.reshape(o.A, o.nq)[fid]
            f, fid = _table_step(o, slot, src, bufs, cur)
            o.boundary_condition()
            o.propagation()
            o.post_processing()
            assert np.array_equal(f, o.f.reshape(o.A, o.nq)[fid]), (name, step)
        lat.close()
        o.close()


@pytest.mark.parametrize("name", ["tiny2d", "tiny3d", "inletbody2d", "wallsphere3d", "ysphere3d", "outletsphere3d",
                                  "ragged2d", "ragged3d", "cyl400x100", "input.in"])
def test_bc_table_matches_oracle_boundary_condition(name):
    lat, o = _host_ctx(get_deck(name))
    stats = _check_bc_table(lat, o)
    if name != "wallsphere3d":  # that body touches the walls: order-dependent reads exist there
        assert stats["order_sensitive"] == 0
    lat.close()
    o.close()


@pytest.mark.slow
def test_bc_table_reference_3d_input():
    lat, o = _host_ctx(I.deck("input_3d.in"))
    stats = _check_bc_table(lat, o)
    assert stats["order_sensitive"] == 0
    lat.close()
    o.close()


def test_error_codes():
    with pytest.raises(I.ImexlbError) as e:
        I.Lattice(4, 10, 10, 10, device=-1)
    assert e.value.code == _lib.E_ARG
    with pytest.raises(I.ImexlbError) as e:
        I.Lattice(2, 1, 10, device=-1)  # fewer than 3 nodes (cart_mpi.f90:107-110)
    assert e.value.code == _lib.E_ARG
    lat = I.Lattice(2, 10, 10, device=-1)
    with pytest.raises(I.ImexlbError) as e:
        lat.geometry_list("fluid_id")  # before SetGeometry
    assert e.value.code == _lib.E_STATE
    lat.set_params(1.5, 1.0, 0.01)
    lat.set_geometry(4.0, 2.0)
    for call in (lat.collision, lat.init_pdf, lat.monitor, lambda: lat.advance(1)):
        with pytest.raises(I.ImexlbError) as e:
            call()
        assert e.value.code == _lib.E_NODEV
    with pytest.raises(I.ImexlbError):
        lat.set_params(1.5, 0.0, 0.01)
    lat.close()


def test_info_sizes():
    lat = I.Lattice(3, 24, 12, 12, device=-1)
    lat.set_params(1.5, 1.0, 0.01)
    lat.set_geometry(6.0, 3.0)
    info = lat.info()
    assert info["nodes"] == 25 * 13 * 13 and info["value_bytes"] == 8
    assert info["fluid"] == len(lat.geometry_list("fluid_id"))
    lat.close()


def test_state_comparison_catches_one_differing_value(r_vec):
    """The bitwise checker used by the GPU parity tests: a stand-in lattice
    that returns the oracle's own arrays passes (in small chunks too), and
    one changed population, pressure or velocity on one interior cell
    fails."""
    from cases import assert_state_equal, oracle_setup
    o, _ = oracle_setup(get_deck("tiny3d"), r_vec)
    o.run(0, 3, 0)

    class Stand:
        def __init__(self):
            self.f, self.p, self.u = o.f.copy(), o.p.copy(), o.u.copy()

        def download_pdf(self, which, layout):
            return self.f.copy()

        def download_fields(self, layout):
            return self.p.copy(), self.u.copy()

    s = Stand()
    assert_state_equal(s, o, chunk=97)
    cell = int(np.nonzero(interior_mask(o))[0][1234])
    for name, arr, idx in (("f", s.f, cell * o.nq + 5), ("p", s.p, cell), ("u", s.u, cell * o.dim + 2)):
        keep = arr[idx]
        arr[idx] = np.nextafter(keep, np.inf)
        with pytest.raises(AssertionError, match=name):
            assert_state_equal(s, o, chunk=97)
        arr[idx] = keep
    o.close()


@pytest.mark.parametrize("aspect,angle", [(2.0, 0.0), (0.6, 40.0), (1.5, 90.0)])
def test_ellipse_geometry_matches_oracle(aspect, angle):
    """The library's elliptic mask, lists and body scan equal the oracle's."""
    nx, ny, pos, radius = 80, 40, 24.0, 7.0
    lat = I.Lattice(2, nx, ny, 0, device=-1)
    lat.set_params(1.5, 1.0, 0.05)
    lat.set_geometry(pos, radius, aspect, angle)
    o = O.Oracle(2, nx, ny, 0)
    o.set_params(1.5, 1.0, 0.05)
    o.set_geometry(pos, radius, aspect, angle)
    assert np.array_equal(lat.geometry_geo(), o.geo)
    for name in ("fluid_id", "b_user", "bl", "br", "bu", "bd"):
        assert np.array_equal(lat.geometry_list(name), o.list(name)), name
    lat.close()
    o.close()


def test_ellipse_rejected_in_3d():
    lat = I.Lattice(3, 8, 8, 8, device=-1)
    lat.set_params(1.5, 1.0, 0.05)
    with pytest.raises(I.ImexlbError):
        lat.set_geometry(4.0, 2.0, 2.0, 0.0)
    lat.close()
