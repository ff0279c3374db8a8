"""Input decks in the reference's 4-record format and DataInput's derived values.

Format (input.in / input_3d.in, read at initialization.f90:61-75 and
initialization_3d.f90:62-76 with list-directed ``read``):

    nx ny [nz]          largest node index per axis (nodes = n+1)
    L  Xpos  Re         body diameter and x-position in units of nx, Reynolds number
    ome  dt  rho0       relaxation parameter, time increment, density
    max_t  inter        run length and monitor interval in units of t_intv

Anything after the values a record needs is ignored, as a list-directed
read skips the rest of its record.
"""
from __future__ import annotations

import math
from dataclasses import dataclass


@dataclass(frozen=True)
class Deck:
    dim: int
    n: tuple            # (nx, ny, nz); nz = 0 in 2D
    Lc: float           # L
    Xpos: float
    Re: float
    ome: float
    dt: float
    rho0: float
    max_t: float
    inter: float


@dataclass(frozen=True)
class Derived:
    charlength: float
    pos: float
    nu: float
    t_intv: float
    uu: float
    tau: float
    ma: float
    radius: float
    interv: int
    max_step: int


def _num(tok: str) -> float:
    return float(tok.replace("d", "e").replace("D", "e"))


def parse_deck(text: str, dim: int) -> Deck:
    """Read the four records of an input deck (list-directed semantics)."""
    lines = [ln.replace(",", " ").split() for ln in text.splitlines()]
    it = iter(lines)

    def record(k):
        vals = []
        while len(vals) < k:
            try:
                toks = next(it)
            except StopIteration:
                raise ValueError("input deck ends before all four records were read")
            vals.extend(toks)
            if len(vals) >= k:
                break
        return [_num(t) for t in vals[:k]]

    n = [int(v) for v in record(dim)]
    Lc, Xpos, Re = record(3)
    ome, dt, rho0 = record(3)
    max_t, inter = record(2)
    nt = (n[0], n[1], n[2] if dim == 3 else 0)
    return Deck(dim, nt, Lc, Xpos, Re, ome, dt, rho0, max_t, inter)


def read_deck(path: str, dim: int) -> Deck:
    with open(path) as fh:
        return parse_deck(fh.read(), dim)


def _trunc_i32(v: float) -> int:
    """Fortran default-integer assignment of a real: truncation toward zero."""
    return int(math.trunc(v))


def derive(deck: Deck) -> Derived:
    """DataInput's derived quantities (initialization.f90:65-81), IEEE double."""
    nx = deck.n[0]
    charlength = deck.Lc * nx
    pos = deck.Xpos * nx
    nu = (2.0 - deck.ome) * deck.dt / (6.0 * deck.ome)
    t_intv = charlength * charlength / nu
    interv = _trunc_i32(deck.inter * t_intv)
    max_step = _trunc_i32(deck.max_t * t_intv)
    uu = deck.Re * nu / charlength
    tau = 3 * nu
    ma = uu * math.sqrt(3.0)
    radius = charlength * 0.5
    return Derived(charlength, pos, nu, t_intv, uu, tau, ma, radius, interv, max_step)


# Named decks.  INPUT_2D / INPUT_3D are the reference's example problems;
# the others are the benchmark configurations written in the same format
# (cylinder 400x100 at Re 100, sphere 512x256x256 at Re 200, cylinder
# 8192x2048 at Re 1000), with the reference's own boundary semantics.
DECKS = {
    "input.in": (2, "512 128\n0.03125 0.125 100\n1.953125 1 1\n0.5 0.01\n"),
    "input_3d.in": (3, "512 128 128\n0.03125 0.125 300\n1.953125 1 1\n0.1 0.01\n"),
    "cyl400x100": (2, "400 100\n0.05 0.2 100\n1.8867924528301887 1 1\n0.49999 0.0025001\n"),
    "sphere512x256x256": (3, "512 256 256\n0.078125 0.25 200\n1.9083969465648855 1 1\n0.024997 0.0005001\n"),
    "cyl8192x2048": (2, "8192 2048\n0.015625 0.125 1000\n1.926040061633282 1 1\n0.0195312 0.000039063\n"),
}


def deck(name: str) -> Deck:
    dim, text = DECKS[name]
    return parse_deck(text, dim)


def nodes(d: Deck) -> int:
    nx, ny, nz = d.n
    return (nx + 1) * (ny + 1) * ((nz + 1) if d.dim == 3 else 1)
