"""imexlb: MI355X lattice Boltzmann time step behind the IMEXLB-1.0 interface.

The compute path is libimexlb.so (hand-written gfx950 HIP kernels behind the
C ABI in include/imexlb.h); this package is a thin ctypes handle on it used
by the tests and the benchmark.  The Fortran drop-in modules live in
imexlb-1.0_amd/fortran.
"""
from ._lib import (BF16_F32, BF16SR_F32, E_ARG, E_HIP, E_NAN, E_NODEV, E_OOM, E_STATE, E_UB, F16_F32, F32, F64,
                   LAYOUT_INTERIOR, LAYOUT_REF, PDF_F, PDF_FB, ImexlbError, device_count, lib)
from .inputs import DECKS, Deck, Derived, deck, derive, nodes, parse_deck, read_deck
from .lattice import Lattice, setup

__all__ = [
    "F64", "F32", "F16_F32", "BF16_F32", "BF16SR_F32", "E_ARG", "E_HIP", "E_OOM", "E_STATE", "E_UB", "E_NODEV", "E_NAN",
    "LAYOUT_REF", "LAYOUT_INTERIOR", "PDF_F", "PDF_FB", "ImexlbError",
    "device_count", "lib", "DECKS", "Deck", "Derived", "deck", "derive", "nodes", "parse_deck", "read_deck",
    "Lattice", "setup",
]
