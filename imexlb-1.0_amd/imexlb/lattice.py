"""Python handle on one lattice context of libimexlb.so.

Method names follow the reference's procedures (lbm.f90, initialization.f90):
``collision``, ``boundary_condition``, ``propagation``, ``post_processing``,
``monitor``, ``init_up``, ``init_pdf``.  Every call goes through the C ABI;
there is no host fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib
from .inputs import Deck, derive

_DP = C.POINTER(C.c_double)


def _dptr(a):
    return a.ctypes.data_as(_DP)


class Lattice:
    """One-rank lattice: D2Q9 (dim=2) or D3Q27 (dim=3)."""

    def __init__(self, dim, nx, ny, nz=0, device=0, precision=_lib.F64):
        self.dim = dim
        self.nq = 9 if dim == 2 else 27
        self.n = (nx, ny, nz if dim == 3 else 0)
        self.precision = precision
        self.device = device
        h = C.c_void_p()
        check(lib().imexlb_create(C.byref(h), dim, nx, ny, nz if dim == 3 else 0, device, precision),
              "imexlb_create")
        self._h = h
        self.A = (nx + 7) * (ny + 7) * ((nz + 7) if dim == 3 else 1)
        self.N = (nx + 1) * (ny + 1) * ((nz + 1) if dim == 3 else 1)

    # ---------------------------------------------------------------- life
    def close(self):
        if self._h:
            lib().imexlb_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- setup
    def set_params(self, ome, rho0, uu):
        check(lib().imexlb_set_params(self._h, ome, rho0, uu), "imexlb_set_params")

    def set_geometry(self, pos, radius, aspect=1.0, angle_deg=0.0):
        if aspect == 1.0 and angle_deg == 0.0:
            check(lib().imexlb_set_geometry(self._h, pos, radius), "imexlb_set_geometry")
        else:
            check(lib().imexlb_set_geometry_shape(self._h, pos, radius, aspect, angle_deg),
                  "imexlb_set_geometry_shape")

    def geometry_list(self, name):
        n = C.c_int64()
        w = _lib.LISTS[name]
        check(lib().imexlb_geometry_size(self._h, w, C.byref(n)), "imexlb_geometry_size")
        out = np.zeros(max(n.value, 1), dtype=np.int32)
        check(lib().imexlb_geometry_list(self._h, w, out.ctypes.data_as(C.POINTER(C.c_int32)), out.size),
              "imexlb_geometry_list")
        return out[:n.value]

    def geometry_geo(self):
        out = np.zeros(self.A, dtype=np.int32)
        check(lib().imexlb_geometry_geo(self._h, out.ctypes.data_as(C.POINTER(C.c_int32)), out.size),
              "imexlb_geometry_geo")
        return out

    def bc_table(self):
        stats = (C.c_int64 * 2)()
        check(lib().imexlb_bc_table(self._h, None, None, 0, stats), "imexlb_bc_table")
        n = stats[0]
        slot = np.zeros(max(n, 1), dtype=np.int64)
        src = np.zeros(max(n, 1), dtype=np.int64)
        P = C.POINTER(C.c_int64)
        check(lib().imexlb_bc_table(self._h, slot.ctypes.data_as(P), src.ctypes.data_as(P), slot.size, stats),
              "imexlb_bc_table")
        return slot[:n], src[:n], {"slots": stats[0], "order_sensitive": stats[1]}

    def init_up(self, r_vec):
        r = np.ascontiguousarray(r_vec, dtype=np.float64)
        if r.shape != (100,):
            raise ValueError("r_vec must hold 100 values")
        check(lib().imexlb_init_up(self._h, _dptr(r)), "imexlb_init_up")

    def upload_fields(self, p, u, layout=_lib.LAYOUT_INTERIOR):
        p = np.ascontiguousarray(p, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        check(lib().imexlb_upload_fields(self._h, _dptr(p), _dptr(u), layout), "imexlb_upload_fields")

    def init_fields_f32(self, fields, p_scale, u_scale, r_vec=None):
        """p, u from a device FP32 stack (3, H, W) (a torch tensor on this
        lattice's GPU): p = p_scale f0, u_x = u_scale f1 (times InitUP's
        perturbation with r_vec), u_y = u_scale f2; 0 on body nodes."""
        if fields.dim() != 3 or fields.shape[0] != 3 or not fields.is_contiguous():
            raise ValueError("fields: contiguous (3, H, W)")
        rv = None if r_vec is None else np.ascontiguousarray(r_vec, dtype=np.float64)
        check(lib().imexlb_init_fields_f32(self._h, C.c_void_p(fields.data_ptr()), fields.shape[1], fields.shape[2],
                                           p_scale, u_scale, None if rv is None else _dptr(rv)),
              "imexlb_init_fields_f32")

    def body_mask_f32(self, out):
        """Body mask (1 = body node) of node rows < H, columns < W into the
        device FP32 tensor out (H, W)."""
        if out.dim() != 2 or not out.is_contiguous():
            raise ValueError("out: contiguous (H, W)")
        check(lib().imexlb_body_mask_f32(self._h, C.c_void_p(out.data_ptr()), out.shape[0], out.shape[1]),
              "imexlb_body_mask_f32")

    def init_pdf(self):
        check(lib().imexlb_init_pdf(self._h), "imexlb_init_pdf")

    # ---------------------------------------------------------------- step
    def collision(self):
        check(lib().imexlb_collision(self._h), "imexlb_collision")

    def boundary_condition(self):
        check(lib().imexlb_boundary_condition(self._h), "imexlb_boundary_condition")

    def propagation(self):
        check(lib().imexlb_propagation(self._h), "imexlb_propagation")

    def post_processing(self):
        check(lib().imexlb_post_processing(self._h), "imexlb_post_processing")

    def set_fused(self, fused):
        check(lib().imexlb_set_fused(self._h, 1 if fused else 0), "imexlb_set_fused")

    def advance(self, nsteps):
        check(lib().imexlb_advance(self._h, int(nsteps)), "imexlb_advance")

    def monitor(self):
        um = C.c_double()
        check(lib().imexlb_monitor(self._h, C.byref(um)), "imexlb_monitor")
        return um.value

    def mass(self):
        """Sum over fluid nodes of sum_q f_q (FP64)."""
        m = C.c_double()
        check(lib().imexlb_mass(self._h, C.byref(m)), "imexlb_mass")
        return m.value

    def forces(self):
        """Momentum-exchange force on the body [drag, lift(, lift_z)], lattice units."""
        F = (C.c_double * 3)()
        check(lib().imexlb_forces(self._h, F), "imexlb_forces")
        return np.array(F[:self.dim])

    def synchronize(self):
        check(lib().imexlb_synchronize(self._h), "imexlb_synchronize")

    # ---------------------------------------------------------------- data
    def download_fields(self, layout=_lib.LAYOUT_INTERIOR):
        if layout == _lib.LAYOUT_REF:
            p = np.zeros(self.A)
            u = np.zeros(self.A * self.dim)
        else:
            p = np.zeros(self.N)
            u = np.zeros(self.N * self.dim)
        check(lib().imexlb_download_fields(self._h, _dptr(p), _dptr(u), layout), "imexlb_download_fields")
        return p, u

    def download_pdf(self, which=_lib.PDF_F, layout=_lib.LAYOUT_INTERIOR):
        n = (self.A if layout == _lib.LAYOUT_REF else self.N) * self.nq
        out = np.zeros(n)
        check(lib().imexlb_download_pdf(self._h, which, _dptr(out), layout), "imexlb_download_pdf")
        return out

    # ---------------------------------------------------------------- timing
    def profile(self, enable):
        check(lib().imexlb_profile(self._h, 1 if enable else 0), "imexlb_profile")

    def profile_read(self):
        n = (C.c_int64 * 3)()
        ms = (C.c_double * 3)()
        check(lib().imexlb_profile_read(self._h, n, ms), "imexlb_profile_read")
        return [n[i] for i in range(3)], [ms[i] for i in range(3)]

    def info(self):
        out = (C.c_int64 * 6)()
        check(lib().imexlb_info(self._h, out), "imexlb_info")
        keys = ("nodes", "fluid", "plane_bytes", "device_bytes", "value_bytes", "bc_entries")
        return dict(zip(keys, list(out)))


def setup(deck: Deck, r_vec, device=0, precision=_lib.F64):
    """DataInput .. InitPDF of main.f90:19-40 for one deck; returns (Lattice, Derived)."""
    d = derive(deck)
    lat = Lattice(deck.dim, *deck.n, device=device, precision=precision)
    lat.set_params(deck.ome, deck.rho0, d.uu)
    lat.set_geometry(d.pos, d.radius)
    if device >= 0:
        lat.init_up(r_vec)
        lat.init_pdf()
    return lat, d
