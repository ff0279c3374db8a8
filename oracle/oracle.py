"""ctypes binding of the oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

The oracle is the plain-C FP64 restatement of the IMEXLB-1.0 time step in
``imexlb_oracle.c``.  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
CPU-baseline leg may import this module, and only to check or time the
product; the product library never calls it.

Arrays are exposed as numpy views in the reference's own layout: flat arrays
with a 3-wide ghost shell, populations AoS ``f[id*nq+iq]``, velocity AoS
``u[id*dim+d]`` (lbm.f90:44; initialization.f90:127-144).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

LIST_NAMES = ("fluid_id", "bl", "br", "bu", "bd", "bf", "bb", "b_user")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        P = C.c_void_p
        L.orc_create.restype = P
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_destroy.argtypes = [P]
        L.orc_set_params.argtypes = [P, C.c_double, C.c_double, C.c_double]
        L.orc_set_geometry.argtypes = [P, C.c_double, C.c_double]
        L.orc_set_geometry_shape.argtypes = [P, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_init_up.argtypes = [P, C.POINTER(C.c_double)]
        for name in ("orc_init_pdf", "orc_collision", "orc_boundary_condition",
                     "orc_propagation", "orc_post_processing", "orc_moments", "orc_post_bc"):
            getattr(L, name).argtypes = [P]
        L.orc_pass_f.argtypes = [P, P]
        L.orc_pass_int.argtypes = [P, P]
        L.orc_monitor.restype = C.c_double
        L.orc_monitor.argtypes = [P]
        L.orc_forces.argtypes = [P, C.POINTER(C.c_double)]
        L.orc_run.restype = C.c_int64
        L.orc_run.argtypes = [P, C.c_int64, C.c_int64, C.c_int64, C.POINTER(C.c_double), C.c_int64]
        for name in ("orc_ptr_f", "orc_ptr_fb", "orc_ptr_p", "orc_ptr_u"):
            getattr(L, name).restype = C.POINTER(C.c_double)
            getattr(L, name).argtypes = [P]
        L.orc_ptr_geo.restype = C.POINTER(C.c_int32)
        L.orc_ptr_geo.argtypes = [P]
        L.orc_ptr_list.restype = C.POINTER(C.c_int64)
        L.orc_ptr_list.argtypes = [P, C.c_int]
        L.orc_list_size.restype = C.c_int64
        L.orc_list_size.argtypes = [P, C.c_int]
        L.orc_array_size.restype = C.c_int64
        L.orc_array_size.argtypes = [P]
        L.orc_ub_reads.restype = C.c_int64
        L.orc_ub_reads.argtypes = [P]
        L.orc_reset_ub_reads.argtypes = [P]
        L.orc_ptr_e.restype = C.POINTER(C.c_int)
        L.orc_ptr_e.argtypes = [P]
        L.orc_ptr_t.restype = C.POINTER(C.c_double)
        L.orc_ptr_t.argtypes = [P]
        L.orc_data_input.argtypes = [C.c_int] + [C.c_double] * 8 + [C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.orc_init_lattice.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        _LIB = L
    return _LIB


@dataclass
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


def data_input(nx, Lc, Xpos, Re, ome, dt, rho0, max_t, inter) -> Derived:
    """a1: derived quantities of DataInput (initialization.f90:59-81)."""
    out = (C.c_double * 8)()
    iout = (C.c_int64 * 2)()
    lib().orc_data_input(int(nx), float(Lc), float(Xpos), float(Re), float(ome), float(dt),
                         float(rho0), float(max_t), float(inter), out, iout)
    return Derived(*[out[i] for i in range(8)], int(iout[0]), int(iout[1]))


def init_lattice(dim):
    """a2: (e[nq,dim], t[nq]) as InitLattice builds them."""
    nq = 9 if dim == 2 else 27
    e = (C.c_int * (nq * dim))()
    t = (C.c_double * nq)()
    lib().orc_init_lattice(dim, e, t)
    return np.array(e[:], dtype=np.int64).reshape(nq, dim), np.array(t[:])


class Oracle:
    """One-rank restatement of the reference lattice (2D D2Q9 or 3D D3Q27)."""

    def __init__(self, dim, nx, ny, nz=0):
        self.dim = dim
        self.nq = 9 if dim == 2 else 27
        self.n = (nx, ny, nz if dim == 3 else 0)
        self._h = lib().orc_create(dim, nx, ny, nz if dim == 3 else 0)
        if not self._h:
            raise MemoryError("orc_create failed")
        self.A = lib().orc_array_size(self._h)
        self.dyy = nx + 7
        self.dzz = (nx + 7) * (ny + 7) if dim == 3 else 0
        ep = lib().orc_ptr_e(self._h)
        tp = lib().orc_ptr_t(self._h)
        self.e = np.array([ep[i] for i in range(self.nq * dim)], dtype=np.int64).reshape(self.nq, dim)
        self.t = np.array([tp[i] for i in range(self.nq)])

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ----- views -----
    def _view(self, ptr, n, dtype):
        return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype) if n else np.zeros(0, dtype)

    @property
    def f(self):
        return np.ctypeslib.as_array(lib().orc_ptr_f(self._h), shape=(self.A * self.nq,))

    @property
    def fb(self):
        return np.ctypeslib.as_array(lib().orc_ptr_fb(self._h), shape=(self.A * self.nq,))

    @property
    def p(self):
        return np.ctypeslib.as_array(lib().orc_ptr_p(self._h), shape=(self.A,))

    @property
    def u(self):
        return np.ctypeslib.as_array(lib().orc_ptr_u(self._h), shape=(self.A * self.dim,))

    @property
    def geo(self):
        return np.ctypeslib.as_array(lib().orc_ptr_geo(self._h), shape=(self.A,))

    def list(self, name):
        w = LIST_NAMES.index(name)
        n = lib().orc_list_size(self._h, w)
        if n <= 0:
            return np.zeros(0, dtype=np.int64)
        return np.ctypeslib.as_array(lib().orc_ptr_list(self._h, w), shape=(n,)).copy()

    def id(self, x, y, z=0):
        i = (x + 3) + self.dyy * (y + 3)
        if self.dim == 3:
            i += self.dzz * (z + 3)
        return i

    @property
    def ub_reads(self):
        return lib().orc_ub_reads(self._h)

    # ----- phases -----
    def set_params(self, ome, rho0, uu):
        lib().orc_set_params(self._h, ome, rho0, uu)

    def set_geometry(self, pos, radius, aspect=1.0, angle_deg=0.0):
        if aspect == 1.0 and angle_deg == 0.0:
            lib().orc_set_geometry(self._h, pos, radius)
        else:
            lib().orc_set_geometry_shape(self._h, pos, radius, aspect, angle_deg)

    def init_up(self, r_vec):
        r = np.ascontiguousarray(r_vec, dtype=np.float64)
        assert r.shape == (100,)
        lib().orc_init_up(self._h, r.ctypes.data_as(C.POINTER(C.c_double)))

    def init_pdf(self):
        lib().orc_init_pdf(self._h)

    def collision(self):
        lib().orc_collision(self._h)

    def pass_f(self):
        lib().orc_pass_f(self._h, lib().orc_ptr_fb(self._h))

    def boundary_condition(self):
        lib().orc_boundary_condition(self._h)

    def propagation(self):
        lib().orc_propagation(self._h)

    def post_processing(self):
        lib().orc_post_processing(self._h)

    def moments(self):
        """PostProcessing's moment pass alone (no inlet/outlet conditions)."""
        lib().orc_moments(self._h)

    def monitor(self):
        return lib().orc_monitor(self._h)

    def forces(self):
        """Momentum-exchange force on the body (call after boundary_condition)."""
        F = (C.c_double * 3)()
        lib().orc_forces(self._h, F)
        return np.array(F[:self.dim])

    def run(self, iter0, iter1, interv, cap=100000):
        buf = (C.c_double * cap)()
        k = lib().orc_run(self._h, iter0, iter1, interv, buf, cap)
        return np.array(buf[:min(k, cap)])

    def interior(self, arr, ncomp):
        """Reshape a flat reference-layout array to [z][y][x][comp] interior."""
        nx, ny, nz = self.n
        if self.dim == 2:
            a = arr.reshape(ny + 7, nx + 7, ncomp)
            return a[3:ny + 4, 3:nx + 4, :]
        a = arr.reshape(nz + 7, ny + 7, nx + 7, ncomp)
        return a[3:nz + 4, 3:ny + 4, 3:nx + 4, :]


def setup_case(dim, n, Lc, Xpos, Re, ome, dt, rho0, max_t, inter, r_vec):
    """DataInput..InitPDF of main.f90:19-40 for one case; returns (Oracle, Derived)."""
    d = data_input(n[0], Lc, Xpos, Re, ome, dt, rho0, max_t, inter)
    o = Oracle(dim, *n)
    o.set_params(ome, rho0, d.uu)
    o.set_geometry(d.pos, d.radius)
    o.init_up(r_vec)
    o.init_pdf()
    return o, d
