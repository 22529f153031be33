"""ctypes access to the Navier-Stokes part of the CPU oracle
(oracle/ns_oracle.c in oracle/_build/libsphere_oracle.so).

Test infrastructure: the oracle is the checker, never the product.
"""
import ctypes

import numpy

from sphere_amd import _abi
from _oracle import lib as _oracle_lib

_pd = ctypes.POINTER(ctypes.c_double)


def _lib():
    L = _oracle_lib()
    if not getattr(L, "_ns_ready", False):
        L.nso_create.restype = ctypes.c_void_p
        L.nso_create.argtypes = [ctypes.POINTER(_abi.SphNsParams), _pd, _pd]
        L.nso_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.nso_read.argtypes = [ctypes.c_void_p, _pd, _pd, _pd, _pd]
        L.nso_field.argtypes = [ctypes.c_void_p, ctypes.c_int, _pd]
        L.nso_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(_abi.SphNsStats)]
        L.nso_destroy.argtypes = [ctypes.c_void_p]
        L.nso_set_grains.argtypes = [ctypes.c_void_p, ctypes.c_uint32, _pd, _pd, _pd, ctypes.c_int]
        L.nso_grain_forces.argtypes = [ctypes.c_void_p, _pd, _pd, _pd, _pd]
        L._ns_ready = True
    return L


def _p(a):
    return a.ctypes.data_as(_pd)


class NsOracle:
    """Same interface as sphere_amd.NavierStokes: arrays [x, y, z(, 3)]."""

    # field ids of nso_field
    FACES, EPS, PRED_FACES, DIV_TAU, DIV_PHI_VV, FORCING = 0, 3, 4, 7, 10, 13

    def __init__(self, params, v, p):
        self.L = _lib()
        self.P = params
        self.shape = (params.nx, params.ny, params.nz)
        vc = numpy.ascontiguousarray(numpy.asarray(v, dtype=numpy.float64).transpose(2, 1, 0, 3))
        pc = numpy.ascontiguousarray(numpy.asarray(p, dtype=numpy.float64).transpose(2, 1, 0))
        self.h = self.L.nso_create(ctypes.byref(params), _p(vc), _p(pc))
        if not self.h:
            raise ValueError("oracle refused the Navier-Stokes parameters")

    def step(self, n=1):
        rc = self.L.nso_step(self.h, int(n))
        if rc != 0:
            raise _abi.SphereError(rc, "oracle Navier-Stokes step failed")

    def read(self):
        nx, ny, nz = self.shape
        v = numpy.empty((nz, ny, nx, 3))
        p, phi, dphi = (numpy.empty((nz, ny, nx)) for _ in range(3))
        self.L.nso_read(self.h, _p(v), _p(p), _p(phi), _p(dphi))
        return (v.transpose(2, 1, 0, 3).copy(), p.transpose(2, 1, 0).copy(),
                phi.transpose(2, 1, 0).copy(), dphi.transpose(2, 1, 0).copy())

    def field(self, which):
        nx, ny, nz = self.shape
        faces = which < 3 or 4 <= which < 10
        n = (nx + 3) * (ny + 3) * (nz + 3) if faces else (nx + 2) * (ny + 2) * (nz + 2)
        out = numpy.empty(n)
        self.L.nso_field(self.h, int(which), _p(out))
        return out

    def cells(self, which):
        """A ghosted cell field as [z + 1, y + 1, x + 1] (index 0 the ghost)."""
        nx, ny, nz = self.shape
        return self.field(which).reshape(nz + 2, ny + 2, nx + 2)

    def faces(self, which):
        """A ghosted face field as [z + 1, y + 1, x + 1]."""
        nx, ny, nz = self.shape
        return self.field(which).reshape(nz + 3, ny + 3, nx + 3)

    def set_grains(self, x, vel, origo=(0.0, 0.0, 0.0), periodic=1):
        x = numpy.ascontiguousarray(x, dtype=numpy.float64).reshape(-1, 4)
        v = numpy.zeros_like(x)
        vel = numpy.asarray(vel, dtype=numpy.float64).reshape(len(x), -1)
        v[:, :3] = vel[:, :3]
        o = numpy.ascontiguousarray(origo, dtype=numpy.float64)
        self.np = len(x)
        if self.L.nso_set_grains(self.h, len(x), _p(x), _p(v), _p(o), int(periodic)) != 0:
            raise ValueError("a grain lies outside the grid")

    def grain_forces(self):
        n = getattr(self, "np", 0)
        out = [numpy.zeros((n, 3)) for _ in range(4)]
        self.L.nso_grain_forces(self.h, *[_p(a) for a in out])
        return tuple(out)

    def stats(self):
        st = _abi.SphNsStats()
        self.L.nso_stats(self.h, ctypes.byref(st))
        return st

    def close(self):
        if self.h:
            self.L.nso_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
