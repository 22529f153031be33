"""ctypes access to the CPU oracle (oracle/_build/libsphere_oracle.so).

Test infrastructure: the oracle is the checker, never the product.
"""
import ctypes
import os

import numpy

from sphere_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libsphere_oracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = ctypes.CDLL(ORACLE_SO)
        vp = ctypes.c_void_p
        S = ctypes.POINTER(_abi.SphState)
        pu = ctypes.POINTER(ctypes.c_uint32)
        pd = ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [S, ctypes.c_char_p, ctypes.c_int]
        L.orc_destroy.argtypes = [vp]
        L.orc_error.restype = ctypes.c_char_p
        L.orc_error.argtypes = [vp]
        L.orc_step.argtypes = [vp, ctypes.c_int]
        L.orc_get_header.argtypes = [vp, S]
        L.orc_read_state.argtypes = [vp, S]
        L.orc_read_sort.argtypes = [vp, pu, pu, pu, pu]
        L.orc_read_contacts.argtypes = [vp, pu, pd, pd]
        L.orc_write_contacts.argtypes = [vp, pu, pd, pd]
        L.orc_darcy_stats.argtypes = [vp, ctypes.POINTER(_abi.SphDarcyStats)]
        L.orc_conv_log.argtypes = [vp, pu, ctypes.c_int]
        L.orc_read_darcy_field.argtypes = [vp, ctypes.c_int, pd]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_contact_slots.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_attach_ns.argtypes = [vp, vp]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("oracle error %d: %s" % (code, msg))
        self.code = code


class Oracle:
    """Same surface as sphere_amd.Engine, backed by the CPU oracle."""

    def __init__(self, buf, threads=None, nc=32):
        L = lib()
        if threads is not None:
            L.orc_set_threads(int(threads))
        self.nc = int(nc)
        L.orc_set_contact_slots(self.nc)
        err = ctypes.create_string_buffer(512)
        self.h = L.orc_create(ctypes.byref(buf.state), err, 512)
        if not self.h:
            raise OracleError(-3, err.value.decode())
        self.L = L

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise OracleError(rc, self.L.orc_error(self.h).decode())

    def attach_ns(self, nso):
        """Couple an NsOracle fluid (tests/_ns_oracle.py) to this oracle's
        grains (None detaches)."""
        self.L.orc_attach_ns(self.h, nso.h if nso is not None else None)
        if nso is not None:
            h = _abi.SphState()
            self.L.orc_get_header(self.h, ctypes.byref(h))
            nso.np = int(h.np)  # rows of grain_forces()
        self._ns = nso

    def step(self, n=1):
        self._check(self.L.orc_step(self.h, int(n)))

    def header(self):
        h = _abi.SphState()
        self.L.orc_get_header(self.h, ctypes.byref(h))
        return h

    def read(self):
        buf = _abi.StateBuffers(self.header())
        self._check(self.L.orc_read_state(self.h, ctypes.byref(buf.state)))
        return buf

    def read_sort(self):
        h = self.header()
        ncell = h.num[0] * h.num[1] * h.num[2]
        cid = numpy.zeros(max(h.np, 1), numpy.uint32)
        idx = numpy.zeros(max(h.np, 1), numpy.uint32)
        cs = numpy.zeros(ncell, numpy.uint32)
        ce = numpy.zeros(ncell, numpy.uint32)
        self.L.orc_read_sort(self.h, _p(cid, ctypes.c_uint32), _p(idx, ctypes.c_uint32),
                             _p(cs, ctypes.c_uint32), _p(ce, ctypes.c_uint32))
        return cid[:h.np], idx[:h.np], cs, ce

    def read_contacts(self):
        h = self.header()
        n = max(h.np, 1) * self.nc
        c = numpy.zeros(n, numpy.uint32)
        dm = numpy.zeros(3 * n)
        dt = numpy.zeros(3 * n)
        self.L.orc_read_contacts(self.h, _p(c, ctypes.c_uint32), _p(dm, ctypes.c_double),
                                 _p(dt, ctypes.c_double))
        m = h.np * self.nc
        return c[:m], dm[:3 * m].reshape(-1, 3), dt[:3 * m].reshape(-1, 3)

    def write_contacts(self, contacts, distmod, delta_t):
        contacts = numpy.ascontiguousarray(contacts, numpy.uint32)
        distmod = numpy.ascontiguousarray(distmod, numpy.float64)
        delta_t = numpy.ascontiguousarray(delta_t, numpy.float64)
        self.L.orc_write_contacts(self.h, _p(contacts, ctypes.c_uint32),
                                  _p(distmod, ctypes.c_double), _p(delta_t, ctypes.c_double))

    def darcy_stats(self):
        s = _abi.SphDarcyStats()
        self.L.orc_darcy_stats(self.h, ctypes.byref(s))
        return s

    def conv_log(self):
        n = self.L.orc_conv_log(self.h, None, 0)
        a = numpy.zeros(max(2 * n, 2), numpy.uint32)
        self.L.orc_conv_log(self.h, _p(a, ctypes.c_uint32), n)
        return a[:2 * n].reshape(-1, 2)

    def darcy_field(self, which):
        h = self.header()
        n = (h.num[0] + 2) * (h.num[1] + 2) * (h.num[2] + 2)
        out = numpy.zeros(n)
        self._check(self.L.orc_read_darcy_field(self.h, int(which), _p(out, ctypes.c_double)))
        return out.reshape(h.num[2] + 2, h.num[1] + 2, h.num[0] + 2)


def run_oracle(s, nsteps, threads=None):
    """Run nsteps of sim object s on the oracle and copy the state back into s."""
    o = Oracle(s.to_state(), threads=threads)
    o.step(nsteps)
    s.from_state(o.read())
    return o
