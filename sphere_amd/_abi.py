"""ctypes mirror of include/sphere_hip.h and the loader of the HIP library.

The reference reaches its CUDA binary through ``subprocess`` (python/sphere.py
:4482-4540).  Here the same run goes through the C ABI of
``libsphere_hip.so`` in process.  The library must have been built
(``__graft_entry__.build()``); there is no fallback path.
"""
import ctypes
import os

import numpy

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsphere_hip.so")

NC = 32
MAXWALLS = 6

SPH_OK = 0
SPH_ERR_ARG = -1
ERRORS = {
    -1: "argument",
    -2: "file",
    -3: "value",
    -4: "hip",
    -5: "grid",
    -6: "nan",
    -7: "overflow",
    -8: "state",
}

K_NAMES = ["hash", "sort", "reorder", "contact", "bonds", "integrate", "walls",
           "darcy_setup", "darcy_sweep", "nlist", "topology"]

_d = ctypes.c_double
_pd = ctypes.POINTER(ctypes.c_double)
_pu = ctypes.POINTER(ctypes.c_uint32)
_pi = ctypes.POINTER(ctypes.c_int32)


class SphState(ctypes.Structure):
    """struct sph_state (include/sphere_hip.h)."""
    _fields_ = [
        ("nd", ctypes.c_int32), ("np", ctypes.c_uint32),
        ("time_dt", _d), ("time_current", _d), ("time_total", _d),
        ("time_file_dt", _d), ("time_step_count", ctypes.c_uint32),
        ("origo", _d * 3), ("L", _d * 3), ("num", ctypes.c_uint32 * 3),
        ("periodic", ctypes.c_int32), ("adaptive", ctypes.c_int32),
        ("x", _pd), ("xyzsum", _pd), ("vel", _pd), ("force", _pd),
        ("angpos", _pd), ("angvel", _pd), ("torque", _pd),
        ("es_dot", _pd), ("es", _pd), ("ev_dot", _pd), ("ev", _pd), ("p", _pd),
        ("g", _d * 3), ("k_n", _d), ("k_t", _d), ("k_r", _d), ("E", _d),
        ("gamma_n", _d), ("gamma_t", _d), ("gamma_r", _d),
        ("mu_s", _d), ("mu_d", _d), ("mu_r", _d), ("gamma_wn", _d),
        ("gamma_wt", _d), ("mu_ws", _d), ("mu_wd", _d), ("rho", _d),
        ("contactmodel", ctypes.c_uint32),
        ("kappa", _d), ("db", _d), ("V_b", _d),
        ("nw", ctypes.c_uint32), ("wmode", ctypes.c_int32 * MAXWALLS),
        ("w_nx", _d * (MAXWALLS * 4)), ("w_mvfd", _d * (MAXWALLS * 4)),
        ("sigma0_A", _d), ("sigma0_f", _d), ("tau_x", _d),
        ("lambda_bar", _d), ("nb0", ctypes.c_uint32),
        ("sigma_b", _d), ("tau_b", _d),
        ("bonds", _pu), ("bonds_delta_n", _pd), ("bonds_delta_t", _pd),
        ("bonds_omega_n", _pd), ("bonds_omega_t", _pd),
        ("fluid", ctypes.c_int32), ("cfd_solver", ctypes.c_int32), ("mu", _d),
        ("v_f", _pd), ("p_f", _pd), ("phi", _pd), ("dphi", _pd),
        ("rho_f", _d), ("p_mod_A", _d), ("p_mod_f", _d), ("p_mod_phi", _d),
        ("bc_xn", ctypes.c_int32), ("bc_xp", ctypes.c_int32),
        ("bc_yn", ctypes.c_int32), ("bc_yp", ctypes.c_int32),
        ("bc_bot", ctypes.c_int32), ("bc_top", ctypes.c_int32),
        ("free_slip_bot", ctypes.c_int32), ("free_slip_top", ctypes.c_int32),
        ("bc_bot_flux", _d), ("bc_top_flux", _d),
        ("p_constant", _pi), ("tolerance", _d),
        ("maxiter", ctypes.c_uint32), ("ndem", ctypes.c_uint32),
        ("c_phi", _d), ("f_p", _pd), ("beta_f", _d), ("k_c", _d),
        ("color", _pi),
    ]


class SphRunOpts(ctypes.Structure):
    _fields_ = [("input_path", ctypes.c_char_p), ("output_dir", ctypes.c_char_p),
                ("fluid", ctypes.c_int32), ("verbose", ctypes.c_int32),
                ("dry", ctypes.c_int32), ("device", ctypes.c_int32)]


class SphDarcyStats(ctypes.Structure):
    _fields_ = [("fluid_steps", ctypes.c_uint64), ("sweeps", ctypes.c_uint64),
                ("nonconverged", ctypes.c_uint64), ("last_nijac", ctypes.c_uint32),
                ("pad", ctypes.c_uint32), ("last_max_residual", _d),
                ("solve_ms", _d), ("corrections", ctypes.c_uint64)]


# Names of the array members and their (dtype, per-item width, count key).
PARTICLE_ARRAYS = [("x", numpy.float64, 4), ("xyzsum", numpy.float64, 3),
                   ("vel", numpy.float64, 4), ("force", numpy.float64, 3),
                   ("angpos", numpy.float64, 3), ("angvel", numpy.float64, 3),
                   ("torque", numpy.float64, 3), ("es_dot", numpy.float64, 1),
                   ("es", numpy.float64, 1), ("ev_dot", numpy.float64, 1),
                   ("ev", numpy.float64, 1), ("p", numpy.float64, 1),
                   ("color", numpy.int32, 1)]
BOND_ARRAYS = [("bonds", numpy.uint32, 2), ("bonds_delta_n", numpy.float64, 1),
               ("bonds_delta_t", numpy.float64, 3), ("bonds_omega_n", numpy.float64, 1),
               ("bonds_omega_t", numpy.float64, 3)]
CELL_ARRAYS = [("v_f", numpy.float64, 3), ("p_f", numpy.float64, 1),
               ("phi", numpy.float64, 1), ("dphi", numpy.float64, 1),
               ("p_constant", numpy.int32, 1)]
FLUID_PARTICLE_ARRAYS = [("f_p", numpy.float64, 3)]

SCALARS = [f[0] for f in SphState._fields_
           if f[0] not in {a[0] for a in PARTICLE_ARRAYS + BOND_ARRAYS + CELL_ARRAYS
                           + FLUID_PARTICLE_ARRAYS}]


def _ptr(arr, ctype):
    return arr.ctypes.data_as(ctypes.POINTER(ctype))


_CT = {numpy.float64: ctypes.c_double, numpy.uint32: ctypes.c_uint32,
       numpy.int32: ctypes.c_int32}


class SphNeighborStats(ctypes.Structure):
    _fields_ = [("enabled", ctypes.c_uint32), ("range", ctypes.c_uint32),
                ("cap", ctypes.c_uint32), ("overflow", ctypes.c_uint32),
                ("skin", ctypes.c_double), ("rebuilds", ctypes.c_uint64),
                ("entries", ctypes.c_uint64), ("longest", ctypes.c_uint32),
                ("full_steps", ctypes.c_uint32)]


class StateBuffers:
    """Owns numpy arrays and a SphState that points at them."""

    def __init__(self, hdr):
        self.state = SphState()
        for name in SCALARS:
            setattr(self.state, name, getattr(hdr, name))
        self.arrays = {}
        np_ = hdr.np
        nb = hdr.nb0
        cells = hdr.num[0] * hdr.num[1] * hdr.num[2]
        specs = [(a, np_) for a in PARTICLE_ARRAYS] + [(a, nb) for a in BOND_ARRAYS]
        if hdr.fluid:
            specs += [(a, cells) for a in CELL_ARRAYS]
            specs += [(a, np_) for a in FLUID_PARTICLE_ARRAYS]
        for (name, dt, w), n in specs:
            arr = numpy.zeros(max(n * w, 1), dtype=dt)
            self.arrays[name] = arr
            setattr(self.state, name, _ptr(arr, _CT[dt]))

    def get(self, name, n, w):
        a = self.arrays[name][:n * w]
        return a.reshape(n, w) if w > 1 else a


_lib = None


class SphEnsembleInfo(ctypes.Structure):
    _fields_ = [("boxes", ctypes.c_uint32), ("multi_boxes", ctypes.c_uint32),
                ("launches_per_step", ctypes.c_uint32), ("graph", ctypes.c_uint32),
                ("multi_steps", ctypes.c_uint64), ("graph_replays", ctypes.c_uint64),
                ("fluid_boxes", ctypes.c_uint32), ("pad", ctypes.c_uint32),
                ("fluid_steps", ctypes.c_uint64)]


class SphNsParams(ctypes.Structure):
    """sph_ns_params of include/sphere_hip.h."""
    _fields_ = [("nx", ctypes.c_uint32), ("ny", ctypes.c_uint32), ("nz", ctypes.c_uint32),
                ("L", ctypes.c_double * 3), ("dt", ctypes.c_double), ("t0", ctypes.c_double),
                ("mu", ctypes.c_double), ("rho_f", ctypes.c_double), ("g", ctypes.c_double * 3),
                ("beta", ctypes.c_double), ("theta", ctypes.c_double), ("gamma", ctypes.c_double),
                ("tolerance", ctypes.c_double), ("c_v", ctypes.c_double),
                ("maxiter", ctypes.c_uint32), ("ndem", ctypes.c_uint32),
                ("bc_bot", ctypes.c_int32), ("bc_top", ctypes.c_int32),
                ("p_mod_A", ctypes.c_double), ("p_mod_f", ctypes.c_double),
                ("p_mod_phi", ctypes.c_double), ("c_phi", ctypes.c_double)]


class SphNsStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_uint64), ("solves", ctypes.c_uint64),
                ("iterations", ctypes.c_uint64), ("nonconverged", ctypes.c_uint64),
                ("last_nijac", ctypes.c_uint32), ("last_max_norm", ctypes.c_double)]


def load_library(path=None):
    """Load libsphere_hip.so; raises if it was not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            "libsphere_hip.so is not built (%s); run __graft_entry__.build()" % p)
    lib = ctypes.CDLL(p)
    vp = ctypes.c_void_p
    sig = {
        "sph_create": (vp, [ctypes.c_int]),
        "sph_destroy": (None, [vp]),
        "sph_last_error": (ctypes.c_char_p, [vp]),
        "sph_set_contact_slots": (ctypes.c_int, [vp, ctypes.c_int]),
        "sph_set_neighbor_lists": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_int]),
        "sph_set_rebuild_kernel": (ctypes.c_int, [vp, ctypes.c_int]),
        "sph_set_barrier_spin_limit": (ctypes.c_int, [vp, ctypes.c_uint32]),
        "sph_get_neighbor_stats": (ctypes.c_int, [vp, ctypes.POINTER(SphNeighborStats)]),
        "sph_load_state": (ctypes.c_int, [vp, ctypes.POINTER(SphState)]),
        "sph_load_file": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_int]),
        "sph_get_header": (ctypes.c_int, [vp, ctypes.POINTER(SphState)]),
        "sph_read_state": (ctypes.c_int, [vp, ctypes.POINTER(SphState)]),
        "sph_write_file": (ctypes.c_int, [vp, ctypes.c_char_p]),
        "sph_step": (ctypes.c_int, [vp, ctypes.c_int]),
        "sph_run": (ctypes.c_int, [ctypes.POINTER(SphRunOpts)]),
        "sph_read_sort": (ctypes.c_int, [vp, _pu, _pu, _pu, _pu]),
        "sph_read_contacts": (ctypes.c_int, [vp, _pu, _pd, _pd]),
        "sph_write_contacts": (ctypes.c_int, [vp, _pu, _pd, _pd]),
        "sph_count_contacts": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_uint64)]),
        "sph_write_restart": (ctypes.c_int, [vp, ctypes.c_char_p]),
        "sph_read_restart": (ctypes.c_int, [vp, ctypes.c_char_p]),
        "sph_get_darcy_stats": (ctypes.c_int, [vp, ctypes.POINTER(SphDarcyStats)]),
        "sph_set_profiling": (ctypes.c_int, [vp, ctypes.c_int]),
        "sph_kernel_time": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_uint64)]),
        "sph_net_param_count": (ctypes.c_uint64, []),
        "sph_net_create": (vp, [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_uint32]),
        "sph_net_destroy": (None, [vp]),
        "sph_net_last_error": (ctypes.c_char_p, []),
        "sph_net_forward": (ctypes.c_int, [vp, vp, vp, vp, vp]),
        "sph_net_backward": (ctypes.c_int, [vp, vp, vp, vp, vp]),
        "sph_net_conv3d": (ctypes.c_int, [vp, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_uint32,
                                          ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, vp, vp]),
        "sph_net_conv3d_noalloc": (ctypes.c_int, [vp, ctypes.c_int, vp, vp, ctypes.c_int,
                                                  ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                                  ctypes.c_uint32, vp, vp]),
        "sph_ensemble_create": (vp, [ctypes.POINTER(vp), ctypes.c_int]),
        "sph_ensemble_destroy": (None, [vp]),
        "sph_ensemble_last_error": (ctypes.c_char_p, [vp]),
        "sph_ensemble_step": (ctypes.c_int, [vp, ctypes.c_int]),
        "sph_ensemble_stats": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_uint64)]),
        "sph_ensemble_info": (ctypes.c_int, [vp, ctypes.POINTER(SphEnsembleInfo)]),
        "sph_ensemble_set_graph": (ctypes.c_int, [vp, ctypes.c_int]),
        "sph_ensemble_autotune": (ctypes.c_int, [vp, ctypes.c_int, _pd, ctypes.c_int, _pd, _pd]),
        "sph_set_darcy_net": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_float), ctypes.c_uint64,
                                             ctypes.c_int]),
        "sph_darcy_export": (ctypes.c_int, [vp, ctypes.c_int, vp]),
        "sph_darcy_sweeps": (ctypes.c_int, [vp, vp, ctypes.c_int]),
        "sph_darcy_net_input": (ctypes.c_int, [vp, vp, vp, ctypes.POINTER(ctypes.c_double)]),
        "sph_ns_create": (vp, [ctypes.c_int, ctypes.POINTER(SphNsParams), _pd, _pd]),
        "sph_ns_destroy": (None, [vp]),
        "sph_ns_last_error": (ctypes.c_char_p, []),
        "sph_ns_step": (ctypes.c_int, [vp, ctypes.c_int]),
        "sph_ns_read": (ctypes.c_int, [vp, _pd, _pd, _pd, _pd]),
        "sph_ns_read_field": (ctypes.c_int, [vp, ctypes.c_int, _pd]),
        "sph_ns_get_stats": (ctypes.c_int, [vp, ctypes.POINTER(SphNsStats)]),
        "sph_ns_set_grains": (ctypes.c_int, [vp, ctypes.c_uint32, _pd, _pd, _pd, ctypes.c_int32]),
        "sph_ns_read_grain_forces": (ctypes.c_int, [vp, _pd, _pd, _pd, _pd]),
        "sph_ns_attach": (ctypes.c_int, [vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


class SphereError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("sphere_hip error %d (%s): %s" % (code, ERRORS.get(code, "?"), msg))
        self.code = code


class Engine:
    """One library context (one device, one loaded state)."""

    def __init__(self, device=0, lib=None, nc=32, skin=None, nl_cap=128, fused_rebuild=True):
        """nc: contact slots per particle (32 or 64); skin: neighbour-list
        skin as a fraction of the largest radius (None keeps the library
        default, 0 searches the 27 cells every step like the reference);
        nl_cap: neighbour-list entries per particle; fused_rebuild: list
        rebuilds as one kernel (True) or as separate sort, reorder and list
        launches (False)."""
        self.lib = lib or load_library()
        self.ctx = self.lib.sph_create(device)
        if not self.ctx:
            raise SphereError(-4, self.lib.sph_last_error(None).decode())
        self.nc = int(nc)
        self._check(self.lib.sph_set_contact_slots(self.ctx, self.nc))
        if not fused_rebuild:
            self._check(self.lib.sph_set_rebuild_kernel(self.ctx, 0))
        if skin is not None or nl_cap != 128:
            self._check(self.lib.sph_set_neighbor_lists(
                self.ctx, 0.1 if skin is None else float(skin), int(nl_cap)))

    def _check(self, rc):
        if rc != SPH_OK:
            raise SphereError(rc, self.lib.sph_last_error(self.ctx).decode())

    def close(self):
        if self.ctx:
            self.lib.sph_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, st):
        """Load a StateBuffers image (e.g. sim.to_state())."""
        self._check(self.lib.sph_load_state(self.ctx, ctypes.byref(st.state)))

    def load_file(self, path, fluid):
        self._check(self.lib.sph_load_file(self.ctx, path.encode(), int(fluid)))

    def header(self):
        h = SphState()
        self._check(self.lib.sph_get_header(self.ctx, ctypes.byref(h)))
        return h

    def read(self):
        buf = StateBuffers(self.header())
        self._check(self.lib.sph_read_state(self.ctx, ctypes.byref(buf.state)))
        return buf

    def step(self, n=1):
        self._check(self.lib.sph_step(self.ctx, int(n)))

    def write_file(self, path):
        self._check(self.lib.sph_write_file(self.ctx, path.encode()))

    def read_sort(self):
        h = self.header()
        ncell = h.num[0] * h.num[1] * h.num[2]
        cid = numpy.zeros(max(h.np, 1), numpy.uint32)
        idx = numpy.zeros(max(h.np, 1), numpy.uint32)
        cs = numpy.zeros(ncell, numpy.uint32)
        ce = numpy.zeros(ncell, numpy.uint32)
        self._check(self.lib.sph_read_sort(self.ctx, _ptr(cid, ctypes.c_uint32),
                                           _ptr(idx, ctypes.c_uint32),
                                           _ptr(cs, ctypes.c_uint32), _ptr(ce, ctypes.c_uint32)))
        return cid[:h.np], idx[:h.np], cs, ce

    def read_contacts(self):
        h = self.header()
        n = max(h.np, 1) * self.nc
        c = numpy.zeros(n, numpy.uint32)
        dm = numpy.zeros(n * 3, numpy.float64)
        dt = numpy.zeros(n * 3, numpy.float64)
        self._check(self.lib.sph_read_contacts(self.ctx, _ptr(c, ctypes.c_uint32),
                                               _ptr(dm, ctypes.c_double), _ptr(dt, ctypes.c_double)))
        m = h.np * self.nc
        return c[:m], dm[:3 * m].reshape(-1, 3), dt[:3 * m].reshape(-1, 3)

    def write_contacts(self, contacts, distmod, delta_t):
        contacts = numpy.ascontiguousarray(contacts, numpy.uint32)
        distmod = numpy.ascontiguousarray(distmod, numpy.float64)
        delta_t = numpy.ascontiguousarray(delta_t, numpy.float64)
        self._check(self.lib.sph_write_contacts(self.ctx, _ptr(contacts, ctypes.c_uint32),
                                                _ptr(distmod, ctypes.c_double),
                                                _ptr(delta_t, ctypes.c_double)))

    def write_restart(self, path):
        """Write the restart side car of the current state."""
        self._check(self.lib.sph_write_restart(self.ctx, path.encode()))

    def read_restart(self, path):
        """Restore a side car written with the loaded state."""
        self._check(self.lib.sph_read_restart(self.ctx, path.encode()))

    def count_contacts(self):
        n = ctypes.c_uint64(0)
        self._check(self.lib.sph_count_contacts(self.ctx, ctypes.byref(n)))
        return n.value

    def attach_fluid(self, ns):
        """Couple a NavierStokes fluid on this context's grid to its grains
        (None detaches); each step then runs the fluid and adds its force
        on the grains."""
        self._check(self.lib.sph_ns_attach(ns.h if ns is not None else None, self.ctx))
        if ns is not None:
            ns.np = int(self.header().np)  # rows of grain_forces()
        self._fluid = ns

    def darcy_stats(self):
        s = SphDarcyStats()
        self._check(self.lib.sph_get_darcy_stats(self.ctx, ctypes.byref(s)))
        return s

    def neighbor_stats(self):
        s = SphNeighborStats()
        self._check(self.lib.sph_get_neighbor_stats(self.ctx, ctypes.byref(s)))
        return s

    def set_barrier_spin_limit(self, sleeps):
        """Bound of the list rebuild kernel's grid-barrier waits (0: default)."""
        self._check(self.lib.sph_set_barrier_spin_limit(self.ctx, int(sleeps)))

    def set_profiling(self, on):
        self._check(self.lib.sph_set_profiling(self.ctx, int(on)))

    # learned Darcy correction (device buffers are torch CUDA tensors)
    def set_darcy_net(self, params, max_corrections=8):
        """params: FP32 host array of Net.param_count() values, or None."""
        if params is None:
            self._check(self.lib.sph_set_darcy_net(self.ctx, None, 0, 0))
            return
        p = numpy.ascontiguousarray(params, dtype=numpy.float32)
        self._check(self.lib.sph_set_darcy_net(self.ctx, _ptr(p, ctypes.c_float), p.size,
                                               int(max_corrections)))

    # The context runs on its own stream and returns when its work is done;
    # torch's pending work on the buffers is finished first.
    @staticmethod
    def _torch_done():
        import torch
        torch.cuda.current_stream().synchronize()

    def darcy_export(self, which, dst):
        """Copy the last solve's start (0) or final (1) pressure, ghost layer
        included, into the float64 CUDA tensor dst."""
        self._torch_done()
        self._check(self.lib.sph_darcy_export(self.ctx, int(which), ctypes.c_void_p(dst.data_ptr())))

    def darcy_sweeps(self, p, n):
        self._torch_done()
        self._check(self.lib.sph_darcy_sweeps(self.ctx, ctypes.c_void_p(p.data_ptr()), int(n)))

    def darcy_net_input(self, p, inp):
        """Network input for iterate p (float64 CUDA tensor with ghosts) into
        the float32 CUDA tensor inp [V, 3]; returns the residual scale."""
        self._torch_done()
        s = ctypes.c_double(0)
        self._check(self.lib.sph_darcy_net_input(self.ctx, ctypes.c_void_p(p.data_ptr()),
                                                 ctypes.c_void_p(inp.data_ptr()), ctypes.byref(s)))
        return s.value

    def kernel_time(self, kid):
        t = ctypes.c_double(0)
        n = ctypes.c_uint64(0)
        self._check(self.lib.sph_kernel_time(self.ctx, int(kid), ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value


class Net:
    """The Darcy correction network (sph_net_*) on one device.  Buffers are
    torch CUDA tensors; calls run on torch's current stream."""

    def __init__(self, nx, ny, nz, periodic=0, device=0, lib=None):
        """periodic: bit 0 x, bit 1 y wrap around."""
        self.lib = lib or load_library()
        self.h = self.lib.sph_net_create(int(device), int(nx), int(ny), int(nz), int(periodic))
        if not self.h:
            raise SphereError(-3, self.lib.sph_net_last_error().decode())
        self.shape = (int(nz), int(ny), int(nx))

    @staticmethod
    def param_count(lib=None):
        return int((lib or load_library()).sph_net_param_count())

    def _stream(self):
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def _check(self, rc):
        if rc != SPH_OK:
            raise SphereError(rc, self.lib.sph_net_last_error().decode())

    def forward(self, params, inp, out):
        self._check(self.lib.sph_net_forward(self.h, ctypes.c_void_p(params.data_ptr()),
                                             ctypes.c_void_p(inp.data_ptr()),
                                             ctypes.c_void_p(out.data_ptr()), self._stream()))

    def backward(self, params, grad_out, grad_params):
        self._check(self.lib.sph_net_backward(self.h, ctypes.c_void_p(params.data_ptr()),
                                              ctypes.c_void_p(grad_out.data_ptr()),
                                              ctypes.c_void_p(grad_params.data_ptr()), self._stream()))

    @staticmethod
    def conv3d(inp, w, b, out, shape, periodic=0, lib=None):
        """One convolution of the network's kernels (torch CUDA tensors):
        inp [V, cin], w [27, cout, cin], b [cout] or None, out [V, cout]."""
        import torch
        lib = lib or load_library()
        nz, ny, nx = shape
        torch.cuda.current_stream().synchronize()
        rc = lib.sph_net_conv3d(ctypes.c_void_p(inp.data_ptr()), int(inp.shape[1]),
                                ctypes.c_void_p(w.data_ptr()),
                                ctypes.c_void_p(b.data_ptr()) if b is not None else None,
                                int(w.shape[1]), nx, ny, nz, int(periodic),
                                ctypes.c_void_p(out.data_ptr()),
                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        if rc != SPH_OK:
            raise SphereError(rc, lib.sph_net_last_error().decode())

    def close(self):
        if self.h:
            self.lib.sph_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ns_params(nx, ny, nz, L, dt, mu=8.9e-4, rho_f=1.0e3, g=(0.0, 0.0, 0.0), beta=0.0, theta=1.0,
              gamma=0.0, tolerance=1.0e-3, c_v=1.0, maxiter=10000, ndem=1, bc_bot=0, bc_top=0,
              p_mod_A=0.0, p_mod_f=0.0, p_mod_phi=0.0, t0=0.0, c_phi=1.0):
    """sph_ns_params with the reference's Navier-Stokes defaults
    (python/sphere.py:3395-3508)."""
    P = SphNsParams()
    P.nx, P.ny, P.nz = int(nx), int(ny), int(nz)
    for k in range(3):
        P.L[k] = float(L[k])
        P.g[k] = float(g[k])
    P.dt, P.t0, P.mu, P.rho_f = float(dt), float(t0), float(mu), float(rho_f)
    P.beta, P.theta, P.gamma = float(beta), float(theta), float(gamma)
    P.tolerance, P.c_v = float(tolerance), float(c_v)
    P.maxiter, P.ndem = int(maxiter), int(ndem)
    P.bc_bot, P.bc_top = int(bc_bot), int(bc_top)
    P.p_mod_A, P.p_mod_f, P.p_mod_phi = float(p_mod_A), float(p_mod_f), float(p_mod_phi)
    P.c_phi = float(c_phi)
    return P


def _one_value(a):
    return numpy.asarray(a).ravel()[0]


def ns_params_from_sim(s, grains=False):
    """sph_ns_params of a sim with a Navier-Stokes fluid (cfd_solver = 0):
    its grid, time step and time, fluid and solver fields (the ones the
    reference's binary file carries).  Grains and walls only with
    grains=True (a fluid attached to an Engine that steps them)."""
    if not s.fluid or int(_one_value(s.cfd_solver)) != 0:
        raise SphereError(SPH_ERR_ARG, "the sim has no Navier-Stokes fluid (cfd_solver = 0)")
    if (int(_one_value(s.np)) != 0 or int(_one_value(s.nw)) != 0) and not grains:
        raise SphereError(SPH_ERR_ARG, "grains and walls need the fluid attached to an Engine "
                                       "(grains=True)")
    return ns_params(*[int(n) for n in s.num], L=s.L, dt=_one_value(s.time_dt),
                     mu=_one_value(s.mu), rho_f=_one_value(s.rho_f), g=s.g,
                     beta=_one_value(s.beta), theta=_one_value(s.theta), gamma=_one_value(s.gamma),
                     tolerance=_one_value(s.tolerance), c_v=_one_value(s.c_v),
                     maxiter=_one_value(s.maxiter), ndem=_one_value(s.ndem),
                     bc_bot=_one_value(s.bc_bot), bc_top=_one_value(s.bc_top),
                     p_mod_A=_one_value(s.p_mod_A), p_mod_f=_one_value(s.p_mod_f),
                     p_mod_phi=_one_value(s.p_mod_phi), t0=_one_value(s.time_current),
                     c_phi=_one_value(s.c_phi))


class NavierStokes:
    """A Navier-Stokes fluid without grains (sph_ns_*): the reference's
    cfd_solver = 0 fluid step on the GPU.  Arrays are numpy, interior cells
    indexed [x, y, z] like sim.p_f and sim.v_f."""

    def __init__(self, params, v, p, device=0, lib=None):
        self.lib = lib or load_library()
        self.P = params
        self.shape = (params.nx, params.ny, params.nz)
        vc = numpy.ascontiguousarray(numpy.asarray(v, dtype=numpy.float64).transpose(2, 1, 0, 3))
        pc = numpy.ascontiguousarray(numpy.asarray(p, dtype=numpy.float64).transpose(2, 1, 0))
        self.h = self.lib.sph_ns_create(int(device), ctypes.byref(params), _ptr(vc, ctypes.c_double),
                                        _ptr(pc, ctypes.c_double))
        if not self.h:
            raise SphereError(SPH_ERR_ARG, self.lib.sph_ns_last_error().decode())

    @classmethod
    def from_sim(cls, s, device=0, lib=None):
        """The fluid of a sim with cfd_solver = 0, no grains and no walls."""
        return cls(ns_params_from_sim(s), s.v_f, s.p_f, device=device, lib=lib)

    def _check(self, rc):
        if rc != SPH_OK:
            raise SphereError(rc, self.lib.sph_ns_last_error().decode())

    def step(self, n=1):
        self._check(self.lib.sph_ns_step(self.h, int(n)))

    def read(self):
        """(v [nx, ny, nz, 3], p, phi, dphi [nx, ny, nz])."""
        nx, ny, nz = self.shape
        v = numpy.empty((nz, ny, nx, 3))
        p, phi, dphi = (numpy.empty((nz, ny, nx)) for _ in range(3))
        self._check(self.lib.sph_ns_read(self.h, *[_ptr(a, ctypes.c_double) for a in (v, p, phi, dphi)]))
        return (v.transpose(2, 1, 0, 3).copy(), p.transpose(2, 1, 0).copy(),
                phi.transpose(2, 1, 0).copy(), dphi.transpose(2, 1, 0).copy())

    def to_sim(self, s, steps_taken=None):
        """Write the fluid fields back into a sim (as an output file holds them)."""
        v, p, phi, dphi = self.read()
        s.v_f[:] = v
        s.p_f[:] = p
        s.phi[:] = phi
        s.dphi[:] = dphi

    def field(self, which):
        """Whole ghosted array: 0-2 face velocities, 3 epsilon."""
        nx, ny, nz = self.shape
        n = (nx + 3) * (ny + 3) * (nz + 3) if which < 3 else (nx + 2) * (ny + 2) * (nz + 2)
        out = numpy.empty(n)
        self._check(self.lib.sph_ns_read_field(self.h, int(which), _ptr(out, ctypes.c_double)))
        return out

    def stats(self):
        st = SphNsStats()
        self._check(self.lib.sph_ns_get_stats(self.h, ctypes.byref(st)))
        return st

    def set_grains(self, x, vel, origo=(0.0, 0.0, 0.0), periodic=1):
        """Grains held in place: x [np, 4] (position, radius), vel [np, 4]
        or [np, 3]."""
        x = numpy.ascontiguousarray(x, dtype=numpy.float64).reshape(-1, 4)
        v = numpy.zeros_like(x)
        vel = numpy.asarray(vel, dtype=numpy.float64).reshape(len(x), -1)
        v[:, :3] = vel[:, :3]
        o = numpy.ascontiguousarray(origo, dtype=numpy.float64)
        self.np = len(x)
        self._check(self.lib.sph_ns_set_grains(self.h, len(x), _ptr(x, ctypes.c_double),
                                               _ptr(v, ctypes.c_double), _ptr(o, ctypes.c_double),
                                               int(periodic)))

    def grain_forces(self):
        """(f_d, f_p, f_v, f_sum), each [np, 3], of the last step."""
        n = getattr(self, "np", 0)
        out = [numpy.zeros((n, 3)) for _ in range(4)]
        self._check(self.lib.sph_ns_read_grain_forces(self.h, *[_ptr(a, ctypes.c_double) for a in out]))
        return tuple(out)

    def close(self):
        if self.h:
            self.lib.sph_ns_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """Independent boxes (Engines with loaded states, one device) stepped
    together: boxes on the list path without fluid share one kernel launch
    per stage (task records per box, pairs of steps replayed from a captured
    HIP graph); the others are enqueued box by box on their own streams."""

    def __init__(self, engines, lib=None):
        self.lib = lib or load_library()
        self.engines = list(engines)  # keeps the contexts alive
        arr = (ctypes.c_void_p * len(self.engines))(*[e.ctx for e in self.engines])
        self.h = self.lib.sph_ensemble_create(arr, len(self.engines))
        if not self.h:
            raise SphereError(-1, self.lib.sph_ensemble_last_error(None).decode())

    def step(self, n=1):
        rc = self.lib.sph_ensemble_step(self.h, int(n))
        if rc != SPH_OK:
            raise SphereError(rc, self.lib.sph_ensemble_last_error(self.h).decode())

    def stats(self):
        a = ctypes.c_uint64(0)
        self.lib.sph_ensemble_stats(self.h, ctypes.byref(a))
        return {"steps": a.value}

    def info(self):
        i = SphEnsembleInfo()
        self.lib.sph_ensemble_info(self.h, ctypes.byref(i))
        return {k: getattr(i, k) for k, _ in SphEnsembleInfo._fields_}

    def set_graph(self, on):
        self.lib.sph_ensemble_set_graph(self.h, 1 if on else 0)

    def autotune(self, trial_steps, fractions):
        """Each box's list skin (fraction of its largest radius) chosen by
        timing trial_steps steps with each fraction; every box advances by
        len(fractions) * trial_steps steps.  Returns (chosen per box, ms per
        step [box][fraction], -1 where a fraction does not fit)."""
        f = numpy.ascontiguousarray(fractions, dtype=numpy.float64)
        nb = len(self.engines)
        chosen = numpy.zeros(nb)
        ms = numpy.zeros((nb, len(f)))
        rc = self.lib.sph_ensemble_autotune(self.h, int(trial_steps), _ptr(f, ctypes.c_double),
                                            len(f), _ptr(chosen, ctypes.c_double),
                                            _ptr(ms, ctypes.c_double))
        if rc != SPH_OK:
            raise SphereError(rc, self.lib.sph_ensemble_last_error(self.h).decode())
        return chosen, ms

    def close(self):
        if self.h:
            self.lib.sph_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
