"""Python 3 ``sim`` object: the subset of python/sphere.py that drives the DEM
step and the Darcy solve, with the same names, arguments and file layout.

Format: versioned binary v2.15, written field by field in the order of
python/sphere.py:1355-1553 and read back with the version switches of
python/sphere.py:965-1353 (files down to v2.11 read).  ``run()`` writes
``../input/<sid>.bin`` and executes the run in process through the C ABI of
libsphere_hip.so, which writes ``../output/<sid>.outputNNNNN.bin``,
``../output/<sid>.status.dat`` and ``../output/<sid>-conv.log`` exactly like
``cd ..; ./sphere input/<sid>.bin`` in the reference (python/sphere.py:4482).
"""
import ctypes
import math
import os

import numpy

from . import _abi

VERSION = 2.15  # src/version.h:5


def V_sphere(r):
    """Volume of a sphere (python/sphere.py:7762-7769)."""
    return 4.0 / 3.0 * math.pi * r**3.0


def status(project):
    """Last output file number from ../output/<sid>.status.dat
    (python/sphere.py:7721-7741)."""
    with open("../output/{0}.status.dat".format(project)) as fh:
        return int(fh.read().split()[2])


def cleanup(sim_obj):
    """Remove input/output files of a simulation (python/sphere.py:7743-7760)."""
    import glob
    pats = ["../input/" + sim_obj.sid + ".bin", "../output/" + sim_obj.sid + ".*.bin",
            "../output/" + sim_obj.sid + ".status.dat", "../output/" + sim_obj.sid + "-conv.log"]
    for pat in pats:
        for f in glob.glob(pat):
            os.remove(f)


class _Reader:
    def __init__(self, data):
        self.b = data
        self.o = 0

    def take(self, dtype, count=1):
        dt = numpy.dtype(dtype)
        n = dt.itemsize * int(count)
        if self.o + n > len(self.b):
            raise IOError("binary file is truncated")
        a = numpy.frombuffer(self.b, dtype=dt, count=int(count), offset=self.o).copy()
        self.o += n
        return a


class sim:
    """Class containing all ``sphere`` data (python/sphere.py:33-455)."""

    def __init__(self, sid='unnamed', np=0, nd=3, nw=0, fluid=False, cfd_solver=0):
        self.version = numpy.ones(1, dtype=numpy.float64) * VERSION
        self.nd = numpy.ones(1, dtype=numpy.int32) * int(nd)
        self.np = numpy.ones(1, dtype=numpy.uint32) * int(np)
        self.sid = sid
        self.time_dt = numpy.zeros(1, dtype=numpy.float64)
        self.time_current = numpy.zeros(1, dtype=numpy.float64)
        self.time_total = numpy.zeros(1, dtype=numpy.float64)
        self.time_file_dt = numpy.zeros(1, dtype=numpy.float64)
        self.time_step_count = numpy.zeros(1, dtype=numpy.uint32)
        self.origo = numpy.zeros(self.nd, dtype=numpy.float64)
        self.L = numpy.zeros(self.nd, dtype=numpy.float64)
        self.num = numpy.zeros(self.nd, dtype=numpy.uint32)
        self.periodic = numpy.zeros(1, dtype=numpy.uint32)
        self.adaptive = numpy.zeros(1, dtype=numpy.uint32)
        n = int(self.np[0])
        d = int(self.nd[0])
        self.x = numpy.zeros((n, d), dtype=numpy.float64)
        self.radius = numpy.ones(n, dtype=numpy.float64)
        self.xyzsum = numpy.zeros((n, 3), dtype=numpy.float64)
        self.vel = numpy.zeros((n, d), dtype=numpy.float64)
        self.fixvel = numpy.zeros(n, dtype=numpy.float64)
        self.force = numpy.zeros((n, d), dtype=numpy.float64)
        self.angpos = numpy.zeros((n, d), dtype=numpy.float64)
        self.angvel = numpy.zeros((n, d), dtype=numpy.float64)
        self.torque = numpy.zeros((n, d), dtype=numpy.float64)
        self.es_dot = numpy.zeros(n, dtype=numpy.float64)
        self.es = numpy.zeros(n, dtype=numpy.float64)
        self.ev_dot = numpy.zeros(n, dtype=numpy.float64)
        self.ev = numpy.zeros(n, dtype=numpy.float64)
        self.p = numpy.zeros(n, dtype=numpy.float64)
        self.g = numpy.array([0.0, 0.0, 0.0], dtype=numpy.float64)
        self.k_n = numpy.ones(1, dtype=numpy.float64) * 1.16e9
        self.k_t = numpy.ones(1, dtype=numpy.float64) * 1.16e9
        self.k_r = numpy.zeros(1, dtype=numpy.float64)
        self.E = numpy.zeros(1, dtype=numpy.float64)
        self.gamma_n = numpy.zeros(1, dtype=numpy.float64)
        self.gamma_t = numpy.zeros(1, dtype=numpy.float64)
        self.gamma_r = numpy.zeros(1, dtype=numpy.float64)
        self.mu_s = numpy.ones(1, dtype=numpy.float64) * 0.5
        self.mu_d = numpy.ones(1, dtype=numpy.float64) * 0.5
        self.mu_r = numpy.zeros(1, dtype=numpy.float64)
        self.gamma_wn = numpy.ones(1, dtype=numpy.float64) * 1.0e3
        self.gamma_wt = numpy.zeros(1, dtype=numpy.float64)
        self.mu_ws = numpy.ones(1, dtype=numpy.float64) * 0.5
        self.mu_wd = numpy.ones(1, dtype=numpy.float64) * 0.5
        self.rho = numpy.ones(1, dtype=numpy.float64) * 2600.0
        self.contactmodel = numpy.ones(1, dtype=numpy.uint32) * 2
        self.kappa = numpy.zeros(1, dtype=numpy.float64)
        self.db = numpy.zeros(1, dtype=numpy.float64)
        self.V_b = numpy.zeros(1, dtype=numpy.float64)
        self.nw = numpy.ones(1, dtype=numpy.uint32) * int(nw)
        w = int(self.nw[0])
        self.wmode = numpy.zeros(w, dtype=numpy.int32)
        self.w_n = numpy.zeros((w, d), dtype=numpy.float64)
        if w >= 1:
            self.w_n[0, 2] = -1.0
        if w >= 2:
            self.w_n[1, 0] = -1.0
        if w >= 3:
            self.w_n[2, 0] = 1.0
        if w >= 4:
            self.w_n[3, 1] = -1.0
        if w >= 5:
            self.w_n[4, 1] = 1.0
        self.w_x = numpy.ones(w, dtype=numpy.float64)
        self.w_m = numpy.zeros(w, dtype=numpy.float64)
        self.w_vel = numpy.zeros(w, dtype=numpy.float64)
        self.w_force = numpy.zeros(w, dtype=numpy.float64)
        self.w_sigma0 = numpy.zeros(w, dtype=numpy.float64)
        self.w_sigma0_A = numpy.zeros(1, dtype=numpy.float64)
        self.w_sigma0_f = numpy.zeros(1, dtype=numpy.float64)
        self.w_tau_x = numpy.zeros(1, dtype=numpy.float64)
        self.lambda_bar = numpy.ones(1, dtype=numpy.float64)
        self.nb0 = numpy.zeros(1, dtype=numpy.uint32)
        self.sigma_b = numpy.ones(1, dtype=numpy.float64) * numpy.inf
        self.tau_b = numpy.ones(1, dtype=numpy.float64) * numpy.inf
        self.bonds = numpy.zeros((0, 2), dtype=numpy.uint32)
        self.bonds_delta_n = numpy.zeros(0, dtype=numpy.float64)
        self.bonds_delta_t = numpy.zeros((0, d), dtype=numpy.float64)
        self.bonds_omega_n = numpy.zeros(0, dtype=numpy.float64)
        self.bonds_omega_t = numpy.zeros((0, d), dtype=numpy.float64)
        self.fluid = fluid
        if self.fluid:
            # python/sphere.py:316-451: the constructor always starts from the
            # Navier-Stokes defaults; initFluid(cfd_solver=1) selects Darcy.
            self.cfd_solver = numpy.zeros(1, dtype=numpy.int32)
            self.mu = numpy.zeros(1, dtype=numpy.float64)
            self._alloc_fluid(p=0.0)
            self.rho_f = numpy.ones(1, dtype=numpy.float64) * 1.0e3
            self.gamma = numpy.array(0.0)
            self.theta = numpy.array(1.0)
            self.beta = numpy.array(0.0)
            self.tolerance = numpy.array(1.0e-3)
            self.maxiter = numpy.array(1e4)
            self.ndem = numpy.array(1)
            self.c_phi = numpy.ones(1, dtype=numpy.float64)
            self.c_v = numpy.ones(1, dtype=numpy.float64)
            self.dt_dem_fac = numpy.ones(1, dtype=numpy.float64)
            self.f_d = numpy.zeros((n, d), dtype=numpy.float64)
            self.f_p = numpy.zeros((n, d), dtype=numpy.float64)
            self.f_v = numpy.zeros((n, d), dtype=numpy.float64)
            self.f_sum = numpy.zeros((n, d), dtype=numpy.float64)
        self.color = numpy.zeros(n, dtype=numpy.int32)

    def _alloc_fluid(self, p):
        nn = (int(self.num[0]), int(self.num[1]), int(self.num[2]))
        self.v_f = numpy.zeros(nn + (int(self.nd[0]),), dtype=numpy.float64)
        self.p_f = numpy.ones(nn, dtype=numpy.float64) * p
        self.phi = numpy.zeros(nn, dtype=numpy.float64)
        self.dphi = numpy.zeros(nn, dtype=numpy.float64)
        self.p_mod_A = numpy.zeros(1, dtype=numpy.float64)
        self.p_mod_f = numpy.zeros(1, dtype=numpy.float64)
        self.p_mod_phi = numpy.zeros(1, dtype=numpy.float64)
        self.bc_bot = numpy.zeros(1, dtype=numpy.int32)
        self.bc_top = numpy.zeros(1, dtype=numpy.int32)
        self.free_slip_bot = numpy.ones(1, dtype=numpy.int32)
        self.free_slip_top = numpy.ones(1, dtype=numpy.int32)
        self.bc_bot_flux = numpy.zeros(1, dtype=numpy.float64)
        self.bc_top_flux = numpy.zeros(1, dtype=numpy.float64)
        self.p_f_constant = numpy.zeros(nn, dtype=numpy.int32)

    # ------------------------------------------------------------------
    # comparison (python/sphere.py:456-793; __cmp__ is Python 2 only)
    # ------------------------------------------------------------------
    _CMP_FIELDS = ['version', 'nd', 'np', 'time_dt', 'time_current', 'time_total',
                   'time_file_dt', 'time_step_count', 'origo', 'L', 'num', 'periodic',
                   'adaptive', 'x', 'radius', 'xyzsum', 'vel', 'fixvel', 'force', 'angpos',
                   'angvel', 'torque', 'es_dot', 'es', 'ev_dot', 'ev', 'p', 'g', 'k_n', 'k_t',
                   'k_r', 'E', 'gamma_n', 'gamma_t', 'gamma_r', 'mu_s', 'mu_d', 'mu_r',
                   'gamma_wn', 'gamma_wt', 'mu_ws', 'mu_wd', 'rho', 'contactmodel', 'kappa',
                   'db', 'V_b', 'nw', 'wmode', 'w_n', 'w_x', 'w_m', 'w_vel', 'w_force',
                   'w_sigma0', 'w_sigma0_A', 'w_sigma0_f', 'w_tau_x', 'lambda_bar', 'nb0',
                   'sigma_b', 'tau_b', 'bonds', 'bonds_delta_n', 'bonds_delta_t',
                   'bonds_omega_n', 'bonds_omega_t', 'fluid']
    _CMP_FLUID = ['cfd_solver', 'mu', 'v_f', 'p_f', 'phi', 'dphi', 'rho_f', 'p_mod_A',
                  'p_mod_f', 'p_mod_phi', 'bc_bot', 'bc_top', 'free_slip_bot',
                  'free_slip_top', 'bc_bot_flux', 'bc_top_flux']
    _CMP_DARCY = ['tolerance', 'maxiter', 'ndem', 'c_phi', 'beta_f', 'k_c']

    def differences(self, other, ignore=()):
        """Names of the fields that differ, in the order of the reference's
        __cmp__ (python/sphere.py:456-793)."""
        names = list(self._CMP_FIELDS)
        if self.fluid:
            names += self._CMP_FLUID
            if int(numpy.asarray(self.cfd_solver).ravel()[0]) == 1:
                names += self._CMP_DARCY
        out = []
        for nm in names:
            if nm in ignore:
                continue
            a, b = getattr(self, nm, None), getattr(other, nm, None)
            if isinstance(a, (bool, str)) or isinstance(b, (bool, str)):
                if a != b:
                    out.append(nm)
                continue
            a = numpy.asarray(a)
            b = numpy.asarray(b)
            if a.size != b.size or not numpy.array_equal(a.ravel(), b.ravel(), equal_nan=True):
                out.append(nm)
        return out

    def __eq__(self, other):
        return isinstance(other, sim) and not self.differences(other)

    __hash__ = object.__hash__

    # ------------------------------------------------------------------
    # binary I/O
    # ------------------------------------------------------------------
    def readbin(self, targetbin, verbose=True, bonds=True, sigma0mod=True, esysparticle=False):
        """Read a ``sphere`` binary file (python/sphere.py:965-1353)."""
        if verbose:
            print("Input file: {0}".format(targetbin))
        with open(targetbin, "rb") as fh:
            r = _Reader(fh.read())
        try:
            self.version = r.take(numpy.float64)
            self.nd = r.take(numpy.int32)
            self.np = r.take(numpy.uint32)
            n = int(self.np[0])
            d = int(self.nd[0])
            self.time_dt = r.take(numpy.float64)
            self.time_current = r.take(numpy.float64)
            self.time_total = r.take(numpy.float64)
            self.time_file_dt = r.take(numpy.float64)
            self.time_step_count = r.take(numpy.uint32)
            self.origo = r.take(numpy.float64, d)
            self.L = r.take(numpy.float64, d)
            self.num = r.take(numpy.uint32, d)
            self.periodic = r.take(numpy.int32)
            if self.version >= 2.14:
                self.adaptive = r.take(numpy.int32)
            else:
                self.adaptive = numpy.zeros(1, dtype=numpy.float64)
            xr = r.take(numpy.float64, n * (d + 1)).reshape(n, d + 1)
            self.x = xr[:, :d].copy()
            self.radius = xr[:, d].copy()
            if self.version >= 1.03:
                self.xyzsum = r.take(numpy.float64, n * 3).reshape(n, 3)
            else:
                self.xyzsum = r.take(numpy.float64, n * 2).reshape(n, 2)
            vf = r.take(numpy.float64, n * (d + 1)).reshape(n, d + 1)
            self.vel = vf[:, :d].copy()
            self.fixvel = vf[:, d].copy()
            self.force = r.take(numpy.float64, n * d).reshape(n, d)
            self.angpos = r.take(numpy.float64, n * d).reshape(n, d)
            self.angvel = r.take(numpy.float64, n * d).reshape(n, d)
            self.torque = r.take(numpy.float64, n * d).reshape(n, d)
            if esysparticle:
                return
            self.es_dot = r.take(numpy.float64, n)
            self.es = r.take(numpy.float64, n)
            self.ev_dot = r.take(numpy.float64, n)
            self.ev = r.take(numpy.float64, n)
            self.p = r.take(numpy.float64, n)
            self.g = r.take(numpy.float64, d)
            self.k_n = r.take(numpy.float64)
            self.k_t = r.take(numpy.float64)
            self.k_r = r.take(numpy.float64)
            if self.version >= 2.13:
                self.E = r.take(numpy.float64)
            else:
                self.E = numpy.zeros(1, dtype=numpy.float64)
            for nm in ['gamma_n', 'gamma_t', 'gamma_r', 'mu_s', 'mu_d', 'mu_r', 'gamma_wn',
                       'gamma_wt', 'mu_ws', 'mu_wd', 'rho']:
                setattr(self, nm, r.take(numpy.float64))
            self.contactmodel = r.take(numpy.uint32)
            self.kappa = r.take(numpy.float64)
            self.db = r.take(numpy.float64)
            self.V_b = r.take(numpy.float64)
            self.nw = r.take(numpy.uint32)
            w = int(self.nw[0])
            self.wmode = r.take(numpy.int32, w)
            wn = r.take(numpy.float64, w * (d + 1)).reshape(w, d + 1)
            self.w_n = wn[:, :d].copy()
            self.w_x = wn[:, d].copy()
            wm = r.take(numpy.float64, w * 4).reshape(w, 4)
            self.w_m = wm[:, 0].copy()
            self.w_vel = wm[:, 1].copy()
            self.w_force = wm[:, 2].copy()
            self.w_sigma0 = wm[:, 3].copy()
            if sigma0mod:
                self.w_sigma0_A = r.take(numpy.float64)
                self.w_sigma0_f = r.take(numpy.float64)
            if self.version >= 2.1:
                self.w_tau_x = r.take(numpy.float64)
            else:
                self.w_tau_x = numpy.zeros(1, dtype=numpy.float64)
            if bonds:
                self.lambda_bar = r.take(numpy.float64)
                self.nb0 = r.take(numpy.uint32)
                self.sigma_b = r.take(numpy.float64)
                self.tau_b = r.take(numpy.float64)
                nb = int(self.nb0[0])
                self.bonds = r.take(numpy.uint32, nb * 2).reshape(nb, 2)
                self.bonds_delta_n = r.take(numpy.float64, nb)
                self.bonds_delta_t = r.take(numpy.float64, nb * d).reshape(nb, d)
                self.bonds_omega_n = r.take(numpy.float64, nb)
                self.bonds_omega_t = r.take(numpy.float64, nb * d).reshape(nb, d)
            else:
                self.nb0 = numpy.zeros(1, dtype=numpy.uint32)
            if self.fluid:
                if self.version >= 2.0:
                    self.cfd_solver = r.take(numpy.int32)
                else:
                    self.cfd_solver = numpy.zeros(1, dtype=numpy.int32)
                self.mu = r.take(numpy.float64)
                nx, ny, nz = int(self.num[0]), int(self.num[1]), int(self.num[2])
                cells = r.take(numpy.float64, nx * ny * nz * 6).reshape(nz, ny, nx, 6)
                cells = cells.transpose(2, 1, 0, 3)
                self.v_f = cells[:, :, :, 0:3].copy()
                self.p_f = cells[:, :, :, 3].copy()
                self.phi = cells[:, :, :, 4].copy()
                # stored scaled by dt*ndem, with the ndem held before this read
                # (python/sphere.py:1198-1200)
                self.dphi = cells[:, :, :, 5] / (self.time_dt * self.ndem)
                if self.version >= 0.36:
                    self.rho_f = r.take(numpy.float64)
                    self.p_mod_A = r.take(numpy.float64)
                    self.p_mod_f = r.take(numpy.float64)
                    self.p_mod_phi = r.take(numpy.float64)
                    if self.version >= 2.12 and self.cfd_solver[0] == 1:
                        self.bc_xn = r.take(numpy.int32)
                        self.bc_xp = r.take(numpy.int32)
                        self.bc_yn = r.take(numpy.int32)
                        self.bc_yp = r.take(numpy.int32)
                    self.bc_bot = r.take(numpy.int32)
                    self.bc_top = r.take(numpy.int32)
                    self.free_slip_bot = r.take(numpy.int32)
                    self.free_slip_top = r.take(numpy.int32)
                    if self.version >= 2.11:
                        self.bc_bot_flux = r.take(numpy.float64)
                        self.bc_top_flux = r.take(numpy.float64)
                    else:
                        self.bc_bot_flux = numpy.zeros(1, dtype=numpy.float64)
                        self.bc_top_flux = numpy.zeros(1, dtype=numpy.float64)
                    if self.version >= 2.15:
                        pc = r.take(numpy.int32, nx * ny * nz).reshape(nz, ny, nx)
                        self.p_f_constant = pc.transpose(2, 1, 0).copy()
                    else:
                        self.p_f_constant = numpy.zeros((nx, ny, nz), dtype=numpy.int32)
                if self.version >= 2.0 and self.cfd_solver[0] == 0:
                    self.gamma = r.take(numpy.float64)
                    self.theta = r.take(numpy.float64)
                    self.beta = r.take(numpy.float64)
                    self.tolerance = r.take(numpy.float64)
                    self.maxiter = r.take(numpy.uint32)
                    self.ndem = r.take(numpy.uint32)
                    self.c_phi = r.take(numpy.float64)
                    self.c_v = r.take(numpy.float64)
                    self.dt_dem_fac = r.take(numpy.float64)
                    self.f_d = r.take(numpy.float64, n * d).reshape(n, d)
                    self.f_p = r.take(numpy.float64, n * d).reshape(n, d)
                    self.f_v = r.take(numpy.float64, n * d).reshape(n, d)
                    self.f_sum = r.take(numpy.float64, n * d).reshape(n, d)
                elif self.version >= 2.0 and self.cfd_solver[0] == 1:
                    self.tolerance = r.take(numpy.float64)
                    self.maxiter = r.take(numpy.uint32)
                    self.ndem = r.take(numpy.uint32)
                    self.c_phi = r.take(numpy.float64)
                    self.f_p = r.take(numpy.float64, n * d).reshape(n, d)
                    self.beta_f = r.take(numpy.float64)
                    self.k_c = r.take(numpy.float64)
            if self.version >= 1.02:
                self.color = r.take(numpy.int32, n)
            else:
                self.color = numpy.zeros(n, dtype=numpy.int32)
        finally:
            self.version = numpy.ones(1, dtype=numpy.float64) * VERSION

    def _bytes(self):
        """Serialise in the order of python/sphere.py:1355-1553."""
        out = []

        def w(a, dt):
            out.append(numpy.ascontiguousarray(numpy.asarray(a).astype(dt)).tobytes())

        n = int(self.np[0])
        w(self.version, numpy.float64)
        w(self.nd, numpy.int32)
        w(self.np, numpy.uint32)
        for nm in ['time_dt', 'time_current', 'time_total', 'time_file_dt']:
            w(getattr(self, nm), numpy.float64)
        w(self.time_step_count, numpy.uint32)
        w(self.origo, numpy.float64)
        w(self.L, numpy.float64)
        w(self.num, numpy.uint32)
        w(self.periodic, numpy.uint32)
        w(self.adaptive, numpy.uint32)
        if n > 0:
            w(numpy.column_stack([self.x, self.radius]), numpy.float64)
            w(self.xyzsum, numpy.float64)
            w(numpy.column_stack([self.vel, self.fixvel]), numpy.float64)
            for nm in ['force', 'angpos', 'angvel', 'torque', 'es_dot', 'es', 'ev_dot', 'ev', 'p']:
                w(getattr(self, nm), numpy.float64)
        w(self.g, numpy.float64)
        for nm in ['k_n', 'k_t', 'k_r', 'E', 'gamma_n', 'gamma_t', 'gamma_r', 'mu_s', 'mu_d',
                   'mu_r', 'gamma_wn', 'gamma_wt', 'mu_ws', 'mu_wd', 'rho']:
            w(getattr(self, nm), numpy.float64)
        w(self.contactmodel, numpy.uint32)
        for nm in ['kappa', 'db', 'V_b']:
            w(getattr(self, nm), numpy.float64)
        w(self.nw, numpy.uint32)
        nwl = int(self.nw[0])
        if nwl > 0:
            w(numpy.asarray(self.wmode)[:nwl], numpy.int32)
            w(numpy.column_stack([numpy.asarray(self.w_n)[:nwl], numpy.asarray(self.w_x)[:nwl]]),
              numpy.float64)
            w(numpy.column_stack([numpy.asarray(self.w_m)[:nwl], numpy.asarray(self.w_vel)[:nwl],
                                  numpy.asarray(self.w_force)[:nwl],
                                  numpy.asarray(self.w_sigma0)[:nwl]]), numpy.float64)
        w(self.w_sigma0_A, numpy.float64)
        w(self.w_sigma0_f, numpy.float64)
        w(self.w_tau_x, numpy.float64)
        w(self.lambda_bar, numpy.float64)
        w(self.nb0, numpy.uint32)
        w(self.sigma_b, numpy.float64)
        w(self.tau_b, numpy.float64)
        nb = int(self.nb0[0])
        if nb > 0:
            w(numpy.asarray(self.bonds).reshape(nb, 2), numpy.uint32)
        w(self.bonds_delta_n, numpy.float64)
        w(self.bonds_delta_t, numpy.float64)
        w(self.bonds_omega_n, numpy.float64)
        w(self.bonds_omega_t, numpy.float64)
        if self.fluid:
            w(self.cfd_solver, numpy.int32)
            w(self.mu, numpy.float64)
            cells = numpy.concatenate(
                [self.v_f, self.p_f[..., None], self.phi[..., None],
                 (self.dphi * self.time_dt * self.ndem)[..., None]], axis=3)
            w(cells.transpose(2, 1, 0, 3), numpy.float64)
            for nm in ['rho_f', 'p_mod_A', 'p_mod_f', 'p_mod_phi']:
                w(getattr(self, nm), numpy.float64)
            if self.cfd_solver[0] == 1:
                for nm in ['bc_xn', 'bc_xp', 'bc_yn', 'bc_yp']:
                    w(getattr(self, nm), numpy.int32)
            for nm in ['bc_bot', 'bc_top', 'free_slip_bot', 'free_slip_top']:
                w(getattr(self, nm), numpy.int32)
            w(self.bc_bot_flux, numpy.float64)
            w(self.bc_top_flux, numpy.float64)
            w(numpy.asarray(self.p_f_constant).transpose(2, 1, 0), numpy.int32)
            if self.cfd_solver[0] == 0:
                for nm in ['gamma', 'theta', 'beta', 'tolerance']:
                    w(getattr(self, nm), numpy.float64)
                w(self.maxiter, numpy.uint32)
                w(self.ndem, numpy.uint32)
                for nm in ['c_phi', 'c_v', 'dt_dem_fac', 'f_d', 'f_p', 'f_v', 'f_sum']:
                    w(getattr(self, nm), numpy.float64)
            elif self.cfd_solver[0] == 1:
                w(self.tolerance, numpy.float64)
                w(self.maxiter, numpy.uint32)
                w(self.ndem, numpy.uint32)
                w(self.c_phi, numpy.float64)
                w(self.f_p, numpy.float64)
                w(self.beta_f, numpy.float64)
                w(self.k_c, numpy.float64)
            else:
                raise Exception('Value of cfd_solver not understood (' +
                                str(self.cfd_solver[0]) + ')')
        w(self.color, numpy.int32)
        return b"".join(out)

    def writebin(self, folder="../input/", verbose=True):
        """Write ``<folder>/<sid>.bin`` (python/sphere.py:1355-1553)."""
        targetbin = folder + "/" + self.sid + ".bin"
        if verbose:
            print("Output file: {0}".format(targetbin))
        data = self._bytes()
        with open(targetbin, "wb") as fh:
            fh.write(data)

    def readfirst(self, verbose=True):
        self.readbin('../output/' + self.sid + '.output00000.bin', verbose)

    def readsecond(self, verbose=True):
        self.readbin('../output/' + self.sid + '.output00001.bin', verbose)

    def readstep(self, step, verbose=True):
        self.readbin("../output/{0}.output{1:0=5}.bin".format(self.sid, step), verbose)

    def readlast(self, verbose=True):
        """python/sphere.py:2314-2327"""
        lastfile = status(self.sid)
        self.readbin("../output/{0}.output{1:0=5}.bin".format(self.sid, lastfile), verbose)

    def convergence(self):
        """Iteration counts of the Darcy solver (python/sphere.py:6111-6120)."""
        self.conv = numpy.loadtxt('../output/' + self.sid + '-conv.log', dtype=numpy.int32,
                                  ndmin=2)

    # ------------------------------------------------------------------
    # conversion to and from the C ABI state image
    # ------------------------------------------------------------------
    def to_state(self):
        """A StateBuffers image of this object for sph_load_state."""
        hdr = _abi.SphState()
        hdr.nd = int(self.nd[0])
        hdr.np = int(self.np[0])
        hdr.time_dt = float(self.time_dt[0])
        hdr.time_current = float(self.time_current[0])
        hdr.time_total = float(self.time_total[0])
        hdr.time_file_dt = float(self.time_file_dt[0])
        hdr.time_step_count = int(self.time_step_count[0])
        for i in range(3):
            hdr.origo[i] = float(self.origo[i])
            hdr.L[i] = float(self.L[i])
            hdr.num[i] = int(self.num[i])
        hdr.periodic = int(numpy.asarray(self.periodic).ravel()[0])
        hdr.adaptive = int(numpy.asarray(self.adaptive).ravel()[0])
        for i in range(3):
            hdr.g[i] = float(self.g[i])
        for nm in ['k_n', 'k_t', 'k_r', 'E', 'gamma_n', 'gamma_t', 'gamma_r', 'mu_s', 'mu_d',
                   'mu_r', 'gamma_wn', 'gamma_wt', 'mu_ws', 'mu_wd', 'rho', 'kappa', 'db', 'V_b',
                   'lambda_bar', 'sigma_b', 'tau_b']:
            setattr(hdr, nm, float(numpy.asarray(getattr(self, nm)).ravel()[0]))
        hdr.contactmodel = int(self.contactmodel[0])
        nw = int(self.nw[0])
        if nw > _abi.MAXWALLS:
            raise ValueError("nw = %d > MAXWALLS" % nw)
        hdr.nw = nw
        for i in range(nw):
            hdr.wmode[i] = int(self.wmode[i])
            for c in range(3):
                hdr.w_nx[4 * i + c] = float(self.w_n[i, c])
            hdr.w_nx[4 * i + 3] = float(self.w_x[i])
            hdr.w_mvfd[4 * i + 0] = float(self.w_m[i])
            hdr.w_mvfd[4 * i + 1] = float(self.w_vel[i])
            hdr.w_mvfd[4 * i + 2] = float(self.w_force[i])
            hdr.w_mvfd[4 * i + 3] = float(self.w_sigma0[i])
        hdr.sigma0_A = float(self.w_sigma0_A[0])
        hdr.sigma0_f = float(self.w_sigma0_f[0])
        hdr.tau_x = float(self.w_tau_x[0])
        hdr.nb0 = int(self.nb0[0])
        hdr.fluid = 1 if self.fluid else 0
        if self.fluid:
            hdr.cfd_solver = int(self.cfd_solver[0])
            hdr.mu = float(self.mu[0])
            for nm in ['rho_f', 'p_mod_A', 'p_mod_f', 'p_mod_phi', 'bc_bot_flux', 'bc_top_flux',
                       'c_phi', 'beta_f', 'k_c']:
                setattr(hdr, nm, float(numpy.asarray(getattr(self, nm)).ravel()[0]))
            for nm in ['bc_xn', 'bc_xp', 'bc_yn', 'bc_yp', 'bc_bot', 'bc_top', 'free_slip_bot',
                       'free_slip_top']:
                setattr(hdr, nm, int(numpy.asarray(getattr(self, nm, [2])).ravel()[0]))
            hdr.tolerance = float(numpy.asarray(self.tolerance).ravel()[0])
            hdr.maxiter = int(numpy.asarray(self.maxiter).ravel()[0])
            hdr.ndem = int(numpy.asarray(self.ndem).ravel()[0])
        buf = _abi.StateBuffers(hdr)
        n = hdr.np
        if n:
            buf.get('x', n, 4)[:] = numpy.column_stack([self.x, self.radius])
            buf.get('xyzsum', n, 3)[:] = self.xyzsum
            buf.get('vel', n, 4)[:] = numpy.column_stack([self.vel, self.fixvel])
            for nm in ['force', 'angpos', 'angvel', 'torque']:
                buf.get(nm, n, 3)[:] = getattr(self, nm)
            for nm in ['es_dot', 'es', 'ev_dot', 'ev', 'p']:
                buf.get(nm, n, 1)[:] = getattr(self, nm)
            buf.get('color', n, 1)[:] = self.color
        nb = hdr.nb0
        if nb:
            buf.get('bonds', nb, 2)[:] = numpy.asarray(self.bonds).reshape(nb, 2)
            buf.get('bonds_delta_n', nb, 1)[:] = self.bonds_delta_n
            buf.get('bonds_delta_t', nb, 3)[:] = numpy.asarray(self.bonds_delta_t).reshape(nb, 3)
            buf.get('bonds_omega_n', nb, 1)[:] = self.bonds_omega_n
            buf.get('bonds_omega_t', nb, 3)[:] = numpy.asarray(self.bonds_omega_t).reshape(nb, 3)
        if self.fluid:
            cells = int(self.num[0] * self.num[1] * self.num[2])
            buf.get('v_f', cells, 3)[:] = self.v_f.transpose(2, 1, 0, 3).reshape(-1, 3)
            buf.get('p_f', cells, 1)[:] = self.p_f.transpose(2, 1, 0).ravel()
            buf.get('phi', cells, 1)[:] = self.phi.transpose(2, 1, 0).ravel()
            buf.get('dphi', cells, 1)[:] = (self.dphi * self.time_dt * self.ndem).transpose(2, 1, 0).ravel()
            buf.get('p_constant', cells, 1)[:] = numpy.asarray(self.p_f_constant).transpose(2, 1, 0).ravel()
            if n:
                buf.get('f_p', n, 3)[:] = self.f_p
        return buf

    def from_state(self, buf):
        """Copy a StateBuffers image (e.g. Engine.read()) into this object:
        the inverse of to_state(), header fields included."""
        h = buf.state
        n = h.np
        f64 = numpy.float64
        self.nd = numpy.ones(1, dtype=numpy.int32) * h.nd
        self.np = numpy.ones(1, dtype=numpy.uint32) * n
        self.time_dt = numpy.array([h.time_dt])
        self.time_current = numpy.array([h.time_current])
        self.time_total = numpy.array([h.time_total])
        self.time_file_dt = numpy.array([h.time_file_dt])
        self.time_step_count = numpy.array([h.time_step_count], dtype=numpy.uint32)
        self.origo = numpy.array(list(h.origo), dtype=f64)
        self.L = numpy.array(list(h.L), dtype=f64)
        self.num = numpy.array(list(h.num), dtype=numpy.uint32)
        self.periodic = numpy.array([h.periodic], dtype=numpy.int32)
        self.adaptive = numpy.array([h.adaptive], dtype=numpy.int32)
        self.g = numpy.array(list(h.g), dtype=f64)
        for nm in ['k_n', 'k_t', 'k_r', 'E', 'gamma_n', 'gamma_t', 'gamma_r', 'mu_s', 'mu_d',
                   'mu_r', 'gamma_wn', 'gamma_wt', 'mu_ws', 'mu_wd', 'rho', 'kappa', 'db', 'V_b',
                   'lambda_bar', 'sigma_b', 'tau_b']:
            setattr(self, nm, numpy.array([getattr(h, nm)], dtype=f64))
        self.contactmodel = numpy.array([h.contactmodel], dtype=numpy.uint32)
        xr = buf.get('x', n, 4)
        self.x = xr[:, :3].copy()
        self.radius = xr[:, 3].copy()
        self.xyzsum = buf.get('xyzsum', n, 3).copy()
        vr = buf.get('vel', n, 4)
        self.vel = vr[:, :3].copy()
        self.fixvel = vr[:, 3].copy()
        for nm in ['force', 'angpos', 'angvel', 'torque']:
            setattr(self, nm, buf.get(nm, n, 3).copy())
        for nm in ['es_dot', 'es', 'ev_dot', 'ev', 'p']:
            setattr(self, nm, buf.get(nm, n, 1).copy())
        self.color = buf.get('color', n, 1).astype(numpy.int32).copy()
        nw = h.nw
        self.nw = numpy.array([nw], dtype=numpy.uint32)
        self.wmode = numpy.array([h.wmode[i] for i in range(nw)], dtype=numpy.int32)
        self.w_n = numpy.array([[h.w_nx[4 * i + c] for c in range(3)] for i in range(nw)],
                               dtype=f64).reshape(nw, 3)
        self.w_x = numpy.array([h.w_nx[4 * i + 3] for i in range(nw)], dtype=f64)
        self.w_m = numpy.array([h.w_mvfd[4 * i] for i in range(nw)], dtype=f64)
        self.w_vel = numpy.array([h.w_mvfd[4 * i + 1] for i in range(nw)], dtype=f64)
        self.w_force = numpy.array([h.w_mvfd[4 * i + 2] for i in range(nw)], dtype=f64)
        self.w_sigma0 = numpy.array([h.w_mvfd[4 * i + 3] for i in range(nw)], dtype=f64)
        self.w_sigma0_A = numpy.array([h.sigma0_A], dtype=f64)
        self.w_sigma0_f = numpy.array([h.sigma0_f], dtype=f64)
        self.w_tau_x = numpy.array([h.tau_x], dtype=f64)
        nb = h.nb0
        self.nb0 = numpy.array([nb], dtype=numpy.uint32)
        self.bonds = buf.get('bonds', nb, 2).copy().reshape(nb, 2)
        self.bonds_delta_n = buf.get('bonds_delta_n', nb, 1).copy()
        self.bonds_delta_t = buf.get('bonds_delta_t', nb, 3).copy().reshape(nb, 3)
        self.bonds_omega_n = buf.get('bonds_omega_n', nb, 1).copy()
        self.bonds_omega_t = buf.get('bonds_omega_t', nb, 3).copy().reshape(nb, 3)
        self.fluid = bool(h.fluid)
        if self.fluid:
            nx, ny, nz = h.num[0], h.num[1], h.num[2]
            cells = nx * ny * nz
            self.cfd_solver = numpy.array([h.cfd_solver], dtype=numpy.int32)
            self.mu = numpy.array([h.mu], dtype=f64)
            for nm in ['rho_f', 'p_mod_A', 'p_mod_f', 'p_mod_phi', 'bc_bot_flux', 'bc_top_flux',
                       'c_phi', 'beta_f', 'k_c', 'tolerance']:
                setattr(self, nm, numpy.array([getattr(h, nm)], dtype=f64))
            for nm in ['bc_xn', 'bc_xp', 'bc_yn', 'bc_yp', 'bc_bot', 'bc_top', 'free_slip_bot',
                       'free_slip_top']:
                setattr(self, nm, numpy.array([getattr(h, nm)], dtype=numpy.int32))
            self.maxiter = numpy.array([h.maxiter], dtype=numpy.uint32)
            self.ndem = numpy.array([h.ndem], dtype=numpy.uint32)
            self.v_f = buf.get('v_f', cells, 3).reshape(nz, ny, nx, 3).transpose(2, 1, 0, 3).copy()
            self.p_f = buf.get('p_f', cells, 1).reshape(nz, ny, nx).transpose(2, 1, 0).copy()
            self.phi = buf.get('phi', cells, 1).reshape(nz, ny, nx).transpose(2, 1, 0).copy()
            self.dphi = (buf.get('dphi', cells, 1).reshape(nz, ny, nx).transpose(2, 1, 0)
                         / (h.time_dt * h.ndem))
            self.p_f_constant = buf.get('p_constant', cells, 1).reshape(nz, ny, nx) \
                .transpose(2, 1, 0).astype(numpy.int32).copy()
            self.f_p = buf.get('f_p', n, 3).copy()

    def addParticle(self, x, radius, xyzsum=numpy.zeros(3), vel=numpy.zeros(3),
                    fixvel=numpy.zeros(1), force=numpy.zeros(3), angpos=numpy.zeros(3),
                    angvel=numpy.zeros(3), torque=numpy.zeros(3), es_dot=numpy.zeros(1),
                    es=numpy.zeros(1), ev_dot=numpy.zeros(1), ev=numpy.zeros(1),
                    p=numpy.zeros(1), color=0):
        """Append one particle (python/sphere.py:821-891)."""
        self.np = self.np + 1
        self.x = numpy.append(self.x.reshape(-1, 3), [x], axis=0)
        self.radius = numpy.append(self.radius, radius)
        self.vel = numpy.append(self.vel.reshape(-1, 3), [vel], axis=0)
        self.xyzsum = numpy.append(self.xyzsum.reshape(-1, 3), [xyzsum], axis=0)
        self.fixvel = numpy.append(self.fixvel, fixvel)
        self.force = numpy.append(self.force.reshape(-1, 3), [force], axis=0)
        self.angpos = numpy.append(self.angpos.reshape(-1, 3), [angpos], axis=0)
        self.angvel = numpy.append(self.angvel.reshape(-1, 3), [angvel], axis=0)
        self.torque = numpy.append(self.torque.reshape(-1, 3), [torque], axis=0)
        self.es_dot = numpy.append(self.es_dot, es_dot)
        self.es = numpy.append(self.es, es)
        self.ev_dot = numpy.append(self.ev_dot, ev_dot)
        self.ev = numpy.append(self.ev, ev)
        self.p = numpy.append(self.p, p)
        self.color = numpy.append(self.color, color).astype(numpy.int32)
        if self.fluid:
            self.f_p = numpy.append(self.f_p.reshape(-1, 3), [numpy.zeros(3)], axis=0)

    def deleteParticle(self, i):
        """Delete the particle(s) with index i (python/sphere.py:892-935)."""
        self.np = self.np - numpy.size(i)
        for nm in ['x', 'radius', 'vel', 'xyzsum', 'fixvel', 'force', 'angpos', 'angvel',
                   'torque', 'es_dot', 'es', 'ev_dot', 'ev', 'p', 'color']:
            setattr(self, nm, numpy.delete(getattr(self, nm), i, axis=0))
        if self.fluid:
            self.f_p = numpy.delete(self.f_p, i, axis=0)

    # ------------------------------------------------------------------
    # setup helpers (python/sphere.py:2358-3170)
    # ------------------------------------------------------------------
    def generateRadii(self, psd='logn', mean=440e-6, variance=8.8e-9, histogram=False):
        """python/sphere.py:2358-2416"""
        if psd == 'logn':
            mu = math.log((mean**2) / math.sqrt(variance + mean**2))
            sigma = math.sqrt(math.log(variance / (mean**2) + 1))
            self.radius = numpy.random.lognormal(mu, sigma, self.np)
        elif psd == 'uni':
            radius_min = mean - variance
            radius_max = mean + variance
            self.radius = numpy.random.uniform(radius_min, radius_max, self.np)
        else:
            raise Exception('Particle size distribution type not understood (' + str(psd)
                            + '). Valid values are \'uni\' or \'logn\'')

    def contactModel(self, contactmodel):
        self.contactmodel[0] = contactmodel

    def wall0iz(self):
        if self.nw[0] > 0:
            return int(self.w_x[0] / (self.L[2] / self.num[2]))
        raise Exception('No dynamic top wall present!')

    def normalBoundariesXY(self):
        self.periodic[0] = 0

    def periodicBoundariesXY(self):
        self.periodic[0] = 1

    def periodicBoundariesX(self):
        self.periodic[0] = 2

    def adaptiveGrid(self):
        self.adaptive[0] = 1

    def staticGrid(self):
        self.adaptive[0] = 0

    def defineWorldBoundaries(self, L, origo=[0.0, 0.0, 0.0], dx=-1):
        """python/sphere.py:2607-2653"""
        if dx > 0.0:
            cellsize_min = dx
        else:
            if self.np[0] < 1:
                raise Exception('Error: You need to define dx in defineWorldBoundaries if '
                                'there are no particles in the simulation.')
            cellsize_min = 2.1 * numpy.amax(self.radius)
        self.origo[:] = origo[:]
        self.L[:] = L[:]
        self.num[0] = numpy.ceil((self.L[0] - self.origo[0]) / cellsize_min)
        self.num[1] = numpy.ceil((self.L[1] - self.origo[1]) / cellsize_min)
        self.num[2] = numpy.ceil((self.L[2] - self.origo[2]) / cellsize_min)
        if self.num[0] < 3 or self.num[1] < 3 or self.num[2] < 3:
            raise Exception("Error: The grid must be at least 3 cells in each direction\n"
                            "Grid: x={}, y={}, z={}\n".format(self.num[0], self.num[1], self.num[2])
                            + "Please increase the world size.")

    def initGrid(self, dx=-1):
        """python/sphere.py:2655-2685 (upper wall put at L[0], as there)"""
        if dx > 0.0:
            cellsize_min = dx
        else:
            cellsize_min = 2.1 * numpy.amax(self.radius)
        self.num[0] = numpy.ceil((self.L[0] - self.origo[0]) / cellsize_min)
        self.num[1] = numpy.ceil((self.L[1] - self.origo[1]) / cellsize_min)
        self.num[2] = numpy.ceil((self.L[2] - self.origo[2]) / cellsize_min)
        if self.num[0] < 4 or self.num[1] < 4 or self.num[2] < 4:
            raise Exception("Error: The grid must be at least 3 cells in each direction\n"
                            "Grid: x={}, y={}, z={}".format(self.num[0], self.num[1], self.num[2]))
        if self.nw > 0:
            self.w_x[0] = self.L[0]

    def initGridAndWorldsize(self, margin=2.0):
        """python/sphere.py:2687-2723"""
        r_max = numpy.amax(self.radius)
        self.origo = numpy.array([numpy.amin(self.x[:, 0] - self.radius[:]),
                                  numpy.amin(self.x[:, 1] - self.radius[:]),
                                  numpy.amin(self.x[:, 2] - self.radius[:])]) - margin * r_max
        self.L = numpy.array([numpy.amax(self.x[:, 0] + self.radius[:]),
                              numpy.amax(self.x[:, 1] + self.radius[:]),
                              numpy.amax(self.x[:, 2] + self.radius[:])]) + margin * r_max
        cellsize_min = 2.1 * r_max
        self.num[0] = numpy.ceil((self.L[0] - self.origo[0]) / cellsize_min)
        self.num[1] = numpy.ceil((self.L[1] - self.origo[1]) / cellsize_min)
        self.num[2] = numpy.ceil((self.L[2] - self.origo[2]) / cellsize_min)
        if self.num[0] < 4 or self.num[1] < 4 or self.num[2] < 4:
            raise Exception("Error: The grid must be at least 3 cells in each direction, num = "
                            + str(self.num))
        if self.nw > 0:
            self.w_x[0] = self.L[0]

    def initGridPos(self, gridnum=numpy.array([12, 12, 36])):
        """python/sphere.py:2725-2785"""
        self.num = numpy.asarray(gridnum).astype(numpy.uint32)
        r_max = numpy.amax(self.radius)
        cellsize = 2.1 * r_max
        self.L = self.num * cellsize
        if (self.num[0] * self.num[1] * self.num[2] - (2**3)) < self.np:
            raise NameError('Error! The grid is not sufficiently large.')
        gridpos = numpy.zeros(self.nd, dtype=numpy.uint32)
        if self.periodic[0] == 1:
            self.num[0] -= 1
            self.num[1] -= 1
        if (self.num[0] * self.num[1] * self.num[2] - (2 * 3 * 3)) < self.np:
            raise NameError('Error! The grid is not sufficiently large.')
        for i in range(int(self.np[0])):
            gridpos[0] = (i % (self.num[0]))
            gridpos[1] = numpy.floor(i / (self.num[0])) % (self.num[0])
            gridpos[2] = numpy.floor(i / ((self.num[0]) * (self.num[1])))
            for d in range(int(self.nd[0])):
                self.x[i, d] = gridpos[d] * cellsize + 0.5 * cellsize
            if self.periodic[0] == 1:
                if gridpos[2] % 2:
                    self.x[i, 0] += 0.5 * cellsize
                    self.x[i, 1] += 0.5 * cellsize
        if self.periodic[0] == 1:
            self.num[0] += 1
            self.num[1] += 1

    def initRandomGridPos(self, gridnum=numpy.array([12, 12, 32]), padding=2.1):
        """python/sphere.py:2787-2846"""
        coarsegrid = numpy.floor(numpy.asarray(gridnum) / 2)
        r_max = numpy.amax(self.radius)
        cellsize = padding * r_max * 2
        if ((coarsegrid[0] - 1) * (coarsegrid[1] - 1) * (coarsegrid[2] - 1)) < self.np:
            raise NameError('Error! The grid is not sufficiently large.')
        gridpos = numpy.zeros(self.nd, dtype=numpy.uint32)
        for i in range(int(self.np[0])):
            gridpos[0] = (i % (coarsegrid[0]))
            gridpos[1] = numpy.floor(i / (coarsegrid[0])) % (coarsegrid[1])
            gridpos[2] = numpy.floor(i / ((coarsegrid[0]) * (coarsegrid[1])))
            for d in range(int(self.nd[0])):
                r = self.radius[i] * 1.05
                self.x[i, d] = gridpos[d] * cellsize + ((cellsize - r) - r) * numpy.random.random_sample() + r
        x_max = numpy.max(self.x[:, 0] + self.radius)
        y_max = numpy.max(self.x[:, 1] + self.radius)
        z_max = numpy.max(self.x[:, 2] + self.radius)
        self.num[0] = numpy.ceil(x_max / cellsize)
        self.num[1] = numpy.ceil(y_max / cellsize)
        self.num[2] = numpy.ceil(z_max / cellsize)
        self.L = self.num * cellsize

    def createBondPair(self, i, j, spacing=-0.1):
        """python/sphere.py:2848-2903"""
        x_i = self.x[i]
        r_i = self.radius[i]
        r_j = self.radius[j]
        dist_ij = (r_i + r_j) * (1.0 + spacing)
        dazi = numpy.random.rand(1) * 360.0
        azi = numpy.radians(dazi)
        dang = numpy.random.rand(1) * 180.0 - 90.0
        ang = numpy.radians(dang)
        x_j = numpy.copy(x_i)
        x_j[0] = x_j[0] + dist_ij * numpy.cos(azi[0]) * numpy.cos(ang[0])
        x_j[1] = x_j[1] + dist_ij * numpy.sin(azi[0]) * numpy.cos(ang[0])
        x_j[2] = x_j[2] + dist_ij * numpy.sin(ang[0]) * numpy.cos(azi[0])
        self.x[j] = x_j
        if self.x[j, 0] < self.origo[0]:
            self.x[j, 0] += x_i[0] - x_j[0]
        if self.x[j, 1] < self.origo[1]:
            self.x[j, 1] += x_i[1] - x_j[1]
        if self.x[j, 2] < self.origo[2]:
            self.x[j, 2] += x_i[2] - x_j[2]
        if self.x[j, 0] > self.L[0]:
            self.x[j, 0] -= abs(x_j[0] - x_i[0])
        if self.x[j, 1] > self.L[1]:
            self.x[j, 1] -= abs(x_j[1] - x_i[1])
        if self.x[j, 2] > self.L[2]:
            self.x[j, 2] -= abs(x_j[2] - x_i[2])
        self.bond(i, j)
        x_ij = self.x[i] - self.x[j]
        x_ij_length = numpy.sqrt(x_ij.dot(x_ij))
        if (x_ij_length - dist_ij) > dist_ij * 0.01:
            raise Exception("Error, something went wrong in createBondPair")

    def randomBondPairs(self, ratio=0.3, spacing=-0.1):
        """python/sphere.py:2905-2928 (random_integers(0, np-1) is randint(0, np))"""
        bondparticles = numpy.unique(numpy.random.randint(0, int(self.np[0]),
                                                          size=int(self.np[0] * ratio)))
        if bondparticles.size % 2 > 0:
            bondparticles = bondparticles[:-1].copy()
        bondparticles = bondparticles.reshape(int(bondparticles.size / 2), 2).copy()
        for n in numpy.arange(bondparticles.shape[0]):
            self.createBondPair(bondparticles[n, 0], bondparticles[n, 1], spacing)

    def bond(self, i, j):
        """python/sphere.py:4075-4119"""
        self.lambda_bar[0] = 1.0
        self.bonds = numpy.vstack((numpy.asarray(self.bonds).reshape(-1, 2),
                                   [i, j])).astype(numpy.uint32)
        self.bonds_delta_n = numpy.append(self.bonds_delta_n, [0.0])
        self.bonds_delta_t = numpy.vstack((numpy.asarray(self.bonds_delta_t).reshape(-1, 3),
                                           [0.0, 0.0, 0.0]))
        self.bonds_omega_n = numpy.append(self.bonds_omega_n, [0.0])
        self.bonds_omega_t = numpy.vstack((numpy.asarray(self.bonds_omega_t).reshape(-1, 3),
                                           [0.0, 0.0, 0.0]))
        self.nb0 += 1

    def zeroKinematics(self):
        """python/sphere.py:2930-2947"""
        n = int(self.np[0])
        d = int(self.nd[0])
        self.force = numpy.zeros((n, d))
        self.torque = numpy.zeros((n, d))
        self.vel = numpy.zeros((n, d), dtype=numpy.float64)
        self.angvel = numpy.zeros((n, d), dtype=numpy.float64)
        self.angpos = numpy.zeros((n, d), dtype=numpy.float64)
        self.es = numpy.zeros(n, dtype=numpy.float64)
        self.ev = numpy.zeros(n, dtype=numpy.float64)
        self.xyzsum = numpy.zeros((n, 3), dtype=numpy.float64)

    def adjustUpperWall(self, z_adjust=1.1):
        """python/sphere.py:2949-2970"""
        self.nw = numpy.ones(1, dtype=numpy.uint32)
        self.wmode = numpy.zeros(1, dtype=numpy.int32)
        self.w_n = numpy.zeros((1, int(self.nd[0])), dtype=numpy.float64)
        self.w_n[0, 2] = -1.0
        self.w_vel = numpy.zeros(1)
        self.w_force = numpy.zeros(1)
        self.w_sigma0 = numpy.zeros(1)
        self.w_x = numpy.zeros(1)
        self.w_m = numpy.zeros(1)
        self.adjustWall(idx=0, adjust=z_adjust)

    def adjustWall(self, idx, adjust=1.1, wall_mass_factor=1.0):
        """python/sphere.py:2972-3017"""
        if idx == 0:
            dim = 2
        elif idx == 1 or idx == 2:
            dim = 0
        elif idx == 3 or idx == 4:
            dim = 1
        else:
            raise Exception("adjustWall: idx value not understood")
        xmin = numpy.min(self.x[:, dim] - self.radius)
        xmax = numpy.max(self.x[:, dim] + self.radius)
        cellsize = self.L[0] / self.num[0]
        self.num[dim] = numpy.ceil(((xmax - xmin) * adjust + xmin) / cellsize)
        self.L[dim] = (xmax - xmin) * adjust + xmin
        if idx == 0 or idx == 1 or idx == 3:
            self.w_x[idx] = xmax
        else:
            self.w_x[idx] = xmin
        self.w_m[idx] = self.totalMass()

    def consolidate(self, normal_stress=10e3):
        """python/sphere.py:3019-3048"""
        self.nw[0] = 1
        if normal_stress <= 0.0:
            raise Exception('consolidate() error: The normal stress should be a positive '
                            'value, but is ' + str(normal_stress) + ' Pa')
        self.zeroKinematics()
        self.adjustUpperWall()
        self.wmode = numpy.array([1], dtype=numpy.int32)
        self.w_sigma0 = numpy.ones(1) * normal_stress
        self.w_m[0] = self.totalMass()

    def uniaxialStrainRate(self, wvel=-0.001):
        """python/sphere.py:3050-3066"""
        self.zeroKinematics()
        self.adjustUpperWall()
        self.wmode = numpy.array([2], dtype=numpy.int32)
        self.w_vel = numpy.array([wvel])

    def triaxial(self, wvel=-0.001, normal_stress=10.0e3):
        """python/sphere.py:3068-3095"""
        self.zeroKinematics()
        self.nw[0] = 5
        self.wmode = numpy.array([2, 1, 1, 1, 1], dtype=numpy.int32)
        self.w_vel = numpy.array([1, 0, 0, 0, 0]) * wvel
        self.w_sigma0 = numpy.array([0, 1, 1, 1, 1]) * normal_stress
        self.w_n = numpy.array(([0, 0, -1], [-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0]),
                               dtype=numpy.float64)
        self.w_x = numpy.zeros(5)
        self.w_m = numpy.zeros(5)
        self.w_force = numpy.zeros(5)
        for i in range(5):
            self.adjustWall(idx=i)

    def shear(self, shear_strain_rate=1.0, shear_stress=False):
        """python/sphere.py:3097-3165"""
        self.nw[0] = 1
        z_min = numpy.min(self.x[:, 2] - self.radius)
        z_max = numpy.max(self.x[:, 2] + self.radius)
        cellsize = self.L[0] / self.num[0]
        self.num[2] += 1
        self.L[2] = self.num[2] * cellsize
        self.zeroKinematics()
        self.wmode = numpy.array([1], dtype=numpy.int32)
        d_max_below = numpy.max(self.radius[numpy.nonzero(self.x[:, 2] < (z_max - z_min) * 0.3)]) * 2.0
        I = numpy.nonzero(self.x[:, 2] < (z_min + d_max_below))
        self.fixvel[I] = 1
        self.angvel[I, 0] = 0.0
        self.angvel[I, 1] = 0.0
        self.angvel[I, 2] = 0.0
        self.vel[I, 0] = 0.0
        self.vel[I, 1] = 0.0
        self.color[I] = -1
        d_max_top = numpy.max(self.radius[numpy.nonzero(self.x[:, 2] > (z_max - z_min) * 0.7)]) * 2.0
        I = numpy.nonzero(self.x[:, 2] > (z_max - d_max_top))
        self.fixvel[I] = 1
        self.angvel[I, 0] = 0.0
        self.angvel[I, 1] = 0.0
        self.angvel[I, 2] = 0.0
        if shear_stress is False:
            self.vel[I, 0] = (z_max - z_min) * shear_strain_rate
        else:
            self.vel[I, 0] = 0.0
            self.wmode[0] = 3
            self.w_tau_x[0] = float(shear_stress)
        self.vel[I, 1] = 0.0
        self.color[I] = -1
        self.gamma_wt[0] = 0.0
        self.mu_ws[0] = 0.0
        self.mu_wd[0] = 0.0

    def largestFluidTimeStep(self, safety=0.5, v_max=-1.0):
        """python/sphere.py:3167-3252 (Darcy: the CFL bound)"""
        if self.fluid:
            v_norm = numpy.sqrt(numpy.sum(self.v_f**2, axis=3)).ravel()
            v_max_obs = numpy.amax(v_norm) if v_norm.size else 0.0
            if v_max_obs == 0:
                v_max_obs = 1.0e-7
            if v_max < 0.0:
                v_max = v_max_obs
            dx_min = numpy.min(self.L / self.num)
            dt_min_cfl = dx_min / v_max
            if self.cfd_solver[0] == 0:
                dt_min_von_neumann = 0.5 * dx_min**2 / (self.mu[0] + 1.0e-16)
                return numpy.min([dt_min_von_neumann, dt_min_cfl]) * safety
            elif self.cfd_solver[0] == 1:
                self.cellSize()
                return dt_min_cfl

    def cellSize(self):
        self.dx = self.L / self.num

    def initTemporal(self, total, current=0.0, file_dt=0.05, step_count=0, dt=-1, epsilon=0.01):
        """python/sphere.py:3307-3376"""
        if dt > 0.0:
            self.time_dt[0] = dt
            if self.np[0] > 0:
                print("Warning: Manually specifying the time step length when simulating "
                      "particles may produce instabilities.")
        elif self.np[0] > 0:
            r_min = numpy.min(self.radius)
            m_min = self.rho * 4.0 / 3.0 * numpy.pi * r_min**3
            if self.E > 0.001:
                k_max = numpy.max(numpy.pi / 2.0 * self.E * self.radius)
            else:
                k_max = numpy.max([self.k_n[:], self.k_t[:]])
            self.time_dt[0] = (epsilon / (numpy.sqrt(k_max / m_min))).ravel()[0]
        elif self.fluid is False:
            raise Exception('Error: Could not automatically set a time step.')
        if self.fluid:
            fluid_time_dt = self.largestFluidTimeStep()
            self.time_dt[0] = numpy.min([fluid_time_dt, self.time_dt[0]])
        self.time_current[0] = current
        self.time_total[0] = total
        self.time_file_dt[0] = file_dt
        self.time_step_count[0] = 0

    def dry(self):
        self.fluid = False

    def wet(self):
        self.fluid = True
        self.initFluid()

    def initFluid(self, mu=8.9e-4, rho=1.0e3, p=0.0, hydrostatic=False, cfd_solver=0):
        """python/sphere.py:3395-3508"""
        self.fluid = True
        self.mu = numpy.ones(1, dtype=numpy.float64) * mu
        self.rho_f = numpy.ones(1, dtype=numpy.float64) * rho
        self._alloc_fluid(p)
        self.phi = numpy.ones(self.p_f.shape, dtype=numpy.float64)
        if hydrostatic:
            dz = self.L[2] / self.num[2]
            if self.nw == 1:
                wall0_iz = int(self.w_x[0] / (self.L[2] / self.num[2]))
                self.p_f[:, :, wall0_iz:] = p
                for iz in numpy.arange(wall0_iz - 1):
                    z = dz * iz + 0.5 * dz
                    depth = self.w_x[0] - z
                    self.p_f[:, :, iz] = p + (depth - dz) * rho * -self.g[2]
            else:
                for iz in numpy.arange(self.num[2] - 1):
                    z = dz * iz + 0.5 * dz
                    depth = self.L[2] - z
                    self.p_f[:, :, iz] = p + (depth - dz) * rho * -self.g[2]
        self.cfd_solver = numpy.ones(1) * cfd_solver
        n = int(self.np[0])
        d = int(self.nd[0])
        if self.cfd_solver[0] == 0:
            self.gamma = numpy.array(0.0)
            self.theta = numpy.array(1.0)
            self.beta = numpy.array(0.0)
            self.tolerance = numpy.array(1.0e-3)
            self.maxiter = numpy.array(1e4)
            self.ndem = numpy.array(1)
            self.c_phi = numpy.ones(1, dtype=numpy.float64)
            self.c_v = numpy.ones(1, dtype=numpy.float64)
            self.dt_dem_fac = numpy.ones(1, dtype=numpy.float64)
            self.f_d = numpy.zeros((n, d), dtype=numpy.float64)
            self.f_p = numpy.zeros((n, d), dtype=numpy.float64)
            self.f_v = numpy.zeros((n, d), dtype=numpy.float64)
            self.f_sum = numpy.zeros((n, d), dtype=numpy.float64)
        elif self.cfd_solver[0] == 1:
            self.tolerance = numpy.array(1.0e-3)
            self.maxiter = numpy.array(1e4)
            self.ndem = numpy.array(1)
            self.c_phi = numpy.ones(1, dtype=numpy.float64)
            self.f_d = numpy.zeros((n, d), dtype=numpy.float64)
            self.beta_f = numpy.ones(1, dtype=numpy.float64) * 4.5e-10
            self.f_p = numpy.zeros((n, d), dtype=numpy.float64)
            self.k_c = numpy.ones(1, dtype=numpy.float64) * 4.6e-10
            self.bc_xn = numpy.ones(1, dtype=numpy.int32) * 2
            self.bc_xp = numpy.ones(1, dtype=numpy.int32) * 2
            self.bc_yn = numpy.ones(1, dtype=numpy.int32) * 2
            self.bc_yp = numpy.ones(1, dtype=numpy.int32) * 2
        else:
            raise Exception('Value of cfd_solver not understood (' + str(self.cfd_solver[0]) + ')')

    def currentTime(self, value=-1):
        if value != -1:
            self.time_current[0] = value
        else:
            return self.time_current[0]

    # fluid boundary conditions, python/sphere.py:3527-3686
    def setFluidBottomNoFlow(self):
        self.bc_bot[0] = 1

    def setFluidBottomNoFlowNoSlip(self):
        self.bc_bot[0] = 2

    def setFluidBottomFixedPressure(self):
        self.bc_bot[0] = 0

    def setFluidBottomFixedFlux(self, specific_flux):
        self.bc_bot[0] = 4
        self.bc_bot_flux[0] = specific_flux

    def setFluidTopNoFlow(self):
        self.bc_top[0] = 1

    def setFluidTopNoFlowNoSlip(self):
        self.bc_top[0] = 2

    def setFluidTopFixedPressure(self):
        self.bc_top[0] = 0

    def setFluidTopFixedFlux(self, specific_flux):
        self.bc_top[0] = 4
        self.bc_top_flux[0] = specific_flux

    def setFluidXFixedPressure(self):
        self.bc_xn[0] = 0
        self.bc_xp[0] = 0

    def setFluidXNoFlow(self):
        self.bc_xn[0] = 1
        self.bc_xp[0] = 1

    def setFluidXPeriodic(self):
        self.bc_xn[0] = 2
        self.bc_xp[0] = 2

    def setFluidYFixedPressure(self):
        self.bc_yn[0] = 0
        self.bc_yp[0] = 0

    def setFluidYNoFlow(self):
        self.bc_yn[0] = 1
        self.bc_yp[0] = 1

    def setFluidYPeriodic(self):
        self.bc_yn[0] = 2
        self.bc_yp[0] = 2

    def setPermeabilityGrainSize(self, verbose=True):
        self.setPermeabilityPrefactor(k_c=numpy.mean(self.radius * 2.0)**2.0 / 180.0,
                                      verbose=verbose)

    def setPermeabilityPrefactor(self, k_c, verbose=True):
        if self.cfd_solver[0] == 1:
            self.k_c[0] = k_c
        else:
            raise Exception('setPermeabilityPrefactor() only relevant for the Darcy solver '
                            '(cfd_solver = 1)')

    def findPermeabilities(self):
        """python/sphere.py:3730-3744 (the Python helper's phi^3/(1-phi^2))"""
        if self.cfd_solver[0] == 1:
            phi = numpy.clip(self.phi, 0.1, 0.9)
            self.k = self.k_c * phi**3 / (1.0 - phi**2)
        else:
            raise Exception('findPermeabilities() only relevant for the Darcy solver')

    def setFluidPressureModulation(self, A, f, phi=0.0, plot=False):
        self.p_mod_A[0] = A
        self.p_mod_f[0] = f
        self.p_mod_phi[0] = phi

    def disableFluidPressureModulation(self):
        self.setFluidPressureModulation(A=0.0, f=0.0)

    def setTopWallNormalStressModulation(self, A, f, plot=False):
        self.w_sigma0_A[0] = A
        self.w_sigma0_f[0] = f

    def disableTopWallNormalStressModulation(self):
        self.setTopWallNormalStressModulation(A=0.0, f=0.0)

    def setTolerance(self, tolerance):
        self.tolerance = numpy.asarray(tolerance)

    def setMaxIterations(self, maxiter):
        self.maxiter = numpy.asarray(maxiter)

    def setDEMstepsPerCFDstep(self, ndem):
        self.ndem = numpy.asarray(ndem)

    # material parameters, python/sphere.py:3762-4054
    def defaultParams(self, mu_s=0.5, mu_d=0.5, mu_r=0.0, rho=2600, k_n=1.16e9, k_t=1.16e9,
                      k_r=0, gamma_n=0.0, gamma_t=0.0, gamma_r=0.0, gamma_wn=1.0e4,
                      gamma_wt=1.0e4, capillaryCohesion=0):
        self.rho = numpy.ones(1, dtype=numpy.float64) * rho
        self.k_n = numpy.ones(1, dtype=numpy.float64) * k_n
        self.k_t = numpy.ones(1, dtype=numpy.float64) * k_t
        self.k_r = numpy.ones(1, dtype=numpy.float64) * k_r
        self.gamma_n = numpy.ones(1, dtype=numpy.float64) * gamma_n
        self.gamma_t = numpy.ones(1, dtype=numpy.float64) * gamma_t
        self.gamma_r = numpy.ones(1, dtype=numpy.float64) * gamma_r
        self.mu_s = numpy.ones(1, dtype=numpy.float64) * mu_s
        self.mu_d = numpy.ones(1, dtype=numpy.float64) * mu_d
        self.mu_r = numpy.ones(1, dtype=numpy.float64) * mu_r
        self.gamma_wn[0] = gamma_wn
        self.gamma_wt[0] = gamma_wt
        self.mu_ws = self.mu_s
        self.mu_wd = self.mu_d
        theta = 0.0
        if capillaryCohesion == 1:
            self.kappa[0] = 2.0 * math.pi * gamma_t * numpy.cos(theta)
            self.V_b[0] = 1e-12
        else:
            self.kappa[0] = 0.0
            self.V_b[0] = 0.0
        self.db[0] = (1.0 + theta / 2.0) * self.V_b[0]**(1.0 / 3.0)

    def setStiffnessNormal(self, k_n):
        self.k_n[0] = k_n

    def setStiffnessTangential(self, k_t):
        self.k_t[0] = k_t

    def setYoungsModulus(self, E):
        """python/sphere.py:3895-3908"""
        self.E[0] = E

    def setPoissonsRatio(self, nu):
        """Poisson's ratio of the Hertz-Mindlin model (contactmodel 3).  The
        v2.15 format has no slot for it; it is carried in the k_r field, which
        the reference reads nowhere (python/sphere.py:169-171)."""
        self.k_r[0] = nu

    def setDampingNormal(self, gamma, over_damping=False):
        self.gamma_n[0] = gamma
        critical_gamma = 2.0 * numpy.sqrt(self.smallestMass() * self.k_n[0])
        damping_ratio = gamma / critical_gamma
        if damping_ratio > 1.0 and not over_damping:
            raise Exception('Warning: The system is over-dampened (ratio = ' + str(damping_ratio)
                            + ') in the normal component.')

    def setDampingTangential(self, gamma, over_damping=False):
        self.gamma_t[0] = gamma
        damping_ratio = gamma / (2.0 * numpy.sqrt(self.smallestMass() * self.k_t[0]))
        if damping_ratio > 1.0 and not over_damping:
            raise Exception('Warning: The system is over-dampened (ratio = ' + str(damping_ratio)
                            + ') in the tangential component.')

    def setStaticFriction(self, mu_s):
        self.mu_s[0] = mu_s

    def setDynamicFriction(self, mu_d):
        self.mu_d[0] = mu_d

    def setFluidCompressibility(self, beta_f):
        if self.cfd_solver[0] == 1:
            self.beta_f[0] = beta_f
        else:
            raise Exception('setFluidCompressibility() only relevant for the Darcy solver')

    def setFluidViscosity(self, mu):
        self.mu[0] = mu

    def setFluidDensity(self, rho_f):
        self.rho_f[0] = rho_f

    def scaleSize(self, factor):
        self.L *= factor
        self.x *= factor
        self.radius *= factor
        self.xyzsum *= factor
        self.vel *= factor
        self.force *= factor
        self.torque *= factor
        self.w_x *= factor
        self.w_m *= factor
        self.w_vel *= factor
        self.w_force *= factor

    # ------------------------------------------------------------------
    # energies and masses, python/sphere.py:4142-4380
    # ------------------------------------------------------------------
    def surfaceArea(self, idx):
        return 4.0 * numpy.pi * self.radius[idx]**2

    def volume(self, idx):
        return V_sphere(self.radius[idx])

    def mass(self, idx):
        return self.rho[0] * self.volume(idx)

    def totalMass(self):
        return float(numpy.sum(self.rho[0] * V_sphere(self.radius)))

    def smallestMass(self):
        # the reference returns a volume here (python/sphere.py:4186-4195)
        return V_sphere(numpy.min(self.radius))

    def largestMass(self):
        return V_sphere(numpy.max(self.radius))

    def momentOfInertia(self, idx):
        return 2.0 / 5.0 * self.mass(idx) * self.radius[idx]**2

    def kineticEnergy(self, idx):
        return 0.5 * self.mass(idx) * numpy.sqrt(numpy.dot(self.vel[idx, :], self.vel[idx, :]))**2

    def totalKineticEnergy(self):
        return sum(self.kineticEnergy(i) for i in range(int(self.np[0])))

    def rotationalEnergy(self, idx):
        return 0.5 * self.momentOfInertia(idx) * numpy.sqrt(
            numpy.dot(self.angvel[idx, :], self.angvel[idx, :]))**2

    def totalRotationalEnergy(self):
        return sum(self.rotationalEnergy(i) for i in range(int(self.np[0])))

    def viscousEnergy(self, idx):
        return self.ev[idx]

    def totalViscousEnergy(self):
        return float(numpy.sum(self.ev))

    def frictionalEnergy(self, idx):
        return self.es[idx]

    def totalFrictionalEnergy(self):
        return float(numpy.sum(self.es))

    def energy(self, method):
        """python/sphere.py:4311-4380"""
        if method == 'pot':
            m = numpy.ones(self.np) * 4.0 / 3.0 * math.pi * self.radius**3 * self.rho
            return numpy.sum(m * math.sqrt(numpy.dot(self.g, self.g)) * self.x[:, 2])
        elif method == 'kin':
            m = numpy.ones(self.np) * 4.0 / 3.0 * math.pi * self.radius**3 * self.rho
            return float(numpy.sum(0.5 * m * numpy.sum(self.vel**2, axis=1)))
        elif method == 'rot':
            m = numpy.ones(self.np) * 4.0 / 3.0 * math.pi * self.radius**3 * self.rho
            return float(numpy.sum(0.5 * 2.0 / 5.0 * m * self.radius**2
                                   * numpy.sum(self.angvel**2, axis=1)))
        elif method == 'shear':
            return numpy.sum(self.es)
        elif method == 'shearrate':
            return numpy.sum(self.es_dot)
        elif method == 'visc_n':
            return numpy.sum(self.ev)
        elif method == 'visc_n_rate':
            return numpy.sum(self.ev_dot)
        elif method == 'bondpot':
            if self.nb0 > 0:
                R_bar = self.lambda_bar * numpy.minimum(self.radius[self.bonds[:, 0]],
                                                        self.radius[self.bonds[:, 1]])
                A = numpy.pi * R_bar**2
                I = 0.25 * numpy.pi * R_bar**4
                J = I * 2.0
                bondpot_fn = numpy.sum(0.5 * A * self.k_n * numpy.abs(self.bonds_delta_n)**2)
                bondpot_ft = numpy.sum(0.5 * A * self.k_t * numpy.linalg.norm(self.bonds_delta_t)**2)
                bondpot_tn = numpy.sum(0.5 * J * self.k_t * numpy.abs(self.bonds_omega_n)**2)
                bondpot_tt = numpy.sum(0.5 * I * self.k_n * numpy.linalg.norm(self.bonds_omega_t)**2)
                return bondpot_fn + bondpot_ft + bondpot_tn + bondpot_tt
            return 0.0
        raise Exception('Unknownw energy() method "' + method + '"')

    # ------------------------------------------------------------------
    # execution, python/sphere.py:4482-4540
    # ------------------------------------------------------------------
    def run(self, verbose=True, hideinputfile=False, dry=False, valgrind=False,
            cudamemcheck=False, device=-1):
        """Write ../input/<sid>.bin and run it through libsphere_hip.so, with
        outputs in ../output (the reference's ``cd ..; ./sphere``)."""
        self.writebin(verbose=False)
        if valgrind or cudamemcheck:
            print("Warning: valgrind/cuda-memcheck wrappers are not available; running plainly")
        lib = _abi.load_library()
        opts = _abi.SphRunOpts()
        inp = ("../input/" + self.sid + ".bin").encode()
        outd = b"../output"
        opts.input_path = inp
        opts.output_dir = outd
        opts.fluid = 1 if self.fluid else 0
        opts.verbose = 1 if verbose else 0
        opts.dry = 1 if dry else 0
        opts.device = int(device)
        st = lib.sph_run(ctypes.byref(opts))
        if st != 0:
            print("Warning: the sphere run returned with status " + str(st) + ": "
                  + lib.sph_last_error(None).decode())
        return st

    def status(self):
        return status(self.sid)

    def cleanup(self):
        cleanup(self)


def run(binary, verbose=True, hideinputfile=False, fluid=False, device=-1):
    """Execute a run of an input file (python/sphere.py:7617-7636)."""
    lib = _abi.load_library()
    opts = _abi.SphRunOpts()
    opts.input_path = binary.encode()
    opts.output_dir = b"output"
    opts.fluid = 1 if fluid else 0
    opts.verbose = 1 if verbose else 0
    opts.dry = 0
    opts.device = int(device)
    return lib.sph_run(ctypes.byref(opts))
