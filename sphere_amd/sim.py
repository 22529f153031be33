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
import re
import shutil

import numpy

from . import _abi

VERSION = 2.15  # src/version.h:5


def V_sphere(r):
    """Volume of a sphere (python/sphere.py:7762-7769)."""
    return 4.0 / 3.0 * math.pi * r**3.0


def _restart_side_car(path):
    """The restart side car written beside an output file
    (<sid>.outputNNNNN.bin -> <sid>.contacts.NNNNN.bin), or None."""
    head, name = os.path.split(path)
    m = re.match(r"^(.*)\.output(\d{5})\.bin$", name)
    if not m:
        return None
    side = os.path.join(head, "%s.contacts.%s.bin" % (m.group(1), m.group(2)))
    return side if os.path.exists(side) else None


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


# ----------------------------------------------------------------------
# What a new ``sim`` holds (python/sphere.py:57-455).  These initial values
# are part of the interface: writebin puts them in the file for every field
# a script leaves alone.
# ----------------------------------------------------------------------
_F64, _U32, _I32 = numpy.float64, numpy.uint32, numpy.int32

# Per-particle arrays: name, columns ("nd": one per spatial dimension, None:
# a flat array), dtype, initial value.
_PARTICLE_ARRAYS = (
    ("x", "nd", _F64, 0.0), ("radius", None, _F64, 1.0), ("xyzsum", 3, _F64, 0.0),
    ("vel", "nd", _F64, 0.0), ("fixvel", None, _F64, 0.0), ("force", "nd", _F64, 0.0),
    ("angpos", "nd", _F64, 0.0), ("angvel", "nd", _F64, 0.0), ("torque", "nd", _F64, 0.0),
    ("es_dot", None, _F64, 0.0), ("es", None, _F64, 0.0), ("ev_dot", None, _F64, 0.0),
    ("ev", None, _F64, 0.0), ("p", None, _F64, 0.0), ("color", None, _I32, 0),
)

# One-value fields: name -> (dtype, initial value).
_SCALARS = {
    "time_dt": (_F64, 0.0), "time_current": (_F64, 0.0), "time_total": (_F64, 0.0),
    "time_file_dt": (_F64, 0.0), "time_step_count": (_U32, 0),
    "periodic": (_U32, 0), "adaptive": (_U32, 0),
    # contact law: model 2 (linear-viscous) with stiff springs and friction 0.5
    "k_n": (_F64, 1.16e9), "k_t": (_F64, 1.16e9), "k_r": (_F64, 0.0), "E": (_F64, 0.0),
    "gamma_n": (_F64, 0.0), "gamma_t": (_F64, 0.0), "gamma_r": (_F64, 0.0),
    "mu_s": (_F64, 0.5), "mu_d": (_F64, 0.5), "mu_r": (_F64, 0.0),
    "gamma_wn": (_F64, 1.0e3), "gamma_wt": (_F64, 0.0),
    "mu_ws": (_F64, 0.5), "mu_wd": (_F64, 0.5), "rho": (_F64, 2600.0),
    "contactmodel": (_U32, 2), "kappa": (_F64, 0.0), "db": (_F64, 0.0), "V_b": (_F64, 0.0),
    # top-wall stress modulation and the wmode-3 shear stress
    "w_sigma0_A": (_F64, 0.0), "w_sigma0_f": (_F64, 0.0), "w_tau_x": (_F64, 0.0),
    # bonds: none, unbreakable
    "lambda_bar": (_F64, 1.0), "nb0": (_U32, 0),
    "sigma_b": (_F64, numpy.inf), "tau_b": (_F64, numpy.inf),
}

# Normal of dynamic wall i: the top wall faces down, then the -x, +x, -y and
# +y side walls face outwards along their axis.
_WALL_NORMALS = numpy.array([[0.0, 0.0, -1.0], [-1.0, 0.0, 0.0], [1.0, 0.0, 0.0],
                             [0.0, -1.0, 0.0], [0.0, 1.0, 0.0]])

# Per-wall arrays other than the normals: name -> (dtype, initial value).
_WALL_ARRAYS = {"wmode": (_I32, 0), "w_x": (_F64, 1.0), "w_m": (_F64, 0.0),
                "w_vel": (_F64, 0.0), "w_force": (_F64, 0.0), "w_sigma0": (_F64, 0.0)}

# Fluid fields both solvers share, besides the cell fields.
_FLUID_SCALARS = {
    "rho_f": (_F64, 1.0e3), "p_mod_A": (_F64, 0.0), "p_mod_f": (_F64, 0.0),
    "p_mod_phi": (_F64, 0.0), "bc_bot": (_I32, 0), "bc_top": (_I32, 0),
    "free_slip_bot": (_I32, 1), "free_slip_top": (_I32, 1),
    "bc_bot_flux": (_F64, 0.0), "bc_top_flux": (_F64, 0.0),
}

# Solver-specific fields (cfd_solver 0: Navier-Stokes, 1: Darcy); the
# per-particle fluid forces listed in _SOLVER_FORCES start at zero.
_SOLVER_SCALARS = {
    0: {"gamma": (_F64, 0.0), "theta": (_F64, 1.0), "beta": (_F64, 0.0),
        "tolerance": (_F64, 1.0e-3), "maxiter": (_U32, 10000), "ndem": (_U32, 1),
        "c_phi": (_F64, 1.0), "c_v": (_F64, 1.0), "dt_dem_fac": (_F64, 1.0)},
    1: {"tolerance": (_F64, 1.0e-3), "maxiter": (_U32, 10000), "ndem": (_U32, 1),
        "c_phi": (_F64, 1.0), "beta_f": (_F64, 4.5e-10), "k_c": (_F64, 4.6e-10)},
}
_SOLVER_FORCES = {0: ("f_d", "f_p", "f_v", "f_sum"), 1: ("f_d", "f_p")}
# Darcy's lateral boundaries are periodic unless a script says otherwise
_DARCY_LATERAL = ("bc_xn", "bc_xp", "bc_yn", "bc_yp")


def _one(dtype, value):
    return numpy.full(1, value, dtype=dtype)


class sim:
    """Class containing all ``sphere`` data (python/sphere.py:33-455)."""

    def __init__(self, sid='unnamed', np=0, nd=3, nw=0, fluid=False, cfd_solver=0):
        self.sid = sid
        self.version = _one(_F64, VERSION)
        self.nd = _one(_I32, int(nd))
        self.np = _one(_U32, int(np))
        d = int(nd)
        for name, (dt, v) in _SCALARS.items():
            setattr(self, name, _one(dt, v))
        self.g = numpy.zeros(3)
        self.origo = numpy.zeros(d)
        self.L = numpy.zeros(d)
        self.num = numpy.zeros(d, dtype=_U32)
        self._alloc_particles(int(np))
        self._alloc_walls(int(nw))
        self._alloc_bonds(0)
        self.fluid = fluid
        if fluid:
            # the constructor always starts from the Navier-Stokes fields
            # (its cfd_solver argument is not read there); initFluid selects
            # the solver
            self.mu = _one(_F64, 0.0)
            self._fluid_fields(p=0.0, phi=0.0)
            self._solver_fields(0, lateral=False)

    # -- allocation --------------------------------------------------------
    def _alloc_particles(self, n):
        d = int(self.nd[0])
        for name, cols, dt, v in _PARTICLE_ARRAYS:
            shape = (n,) if cols is None else (n, d if cols == "nd" else cols)
            setattr(self, name, numpy.full(shape, v, dtype=dt))

    def _alloc_walls(self, nw):
        self.nw = _one(_U32, nw)
        for name, (dt, v) in _WALL_ARRAYS.items():
            setattr(self, name, numpy.full(nw, v, dtype=dt))
        self.w_n = numpy.zeros((nw, int(self.nd[0])))
        k = min(nw, len(_WALL_NORMALS))
        self.w_n[:k] = _WALL_NORMALS[:k]

    def _alloc_bonds(self, nb):
        d = int(self.nd[0])
        self.bonds = numpy.zeros((nb, 2), dtype=_U32)
        self.bonds_delta_n = numpy.zeros(nb)
        self.bonds_delta_t = numpy.zeros((nb, d))
        self.bonds_omega_n = numpy.zeros(nb)
        self.bonds_omega_t = numpy.zeros((nb, d))

    def _fluid_fields(self, p, phi):
        """Cell fields on the current grid (pressure p, porosity phi, zero
        velocities and porosity change, no fixed-pressure cells) and the
        fluid scalars both solvers share."""
        cells = tuple(int(v) for v in self.num[:3])
        self.v_f = numpy.zeros(cells + (int(self.nd[0]),))
        self.p_f = numpy.full(cells, p, dtype=_F64)
        self.phi = numpy.full(cells, phi, dtype=_F64)
        self.dphi = numpy.zeros(cells)
        self.p_f_constant = numpy.zeros(cells, dtype=_I32)
        for name, (dt, v) in _FLUID_SCALARS.items():
            setattr(self, name, _one(dt, v))

    def _solver_fields(self, solver, lateral):
        if solver not in _SOLVER_SCALARS:
            raise Exception('Value of cfd_solver not understood (' + str(solver) + ')')
        self.cfd_solver = _one(_I32, solver)
        for name, (dt, v) in _SOLVER_SCALARS[solver].items():
            setattr(self, name, _one(dt, v))
        n, d = int(self.np[0]), int(self.nd[0])
        for name in _SOLVER_FORCES[solver]:
            setattr(self, name, numpy.zeros((n, d)))
        if solver == 1 and lateral:
            for name in _DARCY_LATERAL:
                setattr(self, name, _one(_I32, 2))


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
        self._restart = _restart_side_car(targetbin)
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
            v_f, p_f, phi, dphi, p_c = self._grid_cells()
            cells = numpy.concatenate(
                [v_f, p_f[..., None], phi[..., None],
                 (dphi * self.time_dt * self.ndem)[..., None]], axis=3)
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
            w(p_c.transpose(2, 1, 0), numpy.int32)
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

    def _grid_cells(self):
        """The cell fields over the current grid: like the reference's
        writer, the first num[0] x num[1] x num[2] cells (a helper that
        changes num after initFluid leaves the arrays larger)."""
        nx, ny, nz = (int(v) for v in self.num[:3])
        out = []
        for name in ("v_f", "p_f", "phi", "dphi", "p_f_constant"):
            a = numpy.asarray(getattr(self, name))
            if a.shape[0] < nx or a.shape[1] < ny or a.shape[2] < nz:
                raise ValueError("%s has %s cells, fewer than the grid's %s"
                                 % (name, a.shape[:3], (nx, ny, nz)))
            out.append(a[:nx, :ny, :nz])
        return out

    def writebin(self, folder="../input/", verbose=True):
        """Write ``<folder>/<sid>.bin`` (python/sphere.py:1355-1553)."""
        targetbin = folder + "/" + self.sid + ".bin"
        if verbose:
            print("Output file: {0}".format(targetbin))
        data = self._bytes()
        with open(targetbin, "wb") as fh:
            fh.write(data)

    def id(self, sid=''):
        """The simulation id, or set it when sid is given."""
        if sid == '':
            return self.sid
        self.sid = sid

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
            v_f, p_f, phi, dphi, p_c = self._grid_cells()
            buf.get('v_f', cells, 3)[:] = v_f.transpose(2, 1, 0, 3).reshape(-1, 3)
            buf.get('p_f', cells, 1)[:] = p_f.transpose(2, 1, 0).ravel()
            buf.get('phi', cells, 1)[:] = phi.transpose(2, 1, 0).ravel()
            buf.get('dphi', cells, 1)[:] = (dphi * self.time_dt * self.ndem).transpose(2, 1, 0).ravel()
            buf.get('p_constant', cells, 1)[:] = p_c.transpose(2, 1, 0).ravel()
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

    # ------------------------------------------------------------------
    # particles
    # ------------------------------------------------------------------
    def _particle_fields(self):
        """(name, columns) of every per-particle array, the fluid forces
        included; columns is None for a flat array."""
        d = int(self.nd[0])
        out = [(name, None if cols is None else (d if cols == "nd" else cols))
               for name, cols, _, _ in _PARTICLE_ARRAYS]
        if self.fluid:
            solver = int(numpy.asarray(self.cfd_solver).ravel()[0])
            out += [(f, d) for f in _SOLVER_FORCES.get(solver, ()) if hasattr(self, f)]
        return out

    def addParticle(self, x, radius, xyzsum=None, vel=None, fixvel=0.0, force=None,
                    angpos=None, angvel=None, torque=None, es_dot=0.0, es=0.0, ev_dot=0.0,
                    ev=0.0, p=0.0, color=0):
        """Append one particle; omitted vectors are zero (python/sphere.py
        :821-891)."""
        given = dict(x=x, radius=radius, xyzsum=xyzsum, vel=vel, fixvel=fixvel, force=force,
                     angpos=angpos, angvel=angvel, torque=torque, es_dot=es_dot, es=es,
                     ev_dot=ev_dot, ev=ev, p=p, color=color)
        for name, cols in self._particle_fields():
            old = numpy.asarray(getattr(self, name))
            value = given.get(name)
            shape = (1,) if cols is None else (1, cols)
            row = numpy.zeros(shape, old.dtype) if value is None else \
                numpy.asarray(value, old.dtype).reshape(-1)[:int(numpy.prod(shape))].reshape(shape)
            setattr(self, name, numpy.concatenate([old.reshape((-1,) + shape[1:]), row]))
        self.np = self.np + 1

    def deleteParticle(self, i):
        """Remove the particle(s) with index i (python/sphere.py:892-935)."""
        for name, _ in self._particle_fields():
            setattr(self, name, numpy.delete(getattr(self, name), i, axis=0))
        self.np = self.np - numpy.size(i)

    def generateRadii(self, psd='logn', mean=440e-6, variance=8.8e-9, histogram=False):
        """Draw the radii from numpy's global generator (python/sphere.py
        :2358-2416): 'logn' is the lognormal distribution with the given mean
        and variance, 'uni' the uniform distribution on mean -+ variance."""
        n = int(self.np[0])
        if psd == 'logn':
            # the lognormal whose mean and variance are the given ones:
            # s^2 = ln(1 + var/mean^2), m = ln(mean) - s^2/2
            s2 = math.log(1.0 + variance / mean**2)
            self.radius = numpy.random.lognormal(math.log(mean) - 0.5 * s2, math.sqrt(s2), n)
        elif psd == 'uni':
            self.radius = numpy.random.uniform(mean - variance, mean + variance, n)
        else:
            raise Exception("generateRadii: psd must be 'logn' or 'uni', not %r" % (psd,))

    def contactModel(self, contactmodel):
        self.contactmodel[0] = contactmodel

    def zeroKinematics(self):
        """Zero forces, torques, linear and angular velocities and positions,
        accumulated energies and displacement sums (python/sphere.py:2930-2947);
        es_dot, ev_dot, p and fixvel are kept."""
        n, d = int(self.np[0]), int(self.nd[0])
        for name in ("force", "torque", "vel", "angvel", "angpos"):
            setattr(self, name, numpy.zeros((n, d)))
        self.es = numpy.zeros(n)
        self.ev = numpy.zeros(n)
        self.xyzsum = numpy.zeros((n, 3))

    # ------------------------------------------------------------------
    # grid and boundaries (python/sphere.py:2418-2785)
    # ------------------------------------------------------------------
    def wall0iz(self):
        """Grid layer of the top wall."""
        if self.nw[0] < 1:
            raise Exception('No dynamic top wall present!')
        return int(self.w_x[0] / (self.L[2] / self.num[2]))

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

    def _size_grid(self, cell, smallest):
        """num = ceil((L - origo) / cell) per axis; at least `smallest` cells
        along every axis."""
        self.num[:3] = numpy.ceil((self.L[:3] - self.origo[:3]) / cell)
        if numpy.any(self.num[:3] < smallest):
            raise Exception("Error: the grid has %s cells; it needs at least %d along each axis"
                            % (list(self.num[:3]), smallest))

    def _default_cell(self, dx):
        """dx when positive, else 2.1 times the largest radius."""
        return dx if dx > 0.0 else 2.1 * numpy.amax(self.radius)

    def defineWorldBoundaries(self, L, origo=[0.0, 0.0, 0.0], dx=-1):
        """World from origo to L with cells of at least dx (python/sphere.py
        :2607-2653)."""
        if dx <= 0.0 and self.np[0] < 1:
            raise Exception('defineWorldBoundaries: without particles the cell size dx is '
                            'required')
        cell = self._default_cell(dx)
        self.origo[:] = origo[:]
        self.L[:] = L[:]
        self._size_grid(cell, 3)

    def initGrid(self, dx=-1):
        """Cell counts for the current world; a top wall is put at L[0] (sic)
        (python/sphere.py:2655-2685)."""
        self._size_grid(self._default_cell(dx), 4)
        if self.nw > 0:
            self.w_x[0] = self.L[0]

    def initGridAndWorldsize(self, margin=2.0):
        """World spanning the grains plus `margin` largest radii on every
        side (python/sphere.py:2687-2723)."""
        r_max = numpy.amax(self.radius)
        pad = margin * r_max
        self.origo = numpy.amin(self.x - self.radius[:, None], axis=0) - pad
        self.L = numpy.amax(self.x + self.radius[:, None], axis=0) + pad
        self._size_grid(2.1 * r_max, 4)
        if self.nw > 0:
            self.w_x[0] = self.L[0]

    def initGridPos(self, gridnum=numpy.array([12, 12, 36])):
        """One grain per cell of a regular grid of 2.1 largest radii, filled
        x first (python/sphere.py:2725-2785).  With periodic x and y the grid
        gives up its last column and row to the wrap, and odd layers are
        shifted by half a cell in x and y."""
        self.num = numpy.asarray(gridnum).astype(_U32)
        cell = 2.1 * numpy.amax(self.radius)
        self.L = self.num * cell
        n = int(self.np[0])
        used = self.num.astype(numpy.int64)
        if self.periodic[0] == 1:
            used[:2] -= 1
        if numpy.prod(self.num.astype(numpy.int64)) - 8 < n or numpy.prod(used) - 18 < n:
            raise NameError('Error! The grid is not sufficiently large.')
        i = numpy.arange(n)
        nx = int(used[0])
        # (the row index wraps at the x count, as in the reference)
        g = numpy.stack([i % nx, (i // nx) % nx, i // (nx * int(used[1]))], axis=1)
        self.x[:n, :3] = g * cell + 0.5 * cell
        if self.periodic[0] == 1:
            odd = (g[:, 2] % 2) == 1
            self.x[:n][odd, :2] += 0.5 * cell

    def initRandomGridPos(self, gridnum=numpy.array([12, 12, 32]), padding=2.1):
        """One grain per cell of a grid half as fine as gridnum, at a uniform
        random position that keeps 1.05 radii from the cell faces; the world
        then ends at the cell holding the outermost grain surface
        (python/sphere.py:2787-2846)."""
        coarse = numpy.floor(numpy.asarray(gridnum) / 2).astype(numpy.int64)
        cell = padding * numpy.amax(self.radius) * 2
        n = int(self.np[0])
        if numpy.prod(coarse - 1) < n:
            raise NameError('Error! The grid is not sufficiently large.')
        i = numpy.arange(n)
        g = numpy.stack([i % coarse[0], (i // coarse[0]) % coarse[1],
                         i // (coarse[0] * coarse[1])], axis=1)
        r = 1.05 * self.radius[:n, None]
        u = numpy.random.random_sample((n, int(self.nd[0])))  # drawn grain by grain
        self.x[:n] = g * cell + ((cell - r) - r) * u + r
        top = numpy.max(self.x[:n] + self.radius[:n, None], axis=0)
        self.num = numpy.ceil(top / cell).astype(_U32)
        self.L = self.num * cell

    # ------------------------------------------------------------------
    # bonds (python/sphere.py:2848-2928, 4075-4119)
    # ------------------------------------------------------------------
    def bond(self, i, j):
        """Add an unstrained bond between particles i and j."""
        self.lambda_bar[0] = 1.0
        nb = int(self.nb0[0])
        old = (self.bonds, self.bonds_delta_n, self.bonds_delta_t, self.bonds_omega_n,
               self.bonds_omega_t)
        self._alloc_bonds(nb + 1)
        for new, prev in zip((self.bonds, self.bonds_delta_n, self.bonds_delta_t,
                              self.bonds_omega_n, self.bonds_omega_t), old):
            new[:nb] = numpy.asarray(prev).reshape(new[:nb].shape)
        self.bonds[nb] = (i, j)
        self.nb0 = self.nb0 + 1

    def createBondPair(self, i, j, spacing=-0.1):
        """Move particle j next to particle i, its centre (r_i + r_j)(1 +
        spacing) away along a random direction, and bond the two
        (python/sphere.py:2848-2903).  The direction uses the reference's
        formula (its z component carries the cosine of the azimuth).  A
        coordinate that would leave the world is pulled back towards i."""
        dist = (self.radius[i] + self.radius[j]) * (1.0 + spacing)
        azi = numpy.radians(numpy.random.rand(1) * 360.0)[0]
        ele = numpy.radians(numpy.random.rand(1) * 180.0 - 90.0)[0]
        x_i = self.x[i].copy()
        step = numpy.array([dist * math.cos(azi) * math.cos(ele),
                            dist * math.sin(azi) * math.cos(ele),
                            dist * math.sin(ele) * math.cos(azi)])
        x_j = x_i + step
        new = x_j.copy()
        below = x_j < self.origo[:3]
        new[below] = x_i[below]
        above = new > self.L[:3]
        new[above] -= numpy.abs(x_j - x_i)[above]
        self.x[j] = new
        self.bond(i, j)
        if numpy.linalg.norm(self.x[i] - self.x[j]) - dist > 0.01 * dist:
            raise Exception("createBondPair: the pair ended up too far apart")

    def randomBondPairs(self, ratio=0.3, spacing=-0.1):
        """Bond disjoint random pairs drawn from about ratio * np distinct
        particles (python/sphere.py:2905-2928)."""
        n = int(self.np[0])
        picks = numpy.unique(numpy.random.randint(0, n, size=int(n * ratio)))
        picks = picks[:picks.size - picks.size % 2].reshape(-1, 2)
        for a, b in picks:
            self.createBondPair(a, b, spacing)

    # ------------------------------------------------------------------
    # walls and loading (python/sphere.py:2949-3165)
    # ------------------------------------------------------------------
    _WALL_AXIS = {0: 2, 1: 0, 2: 0, 3: 1, 4: 1}

    def adjustWall(self, idx, adjust=1.1, wall_mass_factor=1.0):
        """Fit wall idx (0 top, 1/2 -x/+x, 3/4 -y/+y) to the grains: the
        world along its axis ends at lo + adjust (hi - lo), lo and hi being
        the outermost grain surfaces; walls 0, 1 and 3 sit at hi, 2 and 4 at
        lo; the wall mass is the total grain mass (python/sphere.py
        :2972-3017; wall_mass_factor is accepted and, as there, not used)."""
        if idx not in self._WALL_AXIS:
            raise Exception("adjustWall: idx value not understood")
        axis = self._WALL_AXIS[idx]
        lo = numpy.min(self.x[:, axis] - self.radius)
        hi = numpy.max(self.x[:, axis] + self.radius)
        cell = self.L[0] / self.num[0]
        end = (hi - lo) * adjust + lo
        self.num[axis] = numpy.ceil(end / cell)
        self.L[axis] = end
        self.w_x[idx] = hi if idx in (0, 1, 3) else lo
        self.w_m[idx] = self.totalMass()

    def adjustUpperWall(self, z_adjust=1.1):
        """One wall, the top one (fixed), fitted with adjustWall."""
        self._alloc_walls(1)
        self.w_x[0] = 0.0
        self.adjustWall(idx=0, adjust=z_adjust)

    def consolidate(self, normal_stress=10e3):
        """Top wall under a constant normal stress (wmode 1)
        (python/sphere.py:3019-3048)."""
        self.nw[0] = 1
        if normal_stress <= 0.0:
            raise Exception('consolidate() error: The normal stress should be a positive '
                            'value, but is ' + str(normal_stress) + ' Pa')
        self._loaded_top_wall(1, sigma0=normal_stress)

    def uniaxialStrainRate(self, wvel=-0.001):
        """Top wall moving at wvel (wmode 2) (python/sphere.py:3050-3066)."""
        self._loaded_top_wall(2, vel=wvel)

    def _loaded_top_wall(self, mode, sigma0=0.0, vel=0.0):
        self.zeroKinematics()
        self.adjustUpperWall()
        self.wmode = numpy.array([mode], dtype=_I32)
        self.w_sigma0 = numpy.array([sigma0])
        self.w_vel = numpy.array([vel])

    def triaxial(self, wvel=-0.001, normal_stress=10.0e3):
        """Five walls: the top one moving at wvel (wmode 2), the four side
        walls under normal_stress (wmode 1), each fitted with adjustWall
        (python/sphere.py:3068-3095)."""
        self.zeroKinematics()
        self._alloc_walls(5)
        top = numpy.arange(5) == 0
        self.wmode[:] = numpy.where(top, 2, 1)
        # products, so that the side walls' velocities carry the sign of
        # zero that 0 * wvel has (it is written to the file)
        self.w_vel[:] = top * wvel
        self.w_sigma0[:] = ~top * normal_stress
        self.w_x[:] = 0.0
        for idx in range(5):
            self.adjustWall(idx=idx)

    def _fix_layer(self, members, vx):
        """Drive grains at a prescribed velocity (fixvel = 1): x velocity vx,
        y velocity 0, no rotation; they are coloured -1."""
        self.fixvel[members] = 1.0
        self.angvel[members, :] = 0.0
        self.vel[members, 0] = vx
        self.vel[members, 1] = 0.0
        self.color[members] = -1

    def shear(self, shear_strain_rate=1.0, shear_stress=False):
        """Simple shear of a consolidated packing (python/sphere.py
        :3097-3165).  The grid grows by one cell layer; kinematics are zeroed;
        the bottom grains (centre below z_min + one largest diameter of the
        grains whose centre lies below 0.3 h) are held still, and the top
        grains (centre above z_max - one largest diameter of those above
        0.7 h) move at h * shear_strain_rate in x, h = z_max - z_min being the
        packing height.  With shear_stress, the top grains are instead driven
        by that shear stress (wmode 3).  Wall friction is switched off."""
        self.nw[0] = 1
        z_lo = numpy.min(self.x[:, 2] - self.radius)
        z_hi = numpy.max(self.x[:, 2] + self.radius)
        h = z_hi - z_lo
        cell = self.L[0] / self.num[0]
        self.num[2] += 1
        self.L[2] = self.num[2] * cell
        self.zeroKinematics()
        self.wmode = numpy.array([1], dtype=_I32)
        z = self.x[:, 2]
        # (the 0.3 h and 0.7 h bounds are heights above z = 0)
        d_bottom = 2.0 * numpy.max(self.radius[z < 0.3 * h])
        self._fix_layer(numpy.nonzero(z < z_lo + d_bottom), 0.0)
        d_top = 2.0 * numpy.max(self.radius[z > 0.7 * h])
        top = numpy.nonzero(z > z_hi - d_top)
        if shear_stress is False:
            self._fix_layer(top, h * shear_strain_rate)
        else:
            self._fix_layer(top, 0.0)
            self.wmode[0] = 3
            self.w_tau_x[0] = float(shear_stress)
        self.gamma_wt[0] = 0.0
        self.mu_ws[0] = 0.0
        self.mu_wd[0] = 0.0

    # ------------------------------------------------------------------
    # time (python/sphere.py:3167-3376)
    # ------------------------------------------------------------------
    def cellSize(self):
        self.dx = self.L / self.num

    def largestFluidTimeStep(self, safety=0.5, v_max=-1.0):
        """Largest stable fluid time step for the current fluid velocities:
        the CFL bound dx_min / v_max (Darcy), or safety times the smaller of
        it and the von Neumann bound 0.5 dx_min^2 / mu (Navier-Stokes).  An
        all-zero velocity field counts as 1e-7 m/s."""
        if not self.fluid:
            return None
        speeds = numpy.sqrt(numpy.sum(numpy.asarray(self.v_f) ** 2, axis=-1))
        observed = float(numpy.max(speeds)) if speeds.size else 0.0
        if observed == 0.0:
            observed = 1.0e-7
        v = observed if v_max < 0.0 else v_max
        dx_min = numpy.min(self.L / self.num)
        cfl = dx_min / v
        if int(self.cfd_solver[0]) == 0:
            return min(0.5 * dx_min ** 2 / (self.mu[0] + 1.0e-16), cfl) * safety
        self.cellSize()
        return cfl

    def initTemporal(self, total, current=0.0, file_dt=0.05, step_count=0, dt=-1, epsilon=0.01):
        """Time step and run length (python/sphere.py:3307-3376).  Without an
        explicit dt the step is epsilon / sqrt(k_max / m_min): m_min the mass
        of the smallest grain, k_max = max(pi/2 E r) when E > 0.001, else the
        larger of k_n and k_t; with fluid, the fluid's CFL step if smaller.
        The output counter restarts at 0 (step_count is not used)."""
        n = int(self.np[0])
        if dt > 0.0:
            self.time_dt[0] = dt
            if n > 0:
                print("Warning: Manually specifying the time step length when simulating "
                      "particles may produce instabilities.")
        elif n > 0:
            m_min = self.rho[0] * V_sphere(numpy.min(self.radius))
            if self.E[0] > 0.001:
                k_max = numpy.max(numpy.pi / 2.0 * self.E[0] * self.radius)
            else:
                k_max = max(self.k_n[0], self.k_t[0])
            self.time_dt[0] = epsilon / math.sqrt(k_max / m_min)
        elif not self.fluid:
            raise Exception('Error: Could not automatically set a time step.')
        if self.fluid:
            self.time_dt[0] = min(self.largestFluidTimeStep(), self.time_dt[0])
        self.time_current[0] = current
        self.time_total[0] = total
        self.time_file_dt[0] = file_dt
        self.time_step_count[0] = 0

    def currentTime(self, value=-1):
        if value == -1:
            return self.time_current[0]
        self.time_current[0] = value

    # ------------------------------------------------------------------
    # fluid (python/sphere.py:3378-3760)
    # ------------------------------------------------------------------
    def dry(self):
        self.fluid = False

    def wet(self):
        self.fluid = True
        self.initFluid()

    def initFluid(self, mu=8.9e-4, rho=1.0e3, p=0.0, hydrostatic=False, cfd_solver=0):
        """Fluid on the current grid (python/sphere.py:3395-3508): pressure p
        (hydrostatic below the top wall or the world top when asked),
        porosity 1, the given viscosity and density, and the defaults of the
        chosen solver (0 Navier-Stokes, 1 Darcy with periodic lateral
        boundaries)."""
        self.fluid = True
        self.mu = _one(_F64, mu)
        self._fluid_fields(p=p, phi=1.0)
        self.rho_f = _one(_F64, rho)
        if hydrostatic:
            self._hydrostatic(p, rho)
        self._solver_fields(int(cfd_solver), lateral=True)

    def _hydrostatic(self, p, rho):
        """Pressure p + (depth - dz) rho |g_z| in layer iz at depth below the
        top wall (one wall) or the world top, for the layers 0 .. top - 2;
        p from the wall's layer up."""
        dz = self.L[2] / self.num[2]
        if self.nw == 1:
            top_iz = int(self.w_x[0] / dz)
            surface = self.w_x[0]
            self.p_f[:, :, top_iz:] = p
        else:
            top_iz = int(self.num[2])
            surface = self.L[2]
        for iz in range(max(top_iz - 1, 0)):
            depth = surface - (dz * iz + 0.5 * dz)
            self.p_f[:, :, iz] = p + (depth - dz) * rho * -self.g[2]

    # fluid boundary conditions, python/sphere.py:3527-3686: bc 0 fixed
    # pressure, 1 no flow, 2 no flow and no slip (bottom/top) or periodic
    # (sides), 4 fixed flux
    def _bc(self, names, code, flux=None, flux_name=None):
        for nm in names:
            getattr(self, nm)[0] = code
        if flux_name is not None:
            getattr(self, flux_name)[0] = flux

    def setFluidBottomNoFlow(self):
        self._bc(("bc_bot",), 1)

    def setFluidBottomNoFlowNoSlip(self):
        self._bc(("bc_bot",), 2)

    def setFluidBottomFixedPressure(self):
        self._bc(("bc_bot",), 0)

    def setFluidBottomFixedFlux(self, specific_flux):
        self._bc(("bc_bot",), 4, specific_flux, "bc_bot_flux")

    def setFluidTopNoFlow(self):
        self._bc(("bc_top",), 1)

    def setFluidTopNoFlowNoSlip(self):
        self._bc(("bc_top",), 2)

    def setFluidTopFixedPressure(self):
        self._bc(("bc_top",), 0)

    def setFluidTopFixedFlux(self, specific_flux):
        self._bc(("bc_top",), 4, specific_flux, "bc_top_flux")

    def setFluidXFixedPressure(self):
        self._bc(("bc_xn", "bc_xp"), 0)

    def setFluidXNoFlow(self):
        self._bc(("bc_xn", "bc_xp"), 1)

    def setFluidXPeriodic(self):
        self._bc(("bc_xn", "bc_xp"), 2)

    def setFluidYFixedPressure(self):
        self._bc(("bc_yn", "bc_yp"), 0)

    def setFluidYNoFlow(self):
        self._bc(("bc_yn", "bc_yp"), 1)

    def setFluidYPeriodic(self):
        self._bc(("bc_yn", "bc_yp"), 2)

    def _darcy_only(self, what):
        if int(numpy.asarray(self.cfd_solver).ravel()[0]) != 1:
            raise Exception(what + '() only relevant for the Darcy solver (cfd_solver = 1)')

    def setPermeabilityGrainSize(self, verbose=True):
        """k_c = d_mean^2 / 180 (Kozeny-Carman with the mean diameter)."""
        self.setPermeabilityPrefactor(k_c=numpy.mean(2.0 * self.radius) ** 2 / 180.0,
                                      verbose=verbose)

    def setPermeabilityPrefactor(self, k_c, verbose=True):
        self._darcy_only('setPermeabilityPrefactor')
        self.k_c[0] = k_c

    def findPermeabilities(self):
        """Permeability of every cell from its porosity with the Python
        helper's form k_c phi^3 / (1 - phi^2), phi clipped to [0.1, 0.9]
        (python/sphere.py:3730-3744; the solver itself uses phi^3/(1-phi)^2)."""
        self._darcy_only('findPermeabilities')
        phi = numpy.clip(self.phi, 0.1, 0.9)
        self.k = self.k_c * phi ** 3 / (1.0 - phi ** 2)

    def setFluidPressureModulation(self, A, f, phi=0.0, plot=False):
        """Top pressure p_top + A sin(2 pi f t + phi) (Darcy)."""
        self.p_mod_A[0] = A
        self.p_mod_f[0] = f
        self.p_mod_phi[0] = phi

    def disableFluidPressureModulation(self):
        self.setFluidPressureModulation(A=0.0, f=0.0)

    def setTopWallNormalStressModulation(self, A, f, plot=False):
        """Top wall stress sigma0 + A sin(2 pi f t)."""
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

    def setFluidCompressibility(self, beta_f):
        self._darcy_only('setFluidCompressibility')
        self.beta_f[0] = beta_f

    def setFluidViscosity(self, mu):
        self.mu[0] = mu

    def setFluidDensity(self, rho_f):
        self.rho_f[0] = rho_f

    # ------------------------------------------------------------------
    # material parameters (python/sphere.py:3762-4054)
    # ------------------------------------------------------------------
    def defaultParams(self, mu_s=0.5, mu_d=0.5, mu_r=0.0, rho=2600, k_n=1.16e9, k_t=1.16e9,
                      k_r=0, gamma_n=0.0, gamma_t=0.0, gamma_r=0.0, gamma_wn=1.0e4,
                      gamma_wt=1.0e4, capillaryCohesion=0):
        """Set the contact parameters (python/sphere.py:3762-3839): the
        arguments as given, wall friction equal to the grain friction, and
        with capillaryCohesion = 1 a liquid bridge of 1e-12 m^3 with
        kappa = 2 pi gamma_t (contact angle 0) and rupture distance
        V_b^(1/3)."""
        values = dict(rho=rho, k_n=k_n, k_t=k_t, k_r=k_r, gamma_n=gamma_n, gamma_t=gamma_t,
                      gamma_r=gamma_r, mu_s=mu_s, mu_d=mu_d, mu_r=mu_r)
        for name, v in values.items():
            setattr(self, name, _one(_F64, v))
        self.gamma_wn[0] = gamma_wn
        self.gamma_wt[0] = gamma_wt
        self.mu_ws = self.mu_s
        self.mu_wd = self.mu_d
        cohesive = capillaryCohesion == 1
        self.kappa[0] = 2.0 * math.pi * gamma_t if cohesive else 0.0
        self.V_b[0] = 1e-12 if cohesive else 0.0
        self.db[0] = self.V_b[0] ** (1.0 / 3.0)

    def setStiffnessNormal(self, k_n):
        self.k_n[0] = k_n

    def setStiffnessTangential(self, k_t):
        self.k_t[0] = k_t

    def setYoungsModulus(self, E):
        """Young's modulus; above 0.001 it sets the contact stiffness
        (python/sphere.py:3895-3908)."""
        self.E[0] = E

    def setPoissonsRatio(self, nu):
        """Poisson's ratio of the Hertz-Mindlin model (contactmodel 3).  The
        v2.15 format has no slot for it; it is carried in the k_r field, which
        the reference reads nowhere (python/sphere.py:169-171)."""
        self.k_r[0] = nu

    def _damping_check(self, gamma, k, component, over_damping):
        ratio = gamma / (2.0 * numpy.sqrt(self.smallestMass() * k))
        if ratio > 1.0 and not over_damping:
            raise Exception('Warning: The system is over-dampened (ratio = ' + str(ratio)
                            + ') in the ' + component + ' component.')

    def setDampingNormal(self, gamma, over_damping=False):
        """Normal damping; above the critical 2 sqrt(m k_n) raises unless
        over_damping (the reference's smallestMass is a volume)."""
        self.gamma_n[0] = gamma
        self._damping_check(gamma, self.k_n[0], 'normal', over_damping)

    def setDampingTangential(self, gamma, over_damping=False):
        self.gamma_t[0] = gamma
        self._damping_check(gamma, self.k_t[0], 'tangential', over_damping)

    def setStaticFriction(self, mu_s):
        self.mu_s[0] = mu_s

    def setDynamicFriction(self, mu_d):
        self.mu_d[0] = mu_d

    def scaleSize(self, factor):
        """Scale lengths, velocities, forces and wall quantities by factor."""
        for name in ("L", "x", "radius", "xyzsum", "vel", "force", "torque", "w_x", "w_m",
                     "w_vel", "w_force"):
            setattr(self, name, getattr(self, name) * factor)

    # ------------------------------------------------------------------
    # masses, energies and bulk measures (python/sphere.py:4142-4437,
    # 4777-4855)
    # ------------------------------------------------------------------
    def _masses(self):
        return self.rho[0] * V_sphere(self.radius)

    def surfaceArea(self, idx):
        return 4.0 * numpy.pi * self.radius[idx] ** 2

    def volume(self, idx):
        return V_sphere(self.radius[idx])

    def mass(self, idx):
        return self.rho[0] * self.volume(idx)

    def totalMass(self):
        return float(numpy.sum(self._masses()))

    def smallestMass(self):
        """The volume of the smallest grain (the reference returns a
        volume here, python/sphere.py:4186-4195)."""
        return V_sphere(numpy.min(self.radius))

    def largestMass(self):
        return V_sphere(numpy.max(self.radius))

    def momentOfInertia(self, idx):
        return 0.4 * self.mass(idx) * self.radius[idx] ** 2

    def kineticEnergy(self, idx):
        return 0.5 * self.mass(idx) * float(numpy.dot(self.vel[idx], self.vel[idx]))

    def totalKineticEnergy(self):
        return self.energy('kin')

    def rotationalEnergy(self, idx):
        return 0.5 * self.momentOfInertia(idx) * float(numpy.dot(self.angvel[idx],
                                                                 self.angvel[idx]))

    def totalRotationalEnergy(self):
        return self.energy('rot')

    def viscousEnergy(self, idx):
        return self.ev[idx]

    def totalViscousEnergy(self):
        return float(numpy.sum(self.ev))

    def frictionalEnergy(self, idx):
        return self.es[idx]

    def totalFrictionalEnergy(self):
        return float(numpy.sum(self.es))

    def _bond_energy(self):
        """Elastic energy of the bonds (python/sphere.py:4360-4376): with
        R = lambda min(r_i, r_j), A = pi R^2, I = pi R^4 / 4, J = 2 I, the sum
        over bonds of A k_n dn^2/2 + J k_t wn^2/2, plus A k_t |dt|^2/2 and
        I k_n |wt|^2/2 where, as in the reference, |dt| and |wt| are the
        norms of the whole tangential arrays (the per-bond norms for one
        bond)."""
        nb = int(self.nb0[0])
        if nb == 0:
            return 0.0
        b = numpy.asarray(self.bonds).reshape(nb, 2)
        R = self.lambda_bar[0] * numpy.minimum(self.radius[b[:, 0]], self.radius[b[:, 1]])
        A = numpy.pi * R ** 2
        I = 0.25 * numpy.pi * R ** 4
        J = 2.0 * I
        dt2 = numpy.sum(numpy.asarray(self.bonds_delta_t) ** 2)
        wt2 = numpy.sum(numpy.asarray(self.bonds_omega_t) ** 2)
        return float(numpy.sum(0.5 * A * self.k_n[0] * numpy.asarray(self.bonds_delta_n) ** 2)
                     + numpy.sum(0.5 * A * self.k_t[0] * dt2)
                     + numpy.sum(0.5 * J * self.k_t[0] * numpy.asarray(self.bonds_omega_n) ** 2)
                     + numpy.sum(0.5 * I * self.k_n[0] * wt2))

    def energy(self, method):
        """Energy sums of the assemblage (python/sphere.py:4311-4380): 'pot'
        m |g| z, 'kin' m |v|^2 / 2, 'rot' (2/5 m r^2) |w|^2 / 2, 'shear' es,
        'shearrate' es_dot, 'visc_n' ev, 'visc_n_rate' ev_dot, 'bondpot' the
        bonds' elastic energy."""
        sums = {'shear': 'es', 'shearrate': 'es_dot', 'visc_n': 'ev', 'visc_n_rate': 'ev_dot'}
        if method in sums:
            return float(numpy.sum(getattr(self, sums[method])))
        m = self._masses()
        if method == 'pot':
            return float(numpy.sum(m * numpy.linalg.norm(self.g) * self.x[:, 2]))
        if method == 'kin':
            return float(numpy.sum(0.5 * m * numpy.sum(self.vel ** 2, axis=1)))
        if method == 'rot':
            return float(numpy.sum(0.2 * m * self.radius ** 2 * numpy.sum(self.angvel ** 2, axis=1)))
        if method == 'bondpot':
            return self._bond_energy()
        raise Exception('Unknown energy() method "' + method + '"')

    def _grain_box_volume(self):
        lo = numpy.min(self.x - self.radius[:, None], axis=0)
        hi = numpy.max(self.x + self.radius[:, None], axis=0)
        return float(numpy.prod(hi - lo))

    def voidRatio(self):
        """(V_world - V_solid) / V_solid with V_world = L0 L1 L2 (no walls)
        or L0 L1 w_x[0] (one wall) (python/sphere.py:4382-4400)."""
        V_s = float(numpy.sum(V_sphere(self.radius)))
        return (self._bulk_volume() - V_s) / V_s

    def _bulk_volume(self):
        if self.nw[0] == 0:
            return float(self.L[0] * self.L[1] * self.L[2])
        if self.nw[0] == 1:
            V = float(self.L[0] * self.L[1] * self.w_x[0])
            if V <= 0.0:
                raise Exception("Could not determine total volume")
            return V
        return 0.0

    def bulkPorosity(self, trim=True):
        """Porosity 1 - V_solid / V of the assemblage (python/sphere.py
        :4402-4436); V is the axis-parallel box around the grains (trim), the
        world (no walls) or the world below the top wall (one wall)."""
        V = self._grain_box_volume() if trim else self._bulk_volume()
        V_s = float(numpy.sum(V_sphere(self.radius)))
        return (V - V_s) / V

    def _driven(self):
        return numpy.nonzero(self.fixvel > 0.0)

    def shearDisplacement(self):
        """x displacement of the driven (fixvel > 0) grains: the largest
        accumulated x movement among them (python/sphere.py:4777-4792)."""
        return float(numpy.max(self.xyzsum[self._driven()][:, 0]))

    def shearVelocity(self):
        """Largest x velocity of the driven grains (python/sphere.py:4794-4808)."""
        return float(numpy.max(self.vel[self._driven()][:, 0]))

    def shearVel(self):
        return self.shearVelocity()

    def shearStrain(self):
        """Shear displacement over the top wall's height (python/sphere.py:4816-4835)."""
        return self.shearDisplacement() / self.w_x[0]

    def shearStrainRate(self):
        """Shear velocity over the top wall's height (python/sphere.py:4837-4853)."""
        return self.shearVelocity() / self.w_x[0]

    # ------------------------------------------------------------------
    # VTK output for ParaView (python/sphere.py:1555-2191).  ASCII XML files;
    # the array names are the reference's, which ParaView state files use.
    # ------------------------------------------------------------------
    def _vtk_point_arrays(self):
        """(name, VTK type, values) of the grain arrays, in the reference's
        order; fluid pressure forces with the Darcy solver."""
        arrays = [("Diameter", "Float32", 2.0 * self.radius),
                  ("Displacement [m]", "Float32", self.xyzsum),
                  ("Velocity [m/s]", "Float32", self.vel)]
        if self.fluid and int(numpy.asarray(self.cfd_solver).ravel()[0]) == 1:
            arrays.append(("Fluid pressure force [N]", "Float32", self.f_p))
        arrays += [("FixedVel", "Float32", self.fixvel), ("Force [N]", "Float32", self.force),
                   ("Angular position[rad]", "Float32", self.angpos),
                   ("Angular velocity [rad/s]", "Float32", self.angvel),
                   ("Torque [Nm]", "Float32", self.torque),
                   ("Shear Energy Rate [J/s]", "Float32", self.es_dot),
                   ("Shear Energy [J]", "Float32", self.es),
                   ("Viscous Energy Rate [J/s]", "Float32", self.ev_dot),
                   ("Viscous Energy [J]", "Float32", self.ev),
                   ("Pressure [Pa]", "Float32", self.p),
                   ("Type color", "Int32", self.color)]
        return arrays

    @staticmethod
    def _vtk_data(fh, name, vtype, values, indent="        "):
        a = numpy.asarray(values)
        comps = a.shape[1] if a.ndim == 2 else 1
        fh.write('%s<DataArray type="%s" Name="%s"%s format="ascii">\n' % (
            indent, vtype, name, ' NumberOfComponents="%d"' % comps if comps > 1 else ""))
        fmt = "%d" if vtype.startswith("Int") or vtype.startswith("UInt") else "%f"
        fh.write(indent + "  " + " ".join(fmt % v for v in a.ravel()) + "\n")
        fh.write("%s</DataArray>\n" % indent)

    def writeVTK(self, folder='../output/', verbose=True):
        """Grains as the points of an unstructured grid, <folder>/<sid>.vtu
        (python/sphere.py:1646-1998)."""
        path = folder + '/' + self.sid + '.vtu'
        if verbose:
            print('Output file: ' + path)
        n = int(self.np[0])
        with open(path, 'w') as fh:
            fh.write('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="0.1" '
                     'byte_order="LittleEndian">\n  <UnstructuredGrid>\n'
                     '    <Piece NumberOfPoints="%d" NumberOfCells="0">\n      <Points>\n' % n)
            self._vtk_data(fh, "Position [m]", "Float32", self.x)
            fh.write('      </Points>\n      <PointData Scalars="Diameter [m]" Vectors="vector">\n')
            for name, vtype, values in self._vtk_point_arrays():
                self._vtk_data(fh, name, vtype, values)
            fh.write('      </PointData>\n      <Cells>\n')
            for name, vtype in (("connectivity", "Int32"), ("offsets", "Int32"), ("types", "UInt8")):
                self._vtk_data(fh, name, vtype, numpy.zeros(0, dtype=numpy.int32))
            fh.write('      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>')

    def writeFluidVTK(self, folder='../output/', cell_centered=True, verbose=True):
        """The fluid grid as image data, <folder>/fluid-<sid>.vti: pressure,
        velocity, porosity, porosity change and, with the Darcy solver, the
        permeability (the Python helper's form) and the fixed-pressure flags;
        values on the cells (cell_centered) or on points at the cell centres
        (python/sphere.py:2000-2191)."""
        path = folder + '/fluid-' + self.sid + '.vti'
        if verbose:
            print('Output file: ' + path)
        v_f, p_f, phi, dphi, p_c = self._grid_cells()
        dx = (self.L - self.origo) / self.num
        if cell_centered:
            origin, dims, where = self.origo, self.num + 1, "CellData"
        else:
            origin, dims, where = self.origo + 0.5 * dx, self.num, "PointData"
        ext = " ".join("0 %d" % (int(d) - 1) for d in dims)

        def flat(a):  # x fastest, as VTK orders image data
            a = numpy.asarray(a)
            return a.transpose(2, 1, 0, 3).reshape(-1, a.shape[3]) if a.ndim == 4 else \
                a.transpose(2, 1, 0).ravel()
        arrays = [("Pressure [Pa]", p_f), ("Velocity [m/s]", v_f), ("Porosity [-]", phi),
                  ("Porosity change [1/s]", dphi)]
        if int(numpy.asarray(self.cfd_solver).ravel()[0]) == 1:
            phic = numpy.clip(phi, 0.1, 0.9)
            arrays += [("Permeability [m*m]", self.k_c[0] * phic ** 3 / (1.0 - phic ** 2)),
                       ("Constant pressure [-]", p_c)]
        with open(path, 'w') as fh:
            fh.write('<?xml version="1.0"?>\n<VTKFile type="ImageData" version="0.1" '
                     'byte_order="LittleEndian">\n  <ImageData WholeExtent="%s" Origin="%s" '
                     'Spacing="%s">\n    <Piece Extent="%s">\n      <%s>\n' % (
                         ext, " ".join("%r" % float(o) for o in origin),
                         " ".join("%r" % float(d) for d in dx), ext, where))
            for name, values in arrays:
                self._vtk_data(fh, name, "Float64", flat(values))
            fh.write('      </%s>\n    </Piece>\n  </ImageData>\n</VTKFile>' % where)

    def writeVTKall(self, cell_centered=True, verbose=True, forces=False):
        """VTK files of every output file of this simulation, 0 to the last
        (status): <sid>.NNNNN.vtu for the grains and fluid-<sid>.NNNNN.vti for
        the fluid (python/sphere.py:1555-1644).  Contact-force files need the
        vtk module in the reference and are not written."""
        last = status(self.sid)
        for i in range(last + 1):
            sb = sim(fluid=self.fluid)
            sb.readbin("../output/{0}.output{1:0=5}.bin".format(self.sid, i), verbose=False)
            sb.sid = self.sid + ".{:0=5}".format(i)
            loud = verbose and i in (0, last)
            if sb.np[0] > 0:
                sb.writeVTK(verbose=loud)
            if self.fluid:
                sb.writeFluidVTK(verbose=loud, cell_centered=cell_centered)

    # ------------------------------------------------------------------
    # execution, python/sphere.py:4482-4540
    # ------------------------------------------------------------------
    def run(self, verbose=True, hideinputfile=False, dry=False, valgrind=False,
            cudamemcheck=False, device=-1):
        """Write ../input/<sid>.bin and run it through libsphere_hip.so, with
        outputs in ../output (the reference's ``cd ..; ./sphere``)."""
        self.writebin(verbose=False)
        # a restart side car of the state read last goes beside the input
        # file, where the run looks for it; a stale one is removed
        side = "../input/" + self.sid + ".contacts.bin"
        src = getattr(self, "_restart", None)
        if src and os.path.exists(src):
            shutil.copyfile(src, side)
        elif os.path.exists(side):
            os.remove(side)
        if valgrind or cudamemcheck:
            print("Warning: valgrind/cuda-memcheck wrappers are not available; running plainly")
        if self.fluid and int(numpy.asarray(self.cfd_solver).ravel()[0]) == 0:
            return _run_navier_stokes("../input/" + self.sid + ".bin", "../output", verbose, dry,
                                      device)
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


def _run_navier_stokes(path, outdir, verbose, dry, device):
    """The time loop of src/device.cu:833-2834 for an input file with a
    Navier-Stokes fluid (cfd_solver = 0), on the GPU through NavierStokes
    (attached to an Engine that steps the grains, if any): output00000 when
    the step count is 0, then
    DEM iterations until time.current passes time.total, the conv.log line
    of each solve (every tenth iteration's, and every one that reaches
    maxiter, src/device.cu:1792-1820), and at every file interval the
    fields, output%05d and status.dat (src/device.cu:2576-2680)."""
    # the simulation id is the input file's name (src/sphere.cpp:16-95)
    sid = os.path.basename(path)
    if sid.endswith(".bin"):
        sid = sid[:-4]
    r = sim(sid=sid, fluid=True)
    r.readbin(path, verbose=False)
    if dry:
        print("  - Number of particles: %d\n  - Time step length: %g s\n  - Total time: %g s\n"
              "  - File output interval: %g s\n  - Fluid grid: %d*%d*%d (Navier-Stokes)"
              % (int(r.np[0]), r.time_dt[0], r.time_total[0], r.time_file_dt[0], *r.num))
        return 0
    grains = int(r.np[0]) > 0
    try:
        if grains:
            # the grains step in a context; the fluid is attached to it and
            # runs inside each of its steps
            P = _abi.ns_params_from_sim(r, grains=True)
            r.fluid = False
            st = r.to_state()
            r.fluid = True
            eng = _abi.Engine(max(0, int(device)))
            eng.load(st)
            ns = _abi.NavierStokes(P, r.v_f, r.p_f, device=max(0, int(device)))
            eng.attach_fluid(ns)
        else:
            eng = None
            ns = _abi.NavierStokes.from_sim(r, device=max(0, int(device)))
    except _abi.SphereError as e:
        print("Warning: the sphere run returned with status %d: %s" % (e.code, e))
        return e.code

    def write(name):
        with open(outdir + "/" + sid + name, "wb") as fh:
            fh.write(r._bytes())

    def status_file():
        with open(outdir + "/" + sid + ".status.dat", "w") as fh:
            fh.write("%2.4e %2.4e %d\n" % (r.time_current[0], 100.0 * r.time_current[0] / r.time_total[0],
                                            int(r.time_step_count[0])))
    if int(r.time_step_count[0]) == 0:
        write(".output00000.bin")
    status_file()
    dt = float(r.time_dt[0])
    maxiter = int(numpy.asarray(r.maxiter).ravel()[0])
    filetimeclock = 0.0
    it = 0
    rc = 0
    with open(outdir + "/" + sid + "-conv.log", "w") as conv:
        while r.time_current[0] <= r.time_total[0]:
            before = ns.stats()
            try:
                if eng is not None:
                    eng.step(1)
                else:
                    ns.step(1)
            except _abi.SphereError as e:
                print("Warning: the sphere run returned with status %d: %s" % (e.code, e))
                rc = e.code
                break
            after = ns.stats()
            if after.solves > before.solves:
                if after.nonconverged > before.nonconverged:
                    conv.write("%d\t%d\n" % (it + 1, maxiter - 1))
                    print("Error, the epsilon solution in the fluid calculations did not "
                          "converge. Try increasing the value of 'ns.maxiter' (%d) or increase "
                          "'ns.tolerance' (%g)." % (maxiter, numpy.asarray(r.tolerance).ravel()[0]))
                elif it % 10 == 0:
                    conv.write("%d\t%d\n" % (it + 1, after.last_nijac))
            r.time_current[0] += dt
            filetimeclock += dt
            it += 1
            if verbose and it % 10000 == 0:
                print("\r  Current time: %g/%g s." % (r.time_current[0], r.time_total[0]), end="")
            if filetimeclock >= r.time_file_dt[0]:
                if eng is not None:
                    # the context's output event has advanced its file
                    # number and refreshed the fluid's host copy of p
                    r.from_state(eng.read())
                    r.fluid = True
                    r.cfd_solver = numpy.zeros(1, dtype=numpy.int32)
                    v, p, phi, dphi = ns.read()
                    r.v_f, r.p_f, r.phi, r.dphi = v, p, phi, dphi
                    r.f_d, r.f_p, r.f_v, r.f_sum = ns.grain_forces()
                else:
                    ns.to_sim(r)
                    r.time_step_count[0] += 1
                write(".output%05d.bin" % int(r.time_step_count[0]))
                status_file()
                filetimeclock = 0.0
    if eng is not None:
        eng.close()
    ns.close()
    if verbose:
        print()
    return rc


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
