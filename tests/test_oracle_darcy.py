"""The oracle's Darcy solver against the reference's own Darcy assertions
(tests/cfd_tests_darcy.py, tests/fluid_particle_interaction_darcy.py) and
closed forms of the porosity and permeability fields (src/darcy.cuh).
CPU only."""
import math

import numpy
import pytest

import sphere_amd
from _cases import darcy_case, darcy_empty, fluid_particle
from _oracle import Oracle


def cells(s):
    return int(s.num[0] * s.num[1] * s.num[2])


def field(st, s, name, w=1):
    nx, ny, nz = (int(v) for v in s.num)
    a = st.get(name, cells(s), w)
    return a.reshape(nz, ny, nx, w) if w > 1 else a.reshape(nz, ny, nx)


def test_no_forcing():
    """No gravity, no gradient: p stays exactly 0, the solver needs no
    iteration in any fluid step, and the fluid does not move
    (cfd_tests_darcy.py:14-45)."""
    s = darcy_empty()
    o = Oracle(s.to_state())
    o.step(11)
    st = o.read()
    assert numpy.all(field(st, s, "p_f") == 0.0)
    conv = o.conv_log()
    assert conv.shape[0] >= 1 and conv[:, 1].sum() == 0
    assert numpy.all(numpy.abs(field(st, s, "v_f", 3)) < 1e-6)


def test_pressure_gradient_steady_profile():
    """p = 1 on the top layer and 0 at the bottom (Dirichlet): the pressure
    relaxes to the linear profile and the fluid flows down
    (cfd_tests_darcy.py:48-69; the reference allows 0.1, here 1e-6)."""
    s = darcy_empty(p_top=1.0)
    o = Oracle(s.to_state())
    o.step(4000)
    st = o.read()
    p = field(st, s, "p_f")
    ideal = numpy.linspace(0.0, 1.0, int(s.num[2]))
    assert numpy.max(numpy.abs(p - ideal[:, None, None])) < 1e-6
    v = field(st, s, "v_f", 3)
    assert numpy.all(v[..., 2] < 0.0)
    assert numpy.all(numpy.abs(v[..., 0]) < 1e-7) and numpy.all(numpy.abs(v[..., 1]) < 1e-7)
    # Convergence rate (2/3), cfd_tests_darcy.py:72-76
    conv = o.conv_log()
    assert numpy.all(conv[:6, 1] < 1000) and numpy.all(conv[6:, 1] < 20)


@pytest.mark.parametrize("test", [1, 2])
def test_grain_pulled_to_low_pressure(test):
    """A grain in a pressure gradient (test 1) or under gravity with the
    hydrostatic fluid (test 2) moves down (fluid_particle_interaction_darcy.py)."""
    s = fluid_particle(test)
    o = Oracle(s.to_state())
    steps = int(round(s.time_total[0] / s.time_dt[0]))
    o.step(min(steps, 3000))
    st = o.read()
    assert st.get("vel", 1, 4)[0, 2] < 0.0
    if test == 1:
        assert st.get("f_p", 1, 3)[0, 2] < 0.0


def porosity_case(xs, r, nxyz=(5, 5, 8), dx=0.1):
    """Grains at rest, no gravity, in a Darcy grid of dx cells (one fluid
    step computes the porosities)."""
    n = len(xs)
    s = sphere_amd.sim(np=n, sid="phi", nw=0, fluid=True)
    s.x[:] = xs
    s.radius[:] = r
    s.defaultParams(k_n=1e6, k_t=1e6)
    s.g[:] = 0.0
    s.periodic[0] = 1
    s.defineWorldBoundaries([nxyz[0] * dx, nxyz[1] * dx, nxyz[2] * dx], dx=dx)
    s.initFluid(cfd_solver=1)
    s.setPermeabilityPrefactor(3.5e-13, verbose=False)
    s.initTemporal(total=1.0, file_dt=1.0, dt=1e-7)
    return s


def test_single_sphere_porosity_and_permeability():
    """A sphere at the centre of a cell weighs 1 there and 0 in every other
    cell (trilinear weight, src/darcy.cuh:282-311): phi = 1 - V_s / V_cell in
    that cell, the 0.9 cap elsewhere, and k = k_c phi^3 / (1 - phi)^2."""
    dx = 0.1
    r = (0.3 * dx ** 3 * 3.0 / (4.0 * math.pi)) ** (1.0 / 3.0)  # V_s = 0.3 V_cell
    centre = numpy.array([[2.5 * dx, 2.5 * dx, 3.5 * dx]])
    s = porosity_case(centre, r)
    o = Oracle(s.to_state())
    o.step(1)
    phi = o.darcy_field(1)  # with ghost layers
    nx, ny, nz = (int(v) for v in s.num)
    phi = phi.reshape(nz + 2, ny + 2, nx + 2)[1:-1, 1:-1, 1:-1]
    vs = 4.0 / 3.0 * math.pi * r ** 3
    expect = numpy.full((nz, ny, nx), 0.9)
    expect[3, 2, 2] = 1.0 - vs / dx ** 3
    assert abs(expect[3, 2, 2] - 0.7) < 1e-12
    assert numpy.allclose(phi[1:nz - 1], expect[1:nz - 1], rtol=0, atol=1e-12)
    k = o.darcy_field(3).reshape(nz + 2, ny + 2, nx + 2)[1:-1, 1:-1, 1:-1]
    kc = 3.5e-13
    kk = kc * expect[3, 2, 2] ** 3 / (1.0 - expect[3, 2, 2]) ** 2
    assert abs(k[3, 2, 2] - kk) < 1e-12 * kk


def test_half_cell_offset_splits_volume():
    """A sphere half a cell off a centre along x counts half in each of the
    two cells (the weight is linear in the offset)."""
    dx = 0.1
    r = (0.6 * dx ** 3 * 3.0 / (4.0 * math.pi)) ** (1.0 / 3.0)  # V_s = 0.6 V_cell
    s = porosity_case(numpy.array([[3.0 * dx, 2.5 * dx, 3.5 * dx]]), r)
    o = Oracle(s.to_state())
    o.step(1)
    nx, ny, nz = (int(v) for v in s.num)
    phi = o.darcy_field(1).reshape(nz + 2, ny + 2, nx + 2)[1:-1, 1:-1, 1:-1]
    assert abs(phi[3, 2, 2] - 0.7) < 1e-12 and abs(phi[3, 2, 3] - 0.7) < 1e-12
    assert abs(phi[3, 2, 1] - 0.9) < 1e-12 and abs(phi[3, 2, 4] - 0.9) < 1e-12
