"""The oracle against closed forms of the mechanics it restates
(src/contactmodels.cuh, src/integration.cuh): two-sphere collisions
(velocity exchange, energy, linear contact time pi*sqrt(m*/k_n), Hertz
contact time 2.87 (m*^2 / (R* E*^2 v))^(1/5)), viscous dissipation equal to
the accumulated ev, the elastic wall rebound, the TY3 integrator on constant
acceleration (with its first-step jolt), periodic wrap, the sigma0-driven
wall, and the sign of the frictional spin.  A dropped term, a wrong sign or a
wrong coefficient in the oracle fails one of these.  CPU only."""
import math

import numpy
import pytest

import sphere_amd
from _oracle import Oracle

RHO = 2600.0


def mass(r):
    return RHO * 4.0 / 3.0 * math.pi * r ** 3


def pair_case(v=0.1, r=0.01, gap=1e-6, model=2, k_n=1e5, E=0.0, nu=0.3, gamma_n=0.0, mu=0.0,
              offset_y=0.0, spin=None, dt=None, L=0.2):
    s = sphere_amd.sim(np=2, sid="pair", nw=0)
    c = L / 2
    s.x[0] = [c - r - 0.5 * gap, c, c]
    s.x[1] = [c + r + 0.5 * gap, c + offset_y, c]
    s.radius[:] = r
    s.vel[0] = [v, 0.0, 0.0]
    if spin is not None:
        s.angvel[0] = spin
    s.defaultParams(mu_s=mu, mu_d=mu, k_n=k_n, k_t=k_n, gamma_n=gamma_n, gamma_t=0.0,
                    gamma_wn=0.0, gamma_wt=0.0, rho=RHO)
    s.contactmodel[0] = model
    if model == 3:
        s.setYoungsModulus(E)
        s.setPoissonsRatio(nu)
    s.g[:] = 0.0
    s.defineWorldBoundaries(L=[L, L, L], dx=2.2 * r)
    s.initTemporal(total=1.0, file_dt=1.0, dt=dt)
    return s


def run_collision(s, steps):
    o = Oracle(s.to_state())
    n = int(s.np[0])
    in_contact = 0
    for _ in range(steps):
        o.step(1)
        rows = o.read_contacts()[0].reshape(n, 32)
        in_contact += int(numpy.any(rows[0] != n))
    return o, in_contact


def kinetic(st, n):
    m = mass(st.get("x", n, 4)[:, 3])
    v = st.get("vel", n, 4)[:, :3]
    return float(numpy.sum(0.5 * m * numpy.sum(v * v, axis=1)))


def test_linear_head_on_exchange_and_duration():
    r, k_n, v = 0.01, 1e5, 0.1
    m_star = mass(r) / 2.0
    t_c = math.pi * math.sqrt(m_star / k_n)
    dt = t_c / 400.0
    s = pair_case(v=v, r=r, k_n=k_n, dt=dt)
    ke0 = 0.5 * mass(r) * v * v
    o, nsteps = run_collision(s, 520)
    st = o.read()
    vel = st.get("vel", 2, 4)[:, :3]
    assert abs(vel[0, 0]) < 2e-3 * v and abs(vel[1, 0] - v) < 2e-3 * v
    assert abs(kinetic(st, 2) - ke0) < 2e-3 * ke0
    assert abs(nsteps * dt - t_c) < 0.01 * t_c


def test_hertz_contact_time():
    r, E, nu, v = 0.01, 1e7, 0.3, 0.1
    m_star = mass(r) / 2.0
    R_star = r / 2.0
    E_star = E / (2.0 * (1.0 - nu * nu))
    t_c = 2.87 * (m_star ** 2 / (R_star * E_star ** 2 * v)) ** 0.2
    dt = t_c / 400.0
    s = pair_case(v=v, r=r, model=3, E=E, nu=nu, dt=dt, k_n=1e5)
    o, nsteps = run_collision(s, 520)
    st = o.read()
    vel = st.get("vel", 2, 4)[:, :3]
    assert abs(vel[1, 0] - v) < 5e-3 * v and abs(vel[0, 0]) < 5e-3 * v
    # 2.87 is the three-figure constant of the Hertz impact time
    assert abs(nsteps * dt - t_c) < 0.01 * t_c


def test_viscous_dissipation_equals_ev():
    """Normal damping: the kinetic energy lost equals the sum of the
    accumulated viscous energy ev of both grains (each grain books half of
    gamma_n v_n^2), and the restitution coefficient is that of a damped
    linear oscillator, exp(-zeta pi / sqrt(1 - zeta^2)).  Both hold up to the
    short stretch before separation where the law clamps the normal force at
    zero, which costs O(zeta^2): zeta = 0.02 here."""
    r, k_n, v = 0.01, 1e5, 0.1
    m_star = mass(r) / 2.0
    t_c = math.pi * math.sqrt(m_star / k_n)
    zeta = 0.02
    gamma = zeta * 2.0 * math.sqrt(m_star * k_n)
    dt = t_c / 400.0
    s = pair_case(v=v, r=r, k_n=k_n, gamma_n=gamma, dt=dt)
    ke0 = 0.5 * mass(r) * v * v
    o, _ = run_collision(s, 600)
    st = o.read()
    lost = ke0 - kinetic(st, 2)
    ev = float(numpy.sum(st.get("ev", 2, 1)))
    e = math.exp(-zeta * math.pi / math.sqrt(1 - zeta * zeta))
    vel = st.get("vel", 2, 4)[:, 0]
    assert abs((vel[1] - vel[0]) / v - e) < 0.002
    # equal masses, one at rest: the relative motion holds half the energy
    assert abs(lost / ke0 - 0.5 * (1.0 - e * e)) < 0.002
    assert abs(lost - ev) < 0.02 * lost


def test_wall_rebound():
    """A grain hitting the fixed bottom at -0.1 m/s leaves it at +0.1 m/s
    after pi sqrt(m / k_n) (no wall damping)."""
    r, k_n, v = 0.01, 1e5, 0.1
    t_c = math.pi * math.sqrt(mass(r) / k_n)
    dt = t_c / 400.0
    s = sphere_amd.sim(np=1, sid="wall", nw=0)
    s.x[0] = [0.1, 0.1, r + 1e-6]
    s.radius[:] = r
    s.vel[0] = [0.0, 0.0, -v]
    s.defaultParams(mu_s=0.0, mu_d=0.0, k_n=k_n, k_t=k_n, gamma_wn=0.0, gamma_wt=0.0, rho=RHO)
    s.contactmodel[0] = 2
    s.g[:] = 0.0
    s.periodic[0] = 1
    s.defineWorldBoundaries(L=[0.2, 0.2, 0.2], dx=0.05)
    s.initTemporal(total=1.0, file_dt=1.0, dt=dt)
    o = Oracle(s.to_state())
    o.step(520)
    vz = o.read().get("vel", 1, 4)[0, 2]
    assert abs(vz - v) < 2e-3 * v


def single_grain(dt, g=-9.81, v0=(0.3, -0.2, 1.0), x0=(0.5, 0.5, 0.5), periodic=0, L=1.0):
    s = sphere_amd.sim(np=1, sid="one", nw=0)
    s.x[0] = x0
    s.radius[:] = 0.01
    s.vel[0] = v0
    s.defaultParams(rho=RHO)
    s.contactmodel[0] = 2
    s.g[:] = [0.0, 0.0, g]
    s.periodic[0] = periodic
    s.defineWorldBoundaries(L=[L, L, L], dx=0.1)
    s.initTemporal(total=1.0, file_dt=1.0, dt=dt)
    return s


def test_ty3_constant_acceleration():
    """TY3 (src/integration.cuh:222-246) integrates a constant acceleration
    exactly once the previous acceleration is known; the first step sees
    acc0 = 0 and adds g dt^2/6 to the position and g dt/2 to the velocity."""
    dt, g, n = 1e-4, -9.81, 60
    x0 = numpy.array([0.5, 0.5, 0.5])
    v0 = numpy.array([0.3, -0.2, 1.0])
    a = numpy.array([0.0, 0.0, g])
    o = Oracle(single_grain(dt, g, v0, x0).to_state())
    o.step(1)
    st = o.read()
    x1 = x0 + v0 * dt + (0.5 + 1.0 / 6.0) * a * dt * dt
    v1 = v0 + 1.5 * a * dt
    assert numpy.allclose(st.get("x", 1, 4)[0, :3], x1, rtol=0, atol=1e-15)
    assert numpy.allclose(st.get("vel", 1, 4)[0, :3], v1, rtol=0, atol=1e-15)
    o.step(n - 1)
    st = o.read()
    t = (n - 1) * dt
    xn = x1 + v1 * t + 0.5 * a * t * t
    vn = v1 + a * t
    assert numpy.allclose(st.get("x", 1, 4)[0, :3], xn, rtol=0, atol=1e-13)
    assert numpy.allclose(st.get("vel", 1, 4)[0, :3], vn, rtol=0, atol=1e-12)
    # xyzsum accumulates the displacement
    assert numpy.allclose(st.get("xyzsum", 1, 3)[0], xn - x0, rtol=0, atol=1e-13)


def test_periodic_wrap():
    """Leaving through x = L0 re-enters at x - L0 (periodic = 1); xyzsum keeps
    the unwrapped displacement."""
    dt = 1e-3
    L = 1.0
    x0 = numpy.array([L - 2e-4, 0.5, 0.5])
    v0 = numpy.array([1.0, 0.0, 0.0])
    o = Oracle(single_grain(dt, 0.0, v0, x0, periodic=1, L=L).to_state())
    o.step(1)
    st = o.read()
    x = st.get("x", 1, 4)[0, 0]
    assert 0.0 <= x < L
    assert abs(x - (x0[0] + v0[0] * dt - L)) < 1e-15
    assert abs(st.get("xyzsum", 1, 3)[0, 0] - v0[0] * dt) < 1e-15


def test_sigma0_wall_acceleration():
    """A wall under normal stress sigma0 touching nothing accelerates at
    -sigma0 L0 L1 / m_w (src/integration.cuh:373-393), integrated by TY3 with
    the same first-step rule as the grains."""
    dt, n = 1e-5, 40
    s = single_grain(dt, 0.0, (0.0, 0.0, 0.0), (0.5, 0.5, 0.1), periodic=1)
    s.nw[0] = 1
    s.wmode = numpy.array([1], dtype=numpy.int32)
    s.w_n = numpy.array([[0.0, 0.0, -1.0]])
    s.w_x = numpy.array([0.9])
    s.w_m = numpy.array([2.0])
    s.w_vel = numpy.array([0.0])
    s.w_force = numpy.array([0.0])
    s.w_sigma0 = numpy.array([1e3])
    s.w_sigma0_A = numpy.array([0.0])
    s.w_sigma0_f = numpy.array([0.0])
    s.w_tau_x = numpy.array([0.0])
    o = Oracle(s.to_state())
    o.step(1)
    a = -1e3 * 1.0 * 1.0 / 2.0
    w1 = 0.9 + (0.5 + 1.0 / 6.0) * a * dt * dt
    v1 = 1.5 * a * dt
    st = o.read()
    assert abs(st.state.w_nx[3] - w1) < 1e-15 and abs(st.state.w_mvfd[1] - v1) < 1e-15
    o.step(n - 1)
    st = o.read()
    t = (n - 1) * dt
    assert abs(st.state.w_nx[3] - (w1 + v1 * t + 0.5 * a * t * t)) < 1e-13
    assert abs(st.state.w_mvfd[1] - (v1 + a * t)) < 1e-12


def test_friction_spin_sign():
    """A grain sliding over a resting one (offset contact, mu > 0) gains spin
    about the axis n x v that rolling on the partner's surface implies."""
    r, k_n = 0.01, 1e6
    s = sphere_amd.sim(np=2, sid="spin", nw=0)
    s.x[0] = [0.1, 0.1, 0.1]
    s.x[1] = [0.1, 0.1, 0.1 - 2 * r + 1e-5]  # below, overlapping
    s.radius[:] = r
    s.vel[0] = [0.05, 0.0, 0.0]
    s.defaultParams(mu_s=0.5, mu_d=0.5, k_n=k_n, k_t=k_n, gamma_wn=0.0, gamma_wt=0.0, rho=RHO)
    s.contactmodel[0] = 2
    s.g[:] = 0.0
    s.defineWorldBoundaries(L=[0.2, 0.2, 0.2], dx=0.05)
    s.initTemporal(total=1.0, file_dt=1.0, dt=1e-6)
    o = Oracle(s.to_state())
    o.step(20)
    w = o.read().get("angvel", 2, 3)
    # grain 0 slides +x on top of grain 1: friction at its bottom points -x,
    # torque r_c x f = (-r e_z) x (-f e_x) = +r f e_y
    assert w[0, 1] > 0.0 and w[1, 1] > 0.0
    assert abs(w[0, 0]) < 1e-12 * abs(w[0, 1]) + 1e-300
    es = o.read().get("es", 2, 1)
    assert numpy.all(es >= 0.0)
