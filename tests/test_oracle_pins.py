"""Closed-form pins of the oracle functions the collision tests leave open.

Each case is set up so that one step of the oracle must produce a value that
mechanics (or the reference's stated expression, where the reference defines
a quantity of its own such as the frictional heat rate) fixes independently
of the oracle's code:

* contact model 3 (Hertz-Mindlin, the declared deviation of the build):
  normal force 4/3 E* sqrt(R*) d^(3/2) with 1/E* = 2 (1 - nu^2)/E and
  1/R* = 1/r_a + 1/r_b; Mindlin tangential stiffness 8 G* sqrt(R* d) with
  1/G* = 2 (2 - nu)/G; normal damping gamma_n d^(1/4) v_n and its power
  gamma_n v_n^2; the Coulomb limit mu_s |f_n| below a sliding speed of
  0.001 m/s and mu_d |f_n| above; the spring reset to the Coulomb length on
  a slip and the heat rate |delta_t0 - delta_t| k_t / dt; rolling resistance
  min(gamma_r r |w|, mu_r r |f_n|);
* contact model 2 in one step (static spring and slip);
* contact model 1 (history free): viscous or Coulomb friction, normal spring,
  capillary force kappa sqrt(r_a r_b) at contact;
* the fixvel branches of integrate (held horizontal velocity, 2-D, frozen),
  the velocity step at t in [10, 15) s, the wmode-3 shear-stress drive and
  the sinusoidal top-wall stress;
* Darcy: the pressure force on a linear pressure field, -grad(p) V - rho_f V g,
  and the forcing of the pressure equation by two grains approaching each
  other symmetrically about a cell centre (exact porosity change, zero mean
  grain velocity).

CPU only."""
import math

import numpy
import pytest

import sphere_amd
from _oracle import Oracle

RHO = 2600.0


def vol(r):
    return 4.0 / 3.0 * math.pi * r ** 3


# ----------------------------------------------------------------------
# one step of a two-grain contact
# ----------------------------------------------------------------------
def pair_step(model, d, r_a=0.01, r_b=0.012, va=(0, 0, 0), vb=(0, 0, 0), wa=(0, 0, 0),
              wb=(0, 0, 0), dt0=None, dt=1e-6, **params):
    """Grains a (left) and b (right) along x overlapping by d (d < 0: apart
    by -d); contact history dt0 in a's slot for b.  Returns the oracle's
    force, torque, es_dot, ev_dot of a and its stored spring after one step."""
    s = sphere_amd.sim(np=2, sid="pin")
    c, h = 0.1, 0.5 * (r_a + r_b - d)
    s.x[:] = [[c - h, c, c], [c + h, c, c]]
    s.radius[:] = [r_a, r_b]
    s.vel[:] = [va, vb]
    s.angvel[:] = [wa, wb]
    s.defaultParams(rho=RHO, gamma_wn=0.0, gamma_wt=0.0)
    for k, v in params.items():
        getattr(s, k)[0] = v
    s.contactmodel[0] = model
    s.g[:] = 0.0
    s.defineWorldBoundaries(L=[0.2, 0.2, 0.2], dx=0.05)
    s.initTemporal(total=1.0, file_dt=1.0, dt=dt)
    o = Oracle(s.to_state())
    if dt0 is not None:
        nc = 32
        cont = numpy.full(2 * nc, 2, numpy.uint32)
        cont[nc - 1], cont[2 * nc - 1] = 1, 0
        dts = numpy.zeros((2 * nc, 3))
        dts[nc - 1] = dt0
        dts[2 * nc - 1] = -numpy.asarray(dt0, float)
        o.write_contacts(cont, numpy.zeros((2 * nc, 3)), dts)
    o.step(1)
    st = o.read()
    out = dict(F=st.get("force", 2, 3)[0].copy(), T=st.get("torque", 2, 3)[0].copy(),
               es_dot=float(st.get("es_dot", 2, 1)[0]), ev_dot=float(st.get("ev_dot", 2, 1)[0]))
    if model != 1:
        c_, _, d_ = o.read_contacts()
        row = numpy.nonzero(c_[:32] == 1)[0]
        out["spring"] = d_[row[0]].copy() if row.size else None
    return out


def hertz_constants(E, nu, r_a, r_b):
    """Hertz and Mindlin constants of two spheres of one material."""
    G = E / (2.0 * (1.0 + nu))
    E_star = 1.0 / ((1.0 - nu ** 2) / E + (1.0 - nu ** 2) / E)
    G_star = 1.0 / ((2.0 - nu) / G + (2.0 - nu) / G)
    R_star = 1.0 / (1.0 / r_a + 1.0 / r_b)
    return 4.0 / 3.0 * E_star * math.sqrt(R_star), 8.0 * G_star * math.sqrt(R_star)


HM = dict(E=7e9, k_r=0.3, mu_s=0.6, mu_d=0.4, gamma_n=0.0, gamma_t=0.0)
R_A, R_B = 0.01, 0.012


def close(a, b, rel=1e-9):
    a = numpy.asarray(a, float)
    b = numpy.asarray(b, float)
    return numpy.max(numpy.abs(a - b)) <= rel * max(numpy.max(numpy.abs(b)), 1e-300)


def test_hertz_normal_and_mindlin_tangential_stiffness():
    d, s = 2e-5, 1e-7
    o = pair_step(3, d, R_A, R_B, dt0=(0.0, s, 0.0), **HM)
    kn, kt = hertz_constants(7e9, 0.3, R_A, R_B)
    Fn = kn * d ** 1.5
    kT = kt * math.sqrt(d)  # Mindlin: 8 G* sqrt(R* d)
    assert kT * s < 0.6 * Fn  # inside the Coulomb cone
    assert close(o["F"], [-Fn, -kT * s, 0.0])
    # torque of the tangential force about a's centre, lever r_a - d/2
    assert close(o["T"], [0.0, 0.0, -(R_A - 0.5 * d) * kT * s])
    assert close(o["spring"], [0.0, s, 0.0], rel=0) and o["es_dot"] == 0.0 and o["ev_dot"] == 0.0


def test_hertz_normal_damping_power():
    d, v, g = 2e-5, 0.05, 40.0
    o = pair_step(3, d, R_A, R_B, va=(v, 0, 0), **dict(HM, gamma_n=g))
    kn, _ = hertz_constants(7e9, 0.3, R_A, R_B)
    Fn = kn * d ** 1.5 + g * d ** 0.25 * v  # approaching: damping adds to the spring
    assert close(o["F"], [-Fn, 0.0, 0.0])
    assert close(o["ev_dot"], g * v * v)  # the reference's gamma_n v_n^2


@pytest.mark.parametrize("vt,mu", [(0.0005, 0.6), (0.002, 0.4)])
def test_hertz_coulomb_limit_slip_reset_and_heat(vt, mu):
    d, S, dt = 2e-5, 1e-5, 1e-6
    o = pair_step(3, d, R_A, R_B, va=(0, vt, 0), dt0=(0.0, S, 0.0), dt=dt, **HM)
    kn, kt = hertz_constants(7e9, 0.3, R_A, R_B)
    Fn = kn * d ** 1.5
    kT = kt * math.sqrt(d)
    assert kT * S > mu * Fn  # outside the cone: the spring slips
    limit = mu * Fn  # static coefficient below 0.001 m/s, dynamic above
    assert close(o["F"], [-Fn, -limit, 0.0])
    reset = limit / kT  # the spring length that carries the Coulomb force
    assert close(o["spring"], [0.0, reset, 0.0])
    # contactmodels.cuh:560: the spring energy released, |d_t0 - d_t| k_t / dt
    assert close(o["es_dot"], (S - reset) * kt / dt, rel=1e-10)


@pytest.mark.parametrize("gamma_r,mu_r", [(1e-4, 10.0), (10.0, 1e-3)])
def test_hertz_rolling_resistance(gamma_r, mu_r):
    d, w, dt = 2e-5, 3.0, 1e-6
    o = pair_step(3, d, R_A, R_B, wa=(0, 0, w), dt=dt, **dict(HM, gamma_r=gamma_r, mu_r=mu_r))
    kn, _ = hertz_constants(7e9, 0.3, R_A, R_B)
    Fn = kn * d ** 1.5
    resist = min(gamma_r * R_A * w, mu_r * R_A * Fn)
    assert close(o["T"], [0.0, 0.0, -resist])
    # no spring yet: the rolling contact point only stretches it
    assert close(o["F"], [-Fn, 0.0, 0.0])
    assert close(o["spring"], [0.0, (R_A - 0.5 * d) * w * dt, 0.0])


def test_linear_static_spring_and_slip():
    d, dt, kn, gt = 1e-5, 1e-6, 2e6, 5.0
    base = dict(k_n=kn, k_t=kn, mu_s=0.5, mu_d=0.3, gamma_n=0.0, gamma_t=gt)
    vt, S = 0.004, 2e-7
    o = pair_step(2, d, va=(0, vt, 0), dt0=(0.0, S, 0.0), dt=dt, **base)
    spring = S + vt * dt
    ft = kn * spring + gt * vt
    assert ft < 0.5 * kn * d
    assert close(o["F"], [-kn * d, -ft, 0.0])
    assert close(o["spring"], [0.0, spring, 0.0])
    assert close(o["ev_dot"], 0.5 * gt * vt * vt)  # contactLinear's half shares
    S = 1e-5
    o = pair_step(2, d, va=(0, vt, 0), dt0=(0.0, S, 0.0), dt=dt, **base)
    Fn = kn * d
    assert close(o["F"], [-Fn, -0.5 * Fn, 0.0])  # static coefficient caps the force
    # the spring resets to the length at which spring and damping together
    # carry the dynamic Coulomb force: k d + gamma_t v_t = mu_d |f_n|
    assert close(o["spring"], [0.0, (0.3 * Fn - gt * vt) / kn, 0.0])
    assert close(o["es_dot"], 0.5 * (0.5 * Fn) * vt, rel=1e-10)


@pytest.mark.parametrize("vt,gt,expect", [(0.002, 1.0, "viscous"), (0.0005, 1e6, "mu_s"),
                                          (0.002, 1e6, "mu_d")])
def test_history_free_friction(vt, gt, expect):
    d, kn = 1e-5, 2e6
    o = pair_step(1, d, va=(0, vt, 0), k_n=kn, k_t=kn, mu_s=0.5, mu_d=0.3, gamma_n=0.0,
                  gamma_t=gt)
    Fn = kn * d
    ft = {"viscous": gt * vt, "mu_s": 0.5 * Fn, "mu_d": 0.3 * Fn}[expect]
    assert close(o["F"], [-Fn, -ft, 0.0])
    assert close(o["T"], [0.0, 0.0, -(0.01 - 0.5 * d) * ft])


def test_history_free_capillary_force():
    """Model 1 with a liquid bridge (Richefeu et al. 2008, as the reference
    writes it, src/cohesion.cuh:237-277): at contact the pair attracts with
    the bridge's largest force kappa sqrt(r_a r_b), kappa = 2 pi gamma
    (contact angle 0), and the force varies with the gap g as exp(-g /
    lambda), lambda = 0.9 h sqrt(V_b / R_har), h = -sqrt(max(r_a/r_b,
    r_b/r_a)), R_har = 2 r_a r_b / (r_a + r_b).  The reference's h is
    negative, so the force grows with the gap until the rupture distance db
    (replicated)."""
    r_a, r_b, V_b, kappa = 0.01, 0.012, 1e-9, 2 * math.pi * 0.07
    c = float(numpy.float32(0.9))  # the reference's single-precision constant 0.9f
    lam = c * -math.sqrt(r_b / r_a) * math.sqrt(V_b / (2 * r_a * r_b / (r_a + r_b)))
    F0 = kappa * math.sqrt(r_a * r_b)
    for gap in (1e-12, 0.5 * abs(lam)):
        o = pair_step(1, -gap, r_a, r_b, k_n=1e6, k_t=1e6, kappa=kappa, V_b=V_b, db=1e-3)
        # a is pulled towards b (+x)
        assert close(o["F"], [F0 * math.exp(-gap / lam), 0.0, 0.0], rel=1e-8)
    o = pair_step(1, -2e-3, r_a, r_b, k_n=1e6, k_t=1e6, kappa=kappa, V_b=V_b, db=1e-3)
    assert not o["F"].any()  # beyond the rupture distance


# ----------------------------------------------------------------------
# integration branches
# ----------------------------------------------------------------------
def free_grains(fixvel, vel, angvel, dt=1e-3, g=(0.0, 0.0, -9.81), t0=0.0, periodic=1, L=2.0):
    n = len(fixvel)
    s = sphere_amd.sim(np=n, sid="free")
    s.radius[:] = 0.01
    s.x[:] = [[0.3 + 0.2 * i, 1.0, 1.0] for i in range(n)]
    s.vel[:] = vel
    s.angvel[:] = angvel
    s.fixvel[:] = fixvel
    s.defaultParams(rho=RHO)
    s.g[:] = g
    s.periodic[0] = periodic
    s.defineWorldBoundaries(L=[L, L, L], dx=0.1)
    s.initTemporal(total=1e3, file_dt=1e3, dt=dt)
    s.time_current[0] = t0
    return s


def test_fixvel_branches():
    v = [[0.3, -0.2, 0.05], [0.3, -0.2, 0.05], [0.3, -0.2, 0.05], [0.3, -0.2, 0.05]]
    w = [[1.0, 2.0, 3.0]] * 4
    # held (0.0001 < fixvel < 10), 2-D (>= 10), frozen (< -0.0001), free (0)
    s = free_grains([1.0, 10.0, -1.0, 0.0], v, w, g=(0.0, -9.81, -9.81), periodic=0)
    o = Oracle(s.to_state())
    steps, dt = 50, 1e-3
    o.step(steps)
    st = o.read()
    vel = st.get("vel", 4, 4)[:, :3]
    # held: horizontal velocity kept, gravity cancelled (no contact force)
    assert numpy.array_equal(vel[0], v[0])
    # frozen: no acceleration at all
    assert numpy.array_equal(vel[2], v[2])
    # 2-D: y held, z falls freely (TY3 with its first-step jolt: v0 + g dt (n + 1/2))
    assert vel[1][1] == v[1][1] and vel[1][0] == v[1][0]
    assert abs(vel[1][2] - (0.05 - 9.81 * dt * (steps + 0.5))) < 1e-12
    # free grain falls in y and z
    assert abs(vel[3][1] - (-0.2 - 9.81 * dt * (steps + 0.5))) < 1e-12
    wv = st.get("angvel", 4, 3)
    # no torques act, so spins are unchanged in every branch
    assert numpy.array_equal(wv, numpy.array(w))


def test_fixvel_holds_against_contact_forces():
    """A held grain pressed by a free neighbour keeps its horizontal
    velocity and its spin; only the normal force's z share would move it."""
    s = sphere_amd.sim(np=2, sid="held")
    s.radius[:] = 0.01
    s.x[:] = [[0.1, 0.1, 0.1], [0.1195, 0.1, 0.1]]  # overlapping along x
    s.vel[:] = [[0.02, 0.01, 0.0], [0.0, 0.0, 0.0]]
    s.fixvel[:] = [1.0, 0.0]
    s.defaultParams(rho=RHO, k_n=1e5, k_t=1e5, mu_s=0.5, mu_d=0.5)
    s.g[:] = 0.0
    s.defineWorldBoundaries(L=[0.2, 0.2, 0.2], dx=0.05)
    s.initTemporal(total=1.0, file_dt=1.0, dt=1e-5)
    o = Oracle(s.to_state())
    o.step(20)
    st = o.read()
    assert numpy.array_equal(st.get("vel", 2, 4)[0, :3], [0.02, 0.01, 0.0])
    assert numpy.array_equal(st.get("angvel", 2, 3)[0], [0.0, 0.0, 0.0])
    assert numpy.any(st.get("force", 2, 3)[0] != 0.0)
    assert st.get("vel", 2, 4)[1, 0] > 0.0  # the free grain is pushed away


def test_velocity_step_between_10_and_15_s():
    """Grains with 5 < fixvel < 10 have their x velocity multiplied by 10 at
    the first step at t >= 10 s and divided by 10 at the first at t >= 15 s
    (src/device.cu:1000-1005, 2461-2469)."""
    dt = 0.5
    s = free_grains([6.0, 1.0], [[0.125, 0.0, 0.0], [0.125, 0.0, 0.0]], [[0, 0, 0]] * 2, dt=dt,
                    g=(0.0, 0.0, 0.0), t0=9.75, L=50.0)
    o = Oracle(s.to_state())
    seen = []
    for _ in range(13):
        o.step(1)
        seen.append(float(o.read().get("vel", 2, 4)[0, 0]))
    t_start = [9.75 + k * dt for k in range(13)]
    expect = [1.25 if 10.0 <= t < 15.0 else 0.125 for t in t_start]
    assert seen == expect
    assert float(o.read().get("vel", 2, 4)[1, 0]) == 0.125  # fixvel 1 is not stepped


def test_wmode3_shear_stress_drive():
    """wmode 3: driven grains in the upper half accelerate in x at
    (tau_x + tau_eff) L0 L1 / m_w; without contacts tau_eff = 0, gravity is
    cancelled and spins held."""
    s = free_grains([1.0, 1.0, 0.0], [[0.0, 0.0, 0.0]] * 3, [[0.0, 0.0, 0.0]] * 3, dt=1e-4)
    s.x[0, 2] = 1.5  # upper half
    s.x[1, 2] = 0.5  # lower half: held like any fixvel grain
    s.nw = numpy.array([1], numpy.uint32)
    s.wmode = numpy.array([3], numpy.int32)
    s.w_n = numpy.array([[0.0, 0.0, -1.0]])
    s.w_x = numpy.array([1.9])
    s.w_m = numpy.array([7.0])
    s.w_vel = numpy.zeros(1)
    s.w_force = numpy.zeros(1)
    s.w_sigma0 = numpy.zeros(1)
    s.w_tau_x[0] = 150.0
    o = Oracle(s.to_state())
    n = 40
    o.step(n)
    v = o.read().get("vel", 3, 4)[:, :3]
    a = 150.0 * 2.0 * 2.0 / 7.0
    assert abs(v[0, 0] - a * 1e-4 * (n + 0.5)) < 1e-12 * a
    assert v[0, 1] == 0.0 and v[0, 2] == 0.0
    assert numpy.array_equal(v[1], [0.0, 0.0, 0.0])
    assert v[2, 2] < 0.0  # the free grain falls


@pytest.mark.parametrize("t0", [0.3, 1.9])
def test_top_wall_stress_modulation(t0):
    """A wall under sigma0 + A sin(2 pi f t) with no grains touching it: the
    first TY3 step from rest gives v = 1.5 a dt, a = -sigma L0 L1 / m_w."""
    s = free_grains([0.0], [[0.0, 0.0, 0.0]], [[0.0, 0.0, 0.0]], dt=1e-4, g=(0.0, 0.0, 0.0),
                    t0=t0)
    s.nw = numpy.array([1], numpy.uint32)
    s.wmode = numpy.array([1], numpy.int32)
    s.w_n = numpy.array([[0.0, 0.0, -1.0]])
    s.w_x = numpy.array([1.9])
    s.w_m = numpy.array([3.0])
    s.w_vel = numpy.zeros(1)
    s.w_force = numpy.zeros(1)
    s.w_sigma0 = numpy.array([1e3])
    s.setTopWallNormalStressModulation(A=400.0, f=0.7)
    o = Oracle(s.to_state())
    o.step(1)
    h = o.read().state
    sigma = 1e3 + 400.0 * math.sin(2 * math.pi * 0.7 * t0)
    a = -sigma * 4.0 / 3.0
    assert abs(h.w_mvfd[1] - 1.5 * a * 1e-4) < 1e-12 * abs(a)


# ----------------------------------------------------------------------
# Darcy coupling
# ----------------------------------------------------------------------
def darcy_box(n_cells=8, L=0.5, np_=1):
    s = sphere_amd.sim(np=np_, sid="darcypin", fluid=True)
    s.defaultParams(rho=RHO, k_n=1e6, k_t=1e6)
    s.defineWorldBoundaries([L, L, L], dx=L / n_cells)
    s.initFluid(mu=1e-3, rho=1000.0, cfd_solver=1)
    s.setFluidCompressibility(1e-8)
    s.setPermeabilityPrefactor(1e-10, verbose=False)
    s.bc_top[0] = 1
    s.bc_bot[0] = 1
    s.setTolerance(1e-8)
    s.setMaxIterations(100000)
    return s


def test_pressure_force_on_linear_pressure():
    """p = p0 + bx x + bz z: the grain feels -(bx, 0, bz) V - rho_f V g,
    whatever its place inside the cell (the trilinear weights of the eight
    surrounding nodes sum to one, and central differences of a linear field
    are exact)."""
    s = darcy_box()
    s.radius[0] = 0.02
    s.x[0] = [0.2431, 0.2702, 0.2855]
    s.g[:] = [0.0, 0.0, -9.81]
    dx = 0.5 / 8
    xc = (numpy.arange(8) + 0.5) * dx
    bx, bz = 300.0, -2000.0
    s.p_f[:] = 1000.0 + bx * xc[:, None, None] + bz * xc[None, None, :]
    s.initTemporal(total=1.0, file_dt=1.0, dt=1e-6)
    o = Oracle(s.to_state())
    o.step(1)
    fp = o.read().get("f_p", 1, 3)[0]
    V = vol(0.02)
    expect = numpy.array([-bx * V, 0.0, -bz * V + 1000.0 * V * 9.81])
    assert numpy.max(numpy.abs(fp - expect)) <= 1e-10 * numpy.max(numpy.abs(expect))


def test_darcy_forcing_of_two_approaching_grains():
    """Two grains centred on a cell's y and z and placed at x = X -+ a approach
    each other at u.  Only this cell and its x neighbours see them; the
    cell's solid volume is 2 V (1 - a/dx), so between fluid steps 0 and 2
    (positions a + u dt and a - u dt around step 1) the porosity change is
    dphi = (phi(step 2 position) - phi(step 0)) / 2 = -(V / V_c) 2 u dt / dx,
    the mean grain velocity in the cell is 0 and the porosity gradient is 0.
    With p = 0 the explicit update is the forcing alone:
    dp_expl = -dphi / (beta_f phi (1 - phi)) (src/darcy.cuh:1425-1598)."""
    n, L = 8, 0.5
    dx = L / n
    s = darcy_box(n, L, np_=2)
    r, a, u, dt = 0.02, 0.375 * dx, 0.0625, 2.0 ** -16
    X = 4 * dx + 0.5 * dx
    s.radius[:] = r
    s.x[:] = [[X - a, X, X], [X + a, X, X]]
    s.vel[:] = [[u, 0.0, 0.0], [-u, 0.0, 0.0]]
    s.g[:] = 0.0
    s.initTemporal(total=1.0, file_dt=1.0, dt=dt)
    o = Oracle(s.to_state())
    o.step(2)
    V, Vc = vol(r), dx ** 3
    a1 = a - u * dt  # the distance at the second fluid step
    phi1 = 1.0 - 2.0 * V * (1.0 - a1 / dx) / Vc
    dphi = -(V / Vc) * 2.0 * u * dt / dx
    phi = o.darcy_field(1)[4 + 1, 4 + 1, 4 + 1]
    assert abs(phi - phi1) < 1e-13 and 0.1 < phi1 < 0.9
    # (the dphi field itself is zeroed after a step without file output)
    expl = o.darcy_field(5)[4 + 1, 4 + 1, 4 + 1]
    want = -dphi / (1e-8 * phi1 * (1.0 - phi1))
    assert want > 0.0  # compression raises the pore pressure
    assert abs(expl - want) < 1e-9 * want


@pytest.mark.parametrize("face", ["top", "bottom"])
def test_darcy_flux_boundary(face):
    """bc 4 (prescribed specific flux q, src/darcy.cuh:270-312): the ghost
    pressure p - (mu / k) q dz makes the boundary cell's Darcy velocity
    -(k/mu)(p_ghost - p_neighbour)/(2 dz)/phi = q / (2 phi) (the central
    difference spans two cells).  The reference never writes the ghost
    permeability of a bc-4 face (setDarcyGhostNodes handles bc 0-2), so the
    face conductance 2 k_g k / (k_g + k) is that of a zero ghost and the
    pressure field does not change (declared: the ghost is held at 0)."""
    q = 1e-3
    s = sphere_amd.sim(np=0, nd=3, nw=0, sid="flux", fluid=True)
    s.defaultParams()
    s.defineWorldBoundaries([0.4, 0.4, 0.8], dx=0.1)
    s.initFluid(mu=1e-3, cfd_solver=1)
    s.setFluidCompressibility(1e-8)
    s.setPermeabilityPrefactor(1e-12, verbose=False)
    if face == "top":
        s.setFluidBottomNoFlow()
        s.setFluidTopFixedFlux(q)
    else:
        s.setFluidTopNoFlow()
        s.setFluidBottomFixedFlux(q)
    s.initTemporal(total=1.0, file_dt=1.0, dt=1e-4)
    o = Oracle(s.to_state())
    o.step(2)
    st = o.read()
    v = st.get("v_f", 4 * 4 * 8, 3).reshape(8, 4, 4, 3)
    p = st.get("p_f", 4 * 4 * 8, 1)
    phi = 0.9  # no grains
    layer = 7 if face == "top" else 0
    want = q / (2.0 * phi) if face == "top" else -q / (2.0 * phi)
    assert numpy.allclose(v[layer, :, :, 2], want, rtol=1e-12, atol=0)
    others = numpy.delete(v[:, :, :, 2], layer, axis=0)
    assert not others.any() and not v[:, :, :, :2].any()
    assert not p.any()
