"""The oracle's bond model (src/cohesion.cuh:8-233) against the closed forms
the reference's bond tests intended (tests/bond_tests.py:54-187, disabled
there): a bonded pair at rest stays at rest; with no damping the kinetic,
rotational and bond energies (python/sphere.py:4360-4376) add up to a
constant (the springs are advanced with the step's start velocities, so the
sum drifts at first order in dt: 20,000 steps per period keep it below
1e-3); a normal perturbation oscillates at sqrt(2 k_n A / m) and a twist at
sqrt(2 k_t J / I); every mode keeps its motion in its own axes.  CPU only."""
import math

import numpy
import pytest

import sphere_amd
from _oracle import Oracle

RHO = 2600.0
R = 0.01
KN = 1e7
KT = 1e7


def pair(vel1=(0.0, 0.0, 0.0), angvel0=(0.0, 0.0, 0.0), angvel1=(0.0, 0.0, 0.0), dt=None):
    s = sphere_amd.sim(np=2, sid="bond", nw=0)
    d = 0.1 * R  # no contact: only the bond acts
    s.x[0] = [0.1, 0.1, 0.1]
    s.x[1] = [0.1 + 2 * R + d, 0.1, 0.1]
    s.radius[:] = R
    s.vel[1] = vel1
    s.angvel[0] = angvel0
    s.angvel[1] = angvel1
    s.defaultParams(mu_s=0.5, mu_d=0.5, k_n=KN, k_t=KT, gamma_n=0.0, gamma_t=0.0,
                    gamma_wn=0.0, gamma_wt=0.0, rho=RHO)
    s.contactmodel[0] = 2
    s.g[:] = 0.0
    s.defineWorldBoundaries(L=[0.2, 0.2, 0.2], dx=0.05)
    s.bond(0, 1)
    m = RHO * 4.0 / 3.0 * math.pi * R ** 3
    A = math.pi * R ** 2
    w_n = math.sqrt(2.0 * KN * A / m)
    s.initTemporal(total=1.0, file_dt=1.0, dt=dt if dt else 2 * math.pi / w_n / 20000.0)
    return s


def energies(st):
    n = 2
    x = st.get("x", n, 4)
    r = x[:, 3]
    m = RHO * 4.0 / 3.0 * math.pi * r ** 3
    v = st.get("vel", n, 4)[:, :3]
    w = st.get("angvel", n, 3)
    kin = float(numpy.sum(0.5 * m * numpy.sum(v * v, axis=1)))
    rot = float(numpy.sum(0.5 * (0.4 * m * r * r) * numpy.sum(w * w, axis=1)))
    Rb = st.state.lambda_bar * min(r)
    A = math.pi * Rb ** 2
    I = 0.25 * math.pi * Rb ** 4
    J = 2.0 * I
    dn = st.get("bonds_delta_n", 1, 1)[0]
    dtv = st.get("bonds_delta_t", 1, 3)[0]
    on = st.get("bonds_omega_n", 1, 1)[0]
    ot = st.get("bonds_omega_t", 1, 3)[0]
    pot = (0.5 * A * KN * dn * dn + 0.5 * A * KT * float(numpy.dot(dtv, dtv)) +
           0.5 * J * KT * on * on + 0.5 * I * KN * float(numpy.dot(ot, ot)))
    return kin, rot, pot


def test_pair_at_rest_stays_at_rest():
    o = Oracle(pair().to_state())
    o.step(500)
    st = o.read()
    assert numpy.all(st.get("vel", 2, 4)[:, :3] == 0.0)
    assert numpy.all(st.get("angvel", 2, 3) == 0.0)


def test_normal_oscillation():
    """Stretching at 1e-4 m/s: the pair oscillates along x at
    w = sqrt(2 k_n A / m) with amplitude v / w; energy is conserved and
    nothing moves off the bond axis."""
    v = 1e-4
    s = pair(vel1=(v, 0.0, 0.0))
    m = RHO * 4.0 / 3.0 * math.pi * R ** 3
    w_n = math.sqrt(2.0 * KN * math.pi * R ** 2 / m)
    dt = float(s.time_dt[0])
    o = Oracle(s.to_state())
    e0 = 0.5 * m * v * v
    sep = []
    steps_quarter = int(round(0.5 * math.pi / w_n / dt))
    for k in range(4 * steps_quarter // 10):
        o.step(10)
        if True:
            st = o.read()
            x = st.get("x", 2, 4)
            sep.append(x[1, 0] - x[0, 0])
            kin, rot, pot = energies(st)
            assert abs(kin + rot + pot - e0) < 1e-3 * e0
    st = o.read()
    assert numpy.all(st.get("vel", 2, 4)[:, 1:3] == 0.0)
    assert numpy.all(st.get("angvel", 2, 3) == 0.0)
    ext = numpy.array(sep) - (2 * R + 0.1 * R)
    assert abs(ext.max() - v / w_n) < 2e-3 * v / w_n
    # the maximum extension comes a quarter period after the start
    k_max = (int(numpy.argmax(ext)) + 1) * 10
    assert abs(k_max - steps_quarter) <= 0.01 * steps_quarter + 10


def test_twist_oscillation():
    """Spinning grain 1 about the bond axis: the twist spring passes the spin
    on (both spin +x), the relative twist oscillates at sqrt(2 k_t J / I)
    and energy is conserved."""
    wv = 1e-2
    s = pair(angvel1=(wv, 0.0, 0.0))
    m = RHO * 4.0 / 3.0 * math.pi * R ** 3
    I_p = 0.4 * m * R * R
    J = 2.0 * 0.25 * math.pi * R ** 4
    w_t = math.sqrt(2.0 * KT * J / I_p)
    dt = float(s.time_dt[0])
    o = Oracle(s.to_state())
    e0 = 0.5 * I_p * wv * wv
    quarter = int(round(0.5 * math.pi / w_t / dt))
    o.step(quarter)
    st = o.read()
    w = st.get("angvel", 2, 3)
    # a quarter period in, the relative spin is zero: both turn at wv / 2
    assert abs(w[0, 0] - 0.5 * wv) < 5e-3 * wv and abs(w[1, 0] - 0.5 * wv) < 5e-3 * wv
    assert numpy.all(w[:, 1:] == 0.0)
    assert numpy.all(st.get("vel", 2, 4)[:, :3] == 0.0)
    kin, rot, pot = energies(st)
    assert abs(kin + rot + pot - e0) < 1e-3 * e0


@pytest.mark.parametrize("mode", ["shear", "bend"])
def test_shear_and_bend_conserve_energy(mode):
    """Transverse velocity (shear) and opposite spins about y (bend): the bond
    carries both grains along (shear: both move +z; bend: both keep spin about
    y) and kinetic + rotational + bond energy stays constant."""
    if mode == "shear":
        s = pair(vel1=(0.0, 0.0, 1e-4))
    else:
        s = pair(angvel0=(0.0, -1e-2, 0.0), angvel1=(0.0, 1e-2, 0.0))
    o = Oracle(s.to_state())
    st0 = o.read()
    e0 = sum(energies(st0))
    for _ in range(20):
        o.step(500)
        assert abs(sum(energies(o.read())) - e0) < 1e-3 * e0
    st = o.read()
    if mode == "shear":
        v = st.get("vel", 2, 4)
        assert v[0, 2] > 0.0 and v[1, 2] > 0.0
        assert numpy.all(v[:, 1] == 0.0)
    else:
        w = st.get("angvel", 2, 3)
        assert w[0, 1] != 0.0 and w[1, 1] != 0.0
