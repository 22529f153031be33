"""GPU parity: the HIP library through its C ABI against the CPU oracle.

Bar: cell IDs, sort permutation, cellStart/cellEnd and contact index lists
bit-exact; positions, velocities, forces and the rest of the FP64 state
within a normwise relative difference of 1e-10 per step (the kernels keep the
reference's operation order without FMA contraction, so the observed
difference is zero).
"""
import numpy
import pytest

import sphere_amd
from _cases import add_bonds, darcy_case, dense_case, settling_case
from _oracle import Oracle

pytestmark = pytest.mark.gpu

TOL_PER_STEP = 1e-10

PARTICLE_FIELDS = [("x", 4), ("vel", 4), ("angvel", 3), ("force", 3), ("torque", 3),
                   ("angpos", 3), ("xyzsum", 3), ("es_dot", 1), ("es", 1), ("ev_dot", 1),
                   ("ev", 1), ("p", 1)]


@pytest.fixture(scope="module")
def engine():
    e = sphere_amd.Engine(0)
    yield e
    e.close()


def rel_diff(a, b):
    a = numpy.asarray(a, dtype=numpy.float64)
    b = numpy.asarray(b, dtype=numpy.float64)
    scale = numpy.max(numpy.abs(b)) if b.size else 0.0
    if scale == 0.0:
        return float(numpy.max(numpy.abs(a))) if a.size else 0.0
    return float(numpy.max(numpy.abs(a - b)) / scale)


def compare_state(g, o, n, steps, fluid=False, label=""):
    report = {}
    for name, w in PARTICLE_FIELDS:
        report[name] = rel_diff(g.get(name, n, w), o.get(name, n, w))
    if fluid:
        h = o.state
        cells = h.num[0] * h.num[1] * h.num[2]
        for name, w in (("p_f", 1), ("phi", 1), ("dphi", 1), ("v_f", 3)):
            report[name] = rel_diff(g.get(name, cells, w), o.get(name, cells, w))
        report["f_p"] = rel_diff(g.get("f_p", n, 3), o.get("f_p", n, 3))
    worst = max(report.values()) if report else 0.0
    assert worst <= TOL_PER_STEP * steps, "%s: state differs: %r" % (label, report)
    for k in range(g.state.nw):
        for c in range(4):
            assert abs(g.state.w_nx[4 * k + c] - o.state.w_nx[4 * k + c]) <= \
                TOL_PER_STEP * steps * max(1.0, abs(o.state.w_nx[4 * k + c]))
            assert abs(g.state.w_mvfd[4 * k + c] - o.state.w_mvfd[4 * k + c]) <= \
                TOL_PER_STEP * steps * max(1.0, abs(o.state.w_mvfd[4 * k + c]))
    return report


def compare_sort(engine, ora):
    cg, ig, sg, eg = engine.read_sort()
    co, io, so, eo = ora.read_sort()
    assert numpy.array_equal(cg, co), "cell IDs differ"
    assert numpy.array_equal(ig, io), "sort permutation differs"
    assert numpy.array_equal(sg, so), "cellStart differs"
    nonempty = so != 0xffffffff
    assert numpy.array_equal(eg[nonempty], eo[nonempty]), "cellEnd differs"


def compare_contacts(engine, ora, steps):
    cg, dg, tg = engine.read_contacts()
    co, do, to = ora.read_contacts()
    assert numpy.array_equal(cg, co), "contact index lists differ"
    occ = co != ora.header().np
    assert numpy.array_equal(dg[occ], do[occ]), "distmod differs"
    assert rel_diff(tg, to) <= TOL_PER_STEP * steps, "delta_t differs"
    return int(numpy.count_nonzero(occ))


def run_pair(engine, s, chunks, per_chunk, fluid=False, label="", check_contacts=True):
    engine.load(s.to_state())
    ora = Oracle(s.to_state(), nc=engine.nc)
    done = 0
    ncont = 0
    for _ in range(chunks):
        engine.step(per_chunk)
        ora.step(per_chunk)
        done += per_chunk
        compare_sort(engine, ora)
        if check_contacts:
            ncont = compare_contacts(engine, ora, done)
        compare_state(engine.read(), ora.read(), int(s.np[0]), done, fluid=fluid, label=label)
    return ora, ncont


def test_settling_lattice(engine):
    s = settling_case(n=2000, seed=5, steps=100)
    run_pair(engine, s, 2, 50, label="settling")


@pytest.mark.parametrize("periodic", [0, 1, 2])
def test_dense_linear(engine, periodic):
    s = dense_case(n=729, seed=11 + periodic, model=2, periodic=periodic, gamma_n=50.0, gamma_t=20.0)
    _, nc = run_pair(engine, s, 3, 10, label="linear p%d" % periodic)
    assert nc > 100


def test_dense_hertz_mindlin(engine):
    s = dense_case(n=729, seed=21, model=3, periodic=1, E=7e9, nu=0.3, gamma_n=0.0)
    _, nc = run_pair(engine, s, 3, 10, label="hertz")
    assert nc > 100


def test_dense_hertz_64_slots():
    """64 contact slots: same contacts, forces and state as the oracle with
    64 slots, and the same slot order and forces as 32 slots (rows never hold
    more than 32 contacts here)."""
    s = dense_case(n=729, seed=22, model=3, periodic=1, E=7e9, nu=0.3)
    e64 = sphere_amd.Engine(0, nc=64)
    try:
        run_pair(e64, s, 2, 10, label="hertz nc64")
        e32 = sphere_amd.Engine(0, nc=32)
        try:
            e32.load(s.to_state())
            e64.load(s.to_state())
            e32.step(20)
            e64.step(20)
            a = e32.read()
            b = e64.read()
            n = int(s.np[0])
            for name, w in (("x", 4), ("vel", 4), ("force", 3), ("torque", 3)):
                assert numpy.array_equal(a.get(name, n, w), b.get(name, n, w)), name
            c32 = e32.read_contacts()[0].reshape(n, 32)
            c64 = e64.read_contacts()[0].reshape(n, 64)
            assert numpy.array_equal(c64[:, 32:], c32)
            assert numpy.all(c64[:, :32] == n)
        finally:
            e32.close()
    finally:
        e64.close()


def test_dense_model1_capillary(engine):
    s = dense_case(n=512, seed=31, model=1, periodic=0, gamma_t=30.0, kappa=0.05)
    run_pair(engine, s, 2, 10, label="model1", check_contacts=False)


def test_bonds(engine):
    s = dense_case(n=512, seed=41, model=2, periodic=1, spacing_factor=0.99)
    add_bonds(s, 0.3, seed=5)
    assert s.nb0[0] > 50
    run_pair(engine, s, 2, 10, label="bonds")


def test_top_wall_sigma0(engine):
    s = dense_case(n=512, seed=51, model=2, periodic=1, nw=1, gravity=True)
    run_pair(engine, s, 3, 10, label="wall")


def test_shear_fixvel(engine):
    s = dense_case(n=729, seed=61, model=3, periodic=1, E=7e9, nu=0.3, nw=1)
    zmin = numpy.min(s.x[:, 2])
    zmax = numpy.max(s.x[:, 2])
    top = s.x[:, 2] > zmax - 0.01
    bot = s.x[:, 2] < zmin + 0.01
    s.fixvel[top] = 1
    s.fixvel[bot] = 1
    s.vel[top, 0] = 0.01
    s.vel[bot, :2] = 0.0
    run_pair(engine, s, 2, 10, label="shear")


def test_darcy_coupled(engine):
    s = darcy_case(n=300, seed=3, nxyz=(6, 6, 12), tolerance=1e-8)
    ora, _ = run_pair(engine, s, 2, 3, fluid=True, label="darcy")
    sg = engine.darcy_stats()
    so = ora.darcy_stats()
    assert sg.sweeps == so.sweeps and sg.last_nijac == so.last_nijac


def test_darcy_wall_neumann_top(engine):
    s = darcy_case(n=250, seed=4, nxyz=(5, 6, 10), bc_top=1, bc_bot=1, wall=True, tolerance=1e-7)
    ora, _ = run_pair(engine, s, 2, 2, fluid=True, label="darcy wall")
    assert engine.darcy_stats().sweeps == ora.darcy_stats().sweeps
