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
from _cases import add_bonds, darcy_case, darcy_empty, dense_case, fluid_particle, settling_case
from _oracle import Oracle, OracleError
from _parity import (PARTICLE_FIELDS, TOL_PER_STEP, compare_contacts, compare_sort,  # noqa: F401
                     compare_state, rel_diff, run_pair)

pytestmark = pytest.mark.gpu

@pytest.fixture(scope="module")
def engine():
    """Default engine: neighbour lists on (sorted arrays current only on
    rebuild steps)."""
    e = sphere_amd.Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def engine_cells():
    """The reference's pipeline: sort and 27-cell search every step."""
    e = sphere_amd.Engine(0, skin=0.0)
    yield e
    e.close()


@pytest.fixture(params=["lists", "cells"])
def any_engine(request, engine, engine_cells):
    return engine if request.param == "lists" else engine_cells


def test_settling_lattice(any_engine):
    s = settling_case(n=2000, seed=5, steps=100)
    run_pair(any_engine, s, 2, 50, label="settling")


@pytest.mark.parametrize("periodic", [0, 1, 2])
def test_dense_linear(any_engine, periodic):
    s = dense_case(n=729, seed=11 + periodic, model=2, periodic=periodic, gamma_n=50.0, gamma_t=20.0)
    _, nc = run_pair(any_engine, s, 3, 10, label="linear p%d" % periodic)
    assert nc > 100


def test_dense_hertz_mindlin(any_engine):
    s = dense_case(n=729, seed=21, model=3, periodic=1, E=7e9, nu=0.3, gamma_n=0.0)
    _, nc = run_pair(any_engine, s, 3, 10, label="hertz")
    assert nc > 100


def test_dense_hertz_rolling_cohesion(engine):
    """Rolling resistance and linear cohesion on (the laws' general branches;
    with gamma_r = mu_r = kappa = 0 the library uses closed forms of the zero
    terms)."""
    s = dense_case(n=729, seed=23, model=3, periodic=1, E=7e9, nu=0.3)
    s.gamma_r[0] = 1e-3
    s.mu_r[0] = 0.1
    s.kappa[0] = 10.0
    _, nc = run_pair(engine, s, 3, 10, label="hertz rolling")
    assert nc > 100


def test_lists_enabled_on_test_grids(engine):
    """Cells narrower than a diameter plus the skin (the dense and settling
    grids) are searched two cells deep at a rebuild, wider ones one deep."""
    engine.load(dense_case(n=729, seed=21, model=3, periodic=1, E=7e9).to_state())
    st = engine.neighbor_stats()
    assert st.enabled == 1 and st.range == 2 and st.cap == 128
    engine.load(settling_case(n=2000, seed=5, steps=100).to_state())
    assert engine.neighbor_stats().range == 2
    s = crowded_case(12)
    s.num[:] = 3  # cells of 0.067 m around grains of up to 0.03 m
    engine.load(s.to_state())
    assert engine.neighbor_stats().range == 1
    s = dense_case(n=512, seed=31, model=1, periodic=0, kappa=0.05)
    engine.load(s.to_state())
    assert engine.neighbor_stats().enabled == 0  # model 1 searches cells


def test_lists_frequent_rebuilds():
    """A skin of 0.2% of the largest radius: the lists are rebuilt every few
    steps, and every step still matches the cell search of the oracle."""
    s = dense_case(n=729, seed=71, model=3, periodic=1, E=7e9, nu=0.3, vmag=0.5)
    e = sphere_amd.Engine(0, skin=0.002)
    try:
        run_pair(e, s, 4, 5, label="frequent rebuilds")
        assert e.neighbor_stats().rebuilds >= 4
    finally:
        e.close()


@pytest.mark.parametrize("n", [729, 64000])
def test_rebuild_kernel_fused_equals_separate(n):
    """The list rebuild as one kernel with grid barriers (729 grains: 3
    blocks; 64,000 grains: 250 blocks) and as separate sort, reorder and list
    launches give bit-identical states, contact rows and sort arrays over
    steps with frequent rebuilds."""
    s = dense_case(n=n, seed=73, model=3, periodic=1, E=7e9, nu=0.3, vmag=0.5)
    out = []
    for fused in (True, False):
        e = sphere_amd.Engine(0, skin=0.002, fused_rebuild=fused)
        try:
            e.load(s.to_state())
            e.step(12)
            st = e.neighbor_stats()
            out.append((e.read(), e.read_contacts(), e.read_sort(), st.rebuilds))
        finally:
            e.close()
    (ga, ca, sa, ra), (gb, cb, sb, rb) = out
    assert ra == rb and ra >= 3, (ra, rb)
    for name, w in PARTICLE_FIELDS:
        assert numpy.array_equal(ga.get(name, n, w), gb.get(name, n, w)), name
    for u, v in zip(ca, cb):
        assert numpy.array_equal(u, v)
    for u, v in zip(sa, sb):
        assert numpy.array_equal(u, v)


@pytest.mark.parametrize("skin", [0.0, 0.1])
def test_write_contacts_matches_oracle(skin):
    """Contact rows written through the ABI (partners, periodic images and
    tangential springs of a box that ran 6 steps) read back unchanged, and
    stepping from them matches the oracle given the same rows."""
    s = dense_case(n=729, seed=75, model=3, periodic=1, E=7e9, nu=0.3, vmag=0.5)
    a = sphere_amd.Engine(0)
    try:
        a.load(s.to_state())
        a.step(6)
        g = a.read()
        c, dm, dt = a.read_contacts()
    finally:
        a.close()
    n = int(s.np[0])
    occ = c != n
    assert occ.sum() > 100
    assert numpy.any(dm[occ] != 0.0), "no contact across the periodic boundary"
    assert numpy.any(dt[occ] != 0.0)
    e = sphere_amd.Engine(0, skin=skin)
    try:
        e.load(g)
        e.write_contacts(c, dm, dt)
        c2, dm2, dt2 = e.read_contacts()
        assert numpy.array_equal(c2, c)
        assert numpy.array_equal(dm2[occ], dm[occ])
        assert numpy.array_equal(dt2, dt)
        ora = Oracle(g, nc=32)
        ora.write_contacts(c, dm, dt)
        e.step(5)
        ora.step(5)
        compare_contacts(e, ora, 5)
        compare_state(e.read(), ora.read(), n, 5, label="written rows")
    finally:
        e.close()


def test_lists_periodic_crossing():
    """Every grain drifts at 2 m/s along x (plus noise) with a layer sitting on
    the periodic boundary: grains and contacts cross it, images change."""
    s = dense_case(n=729, seed=72, model=2, periodic=1, gamma_n=50.0)
    L0 = float(s.L[0])
    s.x[:, 0] = numpy.mod(s.x[:, 0] + 0.5 * L0 - 1e-4, L0)
    s.vel[:, 0] += 2.0
    e = sphere_amd.Engine(0, skin=0.05)
    try:
        _, nc = run_pair(e, s, 4, 5, label="crossing")
        assert nc > 100
        assert e.neighbor_stats().rebuilds >= 2
    finally:
        e.close()


@pytest.mark.parametrize("periodic", [1, 2])
def test_lists_one_cell_deep_crossing(periodic):
    """Cells wider than a diameter plus the skin: the list search takes the
    nearest periodic image without reading cells; grains drift across the
    periodic boundaries (images change) and all results match the oracle."""
    s = dense_case(n=729, seed=74 + periodic, model=3, periodic=periodic, E=7e9, nu=0.3,
                   cell_factor=2.5)
    L0 = float(s.L[0])
    s.x[:, 0] = numpy.mod(s.x[:, 0] + 0.5 * L0 - 1e-4, L0)
    s.vel[:, 0] += 2.0
    if periodic == 1:
        L1 = float(s.L[1])
        s.x[:, 1] = numpy.mod(s.x[:, 1] + 0.5 * L1 - 1e-4, L1)
        s.vel[:, 1] -= 2.0
    e = sphere_amd.Engine(0, skin=0.05)
    try:
        e.load(s.to_state())
        assert e.neighbor_stats().range == 1
        _, nc = run_pair(e, s, 4, 5, label="one deep p%d" % periodic)
        assert nc > 100
        assert e.neighbor_stats().rebuilds >= 2
    finally:
        e.close()


def test_lists_capacity_overflow():
    """Lists of 4 entries overflow: those grains fall back to the cell
    search and the lists are rebuilt every step."""
    s = dense_case(n=729, seed=73, model=3, periodic=1, E=7e9, nu=0.3)
    e = sphere_amd.Engine(0, nl_cap=4)
    try:
        run_pair(e, s, 2, 5, label="list overflow")
        st = e.neighbor_stats()
        assert st.overflow == 1 and st.rebuilds == 10
    finally:
        e.close()


def crowded_case(n_small, r_big=0.03, r_small=0.005, gap=-1e-6, v_out=0.0, arrivals=()):
    """One big grain at the centre of a box with n_small small grains on a
    Fibonacci sphere around it at surface gap `gap` (negative: overlap),
    moving radially outwards at v_out, plus small grains 1e-8 m away from
    the big one along the unit vectors `arrivals`, approaching at 0.05 m/s."""
    arrivals = numpy.asarray(arrivals, dtype=numpy.float64).reshape(-1, 3)
    n = n_small + 1 + arrivals.shape[0]
    s = sphere_amd.sim(np=n, sid="crowded", nw=0)
    L = 0.2
    c = numpy.array([L / 2, L / 2, L / 2])
    k = numpy.arange(n_small) + 0.5
    phi = numpy.arccos(1 - 2 * k / n_small)
    theta = numpy.pi * (1 + 5 ** 0.5) * k
    u = numpy.stack([numpy.cos(theta) * numpy.sin(phi), numpy.sin(theta) * numpy.sin(phi),
                     numpy.cos(phi)], axis=1)
    s.x[0] = c
    s.radius[0] = r_big
    s.x[1:n_small + 1] = c + u * (r_big + r_small + gap)
    s.radius[1:] = r_small
    s.vel[1:n_small + 1] = u * v_out
    if arrivals.shape[0]:
        s.x[n_small + 1:] = c + arrivals * (r_big + r_small + 1e-8)
        s.vel[n_small + 1:] = -arrivals * 0.05
    s.defaultParams(mu_s=0.5, mu_d=0.5, k_n=1.16e7, k_t=1.16e7)
    s.contactmodel[0] = 2
    s.defineWorldBoundaries(L=[L, L, L], dx=2.2 * r_big)
    s.g[:] = 0.0
    m = 2600.0 * 4.0 / 3.0 * numpy.pi * r_small ** 3
    s.initTemporal(total=1.0, file_dt=1.0, dt=0.05 * numpy.sqrt(m / 1.16e7))
    return s


@pytest.mark.parametrize("skin", [0.0, 0.1])
def test_slot_overflow_rows(skin):
    """40 grains touch one: more contacts than 32 slots.  Both sides report
    the overflow and leave identical slot rows (the reference's rule with
    cpos = nc and the early exit from a search cell)."""
    s = crowded_case(40)
    e = sphere_amd.Engine(0, skin=skin)
    try:
        e.load(s.to_state())
        ora = Oracle(s.to_state(), nc=32)
        with pytest.raises(sphere_amd.SphereError) as ei:
            e.step(1)
        assert ei.value.code == -7
        with pytest.raises(OracleError) as eo:
            ora.step(1)
        assert eo.value.code == -7
        cg = e.read_contacts()[0]
        co = ora.read_contacts()[0]
        assert numpy.array_equal(cg, co)
    finally:
        e.close()


@pytest.mark.parametrize("skin", [0.0, 0.1])
def test_full_row_replacement(skin):
    """32 contacts fill the row of the big grain and separate in the next
    step while 3 new grains arrive: with no empty slot the new contacts are
    written to slot nc (the reference's rule), before the stale ones go."""
    s = crowded_case(32, gap=-2e-8, v_out=0.05,
                     arrivals=[[0.6, 0.8, 0.0], [-0.8, 0.0, 0.6], [0.0, -0.6, -0.8]])
    e = sphere_amd.Engine(0, skin=skin)
    try:
        run_pair(e, s, 3, 1, label="full row")
    finally:
        e.close()


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


@pytest.mark.parametrize("mode", ["lists", "cells"])
def test_shear_stress_wall_with_bonds(engine, engine_cells, mode):
    """wmode 3 (stress-driven top layer, src/integration.cuh:116-129,
    457-484) with bonds: the shear-stress sum reads the forces after the bond
    forces, so bonds are gathered before it."""
    e = engine if mode == "lists" else engine_cells
    s = dense_case(n=729, seed=81, model=2, periodic=1, nw=1, vmag=0.0, wmag=0.0,
                   spacing_factor=0.99)
    add_bonds(s, 0.2, seed=3)
    s.shear(shear_stress=2e3)
    assert s.wmode[0] == 3 and s.nb0[0] > 20
    run_pair(e, s, 2, 10, label="wmode 3")


def test_darcy_with_bonds(engine):
    """Bonds under fluid coupling: the pressure force is added after the bond
    forces (src/device.cu:1970-1990)."""
    s = darcy_case(n=300, seed=9, nxyz=(6, 6, 12), tolerance=1e-8)
    add_bonds(s, 1.0, seed=4)
    assert s.nb0[0] > 5
    run_pair(engine, s, 2, 3, fluid=True, label="darcy bonds")


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


def test_darcy_reference_no_forcing(engine):
    """tests/cfd_tests_darcy.py:14-45 through the C ABI: p stays 0 and no
    fluid step iterates."""
    s = darcy_empty()
    engine.load(s.to_state())
    engine.step(11)
    st = engine.read()
    assert numpy.all(st.get("p_f", 48, 1) == 0.0)
    assert engine.darcy_stats().sweeps == 0 or engine.darcy_stats().last_nijac == 0


def test_darcy_reference_gradient(engine):
    """tests/cfd_tests_darcy.py:48-69: the linear steady profile, matched
    against the oracle field by field every 500 steps."""
    s = darcy_empty(p_top=1.0)
    ora, _ = run_pair(engine, s, 4, 500, fluid=True, label="darcy gradient", check_contacts=False)
    p = engine.read().get("p_f", 48, 1).reshape(4, 3, 4)
    assert numpy.max(numpy.abs(p - numpy.linspace(0.0, 1.0, 4)[:, None, None])) < 1e-3
    assert engine.darcy_stats().sweeps == ora.darcy_stats().sweeps


@pytest.mark.parametrize("test", [1, 2])
def test_darcy_reference_grain(engine, test):
    """tests/fluid_particle_interaction_darcy.py: the grain moves to low
    pressure, with the oracle's trajectory."""
    s = fluid_particle(test)
    run_pair(engine, s, 2, 200, fluid=True, label="grain %d" % test, check_contacts=False)
    assert engine.read().get("vel", 1, 4)[0, 2] < 0.0
