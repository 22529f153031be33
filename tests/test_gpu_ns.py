"""The Navier-Stokes fluid step (sph_ns_*, cfd_solver = 0, no grains) on
the GPU against the CPU oracle (oracle/ns_oracle.c): bit-identical cell
velocities, pressures, face velocities, epsilon and Jacobi counts after
every step, over the boundary combinations, projection variants (beta),
relaxation, smoothing, pressure modulation, ndem > 1, a solve that reaches
maxiter, and a read between steps (which refreshes the host copy of p the
step reads)."""
import numpy
import pytest

import sphere_amd
from sphere_amd import _abi
from _ns_oracle import NsOracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def _case(shape, seed, amp=0.02):
    rng = numpy.random.default_rng(seed)
    nx, ny, nz = shape
    X, Y, Z = numpy.meshgrid(*[(numpy.arange(n) + 0.5) / n for n in shape], indexing="ij")
    v = numpy.zeros(shape + (3,))
    v[..., 0] = amp * numpy.sin(2 * numpy.pi * Y) + 0.3 * amp * rng.standard_normal(shape)
    v[..., 1] = amp * numpy.cos(2 * numpy.pi * X) * numpy.sin(numpy.pi * Z)
    v[..., 2] = 0.5 * amp * numpy.sin(2 * numpy.pi * X) + 0.3 * amp * rng.standard_normal(shape)
    p = 500.0 + 100.0 * rng.standard_normal(shape)
    return v, p


CASES = [
    # bc_bot, bc_top, extra parameters
    (1, 0, dict()),
    (0, 0, dict(beta=0.4)),
    (2, 0, dict(gamma=0.25, theta=0.8)),
    (1, 1, dict(maxiter=35)),
    (2, 2, dict(mu=0.0)),
    (0, 1, dict(ndem=2)),
    (1, 2, dict(p_mod_A=80.0, p_mod_f=40.0, p_mod_phi=0.3, beta=0.2)),
    (1, 0, dict(p_mod_A=50.0, p_mod_f=25.0, g=(0.3, -0.2, -9.81))),
]


@pytest.mark.parametrize("k", range(len(CASES)))
def test_ns_matches_oracle(torch, k):
    bb, bt, extra = CASES[k]
    shape = (9, 7, 10)
    L = (0.18, 0.21, 0.15)
    kw = dict(mu=1.0e-3, rho_f=1000.0, g=(0.0, 0.0, -9.81), tolerance=1e-9, maxiter=5000,
              bc_bot=bb, bc_top=bt, t0=0.125)
    kw.update(extra)
    P = _abi.ns_params(*shape, L, 2.0e-4, **kw)
    v, p = _case(shape, 10 + k)
    print("ns case", k, bb, bt, extra, flush=True)
    gpu = sphere_amd.NavierStokes(P, v, p)
    ora = NsOracle(P, v, p)
    for step in range(6):
        gpu.step(1)
        ora.step(1)
        sg, so = gpu.stats(), ora.stats()
        assert (sg.solves, sg.iterations, sg.nonconverged, sg.last_nijac) == \
            (so.solves, so.iterations, so.nonconverged, so.last_nijac), step
        assert sg.last_max_norm == so.last_max_norm, step
        for c in range(3):
            assert numpy.array_equal(gpu.field(c), ora.field(c)), (step, "faces", c)
        assert numpy.array_equal(gpu.field(3), ora.field(NsOracle.EPS)), (step, "epsilon")
        if step in (2, 4):
            # a read refreshes the host copy of p on both sides
            a, b = gpu.read(), ora.read()
            for x, y in zip(a, b):
                assert numpy.array_equal(x, y), step
    a, b = gpu.read(), ora.read()
    for x, y in zip(a, b):
        assert numpy.array_equal(x, y)
    assert numpy.all(numpy.isfinite(a[0])) and numpy.all(numpy.isfinite(a[1]))
    if extra.get("maxiter") == 35:
        assert sg.nonconverged > 0
    else:
        assert sg.nonconverged == 0 and sg.iterations > sg.solves
    gpu.close()


def test_ns_hydrostatic_balance(torch):
    """12 x 12 x 24 from rest under gravity (closed bottom, fixed top): the
    GPU equals the oracle bit for bit and reaches hydrostatic balance."""
    shape = (12, 12, 24)
    nz = shape[2]
    L = (0.3, 0.3, 0.6)
    P = _abi.ns_params(*shape, L, 1.0e-4, mu=1.0e-3, g=(0.0, 0.0, -9.81), bc_bot=1, bc_top=0,
                       tolerance=1e-24, maxiter=400000)
    v = numpy.zeros(shape + (3,))
    p = numpy.zeros(shape)
    print("ns hydrostatic 12x12x24", flush=True)
    gpu = sphere_amd.NavierStokes(P, v, p)
    gpu.step(1)
    ora = NsOracle(P, v, p)
    ora.step(1)
    a, b = gpu.read(), ora.read()
    for x, y in zip(a, b):
        assert numpy.array_equal(x, y)
    assert gpu.stats().iterations == ora.stats().iterations
    dz = L[2] / nz
    exp = 1000.0 * 9.81 * dz * (nz - 1 - numpy.arange(nz))
    assert numpy.max(numpy.abs(a[1] - exp)) <= 1e-8 * exp.max()
    assert numpy.max(numpy.abs(a[0])) <= 1e-8 * 9.81 * 1e-4
    gpu.close()


def test_ns_consolidation_grid(torch):
    """The consolidation box's grid, 48 x 48 x 96: two steps from a random
    velocity field, bit for bit (the solves stop at maxiter = 600)."""
    shape = (48, 48, 96)
    L = (1.1088, 1.1088, 2.2176)
    P = _abi.ns_params(*shape, L, 1.0e-4, mu=1.0e-3, g=(0.0, 0.0, -9.81), bc_bot=1, bc_top=0,
                       tolerance=1e-14, maxiter=600)
    v, p = _case(shape, 77)
    print("ns 48x48x96", flush=True)
    gpu = sphere_amd.NavierStokes(P, v, p)
    ora = NsOracle(P, v, p)
    gpu.step(2)
    ora.step(2)
    a, b = gpu.read(), ora.read()
    for x, y in zip(a, b):
        assert numpy.array_equal(x, y)
    for c in range(3):
        assert numpy.array_equal(gpu.field(c), ora.field(c))
    sg, so = gpu.stats(), ora.stats()
    assert (sg.iterations, sg.nonconverged) == (so.iterations, so.nonconverged)
    gpu.close()


def test_ns_rejects_bad_parameters():
    P = _abi.ns_params(2, 8, 8, (0.1, 0.1, 0.1), 1e-4)
    with pytest.raises(sphere_amd.SphereError):
        sphere_amd.NavierStokes(P, numpy.zeros((2, 8, 8, 3)), numpy.zeros((2, 8, 8)))


def test_ns_from_sim_and_back(torch, tmp_path):
    """A sim's Navier-Stokes fluid steps on the GPU and goes back into the
    sim and its v2.15 file, equal to the oracle's result."""
    s = sphere_amd.sim(sid="nscolumn", np=0, fluid=True, cfd_solver=0)
    s.defineWorldBoundaries([0.2, 0.15, 0.3], dx=0.025)
    s.initFluid(mu=1e-3, cfd_solver=0)
    s.g[2] = -9.81
    s.setFluidBottomNoFlow()
    s.initTemporal(total=1e-2, dt=1e-4)
    s.setTolerance(1e-12)
    s.setMaxIterations(100000)
    rng = numpy.random.default_rng(4)
    s.v_f[:] = 0.01 * rng.standard_normal(s.v_f.shape)
    P = _abi.ns_params_from_sim(s)
    ora = NsOracle(P, s.v_f, s.p_f)
    ora.step(3)
    print("ns from sim", flush=True)
    ns = sphere_amd.NavierStokes.from_sim(s)
    ns.step(3)
    ns.to_sim(s)
    ns.close()
    v, p, phi, dphi = ora.read()
    assert numpy.array_equal(s.v_f, v) and numpy.array_equal(s.p_f, p)
    assert numpy.array_equal(s.phi, phi)
    s.writebin(folder=str(tmp_path), verbose=False)
    r = sphere_amd.sim(fluid=True)
    r.readbin(str(tmp_path / "nscolumn.bin"), verbose=False)
    assert numpy.array_equal(r.v_f, v) and numpy.array_equal(r.p_f, p)


@pytest.fixture
def rundir(tmp_path):
    import os
    for d in ("input", "output", "work"):
        (tmp_path / d).mkdir()
    old = os.getcwd()
    os.chdir(tmp_path / "work")
    yield tmp_path
    os.chdir(old)


def _schedule(dt, total, file_dt):
    t, clock, n, outs = 0.0, 0.0, 0, []
    while t <= total:
        t += dt
        clock += dt
        n += 1
        if clock >= file_dt:
            outs.append(n)
            clock = 0.0
    return n, outs


def test_sim_run_neumann_gravity(torch, rundir):
    """tests/cfd_tests_neumann.py:44-75 through sim.run(): a 3 x 3 x 3
    water box under gravity over a no-flow bottom reaches the hydrostatic
    pressure (tolerance 1e2 Pa, as there) with |v_z| < 1e-4; the last output
    equals the oracle stepped to the same step count."""
    orig = sphere_amd.sim("neumann", fluid=True)
    orig.defineWorldBoundaries([0.3, 0.3, 0.3], dx=0.1)
    orig.initFluid(mu=8.9e-4)
    orig.initTemporal(total=0.05, file_dt=0.005, dt=1.0e-4)
    orig.g[2] = -10.0
    orig.bc_bot[0] = 1
    print("sim.run neumann", flush=True)
    assert orig.run(verbose=False) == 0
    nsteps, outs = _schedule(1.0e-4, 0.05, 0.005)
    py = sphere_amd.sim(sid=orig.sid, fluid=True)
    py.readlast(verbose=False)
    assert int(py.time_step_count[0]) == len(outs)
    ideal = numpy.linspace(orig.p_f[0, 0, 0] + (orig.L[2] - orig.L[2] / orig.num[2]) * orig.rho_f[0] *
                           numpy.abs(orig.g[2]), orig.p_f[0, 0, -1], orig.num[2])
    assert numpy.max(numpy.abs(ideal - py.p_f[0, 0, :])) <= 1.0e2
    assert numpy.all(numpy.abs(py.v_f[:, :, :, 2]) < 1.0e-4)
    # the oracle to the last output's step count
    P = _abi.ns_params_from_sim(orig)
    ora = NsOracle(P, orig.v_f, orig.p_f)
    done = 0
    for n_out in outs:
        ora.step(n_out - done)
        done = n_out
        ora.read()  # the output transfer refreshes the host copy of p
    v, p, _, _ = ora.read()
    assert numpy.array_equal(py.p_f, p) and numpy.array_equal(py.v_f, v)
    conv = numpy.loadtxt("../output/neumann-conv.log")
    assert conv.ndim == 2 and conv.shape[1] == 2
    assert numpy.array_equal(conv[:, 0], numpy.arange(1, nsteps + 1, 10)[:len(conv)])
    st = numpy.loadtxt("../output/neumann.status.dat")
    assert int(st[2]) == len(outs)


def test_sim_run_cfd_simple(torch, rundir):
    """tests/cfd_simple.py through sim.run(): 5 x 5 x 5 cells, 1 s at
    dt = 1e-3 with gravity over a no-flow bottom; the pressure profile of
    the last output within 1e2 Pa of the hydrostatic one.  (That script's
    ideal starts at L rho |g| = 3000 Pa, where cfd_tests_neumann.py uses
    (L - dz) rho |g|: the cell-centred hydrostatic pressure of the bottom
    layer, 2400 Pa here, which the oracle reaches too.)"""
    orig = sphere_amd.sim("cfd_simple", fluid=True)
    orig.defineWorldBoundaries([0.3, 0.3, 0.3], dx=0.06)
    orig.initFluid(mu=8.9e-4)
    orig.initTemporal(total=1.0, file_dt=0.1, dt=1.0e-3)
    orig.bc_bot[0] = 1
    orig.g[2] = -10.0
    print("sim.run cfd_simple", flush=True)
    assert orig.run(verbose=False) == 0
    py = sphere_amd.sim(sid=orig.sid, fluid=True)
    py.readlast(verbose=False)
    ideal = numpy.linspace(orig.p_f[0, 0, 0] + (orig.L[2] - orig.L[2] / orig.num[2]) * orig.rho_f[0] *
                           numpy.abs(orig.g[2]), orig.p_f[0, 0, -1], orig.num[2])
    assert numpy.max(numpy.abs(ideal - py.p_f[0, 0, :])) <= 1.0e2
    assert numpy.all(numpy.isfinite(py.v_f))


def _grains(n, L, seed, rmax=0.02):
    rng = numpy.random.default_rng(seed)
    x = numpy.column_stack([rng.uniform(0, L[0], n), rng.uniform(0, L[1], n),
                            rng.uniform(0.02 * L[2], 0.98 * L[2], n), rng.uniform(0.3 * rmax, rmax, n)])
    vel = 0.01 * rng.standard_normal((n, 3))
    return x, vel


@pytest.mark.parametrize("periodic", [1, 2, 0])
def test_ns_grains_match_oracle_inviscid(torch, periodic):
    """Grains held in a fluid with mu = 0 (the velocity prediction leaves
    out their drag there): porosities, porosity changes, the fluid and the
    pressure and viscous forces on the grains bit-identical to the oracle;
    the drag (pow, exp and log10 of the device's math library against the
    host's) within 1e-12 relative."""
    shape = (10, 8, 12)
    L = (0.3, 0.24, 0.36)
    P = _abi.ns_params(*shape, L, 1.0e-4, mu=0.0, g=(0.0, 0.0, -9.81), bc_bot=1, bc_top=0,
                       tolerance=1e-10, maxiter=20000)
    v, p = _case(shape, 21)
    x, vel = _grains(150, L, 5 + periodic)
    print("ns grains inviscid", periodic, flush=True)
    gpu = sphere_amd.NavierStokes(P, v, p)
    ora = NsOracle(P, v, p)
    gpu.set_grains(x, vel, periodic=periodic)
    ora.set_grains(x, vel, periodic=periodic)
    for step in range(4):
        gpu.step(1)
        ora.step(1)
        a, b = gpu.read(), ora.read()
        for u, w in zip(a, b):
            assert numpy.array_equal(u, w), step
        fg, fo = gpu.grain_forces(), ora.grain_forces()
        assert numpy.array_equal(fg[1], fo[1]) and numpy.array_equal(fg[2], fo[2]), step
        assert numpy.allclose(fg[0], fo[0], rtol=1e-12, atol=1e-12 * numpy.abs(fo[0]).max()), step
    assert a[2].min() < 0.99 and numpy.abs(fo[0]).max() > 0
    gpu.close()


def test_ns_grains_coupled_close_to_oracle(torch):
    """Grains held in a viscous fluid (their drag enters the velocity
    prediction): every field within 1e-9 of its largest magnitude of the
    oracle's after four steps, the porosities bit-identical."""
    shape = (10, 8, 12)
    L = (0.3, 0.24, 0.36)
    P = _abi.ns_params(*shape, L, 1.0e-4, mu=1.0e-3, g=(0.0, 0.0, -9.81), bc_bot=1, bc_top=0,
                       tolerance=1e-12, maxiter=20000)
    v, p = _case(shape, 22)
    x, vel = _grains(150, L, 9)
    print("ns grains coupled", flush=True)
    gpu = sphere_amd.NavierStokes(P, v, p)
    ora = NsOracle(P, v, p)
    gpu.set_grains(x, vel)
    ora.set_grains(x, vel)
    gpu.step(4)
    ora.step(4)
    a, b = gpu.read(), ora.read()
    assert numpy.array_equal(a[2], b[2])
    for u, w in zip(a[:2], b[:2]):
        assert numpy.max(numpy.abs(u - w)) <= 1e-9 * numpy.abs(w).max()
    fg, fo = gpu.grain_forces(), ora.grain_forces()
    for u, w in zip(fg, fo):
        assert numpy.max(numpy.abs(u - w)) <= 1e-9 * numpy.abs(w).max()
    sg, so = gpu.stats(), ora.stats()
    print("iterations", sg.iterations, so.iterations, flush=True)
    assert abs(int(sg.iterations) - int(so.iterations)) <= 10 * sg.solves
    gpu.close()


def test_ns_grain_outside_refused(torch):
    P = _abi.ns_params(4, 4, 4, (0.1, 0.1, 0.1), 1e-4)
    ns = sphere_amd.NavierStokes(P, numpy.zeros((4, 4, 4, 3)), numpy.zeros((4, 4, 4)))
    with pytest.raises(sphere_amd.SphereError):
        ns.set_grains([[0.05, 0.05, 0.2, 0.01]], [[0, 0, 0]])
    ns.close()


def test_ns_fluid_moves_grains(torch):
    """A Navier-Stokes fluid attached to a context (sph_ns_attach): 512
    grains at rest on a loose periodic lattice in water flowing along x
    (0.01 m/s with a 0.002 m/s random part); each step runs the fluid
    between the grains' contact forces and the integration and adds its
    force on them.  Against the oracle with the same coupling
    (orc_attach_ns) over 20 steps: positions, velocities, the fluid and the
    forces on the grains within 1e-9 of their largest magnitudes (the drag
    uses the device's pow, exp and log10), porosities within 1e-12.  (With
    the DEM time step, the porosity-change term rho dphi / (dt^2 phi) of the
    reference's pressure equation is large for fast grains, so the grains
    start at rest here.)"""
    from _cases import dense_case
    from _oracle import Oracle
    s = dense_case(n=512, seed=3, periodic=1, vmag=0.0, wmag=0.0, gravity=False, spacing_factor=1.1)
    st = s.to_state()
    nx, ny, nz = (int(v) for v in s.num)
    P = _abi.ns_params(nx, ny, nz, s.L, float(s.time_dt[0]), mu=1.0e-3, rho_f=1000.0, g=s.g,
                       bc_bot=1, bc_top=0, tolerance=1e-10, maxiter=100000,
                       t0=float(s.time_current[0]))
    rng = numpy.random.default_rng(5)
    v = numpy.zeros((nx, ny, nz, 3))
    v[..., 0] = 0.01
    v += 0.002 * rng.standard_normal(v.shape)
    p = numpy.zeros((nx, ny, nz))
    print("ns fluid moving grains", (nx, ny, nz), flush=True)
    e = sphere_amd.Engine(0)
    e.load(st)
    ns = sphere_amd.NavierStokes(P, v, p)
    e.attach_fluid(ns)
    o = Oracle(st)
    nso = NsOracle(P, v, p)
    o.attach_ns(nso)
    for k in range(4):
        e.step(5)
        o.step(5)
        ge, go = e.read(), o.read()
        for name, w in (("x", 4), ("vel", 4)):
            a, b = ge.get(name, 512, w), go.get(name, 512, w)
            assert numpy.max(numpy.abs(a - b)) <= 1e-9 * numpy.abs(b).max(), (k, name)
        fa, fb = ns.read(), nso.read()
        for a, b in zip(fa[:2], fb[:2]):
            assert numpy.max(numpy.abs(a - b)) <= 1e-9 * max(numpy.abs(b).max(), 1e-30), k
        assert numpy.max(numpy.abs(fa[2] - fb[2])) <= 1e-12, k
        for a, b in zip(ns.grain_forces(), nso.grain_forces()):
            assert numpy.max(numpy.abs(a - b)) <= 1e-9 * max(numpy.abs(b).max(), 1e-30), k
    assert fa[2].min() < 0.9
    assert numpy.abs(ns.grain_forces()[0]).max() > 0
    assert numpy.abs(ge.get("vel", 512, 4)[:, :3]).max() > 0  # the fluid moved the grains
    e.attach_fluid(None)
    with pytest.raises(sphere_amd.SphereError):
        e.attach_fluid(sphere_amd.NavierStokes(_abi.ns_params(3, 3, 3, (0.1, 0.1, 0.1), 1e-4),
                                               numpy.zeros((3, 3, 3, 3)), numpy.zeros((3, 3, 3))))
    e.close()
    ns.close()


def test_ns_fluid_with_top_wall(torch):
    """The attached fluid under a top wall (wmode 1): at each solve the
    layer holding the wall and the one above take the top pressure and that
    minus rho |g| dz (setNSepsilonAtTopWall), the Jacobi iteration holds
    them so, and the predicted and corrected velocities vanish from the
    wall's layer up (src/navierstokes.cuh:448-516, 2443-2444, 2758-2765,
    3013-3014).  Within 1e-9 of the oracle over 10 steps; no flow at or
    above the wall."""
    from _cases import dense_case
    from _oracle import Oracle
    s = dense_case(n=343, seed=4, periodic=1, nw=1, vmag=0.0, wmag=0.0, gravity=True, spacing_factor=1.1)
    st = s.to_state()
    nx, ny, nz = (int(v) for v in s.num)
    iz = int(s.w_x[0] / (s.L[2] / nz))
    assert iz < nz - 1
    P = _abi.ns_params(nx, ny, nz, s.L, float(s.time_dt[0]), mu=1.0e-3, rho_f=1000.0, g=s.g,
                       bc_bot=1, bc_top=0, tolerance=1e-10, maxiter=100000,
                       t0=float(s.time_current[0]))
    rng = numpy.random.default_rng(6)
    v = 0.002 * rng.standard_normal((nx, ny, nz, 3))
    p = numpy.zeros((nx, ny, nz))
    print("ns fluid under a top wall, wall layer", iz, "of", nz, flush=True)
    e = sphere_amd.Engine(0)
    e.load(st)
    ns = sphere_amd.NavierStokes(P, v, p)
    e.attach_fluid(ns)
    o = Oracle(st)
    nso = NsOracle(P, v, p)
    o.attach_ns(nso)
    e.step(10)
    o.step(10)
    ge, go = e.read(), o.read()
    for name in ("x", "vel"):
        a, b = ge.get(name, 343, 4), go.get(name, 343, 4)
        assert numpy.max(numpy.abs(a - b)) <= 1e-9 * numpy.abs(b).max(), name
    fa, fb = ns.read(), nso.read()
    for a, b in zip(fa[:2], fb[:2]):
        assert numpy.max(numpy.abs(a - b)) <= 1e-9 * max(numpy.abs(b).max(), 1e-30)
    faces = [ns.field(k).reshape(nz + 3, ny + 3, nx + 3)[1:nz + 1, 1:ny + 1, 1:nx + 1] for k in range(3)]
    for f in faces:
        assert numpy.all(f[iz:] == 0.0)
    assert numpy.abs(faces[2][:iz]).max() > 0
    e.close()
    ns.close()


def test_sim_run_fluid_with_grains(torch, rundir):
    """sim.run() of a file with a Navier-Stokes fluid, grains and a top
    wall: the grains step in a context with the fluid attached, output
    files at every file interval carry both; the last one within 1e-9 of
    the oracle coupled the same way and stepped to the same step count."""
    from _cases import dense_case
    from _oracle import Oracle
    s = dense_case(n=343, seed=7, periodic=1, nw=1, vmag=0.0, wmag=0.0, gravity=True, spacing_factor=1.1)
    s.sid = "nsgrains"
    dt = float(s.time_dt[0])
    s.initTemporal(total=40 * dt, file_dt=10 * dt, dt=dt)
    s.initFluid(mu=1.0e-3, cfd_solver=0)
    s.bc_bot[0] = 1
    s.setTolerance(1e-10)
    s.setMaxIterations(100000)
    rng = numpy.random.default_rng(8)
    s.v_f[:] = 0.002 * rng.standard_normal(s.v_f.shape)
    P = _abi.ns_params_from_sim(s, grains=True)
    s.fluid = False
    st = s.to_state()
    s.fluid = True
    print("sim.run fluid with grains", flush=True)
    assert s.run(verbose=False) == 0
    nsteps, outs = _schedule(dt, 40 * dt, 10 * dt)
    py = sphere_amd.sim(sid="nsgrains", fluid=True)
    py.readlast(verbose=False)
    assert int(py.time_step_count[0]) == len(outs) and int(py.cfd_solver[0]) == 0
    o = Oracle(st)
    nso = NsOracle(P, s.v_f, s.p_f)
    o.attach_ns(nso)
    o.step(outs[-1])
    go = o.read()
    n = int(s.np[0])
    xo = go.get("x", n, 4)
    xg = numpy.column_stack([py.x, py.radius])
    assert numpy.max(numpy.abs(xg - xo)) <= 1e-9 * numpy.abs(xo).max()
    v, p, phi, _ = nso.read()
    assert numpy.max(numpy.abs(py.p_f - p)) <= 1e-9 * max(numpy.abs(p).max(), 1e-30)
    assert numpy.max(numpy.abs(py.phi - phi)) <= 1e-12
    f_d = nso.grain_forces()[0]
    assert numpy.max(numpy.abs(py.f_d - f_d)) <= 1e-9 * max(numpy.abs(f_d).max(), 1e-30)
