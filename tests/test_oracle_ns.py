"""Pins of the Navier-Stokes oracle (oracle/ns_oracle.c) by what the
mathematics fixes, independent of how the oracle writes its formulas:

* a fluid at rest stays at rest;
* hydrostatic balance: from rest under gravity with a closed bottom and a
  fixed-pressure top, one projection step leaves v = 0 and the pressure
  rho*|g|*dz per cell deeper (the predictor's gravity term, the forcing
  rho/(c_v dt) div(v_p), the Jacobi stencil along z, the Neumann ghost
  below and the velocity update must all agree); the same pressure for
  ndem = 1 and 2, for Jacobi relaxation theta = 1 and 0.7, for velocity
  scaling c_v = 1 and 0.9, and with the
  pressure-dependent projection (beta = 0.5) from the balanced state;
* the projection makes the face velocities divergence free in every
  Jacobi-solved cell (x, y and z parts of the stencil, periodic ghosts);
* the viscous term of a sine wave is the discrete eigenvalue
  mu (2 cos(k h) - 2) / h^2 times the wave;
* the advective term of a linear velocity u(x) (u^2 quadratic) is the half
  backward, half central difference d(u^2)/dx - (dx/2) (u^2)''/2;
* smoothing of the balanced profile changes only the bottom cell, by
  gamma (e1 - e0) / 6;
* the top pressure modulation p_top + A sin(2 pi f t + phi), with the
  reference's host copy of p (refreshed by a read).
"""
import numpy
import pytest

from sphere_amd import _abi
from _ns_oracle import NsOracle


def _params(shape, L, dt=1.0e-4, **kw):
    nx, ny, nz = shape
    return _abi.ns_params(nx, ny, nz, L, dt, **kw)


def _rest(shape, p0=0.0):
    nx, ny, nz = shape
    return numpy.zeros((nx, ny, nz, 3)), numpy.full((nx, ny, nz), p0)


def _faces_interior(o, k):
    """Interior faces [x, y, z] of component k (face index 0..n-1)."""
    nx, ny, nz = o.shape
    f = o.faces(k)[1:nz + 1, 1:ny + 1, 1:nx + 1]
    return f.transpose(2, 1, 0)


def test_rest_stays_at_rest():
    shape = (4, 5, 6)
    v, p = _rest(shape, 1234.5)
    o = NsOracle(_params(shape, (0.1, 0.1, 0.15), bc_bot=1, bc_top=0, tolerance=1e-12,
                         maxiter=1000), v, p)
    o.step(5)
    v1, p1, phi, dphi = o.read()
    assert numpy.all(v1 == 0.0)
    assert numpy.allclose(p1, 1234.5, rtol=1e-14, atol=0)
    assert numpy.all(phi == 1.0) and numpy.all(dphi == 0.0)


def _hydrostatic(ndem=1, theta=1.0, beta=0.0, p=None, steps=1, gamma=0.0, nz=10, c_v=1.0):
    shape = (4, 4, nz)
    L = (0.1, 0.1, 0.025 * nz)
    v, p0 = _rest(shape)
    if p is not None:
        p0 = p
    o = NsOracle(_params(shape, L, g=(0.0, 0.0, -9.81), rho_f=1000.0, bc_bot=1, bc_top=0,
                         tolerance=1e-26, maxiter=200000, ndem=ndem, theta=theta, beta=beta,
                         gamma=gamma, mu=0.0, c_v=c_v), v, p0)
    o.step(steps)
    return o


def _p_hyd(nz):
    z = numpy.arange(nz)
    return 1000.0 * 9.81 * 0.025 * (nz - 1 - z)


def test_hydrostatic_balance():
    o = _hydrostatic()
    v, p, _, _ = o.read()
    st = o.stats()
    assert st.solves == 1 and st.nonconverged == 0
    exp = _p_hyd(10)
    scale = exp.max()
    for x in range(4):
        for y in range(4):
            assert numpy.max(numpy.abs(p[x, y, :] - exp)) <= 1e-9 * scale
    # |v| against the velocity gravity gives in one step
    assert numpy.max(numpy.abs(v)) <= 1e-9 * 9.81 * 1e-4
    assert numpy.max(numpy.abs(_faces_interior(o, 2))) <= 1e-9 * 9.81 * 1e-4


@pytest.mark.parametrize("kw", [dict(ndem=2), dict(theta=0.7), dict(c_v=0.9)])
def test_hydrostatic_same_pressure(kw):
    p_ref = _hydrostatic().read()[1]
    v, p, _, _ = _hydrostatic(**kw).read()
    assert numpy.max(numpy.abs(p - p_ref)) <= 1e-9 * p_ref.max()
    assert numpy.max(numpy.abs(v)) <= 1e-9 * 9.81 * 1e-4


def test_pressure_dependent_projection_keeps_balance():
    nz = 10
    p0 = numpy.broadcast_to(_p_hyd(nz), (4, 4, nz)).copy()
    o = _hydrostatic(beta=0.5, p=p0, steps=3)
    v, p, _, _ = o.read()
    assert numpy.max(numpy.abs(p - p0)) <= 1e-9 * p0.max()
    assert numpy.max(numpy.abs(v)) <= 1e-9 * 9.81 * 1e-4


def test_projection_is_divergence_free():
    shape = (8, 6, 7)
    nx, ny, nz = shape
    L = (0.16, 0.18, 0.105)  # three different cell widths
    rng = numpy.random.default_rng(3)
    X, Y, Z = numpy.meshgrid(*[(numpy.arange(n) + 0.5) / n for n in shape], indexing="ij")
    v = numpy.zeros(shape + (3,))
    v[..., 0] = 0.01 * numpy.sin(2 * numpy.pi * Y) + 0.004 * numpy.cos(2 * numpy.pi * Z)
    v[..., 1] = 0.006 * numpy.sin(2 * numpy.pi * X) * numpy.cos(numpy.pi * Z)
    v[..., 2] = 0.005 * numpy.cos(2 * numpy.pi * X) + 0.002 * rng.standard_normal(shape)
    p = numpy.zeros(shape)
    o = NsOracle(_params(shape, L, bc_bot=1, bc_top=0, tolerance=1e-24, maxiter=400000,
                         mu=1e-3), v, p)
    o.step(1)
    assert o.stats().nonconverged == 0
    vp = [_faces_interior_pred(o, k) for k in range(3)]
    vf = [_faces_interior(o, k) for k in range(3)]
    h = [L[k] / shape[k] for k in range(3)]

    def div(F):
        d = (numpy.roll(F[0], -1, axis=0) - F[0]) / h[0] + (numpy.roll(F[1], -1, axis=1) - F[1]) / h[1]
        dz = numpy.empty_like(F[2])
        dz[:, :, :-1] = (F[2][:, :, 1:] - F[2][:, :, :-1]) / h[2]
        dz[:, :, -1] = numpy.nan
        return d + dz
    before = numpy.abs(div(vp)[:, :, :nz - 1]).max()
    after = numpy.abs(div(vf)[:, :, :nz - 1]).max()
    assert before > 1e-3
    assert after <= 1e-7 * before, (before, after)


def _faces_interior_pred(o, k):
    nx, ny, nz = o.shape
    f = o.faces(NsOracle.PRED_FACES + k)[1:nz + 1, 1:ny + 1, 1:nx + 1]
    return f.transpose(2, 1, 0)


def test_viscous_term_is_the_discrete_eigenvalue():
    shape = (8, 10, 6)
    nx, ny, nz = shape
    L = (0.2, 0.3, 0.15)  # dx != dy
    h = [L[k] / shape[k] for k in range(3)]
    mu = 2.5e-3
    X, Y, Z = numpy.meshgrid(*[numpy.arange(n) + 0.5 for n in shape], indexing="ij")
    kx, ky = 2 * numpy.pi / nx, 2 * numpy.pi / ny
    v = numpy.zeros(shape + (3,))
    v[..., 0] = 0.3 * numpy.sin(ky * Y)
    v[..., 1] = 0.2 * numpy.sin(kx * X)
    o = NsOracle(_params(shape, L, mu=mu, bc_bot=1, bc_top=0, tolerance=1e-3, maxiter=50), v,
                 numpy.zeros(shape))
    o.step(1)
    dtx = o.faces(NsOracle.DIV_TAU)[1:nz + 1, 1:ny + 1, 1:nx + 1].transpose(2, 1, 0)
    dty = o.faces(NsOracle.DIV_TAU + 1)[1:nz + 1, 1:ny + 1, 1:nx + 1].transpose(2, 1, 0)
    # face values are the means of neighbouring cell values: along x for
    # v_x (which does not vary along x) and along y for v_y
    fx = 0.3 * numpy.sin(ky * Y)
    fy = 0.2 * numpy.sin(kx * X)
    ex = mu * (2 * numpy.cos(ky) - 2) / h[1] ** 2 * fx
    ey = mu * (2 * numpy.cos(kx) - 2) / h[0] ** 2 * fy
    # z neighbours of the top and bottom faces are boundary ghosts
    sl = (slice(None), slice(None), slice(1, nz - 1))
    assert numpy.allclose(dtx[sl], ex[sl], rtol=1e-9, atol=1e-12 * numpy.abs(ex).max())
    assert numpy.allclose(dty[sl], ey[sl], rtol=1e-9, atol=1e-12 * numpy.abs(ey).max())


def test_advection_of_a_linear_velocity():
    shape = (14, 4, 5)
    nx, ny, nz = shape
    L = (0.28, 0.08, 0.1)
    dx = L[0] / nx
    alpha, beta_ = 0.5, 2.0
    xc = (numpy.arange(nx) + 0.5) * dx
    v = numpy.zeros(shape + (3,))
    v[..., 0] = (alpha + beta_ * xc)[:, None, None]
    o = NsOracle(_params(shape, L, bc_bot=1, bc_top=0, tolerance=1e-3, maxiter=50), v,
                 numpy.zeros(shape))
    o.step(1)
    d = o.cells(NsOracle.DIV_PHI_VV)[1:nz + 1, 1:ny + 1, 1:nx + 1].transpose(2, 1, 0)
    B, C = 2 * alpha * beta_, beta_ ** 2
    exp = B + 2 * C * xc - C * dx / 2
    # the periodic wrap breaks linearity within two cells of the x ends
    for x in range(3, nx - 3):
        assert numpy.allclose(d[x], exp[x], rtol=1e-9), (x, d[x].ravel()[:3], exp[x])
    dy = o.cells(NsOracle.DIV_PHI_VV + 1)[1:nz + 1, 1:ny + 1, 1:nx + 1]
    dz = o.cells(NsOracle.DIV_PHI_VV + 2)[1:nz + 1, 1:ny + 1, 1:nx + 1]
    assert numpy.all(dy == 0.0) and numpy.all(dz == 0.0)


def test_smoothing_changes_only_the_bottom_cell():
    gamma = 0.3
    o0 = _hydrostatic()
    o1 = _hydrostatic(gamma=gamma)
    e0 = o0.cells(NsOracle.EPS)[1:-1, 1:-1, 1:-1]
    e1 = o1.cells(NsOracle.EPS)[1:-1, 1:-1, 1:-1]
    scale = numpy.abs(e0).max()
    assert numpy.max(numpy.abs(e1[1:] - e0[1:])) <= 1e-9 * scale
    exp = e0[0] + gamma * (e0[1] - e0[0]) / 6.0
    assert numpy.max(numpy.abs(e1[0] - exp)) <= 1e-9 * scale


def test_top_pressure_modulation():
    shape = (4, 4, 6)
    A, f, ph = 250.0, 3.0, 0.4
    t0, dt = 0.25, 1e-3
    p0 = numpy.full(shape, 100.0)
    o = NsOracle(_params(shape, (0.1, 0.1, 0.15), dt=dt, t0=t0, bc_bot=1, bc_top=0,
                         tolerance=1e-20, maxiter=100000, p_mod_A=A, p_mod_f=f, p_mod_phi=ph),
                 *[numpy.zeros(shape + (3,)), p0])
    o.step(1)
    p1 = o.read()[1]
    assert numpy.allclose(p1[:, :, -1], 100.0 + A * numpy.sin(2 * numpy.pi * f * t0 + ph), rtol=1e-14)
    # the host copy of p was refreshed by the read: the next step
    # modulates the current top pressure
    o.step(1)
    p2 = o.read()[1]
    top = p1[0, 0, -1] + A * numpy.sin(2 * numpy.pi * f * (t0 + dt) + ph)
    assert numpy.allclose(p2[:, :, -1], top, rtol=1e-14)
    # without a read in between, the step uses the copy from before
    q = NsOracle(_params(shape, (0.1, 0.1, 0.15), dt=dt, t0=t0, bc_bot=1, bc_top=0,
                         tolerance=1e-20, maxiter=100000, p_mod_A=A, p_mod_f=f, p_mod_phi=ph),
                 *[numpy.zeros(shape + (3,)), p0])
    q.step(2)
    top_q = 100.0 + A * numpy.sin(2 * numpy.pi * f * (t0 + dt) + ph)
    assert numpy.allclose(q.read()[1][:, :, -1], top_q, rtol=1e-14)


def test_nonconvergence_is_counted():
    shape = (4, 4, 6)
    rng = numpy.random.default_rng(1)
    v = 0.01 * rng.standard_normal(shape + (3,))
    o = NsOracle(_params(shape, (0.1, 0.1, 0.15), bc_bot=1, bc_top=1, tolerance=1e-30, maxiter=25),
                 v, numpy.zeros(shape))
    o.step(2)
    st = o.stats()
    assert st.solves == 2 and st.nonconverged == 2 and st.iterations == 50 and st.last_nijac == 24


def test_params_from_a_sim():
    """A sim set up with the reference's calls (Navier-Stokes fluid, no
    grains) gives the parameters and the [x, y, z] fields the step reads:
    the oracle run from them reaches hydrostatic balance."""
    import sphere_amd
    s = sphere_amd.sim(sid="nscolumn", np=0, fluid=True, cfd_solver=0)
    s.defineWorldBoundaries([0.1, 0.1, 0.25], dx=0.025)
    s.initFluid(mu=1e-3, cfd_solver=0)
    s.g[2] = -9.81
    s.setFluidBottomNoFlow()
    s.setFluidTopFixedPressure()
    s.initTemporal(total=1e-3, dt=1e-4)
    s.setTolerance(1e-26)
    s.setMaxIterations(200000)
    P = _abi.ns_params_from_sim(s)
    assert (P.nx, P.ny, P.nz) == (4, 4, 10) and P.bc_bot == 1 and P.bc_top == 0
    assert (P.beta, P.theta, P.gamma, P.c_v, P.ndem) == (0.0, 1.0, 0.0, 1.0, 1)
    o = NsOracle(P, s.v_f, s.p_f)
    o.step(1)
    p = o.read()[1]
    assert numpy.max(numpy.abs(p - _p_hyd(10))) <= 1e-9 * _p_hyd(10).max()
    s2 = sphere_amd.sim(sid="x", np=10, fluid=True, cfd_solver=0)
    with pytest.raises(_abi.SphereError):
        _abi.ns_params_from_sim(s2)


# --- grains held in the fluid (porosity, interaction forces) ---

def _grain_case(shape=(6, 6, 6), L=(0.6, 0.6, 0.6), mu=1e-3, **kw):
    v, p = _rest(shape)
    return NsOracle(_params(shape, L, mu=mu, bc_bot=1, bc_top=0, tolerance=1e-3, maxiter=20, **kw),
                    v, p)


def _lens(R, r, d):
    """Volume of the intersection of spheres of radii R and r at distance d,
    by numerical integration of the cross-section areas along the axis."""
    from scipy.integrate import quad
    def area(x):  # x along the axis from the centre of R
        a2 = R * R - x * x
        b2 = r * r - (x - d) ** 2
        m = min(a2, b2)
        return numpy.pi * m if m > 0 else 0.0
    lo, hi = max(-R, d - r), min(R, d + r)
    return quad(area, lo, hi, points=[(R * R - r * r + d * d) / (2 * d)], epsabs=1e-16, epsrel=1e-12)[0]


def test_porosity_of_one_grain():
    """Porosity 1 - V_overlap / (4/3 pi R^3), R = min(dx, dy, dz) the
    sampling sphere at each cell centre: a grain wholly inside it, one
    partly inside (the lens volume by numerical integration), and the cells
    it does not reach; and dphi 0 on the first step, the change after."""
    o = _grain_case()
    R = 0.1
    r = 0.03
    # grain at the centre of cell (2, 2, 2) (centre 0.25, 0.25, 0.25)
    o.set_grains([[0.25, 0.25, 0.25, r]], [[0.0, 0.0, 0.0]])
    o.step(1)
    _, _, phi, dphi = o.read()
    Vs = 4.0 / 3.0 * numpy.pi * R ** 3
    assert phi[2, 2, 2] == pytest.approx(1.0 - (r / R) ** 3, rel=1e-13)
    d = 0.1  # neighbour cell centre at one cell width
    assert phi[3, 2, 2] == pytest.approx(1.0 - _lens(R, r, d) / Vs, rel=1e-10)
    assert phi[2, 3, 2] == phi[3, 2, 2] == phi[2, 2, 1]
    assert phi[4, 2, 2] == 1.0 and phi[3, 3, 2] == pytest.approx(1.0 - _lens(R, r, d * 2 ** 0.5) / Vs,
                                                                rel=1e-10)
    assert numpy.all(dphi == 0.0)
    o.set_grains([[0.25, 0.25, 0.25, 0.05]], [[0.0, 0.0, 0.0]])
    o.step(1)
    _, _, phi2, dphi2 = o.read()
    assert dphi2[2, 2, 2] == pytest.approx(phi2[2, 2, 2] - phi[2, 2, 2], rel=1e-13)


def test_porosity_wraps_periodically():
    """Moving every grain by one cell along x (wrapping at the periodic
    boundary) moves the porosity field by one cell."""
    rng = numpy.random.default_rng(8)
    n = 40
    x = numpy.column_stack([rng.uniform(0, 0.6, n), rng.uniform(0, 0.6, n), rng.uniform(0.05, 0.55, n),
                            rng.uniform(0.01, 0.04, n)])
    a = _grain_case()
    a.set_grains(x, numpy.zeros((n, 3)))
    a.step(1)
    b = _grain_case()
    y = x.copy()
    y[:, 0] = (y[:, 0] + 0.1) % 0.6
    b.set_grains(y, numpy.zeros((n, 3)))
    b.step(1)
    pa, pb = a.read()[2], b.read()[2]
    assert numpy.allclose(numpy.roll(pa, 1, axis=0), pb, rtol=1e-12, atol=1e-14)
    assert pa.min() < 0.999


def test_interaction_forces():
    """On the first step the forces use the initial fields (they come
    before the solve): pressure force -grad(p) V_s with V_s = (1 - phi)
    dx dy dz (the solid volume of the grain's cell) on a linear pressure;
    the drag on a grain moving through fluid at rest in the Stokes limit,
    2.88 pi mu r v phi^(2 - chi) (the Di Felice correlation with the radius
    where it names the diameter: 0.125 (4.8^2 / Re) rho pi r^2 |v| v); no
    drag on a grain at rest; f_sum = f_d + f_p + f_v."""
    shape = (6, 6, 6)
    mu, rho, g, dz = 20.0, 1000.0, 9.81, 0.1
    zc = (numpy.arange(6) + 0.5) * dz
    p = numpy.broadcast_to(rho * g * (0.6 - zc), shape).copy()
    o = NsOracle(_params(shape, (0.6, 0.6, 0.6), mu=mu, bc_bot=1, bc_top=0, tolerance=1e-3, maxiter=1),
                 numpy.zeros(shape + (3,)), p)
    r, v = 0.01, 2.0e-6
    grains = numpy.array([[0.25, 0.25, 0.25, r], [0.35, 0.45, 0.35, r], [0.41, 0.31, 0.25, r]])
    o.set_grains(grains, [[-v, 0, 0], [0, -v, 0], [0, 0, 0]])
    o.step(1)
    f_d, f_p, f_v, f_sum = o.grain_forces()
    phi = o.read()[2]
    vol = 0.1 ** 3
    for k, (ix, iy, iz) in enumerate([(2, 2, 2), (3, 4, 3), (4, 3, 2)]):
        Vs = vol - phi[ix, iy, iz] * vol
        assert f_p[k, 2] == pytest.approx(rho * g * Vs, rel=1e-12)
        assert f_p[k, 0] == 0.0 and f_p[k, 1] == 0.0
    for k, (ix, iy, iz) in enumerate([(2, 2, 2), (3, 4, 3)]):
        ph = phi[ix, iy, iz]
        Re = rho * r * ph * v / mu
        assert Re < 2e-6
        chi = 3.7 - 0.65 * numpy.exp(-(1.5 - numpy.log10(Re)) ** 2 / 2)
        stokes = 2.88 * numpy.pi * mu * r * v * ph ** (2 - chi)
        assert f_d[k, k] == pytest.approx(stokes, rel=2e-3)
    assert numpy.all(f_d[2] == 0.0)
    assert numpy.all(f_v == 0.0)
    assert numpy.array_equal(f_sum, f_d + f_p + f_v)


def test_reaction_sums_to_the_drag():
    """The drag's reaction on the fluid, summed over the cells times the cell
    volume, is the total drag on the grains, and it enters the velocity
    prediction as -dt/(rho phi) F (mu > 0): with and without the grains'
    velocities the predicted faces differ by that term."""
    rng = numpy.random.default_rng(9)
    n = 30
    x = numpy.column_stack([rng.uniform(0, 0.6, n), rng.uniform(0, 0.6, n), rng.uniform(0.05, 0.55, n),
                            rng.uniform(0.005, 0.02, n)])
    vel = 0.01 * rng.standard_normal((n, 3))
    runs = []
    for vv in (vel, numpy.zeros_like(vel)):
        o = _grain_case(mu=1e-3)
        o.set_grains(x, vv)
        o.step(1)
        runs.append(o)
    f_d = runs[0].grain_forces()[0]
    assert numpy.abs(f_d).max() > 0
    # the prediction's change from the grains' drag: -dt/(rho phi_face) F_face
    dvp = runs[0].faces(NsOracle.PRED_FACES) - runs[1].faces(NsOracle.PRED_FACES)
    nx = ny = nz = 6
    dvx = dvp[1:nz + 1, 1:ny + 1, 1:nx + 1]  # x faces, [z, y, x]
    # -rho phi_face dv / dt summed over the faces times the cell volume is
    # the reaction the fluid receives: the total drag along x with its sign
    # reversed, as -dt/(rho phi) F removes it from the fluid
    phi = runs[0].read()[2].transpose(2, 1, 0)
    phi_face = 0.5 * (phi + numpy.roll(phi, 1, axis=2))
    reaction = -(1000.0 * phi_face * dvx / 1e-4).sum() * 0.1 ** 3
    assert reaction == pytest.approx(f_d[:, 0].sum(), rel=1e-9)


def test_grains_outside_the_grid_are_refused():
    o = _grain_case()
    with pytest.raises(ValueError):
        o.set_grains([[0.25, 0.25, 0.7, 0.01]], [[0.0, 0.0, 0.0]])


def test_coupled_oracle_moves_grains_by_the_fluid_force():
    """The DEM oracle with a fluid attached (orc_attach_ns): grains apart
    from each other, at rest, no gravity, in water flowing along x.  After
    one step each grain's velocity change is its fluid force over its mass
    times one and the same factor of the time step (the force enters the
    integration like a contact force): m dv / f is 1.5 dt for every grain,
    the third-order Taylor update dv = a dt + (a - a0) dt / 2 with a0 = 0
    on the first step (src/integration.cuh:73-81, 225-260); the fluid's
    force points along the flow."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from _cases import dense_case
    from _oracle import Oracle
    s = dense_case(n=64, seed=2, periodic=1, vmag=0.0, wmag=0.0, gravity=False, spacing_factor=1.3)
    st = s.to_state()
    nx, ny, nz = (int(v) for v in s.num)
    dt = float(s.time_dt[0])
    P = _abi.ns_params(nx, ny, nz, s.L, dt, mu=1.0e-3, rho_f=1000.0, g=(0.0, 0.0, 0.0), bc_bot=1,
                       bc_top=0, tolerance=1e-10, maxiter=100000)
    v = numpy.zeros((nx, ny, nz, 3))
    v[..., 0] = 0.05
    o = Oracle(st)
    nso = NsOracle(P, v, numpy.zeros((nx, ny, nz)))
    o.attach_ns(nso)
    o.step(1)
    f_sum = nso.grain_forces()[3]
    vel = o.read().get("vel", 64, 4)[:, :3]
    m = float(s.rho[0]) * 4.0 / 3.0 * numpy.pi * s.radius ** 3
    assert numpy.all(f_sum[:, 0] > 0)
    ratio = m * vel[:, 0] / f_sum[:, 0]
    assert numpy.allclose(ratio, 1.5 * dt, rtol=1e-9)
