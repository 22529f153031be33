"""The ``sim`` setup and analysis helpers against the state their
specification fixes (python/sphere.py names and arguments; every expected
value below is computed from the stated rule, not by the helper's code)."""
import math

import numpy
import pytest

import sphere_amd


def packing(n=300, seed=4, L=(0.3, 0.3, 0.6)):
    rng = numpy.random.default_rng(seed)
    s = sphere_amd.sim(np=n, sid="pack")
    s.radius[:] = rng.uniform(0.008, 0.012, n)
    s.x[:] = rng.uniform(0.02, 1.0, (n, 3)) * numpy.array(L) * 0.9
    s.vel[:] = rng.normal(size=(n, 3))
    s.angvel[:] = rng.normal(size=(n, 3))
    s.force[:] = 1.0
    s.es[:] = 2.0
    s.es_dot[:] = 3.0
    s.defineWorldBoundaries(L=list(L), dx=0.025)
    return s


def test_constructor_defaults():
    s = sphere_amd.sim(np=4, nw=5, sid="d")
    assert s.k_n[0] == 1.16e9 and s.k_t[0] == 1.16e9 and s.contactmodel[0] == 2
    assert s.mu_s[0] == 0.5 and s.mu_d[0] == 0.5 and s.rho[0] == 2600.0
    assert s.gamma_wn[0] == 1e3 and s.gamma_wt[0] == 0.0
    assert math.isinf(s.sigma_b[0]) and math.isinf(s.tau_b[0]) and s.lambda_bar[0] == 1.0
    assert numpy.array_equal(s.radius, numpy.ones(4)) and s.x.shape == (4, 3)
    assert s.color.dtype == numpy.int32 and s.xyzsum.shape == (4, 3)
    # top wall faces down, then -x, +x, -y, +y
    expect = [[0, 0, -1], [-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0]]
    assert numpy.array_equal(s.w_n, numpy.array(expect, dtype=float))
    assert numpy.array_equal(s.w_x, numpy.ones(5))
    f = sphere_amd.sim(np=2, fluid=True, cfd_solver=1)
    # the constructor ignores cfd_solver: Navier-Stokes fields until initFluid
    assert f.cfd_solver[0] == 0 and f.f_sum.shape == (2, 3) and f.theta == 1.0


def test_define_world_and_grid():
    s = packing()
    assert list(s.num) == [math.ceil(0.3 / 0.025), math.ceil(0.3 / 0.025), math.ceil(0.6 / 0.025)]
    with pytest.raises(Exception):
        s.defineWorldBoundaries([0.05, 0.3, 0.3], dx=0.025)  # 2 cells along x
    t = packing()
    t.nw = numpy.array([1], numpy.uint32)
    t.w_x = numpy.array([0.0])
    t.initGridAndWorldsize(margin=2.0)
    rmax = t.radius.max()
    assert numpy.allclose(t.origo, (t.x - t.radius[:, None]).min(axis=0) - 2 * rmax, rtol=0, atol=0)
    assert numpy.allclose(t.L, (t.x + t.radius[:, None]).max(axis=0) + 2 * rmax, rtol=0, atol=0)
    assert list(t.num) == [math.ceil((t.L[d] - t.origo[d]) / (2.1 * rmax)) for d in range(3)]
    assert t.w_x[0] == t.L[0]  # initGrid's (and this one's) top wall at L[0]


def test_init_grid_pos_lattice():
    s = sphere_amd.sim(np=50, sid="g")
    s.radius[:] = 0.01
    s.initGridPos(numpy.array([6, 6, 6]))
    cell = 2.1 * 0.01
    i = numpy.arange(50)
    expect = numpy.stack([i % 6, (i // 6) % 6, i // 36], axis=1) * cell + 0.5 * cell
    assert numpy.array_equal(s.x, expect)
    p = sphere_amd.sim(np=40, sid="g")
    p.radius[:] = 0.01
    p.periodic[0] = 1
    p.initGridPos(numpy.array([6, 6, 6]))
    assert list(p.num) == [6, 6, 6]  # restored after placing
    g = numpy.stack([i[:40] % 5, (i[:40] // 5) % 5, i[:40] // 25], axis=1)
    expect = g * cell + 0.5 * cell
    expect[g[:, 2] % 2 == 1, :2] += 0.5 * cell
    assert numpy.array_equal(p.x, expect)
    with pytest.raises(NameError):
        sphere_amd.sim(np=300).initGridPos(numpy.array([6, 6, 6]))


def test_random_grid_pos_keeps_grains_in_their_cells():
    numpy.random.seed(7)
    s = sphere_amd.sim(np=120, sid="r")
    s.radius[:] = numpy.random.uniform(0.008, 0.01, 120)
    s.initRandomGridPos(gridnum=numpy.array([12, 12, 20]))
    cell = 2.1 * s.radius.max() * 2
    g = numpy.floor(s.x / cell)
    i = numpy.arange(120)
    assert numpy.array_equal(g, numpy.stack([i % 6, (i // 6) % 6, i // 36], axis=1))
    lo = s.x - g * cell
    assert numpy.all(lo >= 1.05 * s.radius[:, None]) and numpy.all(cell - lo >= 1.05 * s.radius[:, None])
    assert numpy.array_equal(s.L, s.num * cell)
    assert numpy.all(s.num == numpy.ceil((s.x + s.radius[:, None]).max(axis=0) / cell))


def test_consolidate_uniaxial_and_wall_fit():
    s = packing()
    s.consolidate(normal_stress=2.5e4)
    lo = numpy.min(s.x[:, 2] - s.radius)
    hi = numpy.max(s.x[:, 2] + s.radius)
    assert s.nw[0] == 1 and s.wmode[0] == 1 and s.w_sigma0[0] == 2.5e4
    assert s.w_x[0] == hi and s.w_n[0, 2] == -1.0
    assert s.w_m[0] == pytest.approx(numpy.sum(2600.0 * 4.0 / 3.0 * math.pi * s.radius ** 3), rel=1e-14)
    assert s.L[2] == (hi - lo) * 1.1 + lo
    assert s.num[2] == math.ceil(s.L[2] / 0.025)
    # kinematics zeroed; es_dot kept
    assert not s.vel.any() and not s.angvel.any() and not s.force.any() and not s.es.any()
    assert numpy.all(s.es_dot == 3.0)
    with pytest.raises(Exception):
        packing().consolidate(normal_stress=0.0)
    u = packing()
    u.uniaxialStrainRate(wvel=-0.03)
    assert u.wmode[0] == 2 and u.w_vel[0] == -0.03 and u.w_sigma0[0] == 0.0


def test_triaxial_walls():
    s = packing()
    s.triaxial(wvel=-0.002, normal_stress=7e3)
    assert s.nw[0] == 5 and list(s.wmode) == [2, 1, 1, 1, 1]
    assert s.w_vel[0] == -0.002
    # side walls at rest with the sign of 0 * wvel (written to the file)
    assert numpy.all(s.w_vel[1:] == 0.0) and numpy.all(numpy.signbit(s.w_vel[1:]))
    assert list(s.w_sigma0) == [0.0, 7e3, 7e3, 7e3, 7e3]
    hi = (s.x + s.radius[:, None]).max(axis=0)
    lo = (s.x - s.radius[:, None]).min(axis=0)
    # walls 0, 1, 3 at the largest grain surface of their axis, 2 and 4 at the smallest
    assert s.w_x[0] == hi[2] and s.w_x[1] == hi[0] and s.w_x[2] == lo[0]
    assert s.w_x[3] == hi[1] and s.w_x[4] == lo[1]
    with pytest.raises(Exception):
        s.adjustWall(idx=7)


def test_shear_layers():
    s = packing(n=600)
    s.consolidate(1e4)
    num2, cell = int(s.num[2]), s.L[0] / s.num[0]
    s.shear(shear_strain_rate=0.2)
    z, r = s.x[:, 2], s.radius
    lo, hi = numpy.min(z - r), numpy.max(z + r)
    h = hi - lo
    bottom = z < lo + 2 * r[z < 0.3 * h].max()
    top = z > hi - 2 * r[z > 0.7 * h].max()
    assert bottom.any() and top.any()
    assert numpy.array_equal(s.fixvel > 0, bottom | top)
    assert numpy.all(s.color[bottom | top] == -1) and numpy.all(s.color[~(bottom | top)] == 0)
    assert numpy.all(s.vel[bottom] == 0.0)
    assert numpy.all(s.vel[top, 0] == h * 0.2) and numpy.all(s.vel[top, 1:] == 0.0)
    assert not s.angvel.any()
    assert s.num[2] == num2 + 1 and s.L[2] == s.num[2] * cell
    assert s.gamma_wt[0] == 0.0 and s.mu_ws[0] == 0.0 and s.mu_wd[0] == 0.0 and s.wmode[0] == 1
    t = packing(n=600)
    t.consolidate(1e4)
    t.shear(shear_stress=3e3)
    assert t.wmode[0] == 3 and t.w_tau_x[0] == 3e3 and not t.vel.any()


def test_init_temporal():
    s = packing()
    s.defaultParams(k_n=2e7, k_t=3e7)
    s.initTemporal(total=2.0, file_dt=0.5, epsilon=0.02)
    m_min = 2600 * 4.0 / 3.0 * math.pi * s.radius.min() ** 3
    assert s.time_dt[0] == pytest.approx(0.02 / math.sqrt(3e7 / m_min), rel=1e-14)
    assert (s.time_total[0], s.time_file_dt[0], s.time_step_count[0]) == (2.0, 0.5, 0)
    s.setYoungsModulus(7e9)
    s.initTemporal(total=1.0)
    k_max = math.pi / 2 * 7e9 * s.radius.max()
    assert s.time_dt[0] == pytest.approx(0.01 / math.sqrt(k_max / m_min), rel=1e-14)
    f = sphere_amd.sim(np=0, fluid=True)
    f.defineWorldBoundaries([0.4, 0.4, 0.4], dx=0.1)
    f.initFluid(cfd_solver=1)
    f.v_f[1, 1, 1, :] = [0.0, 3.0, 4.0]
    f.time_dt[0] = 1.0
    f.initTemporal(total=1.0, dt=10.0)
    assert f.time_dt[0] == pytest.approx(0.1 / 5.0)  # CFL: dx_min / |v|_max
    with pytest.raises(Exception):
        sphere_amd.sim(np=0).initTemporal(total=1.0)


def test_init_fluid_darcy_and_hydrostatic():
    s = sphere_amd.sim(np=3, nw=0, sid="f")
    s.defineWorldBoundaries([0.4, 0.3, 1.0], dx=0.1)
    s.g[2] = -10.0
    s.initFluid(mu=2e-3, rho=900.0, p=50.0, hydrostatic=True, cfd_solver=1)
    assert s.cfd_solver[0] == 1 and s.mu[0] == 2e-3 and s.rho_f[0] == 900.0
    assert s.tolerance == 1e-3 and s.maxiter == 10000 and s.ndem == 1 and s.c_phi == 1.0
    assert s.beta_f[0] == 4.5e-10 and s.k_c[0] == 4.6e-10
    assert (s.bc_xn[0], s.bc_xp[0], s.bc_yn[0], s.bc_yp[0]) == (2, 2, 2, 2)
    assert (s.bc_bot[0], s.bc_top[0], s.free_slip_bot[0], s.free_slip_top[0]) == (0, 0, 1, 1)
    assert numpy.all(s.phi == 1.0) and s.f_p.shape == (3, 3) and not s.dphi.any()
    dz = 0.1
    for iz in range(10):
        want = 50.0 if iz == 9 else 50.0 + (1.0 - (iz + 0.5) * dz - dz) * 900.0 * 10.0
        assert numpy.allclose(s.p_f[:, :, iz], want, rtol=1e-14, atol=0)
    with pytest.raises(Exception):
        s.initFluid(cfd_solver=2)
    s.setFluidTopFixedFlux(3.0)
    s.setFluidBottomNoFlow()
    s.setFluidXNoFlow()
    assert (s.bc_top[0], s.bc_top_flux[0], s.bc_bot[0], s.bc_xn[0], s.bc_xp[0]) == (4, 3.0, 1, 1, 1)


def test_default_params_cohesion():
    s = sphere_amd.sim(np=1)
    s.defaultParams(mu_s=0.6, mu_d=0.4, gamma_t=3.0, capillaryCohesion=1)
    assert s.mu_ws[0] == 0.6 and s.mu_wd[0] == 0.4 and s.gamma_wn[0] == 1e4
    assert s.kappa[0] == 2 * math.pi * 3.0 and s.V_b[0] == 1e-12
    assert s.db[0] == pytest.approx(1e-4, rel=1e-12)
    s.defaultParams()
    assert s.kappa[0] == 0.0 and s.V_b[0] == 0.0 and s.db[0] == 0.0


def test_bonds_and_bond_pairs():
    numpy.random.seed(2)
    s = packing(n=40)
    s.bond(3, 7)
    assert s.nb0[0] == 1 and list(s.bonds[0]) == [3, 7] and not s.bonds_delta_t.any()
    s.createBondPair(0, 1, spacing=-0.1)
    d = numpy.linalg.norm(s.x[0] - s.x[1])
    assert d <= (s.radius[0] + s.radius[1]) * 0.9 * 1.01 and s.nb0[0] == 2
    s.randomBondPairs(ratio=0.5)
    b = s.bonds[2:]
    assert b.shape[0] == s.nb0[0] - 2 and len(numpy.unique(b)) == b.size


def test_energy_closed_forms():
    s = sphere_amd.sim(np=2, sid="e")
    s.radius[:] = [0.01, 0.02]
    s.x[:] = [[0, 0, 1.0], [0, 0, 2.0]]
    s.vel[:] = [[3.0, 0, 0], [0, 4.0, 0]]
    s.angvel[:] = [[0, 0, 10.0], [0, 0, 0]]
    s.g[:] = [0, 0, -9.81]
    m = 2600 * 4.0 / 3.0 * math.pi * numpy.array([0.01, 0.02]) ** 3
    assert s.energy('pot') == pytest.approx(9.81 * (m[0] * 1.0 + m[1] * 2.0), rel=1e-14)
    assert s.energy('kin') == pytest.approx(0.5 * m[0] * 9 + 0.5 * m[1] * 16, rel=1e-14)
    assert s.energy('rot') == pytest.approx(0.5 * 0.4 * m[0] * 0.01 ** 2 * 100, rel=1e-14)
    s.bond(0, 1)
    s.bonds_delta_n[0] = 1e-3
    s.bonds_omega_t[0] = [0, 2e-3, 0]
    R = 0.01
    want = 0.5 * math.pi * R ** 2 * 1.16e9 * 1e-6 + 0.5 * math.pi * R ** 4 / 4 * 1.16e9 * 4e-6
    assert s.energy('bondpot') == pytest.approx(want, rel=1e-13)
    with pytest.raises(Exception):
        s.energy('nope')


def test_bulk_porosity_and_shear_strain():
    s = sphere_amd.sim(np=1, sid="b")
    s.radius[:] = 0.05
    s.x[:] = [[0.5, 0.5, 0.5]]
    s.L[:] = [1.0, 1.0, 1.0]
    assert s.bulkPorosity(trim=True) == pytest.approx(1 - math.pi / 6, rel=1e-14)
    V = 4.0 / 3.0 * math.pi * 0.05 ** 3
    assert s.bulkPorosity(trim=False) == pytest.approx(1 - V, rel=1e-14)
    assert s.voidRatio() == pytest.approx((1 - V) / V, rel=1e-14)
    t = sphere_amd.sim(np=3, nw=1, sid="s")
    t.fixvel[:] = [0.0, 1.0, 1.0]
    t.xyzsum[:, 0] = [9.0, 0.1, 0.3]
    t.vel[:, 0] = [5.0, 0.0, 0.02]
    t.w_x[0] = 0.6
    assert t.shearDisplacement() == 0.3 and t.shearStrain() == pytest.approx(0.5)
    assert t.shearVelocity() == 0.02 and t.shearStrainRate() == pytest.approx(0.02 / 0.6)


def test_add_and_delete_particles():
    s = sphere_amd.sim(np=2, sid="a")
    s.addParticle([1.0, 2.0, 3.0], 0.5, vel=[0.1, 0.2, 0.3], fixvel=1.0, color=4)
    assert s.np[0] == 3 and list(s.x[2]) == [1.0, 2.0, 3.0] and s.radius[2] == 0.5
    assert list(s.vel[2]) == [0.1, 0.2, 0.3] and s.fixvel[2] == 1.0 and s.color[2] == 4
    assert not s.angvel[2].any() and s.xyzsum.shape == (3, 3)
    s.deleteParticle([0, 1])
    assert s.np[0] == 1 and s.x.shape == (1, 3) and s.radius[0] == 0.5


def test_vtk_files(tmp_path):
    """writeVTK / writeFluidVTK / writeVTKall write well-formed VTK XML with
    the grains' and cells' values (read back with the XML parser)."""
    import xml.etree.ElementTree as ET
    for d in ("input", "output", "work"):
        (tmp_path / d).mkdir()
    s = packing(n=20)
    s.sid = "vtk"
    s.time_dt[0] = 1e-6
    s.initFluid(cfd_solver=1)
    s.p_f[:] = numpy.arange(s.p_f.size).reshape(s.p_f.shape)
    s.writebin(folder=str(tmp_path / "output"), verbose=False)
    import os
    os.rename(str(tmp_path / "output" / "vtk.bin"), str(tmp_path / "output" / "vtk.output00000.bin"))
    (tmp_path / "output" / "vtk.status.dat").write_text("0.0 0.0 0\n")
    old = os.getcwd()
    os.chdir(str(tmp_path / "work"))
    try:
        s.writeVTKall(verbose=False)
    finally:
        os.chdir(old)
    root = ET.parse(str(tmp_path / "output" / "vtk.00000.vtu")).getroot()
    arrays = {a.get("Name"): a for a in root.iter("DataArray")}
    pos = numpy.array(arrays["Position [m]"].text.split(), float).reshape(-1, 3)
    assert numpy.allclose(pos, s.x, atol=1e-6)
    assert numpy.allclose(numpy.array(arrays["Diameter"].text.split(), float), 2 * s.radius, atol=1e-6)
    assert "Fluid pressure force [N]" in arrays and "Type color" in arrays
    img = ET.parse(str(tmp_path / "output" / "fluid-vtk.00000.vti")).getroot()
    p = {a.get("Name"): a for a in img.iter("DataArray")}["Pressure [Pa]"]
    vals = numpy.array(p.text.split(), float)
    assert numpy.array_equal(vals, s.p_f.transpose(2, 1, 0).ravel())
