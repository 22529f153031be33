"""Seeded synthetic workloads of the benchmark configurations.

* settling(n)       -- dry settling, linear contacts, frictional walls.
* shear_box(n)      -- dense periodic shear box, Hertz-Mindlin, 10 kPa top
                       wall, top layer driven at 0.01 m/s, 5% bonded pairs.
* ensemble_boxes(count)
                    -- independent shear boxes of log-uniform size (5,000 to
                       60,000 grains) and drawn friction, load and shear
                       velocity, grown together through an Ensemble.
* consolidation_base(n) / consolidation_finish(s)
                    -- Darcy-coupled consolidation: 200,000 grains under a
                       20 kPa wall on a 48 x 48 x 96 Darcy grid.
* prepare_shear_box(engine, n, seed), prepare_consolidation(engine, n, seed)
                    -- the two benchmark boxes, grown on the device.

Dense packings are made by radius growth: grains start on a jittered cubic
lattice with radii scaled down until no two touch, then the radii grow back
to their drawn values in stages while the DEM (the stepper given) relaxes
the overlaps; velocities are zeroed at the start of every stage.
"""
import time

import numpy

from .sim import sim


def lognormal_radii(n, rng, median=0.01, sigma=0.2):
    return median * numpy.exp(sigma * rng.standard_normal(n))


def lattice(n, spacing, nx, ny, jitter, rng, z0=0.0):
    """n sites of a cubic lattice nx*ny per layer, z-major order, jittered by
    +-jitter*spacing per axis."""
    k = numpy.arange(n)
    ix = k % nx
    iy = (k // nx) % ny
    iz = k // (nx * ny)
    x = numpy.stack([ix, iy, iz], axis=1) * spacing + 0.5 * spacing
    x = x + rng.uniform(-jitter, jitter, size=x.shape) * spacing
    x[:, 2] += z0
    return x


class EngineStepper:
    """Adapter of sphere_amd.Engine for grow_packing.  Packing preparation
    passes through over-compressed states where a large grain can touch more
    than 32 small ones; those steps are kept (the extra contacts are left out
    of the slot row) and counted in ``overflow_steps``."""

    def __init__(self, engine):
        self.engine = engine
        self.overflow_steps = 0

    def start(self, buf):
        self.engine.load(buf)

    def step(self, n):
        from ._abi import SphereError
        try:
            self.engine.step(n)
        except SphereError as e:
            if e.code != -7:
                raise
            self.overflow_steps += 1

    def count_contacts(self):
        return self.engine.count_contacts()

    def read(self):
        return self.engine.read()


def grow_packing(s, stepper, f0, stages, steps_per_stage, relax_steps, log=None):
    """Grow the radii of sim ``s`` from f0 to 1 times their values."""
    r_final = s.radius.copy()
    n = int(s.np[0])
    for k in range(1, stages + 2):
        # stages crowd towards f = 1, where the packing jams
        u = min(1.0, k / stages)
        f = min(1.0, f0 + (1.0 - f0) * (1.0 - (1.0 - u) ** 2))
        s.radius = r_final * f
        s.vel[:] = 0.0
        s.angvel[:] = 0.0
        stepper.start(s.to_state())
        stepper.step(steps_per_stage if k <= stages else relax_steps)
        buf = stepper.read()
        s.x = buf.get("x", n, 4)[:, :3].copy()
        s.vel = buf.get("vel", n, 4)[:, :3].copy()
        if log:
            log("growth stage %d/%d f=%.3f" % (k, stages + 1, f))
    s.radius = r_final.copy()
    s.vel[:] = 0.0
    s.angvel[:] = 0.0
    s.force[:] = 0.0
    s.torque[:] = 0.0
    return s


def grow_to_jamming(s, stepper, f0, z_target=4.5, steps=300, coarse_df=0.02, fine_df=0.0025,
                    z_fine=1.0, max_stages=400, log=None):
    """Grow radii r*f from f = f0: by coarse_df per stage while the mean
    coordination number is below z_fine, then by fine_df until it reaches
    z_target (a jammed packing).  Positions, box and walls are then rescaled
    by 1/f so that the radii are the drawn ones; Hertzian stress depends on
    the relative overlap only, so the rescaled packing carries the same
    stress.  Each stage zeroes the velocities and runs ``steps`` steps.
    Returns (f, mean coordination number)."""
    r_final = s.radius.copy()
    n = int(s.np[0])
    f = f0
    z = 0.0
    for k in range(max_stages):
        s.radius = r_final * f
        s.vel[:] = 0.0
        s.angvel[:] = 0.0
        stepper.start(s.to_state())
        stepper.step(steps)
        buf = stepper.read()
        s.x = buf.get("x", n, 4)[:, :3].copy()
        z = stepper.count_contacts() / n
        if log:
            log("stage %d f=%.6f Z=%.3f" % (k, f, z))
        if z >= z_target:
            break
        f *= 1.0 + (coarse_df if z < z_fine else fine_df)
    scale = 1.0 / f
    s.x *= scale
    s.L = s.L * scale
    s.w_x = s.w_x * scale
    s.radius = r_final.copy()
    s.vel[:] = 0.0
    s.angvel[:] = 0.0
    s.force[:] = 0.0
    s.torque[:] = 0.0
    return f, z


def cap_overlaps(s, pairs, cap_rel, iters, relax=0.6):
    """Geometric overlap capping: each pair (i, j) overlapping by more than
    cap_rel * min(r_i, r_j) is pushed apart along the centre line by a
    fraction ``relax`` of the excess, shared equally, for ``iters`` sweeps
    (Jacobi style, periodic images in x and y when s.periodic == 1; the
    bottom wall at origo and the top wall keep the grains inside).  Bounds
    the Hertzian stress of a grown packing at E* cap_rel^(3/2)."""
    x = s.x
    r = s.radius
    i = pairs[:, 0]
    j = pairs[:, 1]
    n = x.shape[0]
    L = numpy.asarray(s.L, dtype=numpy.float64)
    cap = cap_rel * numpy.minimum(r[i], r[j])
    z_lo = s.origo[2] + r * (1.0 - cap_rel)
    z_hi = (s.w_x[0] if s.nw[0] > 0 else s.L[2]) - r * (1.0 - cap_rel)
    for _ in range(iters):
        d = x[i] - x[j]
        if int(s.periodic[0]) == 1:
            d[:, 0] -= L[0] * numpy.round(d[:, 0] / L[0])
            d[:, 1] -= L[1] * numpy.round(d[:, 1] / L[1])
        dist = numpy.sqrt(numpy.sum(d * d, axis=1))
        excess = numpy.maximum(r[i] + r[j] - dist - cap, 0.0)
        if not numpy.any(excess > 0.0):
            break
        m = (0.5 * relax * excess / dist)[:, None] * d
        for c in range(3):
            x[:, c] += numpy.bincount(i, weights=m[:, c], minlength=n)
            x[:, c] -= numpy.bincount(j, weights=m[:, c], minlength=n)
        x[:, 2] = numpy.clip(x[:, 2], z_lo, z_hi)
        if int(s.periodic[0]) == 1:
            x[:, 0] %= L[0]
            x[:, 1] %= L[1]
    s.x = x
    return s


def relax_packing(s, stepper, rounds, steps, log=None):
    """Quenched relaxation: every round zeroes the velocities and runs
    ``steps`` steps (a damped descent to a static, gravity-loaded state)."""
    n = int(s.np[0])
    for k in range(rounds):
        s.vel[:] = 0.0
        s.angvel[:] = 0.0
        stepper.start(s.to_state())
        stepper.step(steps)
        buf = stepper.read()
        s.x = buf.get("x", n, 4)[:, :3].copy()
        if log:
            log("relax round %d/%d" % (k + 1, rounds))
    s.vel[:] = 0.0
    s.angvel[:] = 0.0
    s.force[:] = 0.0
    s.torque[:] = 0.0
    return s


def touching_pairs(contacts, np_):
    """Unique (i, j), i < j, of a contact list laid out as i * 32 + slot."""
    c = contacts.reshape(np_, -1)
    i = numpy.repeat(numpy.arange(np_, dtype=numpy.int64), c.shape[1])
    j = c.ravel().astype(numpy.int64)
    keep = (j != np_) & (i < j)
    return numpy.stack([i[keep], j[keep]], axis=1)


def bond_pairs(s, pairs, fraction, rng):
    """Bond a seeded fraction of the given pairs (sphere's bond(), vectorised)."""
    pick = pairs[rng.random(pairs.shape[0]) < fraction]
    nb = pick.shape[0]
    s.lambda_bar[0] = 1.0
    s.bonds = pick.astype(numpy.uint32)
    s.bonds_delta_n = numpy.zeros(nb)
    s.bonds_delta_t = numpy.zeros((nb, 3))
    s.bonds_omega_n = numpy.zeros(nb)
    s.bonds_omega_t = numpy.zeros((nb, 3))
    s.nb0 = numpy.ones(1, dtype=numpy.uint32) * nb
    return nb


def shear_box_base(n=1_000_000, seed=2, phi=0.62, footprint=None):
    """Grains of the dense periodic shear box on their growth lattice.

    Radii: lognormal, median 0.01 m, sigma_ln 0.2.  Box: periodic in x and
    y, fixed walls at the bottom (z = 0) and top during growth; volume set so
    the grown grains fill it at solid fraction phi."""
    rng = numpy.random.default_rng(seed)
    r = lognormal_radii(n, rng)
    v_solid = float(numpy.sum(4.0 / 3.0 * numpy.pi * r ** 3))
    spacing = (v_solid / phi / n) ** (1.0 / 3.0)
    if footprint is None:
        nx = int(round((n / 1.05) ** (1.0 / 3.0)))
    else:
        nx = int(round(footprint / spacing))
    ny = nx
    nz = (n + nx * ny - 1) // (nx * ny)
    order = rng.permutation(n)  # radii placed at random sites
    x = lattice(n, spacing, nx, ny, 0.05, rng)
    s = sim(np=n, sid="shear_box", nw=1)
    s.radius[:] = r[order]
    s.x[:] = x
    s.periodic[0] = 1
    s.contactmodel[0] = 3
    s.setYoungsModulus(7e9)
    s.setPoissonsRatio(0.3)
    s.defaultParams(mu_s=0.0, mu_d=0.0, rho=2600, k_n=1.16e9, k_t=1.16e9, k_r=0.3,
                    gamma_n=0.0, gamma_wn=0.0, gamma_wt=0.0)
    s.g[:] = 0.0
    L = [nx * spacing, ny * spacing, nz * spacing]
    r_max = float(numpy.max(r))
    # cells fit grains grown to 1.1 times their drawn radii
    s.defineWorldBoundaries(L=[L[0], L[1], L[2] + 4.0 * r_max], dx=2.1 * 1.1 * r_max)
    s.w_n[0, :] = [0.0, 0.0, -1.0]
    s.w_x[0] = L[2]
    s.wmode[0] = 0
    s.w_m[0] = 1.0
    r_min = float(numpy.min(r))
    m_min = 2600.0 * 4.0 / 3.0 * numpy.pi * r_min ** 3
    k_max = numpy.pi / 2.0 * 7e9 * r_max
    s.time_dt[0] = 0.17 * numpy.sqrt(m_min / k_max)
    s.time_total[0] = 1.0
    s.time_file_dt[0] = 1.0
    s.gamma_n[0] = 2e2  # damping during growth only
    f0 = 0.9 * spacing / (2.0 * r_max)
    return s, f0, rng


def shear_box_finish(s, rng, contacts=None, bond_fraction=0.05, sigma0=10e3, v_top=0.01):
    """Apply the shear-box run configuration to a grown packing: friction
    mu = 0.5, no damping, gravity, a 10 kPa top wall (consolidate()) and a
    driven top layer (shear()), then bond a fraction of the touching pairs."""
    # defaultParams() makes mu_ws/mu_wd aliases of mu_s/mu_d, and shear()
    # zeroes the wall friction: keep them separate arrays
    s.mu_ws = s.mu_ws.copy()
    s.mu_wd = s.mu_wd.copy()
    s.gamma_n[0] = 0.0
    s.g[:] = [0.0, 0.0, -9.81]
    s.consolidate(normal_stress=sigma0)
    z_extent = float(numpy.max(s.x[:, 2] + s.radius) - numpy.min(s.x[:, 2] - s.radius))
    s.shear(shear_strain_rate=v_top / z_extent)
    s.w_sigma0[0] = sigma0
    s.mu_s[0] = 0.5
    s.mu_d[0] = 0.5
    nb = 0
    if contacts is not None and bond_fraction > 0.0:
        nb = bond_pairs(s, touching_pairs(contacts, int(s.np[0])), bond_fraction, rng)
    s.time_current[0] = 0.0
    s.time_total[0] = 1.0
    s.time_file_dt[0] = 1.0
    s.time_step_count[0] = 0
    return nb


def settling(n=10000, seed=1, steps=2000):
    """Dry settling: radii ~U[0.009, 0.011] m on a jittered cubic lattice
    (0.025 m spacing, jitter 10% of the spacing, 10 x 10 per layer) in a
    0.25 x 0.25 m box with frictional walls; linear contacts with
    k_n = k_t = 1.16e7 N/m, gamma_n = 0, mu = 0.5; dt = 0.17 sqrt(m_min/k_n).
    The box is as tall as the lattice (2.5 m for 10,000 grains): 0.5 m cannot
    hold 0.042 m^3 of grains."""
    rng = numpy.random.default_rng(seed)
    per_layer = 100
    nz = (n + per_layer - 1) // per_layer
    spacing = 0.025
    x = lattice(n, spacing, 10, 10, 0.05, rng)
    s = sim(np=n, sid="settling", nw=0)
    s.x[:] = x
    s.radius[:] = rng.uniform(0.009, 0.011, n)
    s.defaultParams(mu_s=0.5, mu_d=0.5, k_n=1.16e7, k_t=1.16e7, gamma_wn=1e4, gamma_wt=1e4)
    s.g[2] = -9.81
    s.defineWorldBoundaries(L=[0.25, 0.25, nz * spacing])
    m_min = 2600.0 * 4.0 / 3.0 * numpy.pi * numpy.min(s.radius) ** 3
    dt = 0.17 * numpy.sqrt(m_min / 1.16e7)
    s.time_dt[0] = dt
    s.time_total[0] = steps * dt
    s.time_file_dt[0] = steps * dt
    return s


# linear contact stiffness of the consolidation box (sphere's defaultParams)
K_N_CONSOLIDATION = 1.16e9


def consolidation_base(n=200_000, seed=3, phi=0.6, cells_xy=48):
    """Grains of the Darcy consolidation box on their growth lattice.

    Radii ~U[0.009, 0.011] m; periodic in x and y; a fixed bottom (z = 0) and
    a fixed top wall during growth; linear contacts k_n = k_t = 1.16e9 N/m
    (sphere's default stiffness), frictionless and damped while growing.  The lattice footprint is chosen
    so that, after the packing is rescaled to the drawn radii, cells_xy cells
    of about 2.1 r_max span it."""
    rng = numpy.random.default_rng(seed)
    r = rng.uniform(0.009, 0.011, n)
    r_max = float(numpy.max(r))
    v_solid = float(numpy.sum(4.0 / 3.0 * numpy.pi * r ** 3))
    spacing = (v_solid / phi / n) ** (1.0 / 3.0)
    nx = int(round(cells_xy * 2.1 * r_max / spacing))
    ny = nx
    nz = (n + nx * ny - 1) // (nx * ny)
    x = lattice(n, spacing, nx, ny, 0.05, rng)
    s = sim(np=n, sid="consolidation", nw=1)
    s.radius[:] = r
    s.x[:] = x
    s.periodic[0] = 1
    s.contactmodel[0] = 2
    s.defaultParams(mu_s=0.0, mu_d=0.0, rho=2600, k_n=K_N_CONSOLIDATION, k_t=K_N_CONSOLIDATION,
                    gamma_n=3e3, gamma_wn=1e4, gamma_wt=0.0)
    s.g[:] = 0.0
    L = [nx * spacing, ny * spacing, nz * spacing]
    s.defineWorldBoundaries(L=[L[0], L[1], L[2] + 4.0 * r_max], dx=2.1 * 1.1 * r_max)
    s.w_n[0, :] = [0.0, 0.0, -1.0]
    s.w_x[0] = L[2]
    s.wmode[0] = 0
    s.w_m[0] = 1.0
    m_min = 2600.0 * 4.0 / 3.0 * numpy.pi * float(numpy.min(r)) ** 3
    s.time_dt[0] = 0.17 * numpy.sqrt(m_min / K_N_CONSOLIDATION)
    s.time_total[0] = 1.0
    s.time_file_dt[0] = 1.0
    f0 = 0.9 * spacing / (2.0 * r_max)
    return s, f0, rng


def consolidation_finish(s, sigma0=2e4, cells=(48, 48, 96), mu_f=1.04e-7, beta_f=1.43e-8,
                         k_c=3.5e-13, tolerance=1e-8, maxiter=200000):
    """Apply the consolidation run configuration to a settled packing.

    consolidate() puts the top wall on the grains with sigma0 = 20 kPa and
    the grains' total mass; friction mu = 0.5, no contact damping, gravity.
    The Darcy grid is the DEM grid, cells[0] x cells[1] x cells[2] cells of
    side L_x / cells[0] (at least 2 r_max).  Fluid: Darcy solver, viscosity
    1.04e-7 Pa s, compressibility 1.43e-8 1/Pa (1/K_w of
    python/halfshear-darcy-starter.py:41-44,57), Kozeny-Carman prefactor
    3.5e-13 m^2, p = 0 at the top, no flow through the base, periodic in x
    and y, Jacobi to a relative change of 1e-8, one DEM step per fluid step.
    dt = 0.17 sqrt(m_min / k_n) with k_n = 1.16e9 N/m.

    The Jacobi sweep p <- 0.95 p + 0.05 (p_old + (dp_expl + dp_impl(p))/2)
    contracts while c = k dt / (beta_f phi mu dx^2) stays below about 6.5 in
    the iterated cells; with these values that holds for phi below about
    0.84 (the cells above the top wall are not iterated)."""
    s.gamma_n[0] = 0.0
    s.mu_s[0] = 0.5
    s.mu_d[0] = 0.5
    s.g[:] = [0.0, 0.0, -9.81]
    s.consolidate(normal_stress=sigma0)
    s.gamma_wn[0] = 1e4
    cell = float(s.L[0]) / cells[0]
    r_max = float(numpy.max(s.radius))
    if cell < 2.0 * r_max:
        raise ValueError("Darcy cell %.5g m is below 2 r_max = %.5g m" % (cell, 2.0 * r_max))
    s.num = numpy.array(cells, dtype=numpy.uint32)
    s.L = numpy.array([cells[0] * cell, cells[1] * cell, cells[2] * cell])
    if float(s.w_x[0]) >= s.L[2]:
        raise ValueError("the packing is taller than the Darcy grid")
    s.initFluid(mu=mu_f, rho=1000.0, p=0.0, cfd_solver=1)
    s.setFluidCompressibility(beta_f)
    s.setPermeabilityPrefactor(k_c, verbose=False)
    s.setFluidBottomNoFlow()
    s.setFluidTopFixedPressure()
    s.setTolerance(tolerance)
    s.setMaxIterations(maxiter)
    s.setDEMstepsPerCFDstep(1)
    m_min = 2600.0 * 4.0 / 3.0 * numpy.pi * float(numpy.min(s.radius)) ** 3
    s.time_dt[0] = 0.17 * numpy.sqrt(m_min / K_N_CONSOLIDATION)
    s.time_current[0] = 0.0
    s.time_total[0] = 1.0
    s.time_file_dt[0] = 1.0
    s.time_step_count[0] = 0
    return s


def ensemble_boxes(count=64, seed=5, n_min=5000, n_max=60000, steps=300, z_target=4.2, relax_steps=400,
                   device=0, log=None):
    """Seeded shear boxes for the ensemble benchmark.  Box b: N_b log-uniform
    in [n_min, n_max], mu ~ U[0.3, 0.7], sigma0 ~ U[5, 80] kPa, top-layer
    velocity ~ U[1e-3, 1e-1] m/s; the shear box's radii, contact law and
    preparation (growth to jamming, relaxation under gravity, 10 kPa-style
    wall and driven layer with the drawn values), all boxes stepped together
    through one Ensemble per stage.  Returns the sims and their parameters."""
    from ._abi import Engine, Ensemble
    rng = numpy.random.default_rng(seed)
    sizes = numpy.exp(rng.uniform(numpy.log(n_min), numpy.log(n_max), count)).astype(int)
    mus = rng.uniform(0.3, 0.7, count)
    sig = rng.uniform(5e3, 80e3, count)
    vel = rng.uniform(1e-3, 1e-1, count)
    boxes = []
    for b in range(count):
        s, f0, brng = shear_box_base(n=int(sizes[b]), seed=int(seed * 1000 + b), phi=0.60)
        s.gamma_n[0] = 5e3
        r_final = s.radius.copy()
        boxes.append({"s": s, "f": f0, "r": r_final, "z": 0.0, "done": False, "rng": brng})
    _grow_together(boxes, steps, z_target, relax_steps, device, log)
    for k, bx in enumerate(boxes):
        s = bx["s"]
        shear_box_finish(s, bx["rng"], contacts=None, bond_fraction=0.0, sigma0=float(sig[k]),
                         v_top=float(vel[k]))
        s.mu_s[0] = mus[k]
        s.mu_d[0] = mus[k]
    params = [{"grains": int(sizes[k]), "mu": float(mus[k]), "sigma0": float(sig[k]),
               "v_top": float(vel[k])} for k in range(count)]
    return [bx["s"] for bx in boxes], params


def _grow_together(boxes, steps, z_target, relax_steps, device=0, log=None):
    """Radius growth to jamming (each box from its own lattice and scale f),
    then relaxation under gravity, of all boxes stepped as one Ensemble per
    stage; each box's positions end in its sim, at rest."""
    from ._abi import Engine, Ensemble
    engines = [Engine(device, nc=64) for _ in boxes]
    ens = Ensemble(engines)
    for stage in range(400):
        active = [k for k, bx in enumerate(boxes) if not bx["done"]]
        if not active:
            break
        for k in active:
            bx = boxes[k]
            s = bx["s"]
            s.radius = bx["r"] * bx["f"]
            s.vel[:] = 0.0
            s.angvel[:] = 0.0
            engines[k].load(s.to_state())
        ens.step(steps)
        for k in active:
            bx = boxes[k]
            s = bx["s"]
            n = int(s.np[0])
            s.x = engines[k].read().get("x", n, 4)[:, :3].copy()
            bx["z"] = engines[k].count_contacts() / n
            if bx["z"] >= z_target:
                bx["done"] = True
            else:
                bx["f"] *= 1.0 + (0.02 if bx["z"] < 1.0 else 0.005)
        if log:
            log("ensemble growth stage %d: %d boxes still growing" % (stage, len(active)))
    for bx in boxes:
        s = bx["s"]
        scale = 1.0 / bx["f"]
        s.x *= scale
        s.L = s.L * scale
        s.w_x = s.w_x * scale
        s.radius = bx["r"].copy()
        s.w_x[0] = s.L[2]
        s.g[2] = -9.81
        s.vel[:] = 0.0
        s.angvel[:] = 0.0
    for k, bx in enumerate(boxes):
        engines[k].load(bx["s"].to_state())
    ens.step(relax_steps)
    for k, bx in enumerate(boxes):
        s = bx["s"]
        n = int(s.np[0])
        s.x = engines[k].read().get("x", n, 4)[:, :3].copy()
        s.vel[:] = 0.0
        s.angvel[:] = 0.0
        s.force[:] = 0.0
        s.torque[:] = 0.0
    ens.close()
    for e in engines:
        e.close()


def darcy_cells_xy(n):
    """Horizontal Darcy cells of a consolidation-type box of n grains: the
    smallest multiple of 8 (the correction network halves the grid three
    times), at least 16, that leaves the settled packing below about 60% of
    the c x c x 2c grid's height (packing height / grid height is about
    0.283 n / c^3 for cells of 2.1 r_max at a packing fraction of 0.6)."""
    c = int(numpy.ceil((0.283 * n / 0.6) ** (1.0 / 3.0) / 8.0)) * 8
    return max(16, c)


def darcy_ensemble_boxes(count=64, seed=7, n_min=5000, n_max=60000, steps=300, z_target=4.2,
                         relax_steps=400, device=0, log=None):
    """Seeded Darcy-coupled shear boxes for the ensemble benchmark.  Box b:
    N_b log-uniform in [n_min, n_max] grains of the consolidation box (radii
    U[0.009, 0.011] m, linear contacts, the same fluid), on a c x c x 2c Darcy
    grid (the DEM grid) with c from darcy_cells_xy; friction mu ~ U[0.3,
    0.7], top-wall stress sigma0 ~ U[5, 80] kPa, and the grains of the top
    layer (within one largest diameter of the highest grain) driven at
    v ~ U[1e-3, 1e-1] m/s along x.  Grown to jamming and relaxed together
    through one Ensemble per stage.  Returns the sims and their parameters."""
    rng = numpy.random.default_rng(seed)
    sizes = numpy.exp(rng.uniform(numpy.log(n_min), numpy.log(n_max), count)).astype(int)
    mus = rng.uniform(0.3, 0.7, count)
    sig = rng.uniform(5e3, 80e3, count)
    vel = rng.uniform(1e-3, 1e-1, count)
    boxes, cells = [], []
    for b in range(count):
        c = darcy_cells_xy(int(sizes[b]))
        s, f0, brng = consolidation_base(n=int(sizes[b]), seed=int(seed * 1000 + b), cells_xy=c)
        boxes.append({"s": s, "f": f0, "r": s.radius.copy(), "z": 0.0, "done": False, "rng": brng})
        cells.append((c, c, 2 * c))
    _grow_together(boxes, steps, z_target, relax_steps, device, log)
    for k, bx in enumerate(boxes):
        s = bx["s"]
        consolidation_finish(s, sigma0=float(sig[k]), cells=cells[k])
        s.mu_s[0] = mus[k]
        s.mu_d[0] = mus[k]
        top = s.x[:, 2] > float(numpy.max(s.x[:, 2])) - 2.0 * float(numpy.max(s.radius))
        s.fixvel[top] = 1.0
        s.vel[top, 0] = float(vel[k])
    params = [{"grains": int(sizes[k]), "mu": float(mus[k]), "sigma0": float(sig[k]),
               "v_top": float(vel[k]), "grid": list(cells[k])} for k in range(count)]
    return [bx["s"] for bx in boxes], params


def prepare_shear_box(engine, n, seed, stage_steps=400, z_target=4.2, relax_rounds=4, relax_steps=400,
                      log=None):
    """The dense periodic shear box of the benchmark: grown to jamming on the
    device, relaxed under gravity, then given its run configuration (Hertz-
    Mindlin, 10 kPa top wall, driven top layer, 5% of the touching pairs
    bonded).  Returns the sim object."""
    log = log or (lambda m: None)
    s, f0, rng = shear_box_base(n=n, seed=seed, phi=0.60)
    s.gamma_n[0] = 5e3
    st = EngineStepper(engine)
    t0 = time.time()
    f, z = grow_to_jamming(s, st, f0, z_target=z_target, steps=stage_steps, fine_df=0.005)
    log("growth to jamming done in %.1f s: f=%.4f, Z=%.2f" % (time.time() - t0, f, z))
    s.w_x[0] = s.L[2]
    s.g[2] = -9.81
    t0 = time.time()
    relax_packing(s, st, relax_rounds, relax_steps)
    log("relaxation done in %.1f s (%d rounds x %d steps); stages with slot overflow: %d" %
        (time.time() - t0, relax_rounds, relax_steps, st.overflow_steps))
    # contacts of the final packing: one step with the run's contact law
    s.gamma_n[0] = 0.0
    engine.load(s.to_state())
    engine.step(1)
    contacts = engine.read_contacts()[0]
    per = numpy.count_nonzero(contacts.reshape(n, -1) != n, axis=1)
    log("contacts per grain after preparation: mean %.2f max %d" % (per.mean(), per.max()))
    nb = shear_box_finish(s, rng, contacts=contacts, bond_fraction=0.05)
    log("packing: %d grains, %d bonds, grid %s, box %s" % (n, nb, [int(v) for v in s.num],
                                                          [round(float(v), 4) for v in s.L]))
    return s


def prepare_consolidation(engine, n, seed, stage_steps=300, relax_rounds=3, relax_steps=400, log=None):
    """The Darcy-coupled consolidation box of the benchmark (200,000 grains
    on a 48 x 48 x 96 grid by default): grown to jamming, relaxed, then
    loaded by a 20 kPa wall with the fluid set up.  Returns the sim object."""
    log = log or (lambda m: None)
    s, f0, rng = consolidation_base(n=n, seed=seed)
    st = EngineStepper(engine)
    t0 = time.time()
    f, z = grow_to_jamming(s, st, f0, z_target=4.2, steps=stage_steps, fine_df=0.005)
    log("consolidation box: growth to jamming done in %.1f s: f=%.4f, Z=%.2f" %
        (time.time() - t0, f, z))
    s.w_x[0] = s.L[2]
    s.g[2] = -9.81
    t0 = time.time()
    relax_packing(s, st, relax_rounds, relax_steps)
    log("consolidation box: relaxation done in %.1f s" % (time.time() - t0))
    consolidation_finish(s)
    log("consolidation box: %d grains, grid %s, box %s, wall at %.4f m" %
        (n, [int(v) for v in s.num], [round(float(v), 4) for v in s.L], float(s.w_x[0])))
    return s
