"""Every benchmark configuration at its real workload against the oracle on
the GPU, and grains pressed on walls.

* Config 1: dry settling, 10,000 grains, all 2,000 steps.
* Config 2: the benchmark's 1,000,000-grain Hertz-Mindlin shear box (64
  contact slots, 5% bonds, driven layers, 10 kPa top wall), 10 steps.
* Config 3: the 200,000-grain consolidation box on its 48 x 48 x 96 Darcy
  grid, 3 fluid steps with equal Jacobi sweep counts.
* Config 4: a learned Darcy solve on 96 x 96 x 192 that ends at the plain
  solve's pressure (the network's forward pass on the 48 x 48 x 96 grid
  against the FP64 network oracle is in test_gpu_net.py).
* Config 5: 16 boxes of 5,000 to 60,000 grains stepped as one ensemble, each
  bit-identical to its own run.
* Walls: grains pressed on the bottom, the four side boundaries and a top
  wall in each wall mode, the five walls of triaxial(), and the adaptive
  grid that follows the top wall.

Bar (the parity bar of all GPU tests): contact index lists and sort arrays
bit-exact; FP64 state within 1e-10 relative per step; the worst relative
difference is printed as the drift of each case."""
import os

import numpy
import pytest

import sphere_amd
from sphere_amd import workloads as W
from _cases import dense_case, pressed_case, settling_case
from _oracle import Oracle
from _parity import compare_contacts, compare_sort, compare_state

pytestmark = pytest.mark.gpu


def threads():
    return max(1, min(16, os.cpu_count() or 1))


def run_compare(eng, s, chunks, per_chunk, nc=32, fluid=False, label="", sort=False,
                contacts=True, each=None):
    eng.load(s.to_state())
    ora = Oracle(s.to_state(), threads=threads(), nc=nc)
    done, drift = 0, 0.0
    for _ in range(chunks):
        if each is not None and done == per_chunk:
            each(ora)
        eng.step(per_chunk)
        ora.step(per_chunk)
        done += per_chunk
        if sort:
            compare_sort(eng, ora)
        if contacts:
            compare_contacts(eng, ora, done)
        rep = compare_state(eng.read(), ora.read(), int(s.np[0]), done, fluid=fluid, label=label)
        drift = max(drift, max(rep.values()))
    print("%s: %d steps, largest relative difference from the oracle %.3g" % (label, done, drift))
    return ora


def test_config1_settling_10k_full_run():
    """Config 1 as specified: 10,000 grains, 2,000 steps, compared every 500."""
    s = settling_case(n=10000, seed=1, steps=2000)
    eng = sphere_amd.Engine(0)
    run_compare(eng, s, 4, 500, label="config 1 settling")
    eng.close()


def test_config2_shear_box_1m():
    """Config 2: the benchmark's packing, built as the benchmark builds it."""
    prep = sphere_amd.Engine(0, nc=64)
    s = W.prepare_shear_box(prep, 1_000_000, 2)
    prep.close()
    assert s.nb0[0] > 40000 and s.contactmodel[0] == 3 and s.periodic[0] == 1
    assert numpy.count_nonzero(s.fixvel > 0) > 1000 and s.wmode[0] == 1
    eng = sphere_amd.Engine(0, nc=64)
    ora = run_compare(eng, s, 2, 5, nc=64, label="config 2 shear box")
    h = ora.read().state
    assert h.w_mvfd[2] != 0.0  # the top wall carries load
    eng.close()


def test_config3_consolidation_200k():
    """Config 3: three fluid steps of the consolidation box, field by field,
    with the oracle's sweep counts."""
    prep = sphere_amd.Engine(0, nc=32)
    s = W.prepare_consolidation(prep, 200_000, 3)
    prep.close()
    assert list(s.num) == [48, 48, 96]
    eng = sphere_amd.Engine(0, nc=32)
    ora = run_compare(eng, s, 3, 1, fluid=True, sort=True, label="config 3 consolidation")
    g, o = eng.darcy_stats(), ora.darcy_stats()
    assert g.sweeps == o.sweeps and g.fluid_steps == 3 and g.sweeps > 30
    eng.close()


def test_config4_size_check_96x96x192():
    """The learned correction on a 96 x 96 x 192 grid (a pressure pulse
    relaxing under fixed top and bottom pressures, no grains): the solve
    with the network ends at the plain solve's pressure within the
    convergence tolerance."""
    s = sphere_amd.sim(np=0, nd=3, nw=0, sid="size96", fluid=True)
    s.defaultParams()
    s.defineWorldBoundaries([0.96, 0.96, 1.92], dx=0.01)
    assert list(s.num) == [96, 96, 192]
    s.initFluid(mu=1e-3, cfd_solver=1)
    s.setFluidCompressibility(1e-8)
    s.setPermeabilityPrefactor(1e-12, verbose=False)
    s.setTolerance(1e-8)
    s.setMaxIterations(200000)
    z = (numpy.arange(192) + 0.5) * 0.01
    s.p_f[:] = 1e3 * numpy.exp(-((z - 0.9) / 0.1) ** 2)[None, None, :]
    s.p_f[:, :, -1] = 0.0
    s.initTemporal(total=1.0, file_dt=1.0, dt=1e-4)
    P = sphere_amd.Net.param_count()
    params = numpy.random.default_rng(6).normal(0.0, 0.02, P).astype(numpy.float32)
    res = []
    for use in (False, True):
        e = sphere_amd.Engine(0, nc=32)
        if use:
            e.set_darcy_net(params, 4)
        e.load(s.to_state())
        e.step(2)
        res.append((e.read().get("p_f", 96 * 96 * 192, 1).copy(), e.darcy_stats()))
        e.close()
    (p0, st0), (p1, st1) = res
    assert st1.corrections > 0 and st0.nonconverged == 0 and st1.nonconverged == 0
    err = numpy.max(numpy.abs(p1 - p0)) / numpy.max(numpy.abs(p0))
    print("config 4 size check 96x96x192: sweeps plain %d, with the network %d; pressure "
          "difference %.3g" % (st0.sweeps, st1.sweeps, err))
    assert err < 1e-5


def test_config5_ensemble_16_boxes():
    """Config 5's structure at 16 boxes of 5,000 to 60,000 grains (log
    uniform), drawn friction, wall stress and driven-layer velocity: stepped
    as one ensemble, each box equals its single-box run bit for bit."""
    rng = numpy.random.default_rng(7)
    sizes = numpy.exp(rng.uniform(numpy.log(5000), numpy.log(60000), 16)).astype(int)
    boxes = []
    for k, n in enumerate(sizes):
        s = dense_case(n=int(n), seed=100 + k, model=3, periodic=1, E=7e9, nu=0.3, nw=1,
                       mu=float(rng.uniform(0.3, 0.7)))
        s.w_sigma0[0] = float(rng.uniform(5e3, 8e4))
        top = s.x[:, 2] > numpy.max(s.x[:, 2]) - 0.01
        s.fixvel[top] = 1.0
        s.vel[top, 0] = float(rng.uniform(1e-3, 1e-1))
        boxes.append(s)
    together = [sphere_amd.Engine(0, nc=64) for _ in boxes]
    alone = [sphere_amd.Engine(0, nc=64) for _ in boxes]
    for e, f, s in zip(together, alone, boxes):
        e.load(s.to_state())
        f.load(s.to_state())
    ens = sphere_amd.Ensemble(together)
    ens.step(20)
    for f in alone:
        f.step(20)
    for k, (e, f, s) in enumerate(zip(together, alone, boxes)):
        n = int(s.np[0])
        a, b = e.read(), f.read()
        for name, w in (("x", 4), ("vel", 4), ("angvel", 3), ("force", 3)):
            assert numpy.array_equal(a.get(name, n, w), b.get(name, n, w)), (k, name)
        assert numpy.array_equal(e.read_contacts()[0], f.read_contacts()[0]), k
    ens.close()
    for e in together + alone:
        e.close()


# ----------------------------------------------------------------------
# grains on walls
# ----------------------------------------------------------------------
@pytest.mark.parametrize("wmode", [0, 1, 2])
@pytest.mark.parametrize("mode", ["lists", "cells"])
def test_grains_pressed_on_walls(wmode, mode):
    """Every wall branch of the contact law (normal, viscous, frictional) on
    the bottom, the four side boundaries and the top wall, with the wall
    force partials non-zero from the first step (the wall integration sums
    them)."""
    s = pressed_case(n=1000, seed=wmode, wmode=wmode)
    eng = sphere_amd.Engine(0, skin=None if mode == "lists" else 0.0)
    seen = {}

    def after_first_chunk(ora):
        st = ora.read()
        seen["wall_force"] = st.state.w_mvfd[2]
        f = st.get("force", 1000, 3)
        x = st.get("x", 1000, 4)
        top = x[:, 2] + x[:, 3] > s.w_x[0]
        seen["top_push"] = float(numpy.max(-f[top, 2])) if top.any() else 0.0

    ora = run_compare(eng, s, 4, 25, label="pressed wmode %d (%s)" % (wmode, mode),
                      each=after_first_chunk)
    # the top grains are pushed down by the wall; a moving wall integrates
    # the (non-zero) sum of the grains' wall forces (a fixed wall, mode 0,
    # keeps its force entry, as in the reference)
    assert seen["top_push"] > 0.0
    if wmode != 0:
        assert seen["wall_force"] != 0.0
        assert ora.read().state.w_nx[3] != s.w_x[0]  # the wall moved
    eng.close()


def test_grains_pressed_hertz():
    s = pressed_case(n=729, seed=3, model=3, E=7e9, wmode=1)
    eng = sphere_amd.Engine(0)
    run_compare(eng, s, 2, 25, label="pressed hertz")
    eng.close()


def test_triaxial_five_walls():
    """triaxial(): the top wall moves down at a set velocity and four side
    walls hold a normal stress, all fitted to the grains."""
    s = pressed_case(n=1000, seed=9, wmode=0)
    s.triaxial(wvel=-0.05, normal_stress=2e4)
    assert s.nw[0] == 5
    eng = sphere_amd.Engine(0)
    ora = run_compare(eng, s, 4, 25, label="triaxial")
    h = ora.read().state
    assert h.w_nx[3] < s.w_x[0]  # the top wall moved down
    assert any(h.w_nx[4 * k + 3] != s.w_x[k] for k in range(1, 5))  # side walls moved
    eng.close()


def test_adaptive_grid_follows_wall():
    """adaptive = 1: the grid's height follows the top wall every step
    (src/device.cu:370-400)."""
    s = pressed_case(n=1000, seed=4, wmode=2, wall_vel=-0.05)
    s.adaptive[0] = 1
    eng = sphere_amd.Engine(0)
    ora = run_compare(eng, s, 2, 25, sort=True, label="adaptive")
    assert ora.read().state.L[2] != s.L[2]
    eng.close()
