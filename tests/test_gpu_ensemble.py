"""Ensembles: boxes of different sizes, contact laws and boundaries stepped
together give, box by box, exactly the state that sph_step gives for that
box alone (bit for bit, every array).  Boxes on the list path share one
launch per stage (pairs of steps replayed from a captured HIP graph, or
enqueued step by step); the others run box by box beside them.  The skin
autotuner changes no result."""
import numpy
import pytest

import sphere_amd
from _cases import add_bonds, dense_case, settling_case

pytestmark = pytest.mark.gpu

FIELDS = [("x", 4), ("vel", 4), ("angvel", 3), ("force", 3), ("torque", 3), ("angpos", 3),
          ("xyzsum", 3), ("es_dot", 1), ("es", 1), ("ev_dot", 1), ("ev", 1), ("p", 1)]


def _boxes():
    out = []
    s = dense_case(n=512, seed=0, model=3, periodic=1, nw=1, E=7e9)
    out.append(s)
    s = dense_case(n=729, seed=1, model=2, periodic=0, nw=0)
    add_bonds(s, 0.2, 3)
    out.append(s)
    s = dense_case(n=343, seed=2, model=3, periodic=2, nw=1, E=1e9, mu=0.3)
    out.append(s)
    out.append(settling_case(n=1500, seed=4))
    for s in out:
        s.time_file_dt[0] = 1.0
        s.time_total[0] = 1.0
    return out


def _same(a, b, n, label):
    for name, w in FIELDS:
        assert numpy.array_equal(a.get(name, n, w), b.get(name, n, w)), (label, name)


def test_ensemble_matches_single_boxes():
    """Four list-path boxes and one box searched cell by cell (stepped box
    by box beside the multi-box launches)."""
    boxes = _boxes()
    boxes.append(boxes[1])
    skins = [None] * 4 + [0.0]
    singles = []
    for s, sk in zip(boxes, skins):
        e = sphere_amd.Engine(0, nc=64, skin=sk)
        e.load(s.to_state())
        singles.append(e)
    engines = []
    for s, sk in zip(boxes, skins):
        e = sphere_amd.Engine(0, nc=64, skin=sk)
        e.load(s.to_state())
        engines.append(e)
    ens = sphere_amd.Ensemble(engines)
    for n in (26, 10, 5, 8):
        print("ensemble steps", n, flush=True)
        ens.step(n)
        for k, (e, ref) in enumerate(zip(engines, singles)):
            ref.step(n)
            npart = int(boxes[k].np[0])
            _same(e.read(), ref.read(), npart, "box %d after %d more steps" % (k, n))
            assert numpy.array_equal(e.read_contacts()[0], ref.read_contacts()[0]), k
            assert e.header().time_current == ref.header().time_current
        if n == 10:
            # a box stepped on its own between ensemble calls (the
            # ensemble then plans its launches again)
            engines[0].step(1)
            singles[0].step(1)
    info = ens.info()
    print("ensemble info", info)
    assert info["boxes"] == 5 and info["multi_boxes"] == 4 and info["graph_replays"] > 0
    assert ens.stats()["steps"] == 26 + 10 + 5 + 8
    ens.close()


def _shear_boxes(n, seed):
    rng = numpy.random.default_rng(seed)
    out = []
    for k in range(n):
        npart = int(rng.integers(400, 1600))
        s = dense_case(n=npart, seed=seed + k, model=3 if k % 3 else 2, periodic=1 + (k % 2),
                       nw=1, E=7e9, mu=float(rng.uniform(0.3, 0.7)))
        top = s.x[:, 2] > numpy.max(s.x[:, 2]) - 0.01
        s.fixvel[top] = 1.0
        s.vel[top, 0] = float(rng.uniform(1e-3, 1e-1))
        s.w_sigma0[0] = float(rng.uniform(5e3, 8e4))
        if k % 4 == 1:
            add_bonds(s, 0.2, seed + k)
        s.time_file_dt[0] = 1.0
        s.time_total[0] = 1.0
        out.append(s)
    return out


def _engines(boxes):
    out = []
    for s in boxes:
        e = sphere_amd.Engine(0, nc=64)
        e.load(s.to_state())
        out.append(e)
    return out


def test_ensemble_graph_and_stepwise_launches_equal():
    """Ten list-path boxes (both laws, periodic 1 and 2, bonds on some):
    replayed graph, step-by-step multi-box launches and single-box runs give
    the same bits."""
    boxes = _shear_boxes(10, 11)
    a, b, c = _engines(boxes), _engines(boxes), _engines(boxes)
    ea, eb = sphere_amd.Ensemble(a), sphere_amd.Ensemble(b)
    eb.set_graph(False)
    for n in (1, 7, 12):
        ea.step(n)
        eb.step(n)
        for e in c:
            e.step(n)
    ia, ib = ea.info(), eb.info()
    print("graph", ia, "stepwise", ib)
    assert ia["multi_boxes"] == 10 and ia["graph"] == 1 and ia["graph_replays"] > 0
    assert ib["graph"] == 0 and ib["multi_steps"] == 20
    assert ia["launches_per_step"] <= 12
    for k, s in enumerate(boxes):
        n = int(s.np[0])
        ra, rb, rc = a[k].read(), b[k].read(), c[k].read()
        _same(ra, rc, n, "graph box %d" % k)
        _same(rb, rc, n, "stepwise box %d" % k)
        assert numpy.array_equal(a[k].read_contacts()[0], c[k].read_contacts()[0]), k
        assert numpy.array_equal(a[k].read_contacts()[2], c[k].read_contacts()[2]), k
    ea.close()
    eb.close()


def test_ensemble_autotune_changes_no_result():
    """The skin autotuner times every box with each fraction; its trial steps
    are ordinary steps, and the tuned ensemble stays bit-identical to the
    boxes run alone with the default skin."""
    boxes = _shear_boxes(6, 23)
    a, c = _engines(boxes), _engines(boxes)
    ens = sphere_amd.Ensemble(a)
    fr = [0.05, 0.1, 0.3]
    chosen, ms = ens.autotune(4, fr)
    print("chosen", chosen, "ms per step", ms)
    assert all(x in fr for x in chosen)
    assert numpy.all(ms > 0.0)
    ens.step(9)
    for e in c:
        e.step(4 * len(fr) + 9)
    for k, s in enumerate(boxes):
        n = int(s.np[0])
        _same(a[k].read(), c[k].read(), n, "tuned box %d" % k)
        assert numpy.array_equal(a[k].read_contacts()[0], c[k].read_contacts()[0]), k
    assert ens.stats()["steps"] == 4 * len(fr) + 9
    ens.close()


FLUID = [("p_f", 1), ("v_f", 3), ("phi", 1), ("dphi", 1), ("f_p", 3)]


def test_ensemble_darcy_boxes_match_single_boxes():
    """Darcy-coupled boxes of different sizes, boundaries and walls stepped
    as one ensemble (one launch per stage of the fluid step, the sweeps of
    all solves batched) give the single-box states bit for bit, fluid
    fields and sweep counts included; two boxes have a learned network
    (their forward passes as multi-box launches) and a DEM box rides
    along."""
    from _cases import darcy_case
    from sphere_amd import learn
    boxes = [darcy_case(n=300, seed=1, nxyz=(6, 6, 12)),
             darcy_case(n=500, seed=2, nxyz=(8, 8, 16), wall=True),
             darcy_case(n=200, seed=3, nxyz=(6, 6, 12), bc_bot=0, bc_top=1),
             darcy_case(n=400, seed=4, nxyz=(8, 6, 12), periodic=2),
             darcy_case(n=250, seed=5, nxyz=(6, 6, 12), bc_top=4, tolerance=1e-7),
             dense_case(n=343, seed=6, model=3, periodic=1, nw=1, E=7e9),
             darcy_case(n=300, seed=7, nxyz=(8, 8, 16))]
    boxes[4].bc_top_flux[0] = 1e-6
    for s in boxes:
        s.time_file_dt[0] = 1.0
        s.time_total[0] = 1.0
    params = learn.init_params(3)

    def make():
        out = []
        for k, s in enumerate(boxes):
            e = sphere_amd.Engine(0, nc=64)
            if k in (1, 6):
                e.set_darcy_net(params, 4)
            e.load(s.to_state())
            out.append(e)
        return out

    together, alone = make(), make()
    ens = sphere_amd.Ensemble(together)
    for n in (5, 4, 3):
        ens.step(n)
        for e in alone:
            e.step(n)
    info = ens.info()
    print("ensemble info", info)
    assert info["fluid_boxes"] == 6 and info["multi_boxes"] == 1 and info["fluid_steps"] == 12
    for k, (a, b, s) in enumerate(zip(together, alone, boxes)):
        n = int(s.np[0])
        ra, rb = a.read(), b.read()
        _same(ra, rb, n, "box %d" % k)
        assert numpy.array_equal(a.read_contacts()[0], b.read_contacts()[0]), k
        if s.fluid:
            V = int(numpy.prod(s.num))
            for name, w in FLUID:
                m = n if name == "f_p" else V
                assert numpy.array_equal(ra.get(name, m, w), rb.get(name, m, w)), (k, name)
            sa, sb = a.darcy_stats(), b.darcy_stats()
            assert (sa.sweeps, sa.fluid_steps, sa.nonconverged, sa.corrections) == (
                sb.sweeps, sb.fluid_steps, sb.nonconverged, sb.corrections), k
            print("box %d: %d sweeps in %d fluid steps" % (k, sa.sweeps, sa.fluid_steps))
    ens.close()
