"""Ensembles: boxes of different sizes, contact laws and boundaries stepped
together (all boxes' kernels enqueued before the host waits) give, box by
box, exactly the state that sph_step gives for that box alone (bit for bit,
every array)."""
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
    boxes = _boxes()
    singles = []
    for s in boxes:
        e = sphere_amd.Engine(0, nc=64)
        e.load(s.to_state())
        singles.append(e)
    engines = []
    for s in boxes:
        e = sphere_amd.Engine(0, nc=64)
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
    assert ens.stats()["steps"] == 26 + 10 + 5 + 8
    ens.close()


def test_ensemble_refuses_fluid_boxes():
    from _cases import darcy_case
    e = sphere_amd.Engine(0)
    e.load(darcy_case(n=50, nxyz=(4, 4, 8)).to_state())
    ens = sphere_amd.Ensemble([e])
    with pytest.raises(sphere_amd.SphereError):
        ens.step(4)
    ens.close()
