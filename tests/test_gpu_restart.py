"""Restart fidelity through the drop-in path.  A run continued from one of
its own output files (readbin of output k, then run()) writes the same
later output files, byte for byte, as the uninterrupted run: the contact
history, TY3's previous accelerations and the time loop's counters come
from the side car sph_run writes beside every output file.  Without the side
car the continued run differs, as the reference's does."""
import os

import numpy
import pytest

import sphere_amd
from _cases import add_bonds, darcy_case, dense_case

pytestmark = pytest.mark.gpu


@pytest.fixture
def rundir(tmp_path):
    for d in ("input", "output", "work"):
        (tmp_path / d).mkdir()
    old = os.getcwd()
    os.chdir(tmp_path / "work")
    yield tmp_path
    os.chdir(old)


def out_bytes(sid, k):
    with open("../output/%s.output%05d.bin" % (sid, k), "rb") as fh:
        return fh.read()


def continue_from(sid, k, new_sid, total, fluid):
    s = sphere_amd.sim(sid=sid, fluid=fluid)
    s.readbin("../output/%s.output%05d.bin" % (sid, k), verbose=False)
    s.time_total[0] = total
    s.id(new_sid)
    assert s.run(verbose=False) == 0
    return s


def check_restart(s, fluid, every, outputs, restart_at):
    dt = float(s.time_dt[0])
    s.time_file_dt[0] = (every - 0.5) * dt
    s.time_total[0] = (every * outputs - 0.5) * dt
    total = float(s.time_total[0])
    assert s.run(verbose=False) == 0
    assert s.status() == outputs
    for k in range(1, outputs + 1):
        assert os.path.exists("../output/%s.contacts.%05d.bin" % (s.sid, k))
    continue_from(s.sid, restart_at, "continued", total, fluid)
    for k in range(restart_at + 1, outputs + 1):
        assert out_bytes("continued", k) == out_bytes(s.sid, k), "output %d differs" % k
    # the same continuation without the side car
    os.remove("../output/%s.contacts.%05d.bin" % (s.sid, restart_at))
    continue_from(s.sid, restart_at, "plain", total, fluid)
    assert out_bytes("plain", outputs) != out_bytes(s.sid, outputs)


def test_restart_dem_bit_identical(rundir):
    """Hertz-Mindlin with friction, bonds, a stress-driven top wall and
    gravity: 5 files of 30 steps, continued from the second."""
    s = dense_case(n=729, seed=91, model=3, periodic=1, E=7e9, nu=0.3, nw=1, gamma_n=0.0)
    add_bonds(s, 0.2, seed=8)
    s.sid = "restart_dem"
    check_restart(s, False, 30, 5, 2)


def test_restart_darcy_bit_identical(rundir):
    """The Darcy-coupled box with a wall: 4 files of 10 fluid steps,
    continued from the second."""
    s = darcy_case(n=300, seed=12, nxyz=(6, 6, 12), wall=True, tolerance=1e-8)
    s.sid = "restart_darcy"
    check_restart(s, True, 10, 4, 2)


def test_engine_restart_roundtrip(rundir):
    """sph_write_restart / sph_read_restart on contexts: a state loaded from
    a written file plus its side car steps like the original."""
    s = dense_case(n=512, seed=93, model=2, periodic=0, nw=1, gamma_n=20.0, gamma_t=10.0)
    a = sphere_amd.Engine(0)
    a.load(s.to_state())
    a.step(37)
    a.write_file("../output/mid.bin")
    a.write_restart("../output/mid.contacts.bin")
    b = sphere_amd.Engine(0)
    b.load_file("../output/mid.bin", 0)
    b.read_restart("../output/mid.contacts.bin")
    a.step(25)
    b.step(25)
    n = 512
    ra, rb = a.read(), b.read()
    for name, w in (("x", 4), ("vel", 4), ("angvel", 3), ("force", 3), ("torque", 3)):
        assert numpy.array_equal(ra.get(name, n, w), rb.get(name, n, w)), name
    assert numpy.array_equal(a.read_contacts()[0], b.read_contacts()[0])
    # a side car of another state is refused
    c = sphere_amd.Engine(0)
    c.load(s.to_state())
    with pytest.raises(sphere_amd.SphereError):
        c.read_restart("../output/mid.contacts.bin")
    for e in (a, b, c):
        e.close()
