"""Binary state format: the sim reader on the reference's own v2.11 output
(tests/golden, copied by make_fixtures.sh), and v2.15 write/read round trips
(python/sphere.py:965-1553, src/file_io.cpp:30-731)."""
import os

import numpy
import pytest

import sphere_amd
from _cases import add_bonds, darcy_case, dense_case

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CUBE = os.path.join(GOLDEN, "cube-init.output00029.bin")


def test_legacy_v211_fixture_parses():
    """cube-init.output00029.bin: 1,000 grains written by the reference
    (format 2.11: no adaptive flag, no E).  Every field lands where the
    version switches put it; the file is consumed exactly."""
    s = sphere_amd.sim()
    s.readbin(CUBE, verbose=False)
    assert s.np[0] == 1000 and s.nd[0] == 3
    assert numpy.allclose(s.L, [0.15112697, 0.15112697, 5.6420735])
    assert list(s.num) == [3, 3, 112]
    assert s.periodic[0] == 1
    assert 0.008 <= s.radius.min() and s.radius.max() <= 0.012
    assert s.k_n[0] == 1e8 and s.contactmodel[0] == 2 and s.mu_s[0] == 0.5
    assert abs(s.time_current[0] - 5.8) < 1e-6
    assert s.nw[0] == 0 and s.nb0[0] == 0
    assert s.adaptive[0] == 0 and s.E[0] == 0.0  # absent before 2.14 / 2.13
    # grains lie inside the box and rest: a settled packing
    assert numpy.all(s.x >= 0.0) and numpy.all(s.x <= s.L)
    assert numpy.max(numpy.abs(s.vel)) < 1e-3
    # reading resets the version to the current one (sphere.py:1351)
    assert s.version[0] == sphere_amd.VERSION


def test_status_fixture_format():
    """<sid>.status.dat holds '%2.4e %2.4e %d' (time, percent, file index),
    src/sphere.cpp:771-780."""
    text = open(os.path.join(GOLDEN, "cube-init.status.dat")).read()
    t, pct, idx = text.split()
    assert "%2.4e %2.4e %d" % (float(t), float(pct), int(idx)) == text.strip()


def roundtrip(s, tmp_path):
    s.writebin(folder=str(tmp_path), verbose=False)
    r = sphere_amd.sim(fluid=bool(s.fluid))
    r.readbin(os.path.join(str(tmp_path), s.sid + ".bin"), verbose=False)
    return r


def test_roundtrip_fixture(tmp_path):
    s = sphere_amd.sim()
    s.readbin(CUBE, verbose=False)
    s.sid = "cube"
    r = roundtrip(s, tmp_path)
    assert r == s, r.differences(s)
    size = os.path.getsize(os.path.join(str(tmp_path), "cube.bin"))
    # v2.15 adds adaptive (4 B) and E (8 B) to the 2.11 layout
    assert size == os.path.getsize(CUBE) + 12


def test_roundtrip_bonds_walls(tmp_path):
    s = dense_case(n=512, seed=3, model=3, periodic=1, nw=1, E=7e9)
    add_bonds(s, 0.2, seed=1)
    assert s.nb0[0] > 10
    r = roundtrip(s, tmp_path)
    assert r == s, r.differences(s)


def test_roundtrip_darcy(tmp_path):
    s = darcy_case(n=200, seed=2, nxyz=(4, 5, 6), wall=True)
    s.p_f[:] = numpy.random.default_rng(1).uniform(0, 1e3, s.p_f.shape)
    r = roundtrip(s, tmp_path)
    assert r == s, r.differences(s)


def test_state_image_roundtrip():
    """sim -> sph_state image -> sim keeps every field (the boundary the
    C ABI reads)."""
    s = dense_case(n=343, seed=4, model=2, periodic=2, nw=1)
    add_bonds(s, 0.3, seed=2)
    buf = s.to_state()
    r = sphere_amd.sim()
    r.from_state(buf)
    assert r == s, r.differences(s)


def test_wrong_file_rejected(tmp_path):
    p = os.path.join(str(tmp_path), "junk.bin")
    with open(p, "wb") as f:
        f.write(b"\x00" * 7)
    s = sphere_amd.sim()
    with pytest.raises(Exception):
        s.readbin(p, verbose=False)
