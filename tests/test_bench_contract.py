"""CPU checks of the benchmark's reporting: the committed traffic summary
parses to the contact stage's bytes, and a run asked for more than one GPU
stops with a message before touching a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_measured_traffic_parses():
    import bench
    t = bench.measured_traffic()
    assert t is not None and 1e8 < t < 1e10, t


def test_contact_bytes_formula():
    import bench
    assert bench.contact_bytes_per_particle(0.0) == 196.0
    assert bench.contact_bytes_per_particle(4.0) == 196.0 + 208.0


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, universal_newlines=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
