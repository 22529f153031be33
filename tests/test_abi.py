"""The C-ABI library: every function include/sphere_hip.h declares is
exported by libsphere_hip.so, and the ctypes layouts match the header.
No GPU call is made (loading the library does not touch the device)."""
import ctypes
import os
import re

import pytest

import sphere_amd
from sphere_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sphere_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(sph_[a-z_0-9]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("sph_create", "sph_load_state", "sph_step", "sph_run", "sph_read_state",
              "sph_write_file", "sph_read_contacts", "sph_set_neighbor_lists"):
        assert n in names


def test_library_exports_every_declared_function():
    lib = _abi.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, "not exported: %s" % missing


def test_ctypes_signatures_cover_header():
    """Every declared function has a ctypes signature in _abi (so none is
    called with default int conversions)."""
    lib = _abi.load_library()
    for n in declared_functions():
        f = getattr(lib, n)
        assert f.argtypes is not None, n


def struct_fields(name):
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    m = re.search(r"typedef struct\s*(?:%s\s*)?\{([^{}]*)\}\s*%s\s*;" % (name, name), text,
                  flags=re.S)
    assert m, name
    fields = []
    for decl in m.group(1).split(";"):
        decl = decl.strip()
        if not decl:
            continue
        typ, rest = decl.split(None, 1) if not decl.startswith("const ") else \
            ("const " + decl.split(None, 2)[1], decl.split(None, 2)[2])
        for part in rest.split(","):
            part = part.strip().lstrip("*").strip()
            fields.append(re.sub(r"\[.*\]", "", part))
    return fields


@pytest.mark.parametrize("cname,pyclass", [("sph_state", _abi.SphState),
                                           ("sph_run_opts", _abi.SphRunOpts),
                                           ("sph_darcy_stats", _abi.SphDarcyStats),
                                           ("sph_neighbor_stats", _abi.SphNeighborStats),
                                           ("sph_ns_params", _abi.SphNsParams),
                                           ("sph_ns_stats", _abi.SphNsStats)])
def test_struct_field_order(cname, pyclass):
    assert struct_fields(cname) == [f[0] for f in pyclass._fields_]


def test_kernel_names_match_ids():
    text = open(HEADER).read()
    count = int(re.search(r"#define SPH_K_COUNT (\d+)", text).group(1))
    assert len(_abi.K_NAMES) == count


def test_engine_fails_loudly_without_device():
    """Without a GPU the engine raises instead of falling back to the CPU."""
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(sphere_amd.SphereError):
        sphere_amd.Engine(0)
