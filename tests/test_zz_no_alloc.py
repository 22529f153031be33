"""sph_net_conv3d_noalloc (no device allocation; scratch in a static device
array) against sph_net_conv3d, through the C ABI: two calls in a row, the
second on a smaller grid with fewer channels, each bit-identical to the
allocating entry point's result (same kernels, same operands); and a call
whose staging would not fit the scratch is refused with an error."""
import ctypes

import numpy
import pytest

from sphere_amd import _abi

pytestmark = pytest.mark.gpu

SCRATCH = 16 << 20


def _conv_pair(torch, lib, shape, cin, cout, wrap, seed):
    rng = numpy.random.default_rng(seed)
    nz, ny, nx = shape
    V = nx * ny * nz
    dev = torch.device("cuda:0")
    x = torch.tensor(rng.standard_normal((V, cin)).astype(numpy.float32), device=dev)
    w = torch.tensor((rng.standard_normal((27, cout, cin)) / numpy.sqrt(27 * cin)).astype(numpy.float32),
                     device=dev)
    b = torch.tensor(rng.standard_normal(cout).astype(numpy.float32), device=dev)
    ref = torch.empty((V, cout), device=dev)
    got = torch.full((V, cout), float("nan"), device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    args = (ctypes.c_void_p(x.data_ptr()), cin, ctypes.c_void_p(w.data_ptr()),
            ctypes.c_void_p(b.data_ptr()), cout, nx, ny, nz, wrap)
    torch.cuda.synchronize()
    print("no-alloc case", shape, cin, cout, wrap, flush=True)
    assert lib.sph_net_conv3d(*args, ctypes.c_void_p(ref.data_ptr()), stream) == 0
    rc = lib.sph_net_conv3d_noalloc(*args, ctypes.c_void_p(got.data_ptr()), stream)
    assert rc == 0, lib.sph_net_last_error().decode()
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_conv_equals_allocating_conv():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    lib = _abi.load_library()
    # a 48 x 48 x 32 grid with 16 input channels (4.7 MB of staged input),
    # then a smaller one with 3 input channels: the second call reuses the
    # scratch the first one filled
    for shape, cin, cout, wrap, seed in (((32, 48, 48), 16, 32, 3, 1), ((8, 16, 12), 3, 16, 1, 2)):
        ref, got = _conv_pair(torch, lib, shape, cin, cout, wrap, seed)
        assert numpy.all(numpy.isfinite(ref))
        assert numpy.array_equal(ref, got), (shape, numpy.max(numpy.abs(ref - got)))


def test_noalloc_conv_refuses_oversized_input():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    lib = _abi.load_library()
    nx = ny = nz = 128  # 2 Mi voxels x 16 channels x 2 B = 64 MiB > 16 MiB
    dev = torch.device("cuda:0")
    x = torch.zeros(1, device=dev)
    rc = lib.sph_net_conv3d_noalloc(ctypes.c_void_p(x.data_ptr()), 16, ctypes.c_void_p(x.data_ptr()),
                                    None, 16, nx, ny, nz, 0, ctypes.c_void_p(x.data_ptr()),
                                    ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == -1  # SPH_ERR_ARG, before any launch
    msg = lib.sph_net_last_error().decode()
    assert "scratch" in msg and str(SCRATCH) in msg
