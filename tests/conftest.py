import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")
    # torch carries its own HIP runtime; the network tests use torch for
    # device memory, so in any run that may hold GPU tests it is initialised
    # before the library is loaded and both then share that runtime
    expr = config.getoption("markexpr", "") or ""
    if "not gpu" not in expr:
        try:
            import torch
            torch.cuda.is_available()
        except ImportError:
            pass
