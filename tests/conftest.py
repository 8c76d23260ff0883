"""Shared pytest setup: the `gpu` marker and loaders for the checker."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def bridge():
    """The product bridge with a live device; GPU tests fail (not skip) without it."""
    from truepix_amd import native_fp2_bridge as nb
    if not nb.AVAILABLE:
        raise RuntimeError("HIP library not active on this GPU test run: " + nb.GPU_FAIL_REASON)
    return nb


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("torch sees no GPU on a GPU test run")
    return torch
