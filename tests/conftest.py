"""Test configuration: the `gpu` marker (tests that need a HIP device and libtnkit_hip.so)."""
import os
import sys

# the oracle's float sequence (and so its matvec counts) is compared with tnkit runs recorded
# on one BLAS thread
for _v in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ.setdefault(_v, "1")

import pytest  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP (MI355X) device and the built library")


@pytest.fixture(scope="session")
def hip():
    """The HIP device; a gpu-marked test fails (not skips) when there is none."""
    import torch
    from tnkit_amd import _lib
    _lib.load()
    assert torch.cuda.is_available(), "gpu test without a HIP device"
    return torch.device("cuda", 0)
