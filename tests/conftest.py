import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device and the built HIP library")
    # autotune picks go to a file of this session, so no test reads picks an earlier run left
    if "ZS_AUTOTUNE_CACHE" not in os.environ:
        import tempfile
        os.environ["ZS_AUTOTUNE_CACHE"] = os.path.join(tempfile.mkdtemp(prefix="zs_tests_"), "autotune.txt")


def _load_pkg():
    import __graft_entry__
    return __graft_entry__.load_package()


_load_pkg()


@pytest.fixture(scope="session")
def zctx():
    """One zs_hip context on device 0 for the whole GPU session."""
    import zig_stark_amd as Z
    ctx = Z.Context(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def torch_dev():
    import torch
    dev = torch.device("cuda:0")
    torch.zeros(1, device=dev)  # raises with the runtime's reason if the device is unusable
    return dev


def to_dev(arr, device):
    """numpy uint32 -> torch int32 tensor on the device (same bits)."""
    import numpy as np
    import torch
    return torch.from_numpy(np.ascontiguousarray(arr, dtype=np.uint32).view(np.int32)).to(device)


def from_dev(t):
    import numpy as np
    return t.cpu().numpy().view(np.uint32)
