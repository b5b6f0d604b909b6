import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "lbfs-ibm-ade-sp_amd")
for p in (REPO, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _ensure_built():
    from lbfs_amd import abi
    from oracle import oracle as orc
    if not os.path.exists(orc.LIB_PATH):
        from oracle import build as ob
        ob.build()
    if not os.path.exists(abi.LIB_PATH):
        from lbfs_amd import build as lb
        lb.build()


_ensure_built()


@pytest.fixture(scope="session")
def has_gpu():
    import torch
    return torch.cuda.is_available()
