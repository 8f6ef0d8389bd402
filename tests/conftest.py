import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "qmtools-pycuda_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
BASIS = os.path.join(PKG, "qmtools", "data", "cc-pvdz.bin")
MOL29766 = os.path.join(GOLDEN, "molecule_29766_0")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built native library")


def pytest_collection_modifyitems(config, items):
    have_gpu = None
    for item in items:
        if "gpu" in item.keywords:
            if have_gpu is None:
                try:
                    import torch
                    have_gpu = torch.cuda.is_available()
                except Exception:
                    have_gpu = False
            if not have_gpu:
                item.add_marker(pytest.mark.skip(reason="no HIP device"))


@pytest.fixture(scope="session")
def basis():
    from qmtools import BasisSet
    return BasisSet(BASIS)


@pytest.fixture(scope="session")
def oracle_basis():
    from oracle.basisfile import read_basis
    return read_basis(BASIS)
