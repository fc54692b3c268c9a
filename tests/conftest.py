import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "ref: needs the reference Fortran kernels built into oracle/_ref")


def _ensure_oracle():
    from oracle import binding

    if not binding.ORACLE_SO.exists():
        binding.build(with_ref=False)
    return binding


@pytest.fixture(scope="session")
def oracle():
    binding = _ensure_oracle()
    return binding.Oracle()


@pytest.fixture(scope="session")
def reflibs():
    from oracle import binding

    if not binding.RefLibs.available():
        if Path("/root/reference/kernels/NICAM").is_dir():
            binding.build(with_ref=True)
        else:
            pytest.skip("reference kernels (oracle/_ref) not built and reference tree absent")
    return binding.RefLibs()


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on a real GPU; loads the in-tree libnicam_tt.so."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from icoatmosbenchmark_v1_amd import runtime

    if not runtime.LIB_PATH.exists():
        from icoatmosbenchmark_v1_amd import _build

        _build.build_library()
    runtime.load_library()
    return torch.device("cuda:0")
