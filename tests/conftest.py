import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def has_gpu():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def golden():
    return GOLDEN


@pytest.fixture(scope="session")
def tiny_newick():
    return (GOLDEN / "tiny.nwk").read_text().strip()


@pytest.fixture(scope="session")
def tiny_fasta():
    return str(GOLDEN / "tiny.fa")


@pytest.fixture(scope="session")
def flu_newick():
    return (GOLDEN / "fluA.tree").read_text().strip()


@pytest.fixture(scope="session")
def flu_fasta():
    return str(GOLDEN / "fluA.fa")


@pytest.fixture
def gpu():
    if not has_gpu():
        pytest.skip("no HIP device")
    import torch

    return torch.device("cuda", 0)
