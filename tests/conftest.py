import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FIXTURES = os.path.join(ROOT, "tests", "fixtures")
REFERENCE_MODELS = ["cellcycle", "metastasis", "sizek", "Montagud2022_Prostate_Cancer"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (gfx950) and the built libmbg.so")


def model_text(name):
    with open(os.path.join(FIXTURES, name + ".bnd")) as f:
        b = f.read()
    with open(os.path.join(FIXTURES, name + ".cfg")) as f:
        c = f.read()
    return b, c


@pytest.fixture(scope="session")
def gpu_ctx():
    from upmabossg_amd import Context

    ctx = Context(0)
    yield ctx
    ctx.close()
