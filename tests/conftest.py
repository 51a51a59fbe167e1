import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through librte.so on the device)")


@pytest.fixture(scope="session")
def ctx():
    from doi_amd import rte
    c = rte.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def golden_ref():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "ref_cpu_uniform.json")) as f:
        return json.load(f)["cases"]


@pytest.fixture(scope="session")
def golden_mod():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "ref_cpu_modulated.json")) as f:
        return json.load(f)["cases"]
