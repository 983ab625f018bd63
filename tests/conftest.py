import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle import bindings

    if not os.path.exists(bindings.ORACLE_SO):
        bindings.ensure_built(ref=False)
    return bindings.Oracle()


@pytest.fixture(scope="session")
def ref():
    from oracle import bindings

    if not bindings.ref_available():
        if os.path.isdir(os.path.join(bindings.REF_SRC, "src")):
            bindings.ensure_built(ref=True)
    if not bindings.ref_available():
        pytest.skip("compiled reference (oracle/_ref) not available")
    return bindings.Ref()


@pytest.fixture(scope="session")
def tables():
    import json

    with open(os.path.join(ROOT, "tests", "golden", "reference_tables.json")) as f:
        return json.load(f)
