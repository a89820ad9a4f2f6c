import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

REF = os.path.join(ROOT, "tests", "golden", "reference")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long CPU-only check, run with PAMPA_SLOW=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("PAMPA_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow oracle check; set PAMPA_SLOW=1 to run")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def golden_keff():
    """keff golden values from the reference's check_ref.txt, by case."""
    import re
    text = open(os.path.join(REF, "check_ref.txt")).read()
    out = {}
    for block in text.split("\nRun ")[1:]:
        name = block.split("...")[0]
        m = re.search(r"Effective multiplication factor: ([0-9.]+)\.", block)
        if m:
            out[name] = float(m.group(1))
    return out


@pytest.fixture(scope="session")
def pk():
    import pampa_amd
    pampa_amd.lib()
    pampa_amd.initialize(0)
    return pampa_amd
