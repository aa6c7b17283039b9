import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fastmc-1.4.0_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        import subprocess
        subprocess.check_call(["make", "-s", "-C", ORACLE])
    return pyoracle


@pytest.fixture(scope="session")
def restart_example():
    """The reference's completed example run (data fixtures)."""
    from fastmc_amd import inputs
    import json
    d = os.path.join(GOLDEN, "restart_example")
    field = inputs.read_field(os.path.join(d, "FIELD"))
    config = inputs.read_config(os.path.join(d, "CONFIG"), field)
    control = inputs.read_control(os.path.join(d, "CONTROL"))
    kat = json.load(open(os.path.join(d, "known_answers.json")))
    rev = inputs.read_revive(os.path.join(d, "REVIVE"), field)
    return dict(field=field, config=config, control=control, kat=kat,
                revive=rev, dir=d)


def revive_config(ex):
    """CONFIG of the example with the REVIVE guest coordinates loaded, and
    the REVIVE's accumulated energy."""
    from fastmc_amd import inputs
    field, config = inputs.apply_revive(ex["field"], ex["config"], ex["revive"])
    return field, config, ex["revive"].energy[0]
