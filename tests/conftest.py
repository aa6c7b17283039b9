import os
import struct
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


def read_fortran_unformatted(path):
    raw = open(path, "rb").read()
    recs, p = [], 0
    while p < len(raw):
        n = struct.unpack_from("<i", raw, p)[0]
        recs.append(raw[p + 4:p + 4 + n])
        p += 8 + n
    return recs


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
    recs = read_fortran_unformatted(os.path.join(d, "REVIVE"))
    return dict(field=field, config=config, control=control, kat=kat,
                revive=recs, dir=d)


def revive_config(ex):
    """CONFIG of the example with the REVIVE guest coordinates loaded."""
    import copy
    field = copy.deepcopy(ex["field"])
    config = copy.deepcopy(ex["config"])
    recs = ex["revive"]
    nummols = np.frombuffer(recs[1], "<i4")
    gx, gy, gz = [np.frombuffer(recs[4 + i], "<f8") for i in range(3)]
    field.molecules[0].nummols = int(nummols[0])
    config.coords[0] = list(zip(gx, gy, gz))
    energy = float(np.frombuffer(recs[-1], "<f8")[0])
    return field, config, energy
