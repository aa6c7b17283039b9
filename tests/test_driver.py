"""The Fortran driver (fmc_gcmc, the gcmc.f move loop) over the C ABI.

CPU: the driver linked against the oracle ABI reproduces, move for move, the
C oracle's own restatement of the move loop (same duni seed): identical
move kinds, accept/reject decisions and loadings, identical estep.
GPU: the driver linked against libfastmc_amd.so reproduces the oracle-linked
driver's accept/reject sequence exactly, with every estep within 1e-10
relative."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ORACLE, PKG

DRV_ORACLE = os.path.join(ORACLE, "_build", "fmc_gcmc_oracle")
DRV_GPU = os.path.join(PKG, "bin", "fmc_gcmc")


def _workload(name):
    from fastmc_amd import synth
    import numpy as np_
    if name == "c1":
        return synth.c1(), [], 2000, 1000
    # a porous two-guest system with translate/rotate-only moves
    rng = np_.random.default_rng(21)
    fw = synth.make_framework(rng, (26.0, 26.0, 26.0), 260)
    w = synth.Workload("mix", fw, ["co2", "n2"], [4.0, 8.0], 298.0, (8, 8, 8))
    extra = ["move translate 0.5", "move rotate 0.5"]
    return w, extra, 1200, 400


def _write(tmp, name):
    w, extra, steps, eq = _workload(name)
    w.write(str(tmp), steps=steps, eqsteps=eq)
    p = os.path.join(str(tmp), "CONTROL")
    s = open(p).read()
    lines = extra + ["seed 12 34 56 78", "trace", "finish"]
    s = s.replace("finish", "\n".join(lines))
    open(p, "w").write(s)
    return w


def _run(exe, d):
    r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    return np.loadtxt(os.path.join(d, "branch01", "trace.out"), ndmin=2)


@pytest.fixture(scope="module")
def oracle_driver(oracle):
    if not os.path.exists(DRV_ORACLE):
        subprocess.check_call(["make", "-s", "-C", ORACLE])
    return DRV_ORACLE


@pytest.mark.parametrize("name", ["c1", "mix"])
def test_driver_matches_oracle_chain(oracle_driver, oracle, tmp_path, name):
    from fastmc_amd import inputs
    _write(tmp_path, name)
    tr = _run(oracle_driver, str(tmp_path))
    f = inputs.read_field(str(tmp_path / "FIELD"))
    c = inputs.read_config(str(tmp_path / "CONFIG"), f)
    ct = inputs.read_control(str(tmp_path / "CONTROL"))
    s, _, _ = oracle.system_from_inputs(f, c, ct)
    gids = sorted(ct.guests)
    ch = oracle.OracleChain(s, ct.temp, [ct.guests[g].press_pa for g in gids],
                            (12, 34, 56, 78),
                            freq=[ct.guests[g].freq for g in gids], delr=ct.delr)
    recs = ch.run(len(tr))
    kind = np.array([r.kind for r in recs])
    acc = np.array([r.accepted for r in recs])
    ig = np.array([r.iguest for r in recs]) + 1
    nm = np.array([r.nmol_after for r in recs])
    est = np.array([r.estep for r in recs])
    assert np.array_equal(kind, tr[:, 2])
    assert np.array_equal(ig, tr[:, 1])
    assert np.array_equal(acc, tr[:, 3])
    assert np.array_equal(nm, tr[:, 4])
    assert np.allclose(est, tr[:, 5], rtol=1e-14, atol=1e-12)
    assert acc.sum() > 10
    if name == "mix":
        assert {3, 4, 5} <= set(kind.tolist())


def test_output_reports_energy_and_loading(oracle_driver, tmp_path):
    _write(tmp_path, "c1")
    _run(oracle_driver, str(tmp_path))
    out = open(tmp_path / "OUTPUT").read()
    assert "Energy reported by DL_POLY" in out
    assert "average number of guests" in out
    assert os.path.getsize(tmp_path / "branch01" / "numguests.out") > 0


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c1", "mix"])
def test_gpu_driver_matches_oracle_driver(oracle_driver, tmp_path, name):
    if not os.path.exists(DRV_GPU):
        pytest.fail(f"{DRV_GPU} missing: build first")
    da, db = tmp_path / "gpu", tmp_path / "cpu"
    da.mkdir(); db.mkdir()
    _write(da, name)
    for f in ("CONTROL", "FIELD", "CONFIG"):
        shutil.copy(da / f, db / f)
    tg = _run(DRV_GPU, str(da))
    tc = _run(oracle_driver, str(db))
    assert tg.shape == tc.shape
    assert np.array_equal(tg[:, :5], tc[:, :5])  # kinds, accept/reject, N
    rel = np.abs(tg[:, 5] - tc[:, 5]) / np.maximum(np.abs(tc[:, 5]), 1e-3)
    assert rel.max() < 1e-10, rel.max()
