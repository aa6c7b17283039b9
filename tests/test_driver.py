"""The Fortran driver (fmc_gcmc, the gcmc.f move loop) over the C ABI.

CPU: the driver linked against the oracle ABI reproduces, move for move, the
C oracle's own restatement of the move loop (same duni seed): identical
move kinds, accept/reject decisions and loadings, identical estep.
GPU: the driver linked against libfastmc_amd.so reproduces the oracle-linked
driver's accept/reject sequence exactly, with every estep within 1e-10
relative (on the 3392-atom cell relative to the scale of its components,
which the oracle's own rounding sets there)."""
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
    if name == "c2":
        # the plain FP64 drop-in's benchmarked run: the 3392-atom 2x2x2
        # supercell, CO2/N2 at 0.15/0.85 bar, kmax 12, 400 equilibration and
        # 1600 production moves
        return synth.c2(), [], 1600, 400
    if name == "ar":
        # a one-site uncharged guest (argon) beside CO2, in the FIELD as a
        # one-atom molecule; the default move mix (a displacement also
        # translates, so no move leaves the energy exactly unchanged)
        rng = np_.random.default_rng(21)
        fw = synth.make_framework(rng, (26.0, 26.0, 26.0), 260)
        w = synth.Workload("ar", fw, ["co2", "ar"], [4.0, 8.0], 298.0, (8, 8, 8))
        return w, [], 1200, 400
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


@pytest.mark.parametrize("name", ["c1", "mix", "ar", "c2"])
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
@pytest.mark.parametrize("name", ["c1", "mix", "ar", "c2"])
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
    if name == "c2":
        # 51.6 A and Nk 4192: the oracle's reference-form reciprocal sum
        # a_k (|S_new|^2 - |S_old|^2) carries ~1e-10 internal units of its
        # own rounding (test_reciprocal_change_rounding_on_the_large_cell
        # measures 4e-11 to 5e-10 against an extended-precision sum, the
        # kernels' cancellation-free form 4e-14), so estep is compared at
        # 1e-10 of the scale of its components
        scale = np.abs(tc[:, 6:9]).sum(axis=1) / 418.4
        rel = np.abs(tg[:, 5] - tc[:, 5]) / np.maximum(np.maximum(np.abs(tc[:, 5]), scale), 1e-3)
        assert (tc[:, 3] == 1).sum() > 50
    print(f"\n{name}: {len(tc)} moves, {int((tc[:, 3] == 1).sum())} accepted, worst estep "
          f"{rel.max():.2e}")
    assert rel.max() < 1e-10, rel.max()
    # the end-of-run reports agree to the printed digits
    from test_stats import _report
    rg, rc = _report(da / "OUTPUT"), _report(db / "OUTPUT")
    assert rg.keys() == rc.keys() and "average number of guests:" in rg
    for k in rg:
        a, b = np.atleast_1d(rg[k]), np.atleast_1d(rc[k])
        assert np.allclose(a, b, rtol=0, atol=2e-6, equal_nan=True), (k, a, b)


def test_revcon_after_the_guest_count_changes(oracle_driver, tmp_path):
    """REVCON is rewritten every 1000 moves (gcmc.f:737-740).  The driver
    formats the framework's lines once and renumbers their atoms when the
    guest count moves them, so the REVCON of the last write (three writes
    in a row at different loadings) must equal, byte for byte, the file the
    Python writer makes from the REVIVE written with it.  Both files are
    rewritten in place and cut at their end."""
    from fastmc_amd import inputs
    w, extra, _, _ = _workload("ar")
    w.write(str(tmp_path), steps=3000, eqsteps=500)
    p = tmp_path / "CONTROL"
    p.write_text(p.read_text().replace("finish", "seed 12 34 56 78\ntrace\nfinish"))
    # the files are rewritten in place: older, longer files in their place
    # must be cut at the new end
    (tmp_path / "branch01").mkdir()
    for f in ("REVCON", "REVIVE"):
        (tmp_path / "branch01" / f).write_bytes(b"\x07" * 2_000_000)
    tr = _run(oracle_driver, str(tmp_path))
    field = inputs.read_field(str(tmp_path / "FIELD"))
    config = inputs.read_config(str(tmp_path / "CONFIG"), field)
    ns = [len(m.names) for m in field.molecules if m.guest]

    def guest_atoms(upto):
        # trace rows: move number, guest (1-based), ..., that guest's count
        n = [0] * len(ns)
        for row in tr[tr[:, 0] < upto]:
            n[int(row[1]) - 1] = int(row[4])
        return sum(a * b for a, b in zip(n, ns)), n

    # the files on disk come from the last write, made before that move
    lastw = int(tr[-1, 0]) // 1000 * 1000
    at = [guest_atoms(m) for m in (lastw - 2000, lastw - 1000, lastw)]
    assert len({a for a, _ in at}) == 3, at
    rev = inputs.read_revive(str(tmp_path / "branch01" / "REVIVE"), field)
    assert rev.nummols[:len(ns)] == at[2][1]
    # production, counts, chain statistics, atom count, 3 arrays per
    # molecule type, energies: nothing after them
    recs = inputs.read_fortran_records(str(tmp_path / "branch01" / "REVIVE"))
    assert len(recs) == 4 + 3 * len(field.molecules) + 1
    last = tr[tr[:, 0] < lastw][-1]
    want = tmp_path / "REVCON.want"
    inputs.write_revcon(str(want), config.title, config.imcon, config.cell,
                        inputs.revcon_atoms(field, rev.coords),
                        eng=last[5] if last[3] else 0.0)
    assert (tmp_path / "branch01" / "REVCON").read_bytes() == want.read_bytes()


def _restart_run_dir(d, ex_dir, steps=1000):
    """The reference's restart example set up as a restart: CONFIG, FIELD,
    branch01/REVIVE, and a CONTROL asking for `restart`, a fixed seed and a
    short production run."""
    os.makedirs(os.path.join(d, "branch01"), exist_ok=True)
    for f in ("CONFIG", "FIELD"):
        shutil.copy(os.path.join(ex_dir, f), os.path.join(d, f))
    shutil.copy(os.path.join(ex_dir, "REVIVE"), os.path.join(d, "branch01", "REVIVE"))
    for f in ("CONFIG", "FIELD", os.path.join("branch01", "REVIVE")):
        os.chmod(os.path.join(d, f), 0o644)
    s = open(os.path.join(ex_dir, "CONTROL")).read()
    s = s.replace("steps    100000000", f"steps    {steps}")
    s = s.replace("equilibration    -1500000", "equilibration    0")
    s = s.replace("finish", "seed 12 34 56 78\ntrace\nrestart\nfinish")
    open(os.path.join(d, "CONTROL"), "w").write(s)


def _initial_energy(d):
    for line in open(os.path.join(d, "OUTPUT")):
        if "Initial guest energy" in line:
            return float(line.split()[-1])
    raise AssertionError("no initial energy in OUTPUT")


def test_driver_restart_reads_and_writes_revive(oracle_driver, oracle, restart_example,
                                                tmp_path):
    """The driver restarted from the reference's REVIVE (gcmc.f:349-410)
    starts from known answer 2, and the REVIVE/REVCON it writes at move
    1000 (gcmc.f:737-740) hold the state after 999 moves:
    * molecule count, rolling <N> and the accumulated energy follow the
      move trace;
    * the single_point energy of the written configuration plus the
      starting tail equals the accumulated energy (the move loop's elrc
      starts at zero, gcmc.f:138-139, so the first accepted insert or
      delete adds the whole tail once);
    * the REVCON equals, byte for byte, the one written from the REVIVE
      coordinates by the Python writer that reproduces the reference's
      own REVCON;
    * a second restart from the written REVIVE starts from that state's
      single_point energy."""
    from fastmc_amd import inputs
    from conftest import revive_config
    ex = restart_example
    d = str(tmp_path)
    _restart_run_dir(d, ex["dir"])
    tr = _run(oracle_driver, d)
    kat = ex["kat"]["revive_energy_kcal"]
    assert abs(_initial_energy(d) - kat) < 1e-6
    field, config = ex["field"], ex["config"]
    rev = inputs.read_revive(os.path.join(d, "branch01", "REVIVE"), field)
    pre = tr[tr[:, 0] < 1000]
    assert len(pre) == 999 and rev.is_production
    assert rev.nummols == [int(pre[-1, 4]), 36]
    assert rev.chainstats[0] == 999.0
    assert abs(rev.chainstats[1] - pre[:, 4].mean()) < 1e-9 * pre[:, 4].mean()
    # accumulated energy from the starting single_point
    f0, c0, _ = revive_config(ex)
    s0, _, _ = oracle.system_from_inputs(f0, c0, ex["control"])
    e0 = s0.energy_full()
    e = sum(e0[1:6]) / field.engunit
    for row in pre:
        if row[3]:
            e = e + row[5]
    assert abs(rev.energy[0] - e) < 1e-9
    insdel = pre[(pre[:, 2] <= 2) & (pre[:, 3] == 1)]
    assert len(insdel) > 0
    f1, c1 = inputs.apply_revive(field, config, rev)
    s1, _, _ = oracle.system_from_inputs(f1, c1, ex["control"])
    e1 = s1.energy_full()
    e_sp = sum(e1[1:6]) / field.engunit
    tail0 = e0[5] / field.engunit
    assert abs(e_sp + tail0 - rev.energy[0]) < 1e-7, (e_sp, tail0, rev.energy[0])
    # REVCON (eng = the last move's accepted estep, delE at gcmc.f:737-741)
    last = pre[-1]
    want = tmp_path / "REVCON.want"
    inputs.write_revcon(str(want), config.title, config.imcon, config.cell,
                        inputs.revcon_atoms(field, rev.coords),
                        eng=last[5] if last[3] else 0.0)
    assert open(os.path.join(d, "branch01", "REVCON")).read() == want.read_text()
    # restart from the driver's own REVIVE
    _run(oracle_driver, d)
    assert abs(_initial_energy(d) - e_sp) < 1e-6


@pytest.mark.gpu
def test_gpu_driver_restart_matches_oracle_driver(oracle_driver, restart_example, tmp_path):
    """The restart run on the MI355X library: the same moves, decisions and
    loadings as the oracle-linked driver, estep within 1e-10 relative, and
    the same REVIVE state."""
    from fastmc_amd import inputs
    if not os.path.exists(DRV_GPU):
        pytest.fail(f"{DRV_GPU} missing: build first")
    da, db = tmp_path / "gpu", tmp_path / "cpu"
    da.mkdir(); db.mkdir()
    for d in (da, db):
        _restart_run_dir(str(d), restart_example["dir"])
    tg = _run(DRV_GPU, str(da))
    tc = _run(oracle_driver, str(db))
    assert tg.shape == tc.shape
    assert np.array_equal(tg[:, :5], tc[:, :5])
    rel = np.abs(tg[:, 5] - tc[:, 5]) / np.maximum(np.abs(tc[:, 5]), 1e-3)
    assert rel.max() < 1e-10, rel.max()
    field = restart_example["field"]
    rg = inputs.read_revive(str(da / "branch01" / "REVIVE"), field)
    rc = inputs.read_revive(str(db / "branch01" / "REVIVE"), field)
    assert rg.nummols == rc.nummols
    for a, b in zip(rg.coords, rc.coords):
        for c in range(3):
            assert np.array_equal(a[c], b[c])
    assert abs(rg.energy[0] - rc.energy[0]) < 1e-9
