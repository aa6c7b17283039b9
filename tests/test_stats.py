"""End-of-run statistics (fastmc_amd.stats, the reference's gcmc.f:1969-2031,
2051-2100, 2366-2432): closed forms that fix the window arithmetic, and the
Fortran driver's report against the same restatement."""
import math

import numpy as np
import pytest

T = 303.15


def test_window_quirk_on_a_constant_series():
    """A constant series gives windows 1-4 of nwind+1 samples divided by
    nwind, window 5 one sample short (the index past the end reads zero):
    the reported <N> is c (4 (nwind+1)/nwind + 1) / 5, not c."""
    from fastmc_amd import stats
    c, n = 7.0, 1000
    nd = stats.from_samples([c] * n, [-2.0] * n, T)
    nwind = n // 5
    v = [c * (nwind + 1) / nwind] * 4 + [c]
    assert abs(nd.avg[0] - sum(v) / 5) < 1e-12
    assert abs(nd.std[0] - np.std(v)) < 1e-12
    assert nd.weight == n


def test_windows_and_welford_against_numpy():
    from fastmc_amd import stats
    rng = np.random.default_rng(3)
    n = 2503
    N = rng.integers(0, 40, n).astype(float)
    E = -3.0 * N + rng.normal(0, 5, n)
    sums, nwind = stats.window_sums(N, E)
    assert nwind == 500
    Np = np.concatenate([N, [0.0]])
    Ep = np.concatenate([E, [0.0]])
    win = []
    for x in range(5):
        sl = slice(x * nwind, x * nwind + nwind + 1)
        a, e = Np[sl], Ep[sl]
        w = [a.sum() / nwind, e.sum() / nwind, (a * e).sum() / nwind,
             (a * a).sum() / nwind, (e * e).sum() / nwind]
        w.append(stats.calc_qst(w[4], w[1], w[0], w[3], w[2], T))
        w.append(stats.calc_cv(w[4], w[1], w[0], w[3], w[2], T))
        win.append(w)
        assert np.allclose(sums[x], [a.sum(), e.sum(), (a * e).sum(), (a * a).sum(),
                                     (e * e).sum()], rtol=1e-13)
    win = np.array(win)
    nd = stats.node_stats(sums, nwind, float(n), T)
    assert np.allclose(nd.avg[:5], win.mean(axis=0)[:5], rtol=1e-12)
    assert np.allclose(nd.std[:7], win.std(axis=0), rtol=1e-9)
    assert abs(nd.avg[7] - win[:, 5].mean()) < 1e-9 * abs(win[:, 5].mean())
    a = nd.avg
    assert nd.avg[5] == stats.calc_qst(a[4], a[1], a[0], a[3], a[2], T)


def test_isosteric_heat_of_a_linear_model():
    """E = slope * N exactly in every sample: Q_st = -slope + R T."""
    from fastmc_amd import stats
    rng = np.random.default_rng(5)
    N = rng.integers(1, 30, 4000).astype(float)
    slope = -4.25
    nd = stats.from_samples(N, slope * N, T)
    assert abs(nd.avg[5] - (-slope + stats.RGAS * T)) < 1e-9
    assert nd.std[5] < 1e-8


def test_union_of_replicas():
    """Identical replicas combine to the replica; the union's <N> is the
    sample-weighted mean and its error pools the second moments."""
    from fastmc_amd import stats
    rng = np.random.default_rng(9)
    nodes = []
    for k in range(3):
        N = rng.integers(0, 20, 1500 + 500 * k).astype(float)
        nodes.append(stats.from_samples(N, -2.0 * N + rng.normal(0, 1, len(N)), T))
    one = stats.union([nodes[0]] * 4, T)
    assert np.allclose(one.avg[:5], nodes[0].avg[:5], rtol=1e-13)
    assert abs(one.std[0] - nodes[0].std[0]) < 1e-6
    u = stats.union(nodes, T)
    w = np.array([n.weight for n in nodes])
    a = np.array([n.avg[0] for n in nodes])
    sd = np.array([n.std[0] for n in nodes])
    assert abs(u.avg[0] - (w * a).sum() / w.sum()) < 1e-12
    assert abs(u.std[0] - math.sqrt((w * (sd ** 2 + a ** 2)).sum() / w.sum() - u.avg[0] ** 2)) < 1e-12
    assert u.weight == w.sum()


def _report(path):
    """The final-statistics block of the driver's OUTPUT (the last guest's
    block, and every selectivity pair as ("<S>", j, i))."""
    vals = {}
    pair = None
    for line in open(path):
        if line.startswith("selectivity stats for guest"):
            j, i = line.split("guest")[1].split("/")
            pair = ("<S>", int(j), int(i))
        if line.strip().startswith("<S>:") and pair:
            a, b = line.split("<S>:")[1].split("+/-")
            vals[pair] = (float(a), float(b))
        for key in ("<N>:", "<E>:", "<E*N>:", "<N*N>:", "<E*E>:"):
            if line.strip().startswith(key):
                a, b = line.split(key)[1].split("+/-")
                vals[key] = (float(a), float(b))
        for key in ("average number of guests:", "standard error:",
                    "Isosteric heat of adsorption (kcal/mol):", "Isosteric heat error:",
                    "Heat capacity, Cv (kcal/mol/K):", "Heat capacity error:",
                    "Total steps counted:"):
            if line.strip().startswith(key):
                vals[key] = float(line.split(key)[1])
    return vals


def test_driver_report_matches_the_restatement(oracle, tmp_path):
    """The driver's end-of-run report (linked against the oracle ABI, a
    production run from move one) equals the restatement applied to the
    samples in its move trace, to the printed digits."""
    import os
    import subprocess
    from fastmc_amd import stats, synth
    from conftest import ORACLE
    exe = os.path.join(ORACLE, "_build", "fmc_gcmc_oracle")
    if not os.path.exists(exe):
        subprocess.check_call(["make", "-s", "-C", ORACLE])
    w = synth.c1()
    w.write(str(tmp_path), steps=2000, eqsteps=0)
    p = tmp_path / "CONTROL"
    p.write_text(p.read_text().replace("finish", "seed 12 34 56 78\ntrace\nfinish"))
    r = subprocess.run([exe], cwd=str(tmp_path), capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    tr = np.loadtxt(tmp_path / "branch01" / "trace.out", ndmin=2)
    assert len(tr) == 2400
    nd = stats.union([stats.from_samples(tr[:, 4], tr[:, 10], w.temp)], w.temp)
    got = _report(tmp_path / "OUTPUT")
    assert got["Total steps counted:"] == 2400
    names = ["<N>:", "<E>:", "<E*N>:", "<N*N>:", "<E*E>:"]
    for z, key in enumerate(names):
        assert abs(got[key][0] - nd.avg[z]) <= 6e-7 * max(1.0, abs(nd.avg[z]) * 1e-9 + 1), key
        assert abs(got[key][1] - nd.std[z]) <= 6e-4, key
    assert abs(got["average number of guests:"] - nd.avg[0]) <= 6e-7
    assert abs(got["standard error:"] - nd.std[0]) <= 6e-7
    assert abs(got["Isosteric heat of adsorption (kcal/mol):"] - nd.avg[5]) <= 6e-7
    assert abs(got["Heat capacity, Cv (kcal/mol/K):"] - nd.avg[6]) <= 6e-7
    for key, z in (("Isosteric heat error:", 5), ("Heat capacity error:", 6)):
        if math.isnan(nd.std[z]):
            assert math.isnan(got[key])
        else:
            assert abs(got[key] - nd.std[z]) <= 6e-7, key
    assert nd.std[0] > 0 and 0 < nd.avg[0]


# ---------------------------------------------------------------------------
# selectivity (gcmc.f:1887-1916)
# ---------------------------------------------------------------------------

def test_selectivity_of_loadings_in_constant_ratio():
    """Two guests whose loadings keep the ratio 3:1 in every sample report
    S(1/2) = 3 with zero variance, over full windows and a partial last one
    (its weight only scales a zero change)."""
    from fastmc_amd import stats
    rng = np.random.default_rng(2)
    k = rng.integers(1, 20, 2750).astype(float)
    samples = np.stack([3 * k, k], axis=1)
    sel = stats.selectivity_from_samples(samples, nwind=500)
    s = sel[(0, 1)]
    assert abs(s.mean - 3.0) < 1e-14
    assert abs(s.var) < 1e-12
    assert abs(s.weight - (5 + 250 / 500)) < 1e-15


def test_selectivity_of_a_run_shorter_than_one_window():
    """The reference takes a run shorter than its averaging window as one
    partial window of weight samples / nwind, and its running mean moves
    from zero by that weight: the printed <S> is the ratio times the
    fraction of the window filled."""
    from fastmc_amd import stats
    samples = [[6.0, 2.0]] * 800
    s = stats.selectivity_from_samples(samples, nwind=100000)[(0, 1)]
    assert abs(s.mean - 3.0 * 800 / 100000) < 1e-15
    assert abs(s.var - (9.0 - s.mean ** 2)) < 1e-12


def test_selectivity_windows_against_numpy():
    """Full windows only: the running update is the plain mean of the window
    ratios <N_j>/<N_i> and the printed variance is their population
    variance; three guests give the pairs 1/2, 1/3, 2/3."""
    from fastmc_amd import stats
    rng = np.random.default_rng(8)
    nwind, nw = 400, 7
    samples = rng.integers(1, 30, (nwind * nw, 3)).astype(float)
    sel = stats.selectivity_from_samples(samples, nwind=nwind)
    means = samples.reshape(nw, nwind, 3).mean(axis=1)
    for (j, i), s in sel.items():
        r = means[:, j] / means[:, i]
        assert abs(s.mean - r.mean()) < 1e-13 * r.mean()
        assert abs(s.var - r.var()) < 1e-10 * max(r.var(), 1e-12)
    assert sorted(sel) == [(0, 1), (0, 2), (1, 2)]


def test_selectivity_skips_windows_without_guest_i():
    """A window whose <N_i> is zero updates no pair of guest i but counts
    (avcount), so the next window's change is divided by the larger count:
    windows with ratios 2, (skipped), 4 give 2 + (4 - 2) / 3."""
    from fastmc_amd import stats
    w1 = [[4.0, 2.0]] * 10
    w2 = [[5.0, 0.0]] * 10
    w3 = [[8.0, 2.0]] * 10
    s = stats.selectivity_from_samples(w1 + w2 + w3, nwind=10)[(0, 1)]
    assert abs(s.mean - (2.0 + 2.0 / 3.0)) < 1e-15
    sec = 4.0 + (16.0 - 4.0) / 3.0        # selwin after window 3
    assert abs(s.var - (sec - s.mean ** 2)) < 1e-13


def test_selectivity_union_of_identical_replicas_is_the_replica():
    from fastmc_amd import stats
    rng = np.random.default_rng(4)
    samples = rng.integers(1, 25, (3300, 2)).astype(float)
    one = stats.selectivity_from_samples(samples, nwind=600)
    u = stats.selectivity_union([one, one, one])
    for key in one:
        assert abs(u[key].mean - one[key].mean) < 1e-15
        assert abs(u[key].var - (one[key].second - one[key].mean ** 2)) < 1e-13
        assert abs(u[key].weight - 3 * one[key].weight) < 1e-12


def test_driver_selectivity_matches_the_restatement(oracle, tmp_path):
    """The driver (oracle ABI) on a two-guest system with an averaging
    window of 300 samples: its printed selectivity equals the restatement
    applied to the loadings of its move trace, to the printed digits."""
    import os
    import subprocess
    from fastmc_amd import stats, synth
    from conftest import ORACLE
    exe = os.path.join(ORACLE, "_build", "fmc_gcmc_oracle")
    if not os.path.exists(exe):
        subprocess.check_call(["make", "-s", "-C", ORACLE])
    rng = np.random.default_rng(21)
    fw = synth.make_framework(rng, (26.0, 26.0, 26.0), 260)
    w = synth.Workload("mix", fw, ["co2", "n2"], [4.0, 8.0], 298.0, (8, 8, 8))
    w.write(str(tmp_path), steps=2000, eqsteps=0)
    p = tmp_path / "CONTROL"
    p.write_text(p.read_text().replace(
        "finish", "averaging window 300\nseed 12 34 56 78\ntrace\nfinish"))
    r = subprocess.run([exe], cwd=str(tmp_path), capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    tr = np.loadtxt(tmp_path / "branch01" / "trace.out", ndmin=2)
    n = np.zeros(2)
    samples = []
    for row in tr:
        n[int(row[1]) - 1] = row[4]
        samples.append(n.copy())
    assert len(samples) == 2400 and min(s[1] for s in samples[300:]) >= 0
    sel = stats.selectivity_from_samples(samples, nwind=300)[(0, 1)]
    got = _report(tmp_path / "OUTPUT")[("<S>", 1, 2)]
    print(f"\ndriver <S> {got}, restatement {sel.mean:.6f} +/- {sel.var:.3f}")
    assert abs(got[0] - sel.mean) <= 6e-7
    assert abs(got[1] - sel.var) <= 6e-4
    assert sel.mean > 0
