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
    """The final-statistics block of the driver's OUTPUT."""
    vals = {}
    for line in open(path):
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
