"""GPU parity of the single-chain drop-in path (fmc_* C ABI) against the CPU
oracle and the reference's known answers.  Bar: every energy component
within 1e-10 relative (FP64 path), S(k) equal to a fresh recompute."""
import numpy as np
import pytest

from conftest import revive_config

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _close(a, b, rtol=RTOL, floor=1.0):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), floor))


@pytest.fixture(scope="module")
def fmc():
    import fastmc_amd
    fastmc_amd.lib()
    return fastmc_amd


def test_kat_framework_energy(fmc, oracle, restart_example):
    ex = restart_example
    ctx = fmc.context_from_inputs(ex["field"], ex["config"], ex["control"])
    info = ctx.info()
    assert info["nk"] == ex["kat"]["nk"]
    e = ctx.recip_full(0) / ex["field"].engunit
    assert abs(e - ex["kat"]["framework_recip_plus_self_kcal"]) <= 5e-7
    s, _, _ = oracle.system_from_inputs(ex["field"], ex["config"], ex["control"])
    assert abs(e * ex["field"].engunit - s.recip_full(0)) <= 1e-11 * abs(s.recip_full(0))


def test_kat_revive_energy(fmc, oracle, restart_example):
    ex = restart_example
    field, config, e_rev = revive_config(ex)
    ctx = fmc.context_from_inputs(field, config, ex["control"])
    assert ctx.nmol(0) == 90
    e = ctx.energy_full()
    sp = (e[1] + e[2] + e[3] + e[4] + e[5]) / field.engunit
    assert abs(sp - e_rev) < 1e-7
    s, _, _ = oracle.system_from_inputs(field, config, ex["control"])
    eo = s.energy_full()
    assert _close(e, eo, rtol=1e-10)


def _systems():
    from fastmc_amd import synth
    return {"c1": synth.c1(), "c2": synth.c2()}


def _random_moves(fmc, oracle, A, n_moves, seed, gtypes):
    rng = np.random.default_rng(seed)
    ctx = fmc.Context.from_arrays(A)
    orc = oracle.OracleSystem(A["imcon"], A["cell"], A["alpha"], A["kmax"],
                              A["rcut"], A["delr"], A["ntpatm"], A["vdw"],
                              A["fw"], A["guests"])
    cell = np.asarray(A["cell"])
    L = np.array([cell[0], cell[4], cell[8]])
    worst = 0.0
    for it in range(n_moves):
        g = int(rng.integers(0, gtypes))
        n = orc.nmol(g)
        kind = 1 if n == 0 else int(rng.integers(1, 4))
        tpl = A["guests"][g]
        if kind == 1:
            shift = rng.uniform(0, 1, 3) * L
            x, y, z = [np.asarray(tpl[k]) + shift[i] for i, k in enumerate("xyz")]
            a = ctx.trial(1, g, -1, x, y, z)
            b = orc.trial(1, g, -1, x, y, z)
        elif kind == 2:
            m = int(rng.integers(0, n))
            a = ctx.trial(2, g, m)
            b = orc.trial(2, g, m)
        else:
            m = int(rng.integers(0, n))
            x, y, z = orc.molecule(g, m)
            d = rng.normal(0, 0.7, 3)
            a = ctx.trial(3, g, m, x + d[0], y + d[1], z + d[2])
            b = orc.trial(3, g, m, x + d[0], y + d[1], z + d[2])
        rel = np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0))
        worst = max(worst, rel)
        assert _close(a, b), (it, kind, a, b)
        if rng.uniform() < 0.6 or n == 0:
            ctx.commit()
            orc.commit()
    return ctx, orc, worst


@pytest.mark.parametrize("name,gtypes,n", [("c1", 1, 120), ("c2", 2, 60)])
def test_trial_parity(fmc, oracle, name, gtypes, n):
    A = _systems()[name].system_arrays()
    ctx, orc, worst = _random_moves(fmc, oracle, A, n, seed=5, gtypes=gtypes)
    print(f"{name}: worst component rel diff {worst:.3e}")
    for g in range(gtypes):
        assert ctx.nmol(g) == orc.nmol(g)
        for m in range(ctx.nmol(g)):
            xa = np.concatenate(ctx.molecule(g, m))
            xb = np.concatenate(orc.molecule(g, m))
            assert np.array_equal(xa, xb)
    # running S(k) after many commits equals a fresh recompute
    re, im = ctx.sk()
    e_all = ctx.recip_full(1)
    re2, im2 = ctx.sk()
    scale = max(np.abs(re2).max(), 1.0)
    assert np.abs(re - re2).max() < 1e-11 * scale * 100
    assert np.abs(im - im2).max() < 1e-11 * scale * 100
    assert abs(e_all - orc.recip_full(1)) < 1e-10 * abs(e_all)


@pytest.mark.parametrize("guests", [("ar", "co", "h2o"), ("ch4", "c2h6", "co2")])
def test_trial_parity_molecule_sizes(fmc, oracle, guests):
    """Molecules of 1, 2, 4 sites (and 3) take the trial kernel's unrolled
    forms; 5 and 8 sites the form that goes one position at a time. Random
    inserts, deletes and displacements against the oracle, with commits."""
    from fastmc_amd import synth
    rng = np.random.default_rng(17)
    fw = synth.make_framework(rng, (25.8, 25.8, 25.8), 424)
    w = synth.Workload("sizes", fw, list(guests), [1.0] * len(guests), 298.0, (8, 8, 8))
    A = w.system_arrays()
    ctx, orc, worst = _random_moves(fmc, oracle, A, 150, seed=11, gtypes=len(guests))
    print(f"{guests}: worst component rel diff {worst:.3e}")
    for g in range(len(guests)):
        assert orc.nmol(g) > 0   # every size was inserted and moved
        assert ctx.nmol(g) == orc.nmol(g)


def test_trial_sums_are_each_calls_own(fmc, oracle):
    """A trial returns once its work-groups' sums have reached host memory,
    not when its stream drains. Two different trials alternated without a
    commit must each read their own sums every time, bit for bit, never a
    previous trial's."""
    from fastmc_amd import synth
    A = synth.c2().system_arrays()
    ctx = fmc.Context.from_arrays(A)
    orc = oracle.OracleSystem(A["imcon"], A["cell"], A["alpha"], A["kmax"],
                              A["rcut"], A["delr"], A["ntpatm"], A["vdw"],
                              A["fw"], A["guests"])
    tpl = A["guests"][0]
    p = [np.asarray(tpl[k]) + 7.0 for k in "xyz"]
    r = [np.asarray(tpl[k]) + 19.0 for k in "xyz"]
    a = ctx.trial(1, 0, -1, *p)
    b = ctx.trial(1, 0, -1, *r)
    assert _close(a, orc.trial(1, 0, -1, *p))
    assert _close(b, orc.trial(1, 0, -1, *r))
    assert not np.array_equal(a, b)
    for _ in range(50):
        assert np.array_equal(ctx.trial(1, 0, -1, *p), a)
        assert np.array_equal(ctx.trial(1, 0, -1, *r), b)
    # the last trial (b) commits; S(k) read afterwards includes it
    ctx.commit()
    orc.commit()
    re, im = ctx.sk()
    ctx.recip_full(1)
    re2, im2 = ctx.sk()
    scale = max(np.abs(re2).max(), 1.0)
    assert np.abs(re - re2).max() < 1e-9 * scale
    assert np.abs(im - im2).max() < 1e-9 * scale


def test_triclinic_trial_parity(fmc, oracle, restart_example):
    """imcon 3 (the reference's calf-20 example) with 90 CO2 loaded."""
    ex = restart_example
    field, config, _ = revive_config(ex)
    ctx = fmc.context_from_inputs(field, config, ex["control"])
    s, _, _ = oracle.system_from_inputs(field, config, ex["control"])
    rng = np.random.default_rng(9)
    for it in range(30):
        kind = [1, 2, 3][it % 3]
        if kind == 1:
            fr = rng.uniform(0, 1, 3)
            c = np.asarray(config.cell)
            com = fr[0] * c[0:3] + fr[1] * c[3:6] + fr[2] * c[6:9]
            g = field.molecules[0]
            x, y, z = (np.asarray(g.gx) + com[0], np.asarray(g.gy) + com[1],
                       np.asarray(g.gz) + com[2])
            a = ctx.trial(1, 0, -1, x, y, z); b = s.trial(1, 0, -1, x, y, z)
        elif kind == 2:
            m = int(rng.integers(0, s.nmol(0)))
            a = ctx.trial(2, 0, m); b = s.trial(2, 0, m)
        else:
            m = int(rng.integers(0, s.nmol(0)))
            x, y, z = s.molecule(0, m)
            d = rng.normal(0, 0.5, 3)
            a = ctx.trial(3, 0, m, x + d[0], y + d[1], z + d[2])
            b = s.trial(3, 0, m, x + d[0], y + d[1], z + d[2])
        assert _close(a, b), (it, a, b)
        if it % 2 == 0:
            ctx.commit(); s.commit()


def test_edge_cases(fmc, oracle):
    from fastmc_amd import synth
    A = synth.c1().system_arrays()
    ctx = fmc.Context.from_arrays(A, max_guest_atoms=9)
    with pytest.raises(fmc.FmcError) as e:
        ctx.commit()
    assert e.value.code == -2  # commit without a trial
    with pytest.raises(fmc.FmcError):
        ctx.trial(2, 0, 0)  # delete from an empty framework
    with pytest.raises(fmc.FmcError):
        ctx.trial(1, 3, -1, np.zeros(3), np.zeros(3), np.zeros(3))
    tpl = A["guests"][0]
    for i in range(2):
        ctx.trial(1, 0, -1, np.asarray(tpl["x"]) + 5 + 3 * i,
                  np.asarray(tpl["y"]) + 5, np.asarray(tpl["z"]) + 5)
        ctx.commit()
    # capacity 9 guest atoms: a third CO2 would reach it (reference error 75)
    with pytest.raises(fmc.FmcError) as e:
        ctx.trial(1, 0, -1, np.asarray(tpl["x"]) + 12, np.asarray(tpl["y"]) + 12,
                  np.asarray(tpl["z"]) + 12)
    assert e.value.code == -4
    # deleting the first of two keeps the order of the other
    x1 = ctx.molecule(0, 1)
    ctx.trial(2, 0, 0)
    ctx.commit()
    assert ctx.nmol(0) == 1
    assert np.array_equal(np.concatenate(ctx.molecule(0, 0)), np.concatenate(x1))
    ctx.trial(2, 0, 0)
    ctx.commit()
    assert ctx.nmol(0) == 0
