"""Batched device-side chains against the oracle's restatement of the move
loop.  FP64: the same duni seed gives the same trajectory -- identical move
kinds, accept/reject decisions and loadings, estep within 1e-10 relative.
Mixed precision: the same decisions while the trajectories coincide, with
the estep error bars each test states (tests/test_gpu_mixed_c3.py holds the
bars on the batched workload itself)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fmc():
    import fastmc_amd
    fastmc_amd.lib()
    return fastmc_amd


def _porous(seed=21, guests=("co2", "n2"), press=(4.0, 8.0), n=260, side=26.0):
    from fastmc_amd import synth
    rng = np.random.default_rng(seed)
    fw = synth.make_framework(rng, (side, side, side), n)
    return synth.Workload("porous", fw, list(guests), list(press), 298.0, (8, 8, 8))


def _oracle_recs(oracle, A, temp, press_pa, seed, freq, n):
    s = oracle.OracleSystem(A["imcon"], A["cell"], A["alpha"], A["kmax"], A["rcut"],
                            A["delr"], A["ntpatm"], A["vdw"], A["fw"], A["guests"])
    ch = oracle.OracleChain(s, temp, press_pa, seed, freq=freq, delr=A["delr"])
    recs = ch.run(n)
    arr = np.array([[r.kind, r.iguest, r.accepted, r.nmol_after, r.estep]
                    for r in recs])
    return arr, ch, s


def _compare(log, ref, tol=1e-10):
    assert np.array_equal(log[:, :4], ref[:, :4]), np.nonzero(
        np.any(log[:, :4] != ref[:, :4], axis=1))[0][:5]
    rel = np.abs(log[:, 4] - ref[:, 4]) / np.maximum(np.abs(ref[:, 4]), 1e-3)
    assert rel.max() < tol, rel.max()
    return rel.max()


@pytest.mark.parametrize("tile", [256, 512, 256 | 0x10000])
def test_batch_fp64_matches_oracle_chain(fmc, oracle, tile):
    from fastmc_amd.batch import Batch
    from fastmc_amd import synth
    n = 1500
    cases = [
        (synth.c1().system_arrays(), 298.0, [0.15e5], (12, 34, 56, 78), None),
        (_porous().system_arrays(), 298.0, [4.0e5, 8.0e5], (3, 5, 7, 11),
         [[1, 1, 1, 0, 0, 0, 0.5, 0.5, 0]] * 2),
        (_porous(seed=4, guests=("co2",), press=(20.0,)).system_arrays(), 298.0,
         [20.0e5], (101, 23, 17, 5), None),
    ]
    b = Batch(len(cases))
    ids, refs = [], []
    for (A, T, P, seed, freq) in cases:
        ctx = fmc.Context.from_arrays(A)
        cid = b.add_chain(ctx, T, P, seed, freq=freq, delr=A["delr"])
        b.enable_log(cid, n)
        ids.append(cid)
        refs.append(_oracle_recs(oracle, A, T, P, seed, freq, n)[0])
    b.set_tile(tile)
    b.run(n // 3)
    b.run(n - n // 3)   # state carried across launches
    for cid, ref in zip(ids, refs):
        log = b.get_log(cid, n)
        worst = _compare(log, ref)
        print(f"chain {cid}: {int(ref[:, 2].sum())} accepted, worst estep rel {worst:.2e}")
        st = b.stats(cid)
        assert st["steps"] == n
        assert not st["capacity_fault"]


@pytest.mark.parametrize("kmax", [12, 14])
def test_batch_fp64_large_nk_register_tiles(fmc, oracle, kmax):
    """The register tiles of 12 and 14 k-vectors per thread at 512 threads
    (cubic cell, Nk 4192 and about 6650; at 14 the grouped reciprocal sum
    ends in a partial group of two) reproduce the oracle's trajectory."""
    from fastmc_amd.batch import Batch
    from fastmc_amd import synth
    rng = np.random.default_rng(31 + kmax)
    fw = synth.make_framework(rng, (26.0, 26.0, 26.0), 260)
    w = synth.Workload("porous", fw, ["co2"], [20.0], 298.0, (kmax, kmax, kmax))
    A = w.system_arrays()
    n, seed = 400, (7, 11, 13, 17)
    ctx = fmc.Context.from_arrays(A)
    nk = ctx.info()["nk"]
    assert (4096 < nk <= 6144) if kmax == 12 else (6144 < nk <= 7168), nk
    b = Batch(1)
    cid = b.add_chain(ctx, 298.0, [20.0e5], seed, delr=A["delr"])
    b.enable_log(cid, n)
    b.set_tile(512)
    b.run(n)
    log = b.get_log(cid, n)
    b.close()
    ref = _oracle_recs(oracle, A, 298.0, [20.0e5], seed, None, n)[0]
    worst = _compare(log, ref)
    print(f"kmax {kmax}, Nk {nk}: {int(ref[:, 2].sum())} accepted, worst estep rel {worst:.2e}")


def test_single_site_uncharged_guest(fmc, oracle):
    """A one-site uncharged guest (argon) beside CO2: one trial position per
    set, a centre of mass over one site and zero charges in every Ewald
    term.  The batch chain (FP64) follows the oracle's trajectory and the
    single-chain drop-in matches its trial components.

    Rotating a one-site guest changes no energy: estep is rounding noise of
    order 1e-14 in both codes, and its sign decides whether the acceptance
    draw is taken (gcmc.f:1330-1338), so the streams may part there.  With
    rotate-only moves of argon on, the trajectories agree up to the first
    such move, where both esteps are below 1e-12 kcal/mol; with them off
    they agree throughout."""
    from fastmc_amd.batch import Batch
    from fastmc_amd import synth
    rng = np.random.default_rng(21)
    fw = synth.make_framework(rng, (26.0, 26.0, 26.0), 260)
    w = synth.Workload("porous", fw, ["co2", "ar"], [4.0, 8.0], 298.0, (8, 8, 8))
    A = w.system_arrays()
    ctx = fmc.Context.from_arrays(A)
    seed = (3, 5, 7, 11)
    co2 = [1, 1, 1, 0, 0, 0, 0.5, 0.5, 0]
    for ar, n in (([1, 1, 1, 0, 0, 0, 0.5, 0.5, 0], 300), ([1, 1, 1, 0, 0, 0, 0.5, 0, 0], 900)):
        freq = [co2, ar]
        ref = _oracle_recs(oracle, A, 298.0, [4.0e5, 8.0e5], seed, freq, n)[0]
        assert set(ref[:, 1].astype(int)) == {0, 1} and ref[ref[:, 1] == 1, 2].sum() > 0
        b = Batch(1)
        cid = b.add_chain(ctx, 298.0, [4.0e5, 8.0e5], seed, freq=freq, delr=A["delr"])
        b.enable_log(cid, n)
        b.run(n)
        log = b.get_log(cid, n)
        b.close()
        tie = np.nonzero((ref[:, 0] == 5) & (ref[:, 1] == 1))[0]
        if ar[7] > 0:
            assert len(tie) > 0
            t = tie[0]
            _compare(log[:t], ref[:t])
            assert np.array_equal(log[t, :2], ref[t, :2])
            assert abs(log[t, 4]) < 1e-12 and abs(ref[t, 4]) < 1e-12, (log[t, 4], ref[t, 4])
        else:
            assert len(tie) == 0
            _compare(log, ref)
    # the drop-in: insert, displace and delete one argon atom
    orc = oracle.OracleSystem(A["imcon"], A["cell"], A["alpha"], A["kmax"], A["rcut"],
                              A["delr"], A["ntpatm"], A["vdw"], A["fw"], A["guests"])
    x, y, z = np.array([7.3]), np.array([5.1]), np.array([9.2])
    for kind, m, xyz in ((1, -1, (x, y, z)), (3, 0, (x + 0.4, y - 0.3, z + 0.2)), (2, 0, None)):
        args = xyz if xyz is not None else (None, None, None)
        a = ctx.trial(kind, 1, m, *args)
        r = orc.trial(kind, 1, m, *args)
        assert np.all(np.abs(a - r) <= 1e-10 * np.maximum(np.abs(r), 1.0)), (kind, a, r)
        assert a[1] == 0.0   # no charge: no real-space Coulomb term
        ctx.commit()
        orc.commit()


def test_eight_site_guest(fmc, oracle):
    """The largest rigid guest the device takes (8 sites: all-atom ethane)
    beside CO2: 16 trial positions for a displacement, several positions
    per table-building wave.  FP64 batch chain against the oracle's
    trajectory, mixed precision within the error bars of the small-system
    test, and the drop-in's components for one insert, displace, delete."""
    from fastmc_amd.batch import Batch
    from fastmc_amd.capi import PREC_MIXED
    from fastmc_amd import synth
    rng = np.random.default_rng(8)
    fw = synth.make_framework(rng, (26.0, 26.0, 26.0), 260)
    w = synth.Workload("porous", fw, ["co2", "c2h6"], [4.0, 2.0], 298.0, (8, 8, 8))
    A = w.system_arrays()
    ctx = fmc.Context.from_arrays(A)
    assert list(ctx.nsites) == [3, 8]
    n, seed = 900, (9, 8, 7, 6)
    freq = [[1, 1, 1, 0, 0, 0, 0.5, 0.5, 0]] * 2
    ref = _oracle_recs(oracle, A, 298.0, [4.0e5, 2.0e5], seed, freq, n)[0]
    assert ref[(ref[:, 1] == 1) & (ref[:, 0] == 3), 2].sum() > 0   # accepted ethane displacements
    logs = {}
    for prec in (0, PREC_MIXED):
        b = Batch(1, prec) if prec else Batch(1)
        cid = b.add_chain(ctx, 298.0, [4.0e5, 2.0e5], seed, freq=freq, delr=A["delr"])
        b.enable_log(cid, n)
        b.run(n)
        logs[prec] = b.get_log(cid, n)
        b.close()
    _compare(logs[0], ref)
    m = logs[PREC_MIXED]
    same = np.all(m[:, :4] == ref[:, :4], axis=1)
    first_diff = np.argmin(same) if not same.all() else n
    assert first_diff > 50
    kT = 1.9872041e-3 * 298.0
    a = ref[:first_diff, 4]
    keep = np.abs(a) < 50 * kT
    err = np.abs(m[:first_diff, 4] - a)[keep]
    assert err.max() < 1e-3 * kT
    # the drop-in
    orc = oracle.OracleSystem(A["imcon"], A["cell"], A["alpha"], A["kmax"], A["rcut"],
                              A["delr"], A["ntpatm"], A["vdw"], A["fw"], A["guests"])
    g = A["guests"][1]
    x, y, z = [np.asarray(g[k]) + 7.3 for k in "xyz"]
    for kind, mol, xyz in ((1, -1, (x, y, z)), (3, 0, (x + 0.4, y - 0.3, z + 0.2)), (2, 0, None)):
        args = xyz if xyz is not None else (None, None, None)
        r1 = ctx.trial(kind, 1, mol, *args)
        r2 = orc.trial(kind, 1, mol, *args)
        assert np.all(np.abs(r1 - r2) <= 1e-10 * np.maximum(np.abs(r2), 1.0)), (kind, r1, r2)
        ctx.commit()
        orc.commit()


def test_four_guest_types_one_at_zero_pressure(fmc, oracle):
    """Four guest types, the device maximum (CO2, N2, CH4, argon), one of
    them at zero pressure: the guest-type draw, the per-type atom offsets
    and the 4x4 tail form at full size, and energy_eval's zero-pressure
    insert (utility_pack.f:3049-3095: a zero prefactor, never accepted),
    where no rejection bound applies; with no molecule of that type every
    move drawn for it becomes such an insert (gcmc.f:837-847).  FP64 batch
    chain against the oracle's trajectory."""
    from fastmc_amd.batch import Batch
    from fastmc_amd import synth
    rng = np.random.default_rng(44)
    fw = synth.make_framework(rng, (26.0, 26.0, 26.0), 260)
    w = synth.Workload("porous", fw, ["co2", "n2", "ch4", "ar"], [4.0, 6.0, 3.0, 0.0],
                       298.0, (8, 8, 8))
    A = w.system_arrays()
    ctx = fmc.Context.from_arrays(A)
    press = [4.0e5, 6.0e5, 3.0e5, 0.0]
    n, seed = 1200, (4, 3, 2, 1)
    freq = [[1, 1, 1, 0, 0, 0, 0.5, 0.5, 0]] * 3 + [[1, 1, 1, 0, 0, 0, 0.5, 0, 0]]
    ref = _oracle_recs(oracle, A, 298.0, press, seed, freq, n)[0]
    assert set(ref[:, 1].astype(int)) == {0, 1, 2, 3}
    # the zero-pressure guest is tried and never inserted
    ar = ref[:, 1] == 3
    assert ar.sum() > 0 and not np.any(ref[ar & (ref[:, 0] == 1), 2])
    b = Batch(1)
    cid = b.add_chain(ctx, 298.0, press, seed, freq=freq, delr=A["delr"])
    b.enable_log(cid, n)
    b.run(n)
    log = b.get_log(cid, n)
    b.close()
    _compare(log, ref)


def test_bench_workload_chains_fp64(fmc, oracle):
    """Three chains of the 256-chain bench workload (orthorhombic cells of
    unequal sides, a kmax per axis, 471 to 5445 framework atoms, CO2 at
    10 bar) in one FP64 batch: each follows the oracle's trajectory over
    2000 moves from an empty framework."""
    from fastmc_amd.batch import Batch
    from fastmc_amd import synth
    ws = synth.c3()
    picks = [255, 151, 183]   # frameworks 31, 18, 22 at the highest pressure
    n = 2000                  # the first insertions are accepted after 270-960 moves
    b = Batch(len(picks))
    ids, refs = [], []
    for i in picks:
        w = ws[i]
        A = w.system_arrays()
        ctx = fmc.Context.from_arrays(A)
        press = [p * 1e5 for p in w.press_bar]
        cid = b.add_chain(ctx, w.temp, press, w.seed, delr=w.delr)
        b.enable_log(cid, n)
        ids.append(cid)
        refs.append(_oracle_recs(oracle, A, w.temp, press, w.seed, None, n)[0])
    b.run(n)
    for cid, ref, i in zip(ids, refs, picks):
        worst = _compare(b.get_log(cid, n), ref)
        assert ref[:, 2].sum() > 0
        print(f"chain {i}: {int(ref[:, 2].sum())} accepted, worst estep rel {worst:.2e}")
    b.close()


def test_serving_requests_fp64(fmc, oracle):
    """Three generated serving requests (1-3 of CO2, N2, CH4; 894 to 9875
    framework atoms; their own temperatures and pressures) in one FP64
    batch: each makes the oracle's decisions over 1500 moves.  Energies use
    the bar of the large-cell stress test: each component within 1e-10 of
    max(its magnitude, 100 internal units), and estep within 1e-10 of its
    components' summed magnitude, since estep can be the small difference
    of components of ~10^3 internal units (here -0.003 kcal/mol from
    -1549 + 1549), where summation order alone moves it by ~1e-12 kcal/mol."""
    from fastmc_amd.batch import Batch
    from fastmc_amd import synth
    rng = np.random.default_rng(77)
    ws = [synth.c4_request(rng, i) for i in range(8)]
    picks = [3, 5, 7]
    n = 1500
    b = Batch(len(picks))
    ids, refs = [], []
    for i in picks:
        w = ws[i]
        A = w.system_arrays()
        ctx = fmc.Context.from_arrays(A)
        press = [p * 1e5 for p in w.press_bar]
        cid = b.add_chain(ctx, w.temp, press, w.seed, delr=w.delr)
        b.enable_log(cid, n)
        ids.append(cid)
        s = oracle.OracleSystem(A["imcon"], A["cell"], A["alpha"], A["kmax"], A["rcut"],
                                A["delr"], A["ntpatm"], A["vdw"], A["fw"], A["guests"])
        recs = oracle.OracleChain(s, w.temp, press, w.seed, delr=w.delr).run(n)
        refs.append(np.array([[r.kind, r.iguest, r.accepted, r.nmol_after, r.estep,
                               r.e1, r.e2, r.vdw] for r in recs]))
    b.run(n)
    for cid, ref, i in zip(ids, refs, picks):
        log = b.get_log(cid, n)
        assert np.array_equal(log[:, :4], ref[:, :4]), i
        comp = np.abs(log[:, 5:8] - ref[:, 5:8]) / np.maximum(np.abs(ref[:, 5:8]), 100.0)
        scale = np.abs(ref[:, 5:8]).sum(axis=1) / 418.4
        rel = np.abs(log[:, 4] - ref[:, 4]) / np.maximum(np.maximum(np.abs(ref[:, 4]), scale), 1e-3)
        print(f"request {i} ({len(ws[i].guests)} species, {ws[i].fw.n} atoms, "
              f"{ws[i].temp:.1f} K): {int(ref[:, 2].sum())} accepted, worst component "
              f"{comp.max():.2e}, worst estep {rel.max():.2e}")
        assert comp.max() < 1e-10 and rel.max() < 1e-10, i
    assert sum(r[:, 2].sum() for r in refs) > 100
    b.close()


def test_batch_mixed_precision_error(fmc, oracle):
    """Mixed precision on the same system as the FP64 oracle chain: the
    same decisions for the first 50 moves at least, and while the
    trajectories agree, on the moves with |estep| < 50 kT, an estep error
    below 1e-4 kT and below 1e-4 relative to max(|estep|, 1e-3 kT)."""
    from fastmc_amd.batch import Batch
    from fastmc_amd.capi import PREC_MIXED
    A = _porous(seed=4, guests=("co2",), press=(20.0,)).system_arrays()
    n = 600
    ref, _, _ = _oracle_recs(oracle, A, 298.0, [20.0e5], (101, 23, 17, 5), None, n)
    b = Batch(1, PREC_MIXED)
    ctx = fmc.Context.from_arrays(A)
    cid = b.add_chain(ctx, 298.0, [20.0e5], (101, 23, 17, 5), delr=A["delr"])
    b.enable_log(cid, n)
    b.run(n)
    log = b.get_log(cid, n)
    same = np.all(log[:, :4] == ref[:, :4], axis=1)
    first_diff = np.argmin(same) if not same.all() else n
    assert first_diff > 50
    kT = 1.9872041e-3 * 298.0
    # moves whose |dE| exceeds 50 kT (overlaps) are rejected with certainty
    # in either precision; the error is assessed on the others
    a, m = ref[:first_diff, 4], log[:first_diff, 4]
    rel_moves = np.abs(a) < 50 * kT
    err = np.abs(m - a)[rel_moves]
    av = np.abs(a[rel_moves])
    raw = err / np.maximum(av, 1e-300)
    floored = err / np.maximum(av, 1e-3 * kT)
    print(f"mixed: identical for {first_diff} moves, {rel_moves.sum()} with |dE|<50kT: "
          f"max err {err.max():.2e} kcal/mol = {err.max() / kT:.2e} kT, median "
          f"{np.median(err):.2e}; raw rel max {raw.max():.2e} (within 1e-6: "
          f"{np.mean(raw < 1e-6):.1%}); floored(1e-3 kT) rel max {floored.max():.2e}")
    assert err.max() < 1e-4 * kT
    assert floored.max() < 1e-4


@pytest.mark.parametrize("kmax", [12, 14])
def test_batch_mixed_large_nk_register_tiles(fmc, oracle, kmax):
    """Mixed precision on the 512x12 and 512x14 register tiles: the same
    decisions as the FP64 oracle while the trajectories agree, and the
    estep error bars of the small-system test (below 1e-4 kT, and below
    1e-4 relative to max(|estep|, 1e-3 kT))."""
    from fastmc_amd.batch import Batch
    from fastmc_amd.capi import PREC_MIXED
    from fastmc_amd import synth
    rng = np.random.default_rng(31 + kmax)
    fw = synth.make_framework(rng, (26.0, 26.0, 26.0), 260)
    w = synth.Workload("porous", fw, ["co2"], [20.0], 298.0, (kmax, kmax, kmax))
    A = w.system_arrays()
    n, seed = 400, (7, 11, 13, 17)
    ref = _oracle_recs(oracle, A, 298.0, [20.0e5], seed, None, n)[0]
    b = Batch(1, PREC_MIXED)
    ctx = fmc.Context.from_arrays(A)
    cid = b.add_chain(ctx, 298.0, [20.0e5], seed, delr=A["delr"])
    b.enable_log(cid, n)
    b.set_tile(512)
    b.run(n)
    log = b.get_log(cid, n)
    b.close()
    same = np.all(log[:, :4] == ref[:, :4], axis=1)
    first_diff = np.argmin(same) if not same.all() else n
    assert first_diff > 50
    kT = 1.9872041e-3 * 298.0
    a, m = ref[:first_diff, 4], log[:first_diff, 4]
    keep = np.abs(a) < 50 * kT
    err = np.abs(m - a)[keep]
    floored = err / np.maximum(np.abs(a[keep]), 1e-3 * kT)
    print(f"kmax {kmax}: identical for {first_diff} moves; max err {err.max() / kT:.2e} kT, "
          f"median {np.median(err) / kT:.2e} kT; floored rel max {floored.max():.2e}")
    assert err.max() < 1e-4 * kT
    assert floored.max() < 1e-4


def test_batch_recompute_matches_oracle_energy(fmc, oracle):
    """FP64 recompute after 1500 moves: the full guest energy equals the
    oracle's single_point of the same (identical-trajectory) configuration,
    the running S(k) equals the exact one, and the running energy sum has
    not drifted."""
    from fastmc_amd.batch import Batch
    A = _porous(seed=4, guests=("co2",), press=(20.0,)).system_arrays()
    n = 1500
    ref, ch, s = _oracle_recs(oracle, A, 298.0, [20.0e5], (101, 23, 17, 5), None, n)
    b = Batch(1)
    ctx = fmc.Context.from_arrays(A)
    cid = b.add_chain(ctx, 298.0, [20.0e5], (101, 23, 17, 5), delr=A["delr"])
    st0 = b.stats(cid)
    assert st0["e_full"] == 0.0 and st0["recomputes"] == 1   # admission
    b.run(n)
    skdev, drift = b.recompute()
    st = b.stats(cid)
    assert st["nmol"][0] == ref[-1, 3] and st["nmol"][0] > 5
    eo = s.energy_full()
    e_ref = (eo[1] + eo[2] + eo[3] + eo[4] + eo[5]) / 418.4
    print(f"recompute: e_full {st['e_full']:.10f} oracle {e_ref:.10f} "
          f"drift {st['drift']:.2e} skdev {skdev:.2e}")
    assert abs(st["e_full"] - e_ref) <= 1e-9 * max(1.0, abs(e_ref))
    assert abs(drift) < 1e-8 * max(1.0, abs(e_ref))
    assert skdev < 1e-9


def test_batch_admission_energy_is_revive_kat(fmc, restart_example):
    """A chain admitted with the REVIVE configuration (90 CO2) starts from
    the reference's accumulated energy -731.5072668178669 kcal/mol."""
    from fastmc_amd.batch import Batch
    from conftest import revive_config
    ex = restart_example
    field, config, e_rev = revive_config(ex)
    ctx = fmc.context_from_inputs(field, config, ex["control"])
    for prec in (0, 1):
        b = Batch(2, prec)
        cid = b.add_chain(ctx, 303.15, [0.994808e5], (1, 2, 3, 4), delr=1.0)
        st = b.stats(cid)
        assert abs(st["e_full"] - e_rev) < 1e-7, (prec, st["e_full"], e_rev)
        assert st["energy"][0] == st["e_full"]
        b.close()


def test_graph_cycles_equal_eager(fmc):
    """Replays of the captured (moves + FP64 recompute) cycle, with eager
    work and synchronisation between them, equal the eager sequence on an
    identical batch, replay by replay."""
    from fastmc_amd.batch import Batch
    from fastmc_amd.capi import PREC_MIXED
    ws = [_porous(seed=s, guests=("co2",), press=(p,)) for s, p in ((4, 20.0), (9, 2.0))]
    bs = []
    for _ in range(2):
        b = Batch(2, PREC_MIXED)
        for i, w in enumerate(ws):
            A = w.system_arrays()
            b.add_chain(fmc.Context.from_arrays(A), 298.0, [w.press_bar[0] * 1e5],
                        (11 + i, 22, 33, 44), delr=A["delr"])
        b.set_tile(256)
        bs.append(b)
    g, e = bs
    for rep in range(4):
        g.launch_cycles(1, 150)
        g.sync()
        e.run(150)
        e.recompute()
        for cid in range(2):
            sg, se = g.stats(cid), e.stats(cid)
            for key in ("steps", "nmol", "energy", "e_full", "drift", "recomputes",
                        "accepted"):
                assert sg[key] == se[key], (rep, cid, key, sg[key], se[key])
    # several replays in one call
    g.launch_cycles(3, 150)
    g.sync()
    for _ in range(3):
        e.run(150)
        e.recompute()
    for cid in range(2):
        assert g.stats(cid)["energy"] == e.stats(cid)["energy"]
    ms, nl = g.timing()
    assert nl >= 5 and ms > 0.0


def test_max_moves_and_poll(fmc):
    from fastmc_amd.batch import Batch
    from fastmc_amd.capi import PREC_MIXED
    A = _porous(seed=4, guests=("co2",), press=(20.0,)).system_arrays()
    b = Batch(3, PREC_MIXED)
    ids = [b.add_chain(fmc.Context.from_arrays(A), 298.0, [20.0e5], (5, 6, 7, i + 1),
                       delr=A["delr"], max_moves=m) for i, m in enumerate((120, 0, 300))]
    b.run(200)
    assert b.poll() == [ids[0]]
    b.run(200)
    assert b.poll() == [ids[0], ids[2]]
    assert [b.stats(i)["steps"] for i in ids] == [120, 400, 300]
    b.remove_chain(ids[0])
    b.run(10)
    assert b.stats(ids[1])["steps"] == 410


def test_stream_form_equals_register_form(fmc):
    """Mixed precision: the streaming form (S(k) in HBM) and the
    register-resident form give bit-identical move logs."""
    from fastmc_amd.batch import Batch
    from fastmc_amd.capi import PREC_MIXED
    A = _porous(seed=4, guests=("co2",), press=(20.0,)).system_arrays()
    logs = []
    for tile in (512, 512 | 0x10000):
        b = Batch(1, PREC_MIXED)
        cid = b.add_chain(fmc.Context.from_arrays(A), 298.0, [20.0e5], (3, 1, 4, 1),
                          delr=A["delr"])
        b.set_tile(tile)
        b.enable_log(cid, 800)
        b.run(800)
        assert b.tile() == tile
        logs.append(b.get_log(cid, 800))
        b.close()
    assert np.array_equal(logs[0], logs[1])


def test_batch_fp64_triclinic_revive(fmc, oracle, restart_example):
    """The batch on the reference's triclinic example, started from its
    REVIVE configuration (90 CO2): identical trajectory to the oracle chain,
    and the admission energy is known answer 2."""
    from fastmc_amd.batch import Batch
    from conftest import revive_config
    ex = restart_example
    field, config, e_rev = revive_config(ex)
    n = 400
    s, _, _ = oracle.system_from_inputs(field, config, ex["control"])
    ch = oracle.OracleChain(s, 303.15, [0.994808e5], (7, 8, 9, 10), delr=1.0)
    recs = ch.run(n)
    ref = np.array([[r.kind, r.iguest, r.accepted, r.nmol_after, r.estep] for r in recs])
    ctx = fmc.context_from_inputs(field, config, ex["control"])
    b = Batch(1)
    cid = b.add_chain(ctx, 303.15, [0.994808e5], (7, 8, 9, 10), delr=1.0)
    assert abs(b.stats(cid)["e_full"] - e_rev) < 1e-7
    b.enable_log(cid, n)
    b.run(n)
    worst = _compare(b.get_log(cid, n), ref)
    print(f"triclinic revive: {int(ref[:, 2].sum())} accepted, worst estep rel {worst:.2e}")


def test_batch_binary_example(fmc, oracle, restart_example):
    """The reference's binary example (CO2 + TIP4P-Ew with its massless
    charge site, started from the unary REVCON): FP64 gives the oracle
    chain's trajectory; mixed precision makes the same decisions, with
    estep within 1e-5 kcal/mol of FP64 where |estep| < 100 kcal/mol and
    within 1e-6 relative on overlapping trials (|estep| up to ~1e9, all
    rejected)."""
    from fastmc_amd.batch import Batch
    from test_restart_files import binary_example
    field, config, control = binary_example(restart_example)
    gids = sorted(control.guests)
    press = [control.guests[g].press_pa for g in gids]
    freq = [control.guests[g].freq for g in gids]
    n, seed = 600, (31, 41, 59, 26)
    s, _, _ = oracle.system_from_inputs(field, config, control)
    ch = oracle.OracleChain(s, control.temp, press, seed, freq=freq, delr=control.delr)
    recs = ch.run(n)
    ref = np.array([[r.kind, r.iguest, r.accepted, r.nmol_after, r.estep] for r in recs])
    ctx = fmc.context_from_inputs(field, config, control)
    logs = []
    for prec in (0, 1):
        b = Batch(1, prec)
        b.set_tile(256)
        cid = b.add_chain(ctx, control.temp, press, seed, freq=freq, delr=control.delr)
        b.enable_log(cid, n)
        b.run(n)
        logs.append(b.get_log(cid, n))
        b.close()
    worst = _compare(logs[0], ref)
    same = np.all(logs[1][:, :4] == ref[:, :4], axis=1)
    k = int(np.argmin(same)) if not same.all() else n
    d = np.abs(logs[1][:k, 4] - ref[:k, 4])
    small = np.abs(ref[:k, 4]) < 100.0
    dmix = d[small].max()
    rbig = (d[~small] / np.abs(ref[:k, 4][~small])).max() if (~small).any() else 0.0
    print(f"binary example: {int(ref[:, 2].sum())} accepted, water moves "
          f"{int((ref[:, 1] == 1).sum())}, worst FP64 estep rel {worst:.2e}, "
          f"mixed identical for {k} moves, max |dE| {dmix:.2e} kcal/mol "
          f"(|estep| < 100), max rel {rbig:.2e} on {int((~small).sum())} overlaps")
    assert (ref[:, 1] == 1).sum() > 100
    assert k > 100 and dmix < 1e-5 and rbig < 1e-6


def test_capacity_fault_reported(fmc):
    """Guest capacity exhausted (the reference's error 75, gcmc.f:870): the
    chain stops inserting, flags the fault, and poll() hands it back."""
    from fastmc_amd.batch import Batch
    from fastmc_amd.capi import PREC_MIXED
    A = _porous(seed=4, guests=("co2",), press=(20.0,)).system_arrays()
    ctx = fmc.Context.from_arrays(A, max_guest_atoms=10)
    b = Batch(2, PREC_MIXED)
    full = b.add_chain(ctx, 298.0, [20.0e5], (1, 2, 3, 4), delr=A["delr"])
    ok = b.add_chain(fmc.Context.from_arrays(A), 298.0, [20.0e5], (1, 2, 3, 4),
                     delr=A["delr"], max_moves=10 ** 9)
    b.run(3000)
    st = b.stats(full)
    assert st["capacity_fault"] and st["nmol"][0] <= 3
    assert not b.stats(ok)["capacity_fault"] and b.stats(ok)["nmol"][0] > 3
    assert b.poll() == [full]


def _two_batches(fmc, n_chains=3, **kw):
    from fastmc_amd.batch import Batch
    from fastmc_amd.capi import PREC_MIXED
    out = []
    for _ in range(2):
        b = Batch(n_chains, PREC_MIXED)
        for i in range(n_chains):
            w = _porous(seed=4 + i, guests=("co2",), press=(20.0 / (i + 1),))
            A = w.system_arrays()
            b.add_chain(fmc.Context.from_arrays(A), 298.0, [w.press_bar[0] * 1e5],
                        (9 + i, 8, 7, 6), delr=A["delr"], **kw)
        b.set_tile(512)
        out.append(b)
    return out


def test_inlaunch_recompute_equals_separate(fmc):
    """A recompute every 300 moves inside one 900-move launch gives the same
    trajectories as three 300-move launches each followed by the separate
    recompute kernel (same S(k) bits), and the same bookkeeping."""
    a, b = _two_batches(fmc, recompute_every=300)
    for cid in range(3):
        a.enable_log(cid, 900)
        b.enable_log(cid, 900)
    a.run(900)
    for _ in range(3):
        b.run(300)   # recompute_every is met at the launch ends too ...
    # ... so b also recomputed inside; compare against a third batch that
    # relies on the separate kernel only
    c, _ = _two_batches(fmc)
    for cid in range(3):
        c.enable_log(cid, 900)
    for _ in range(3):
        c.run(300)
        c.recompute()
    for cid in range(3):
        la, lb, lc = a.get_log(cid, 900), b.get_log(cid, 900), c.get_log(cid, 900)
        assert np.array_equal(la, lb) and np.array_equal(la, lc), cid
        sa, sc = a.stats(cid), c.stats(cid)
        for key in ("steps", "nmol", "energy", "e_acc", "accepted", "recomputes"):
            assert sa[key] == sc[key], (cid, key, sa[key], sc[key])
        assert abs(sa["e_full"] - sc["e_full"]) <= 1e-10 * max(1.0, abs(sc["e_full"]))


def test_budget_launches_do_not_change_trajectories(fmc):
    """Time-budgeted launches split each chain's moves at arbitrary points;
    with the recompute cadence kept per chain, the trajectory equals a run
    of the same number of moves in fixed launches."""
    a, b = _two_batches(fmc, recompute_every=250)
    for cid in range(3):
        a.enable_log(cid, 4000)
        b.enable_log(cid, 4000)
    for _ in range(3):
        a.launch_budget(1300, 2000.0)
        a.sync()
    steps = [a.stats(cid)["steps"] for cid in range(3)]
    # every launch ended on its time budget, not on its move cap, so each
    # chain's moves were split where its own clock said (three chains of
    # nearly equal speed may well stop at the same counts)
    assert min(steps) > 0 and max(steps) < 3 * 1300, steps
    for cid in range(3):
        # the same moves for this chain in launches of 111
        bb, _ = _two_batches(fmc, n_chains=3, recompute_every=250)
        bb.enable_log(cid, 4000)
        done = 0
        while done < steps[cid]:
            n = min(111, steps[cid] - done)
            bb.run(n)
            done += n
        assert np.array_equal(a.get_log(cid, steps[cid]), bb.get_log(cid, steps[cid]))
        assert a.stats(cid)["recomputes"] == bb.stats(cid)["recomputes"]
        bb.close()


def test_stats_and_log_calls_wait_for_the_launch_in_flight(fmc):
    """fmc_batch_stats, fmc_batch_enable_log and fmc_batch_get_log read or
    rewrite the chain struct that a running launch holds in LDS and writes
    back at its end: issued while a launch is in flight, they wait for it.
    The log enabled that way stays enabled and records the next launch; the
    stats read meanwhile equal a read after an explicit sync."""
    from fastmc_amd.batch import Batch
    from fastmc_amd import synth
    A = synth.c1().system_arrays()
    b = Batch(2)
    ctx = fmc.Context.from_arrays(A)
    c0 = b.add_chain(ctx, 298.0, [0.15e5], (12, 34, 56, 78), delr=A["delr"])
    c1 = b.add_chain(ctx, 298.0, [0.15e5], (21, 43, 65, 87), delr=A["delr"])
    b.launch(400)
    st_early = b.stats(c0)          # no sync() before it
    b.sync()
    st_synced = b.stats(c0)
    assert st_early["steps"] == st_synced["steps"] == 400
    assert st_early == st_synced
    b.launch(300)
    b.enable_log(c1, 200)           # while the second launch runs
    b.sync()
    b.launch(200)
    log = b.get_log(c1, 200)        # while the third launch runs
    assert b.stats(c1)["steps"] == 900
    assert np.all(log[:, 0] >= 0) and np.any(log[:, 0] != 0)
    b.close()


def test_device_proposals_are_rodrigues_rotations(fmc, oracle):
    """The batch kernel's own proposals (wave-wide propose(), FP64 chain):
    a chain of methane (five sites, so the rotation is fixed completely)
    from an empty framework tries insertions until one is accepted, then
    only displacements.  The molecule after the accepted insertion is the
    template rotated by 2 pi r1 about normalize(2 r2 - 1, 2 r3 - 1, 2 r4 - 1)
    and placed at (ra, rb, rc) in cell coordinates; after each accepted
    displacement it is the previous molecule translated by (2 r - 1) delr /
    2, its centre of mass wrapped into the cell and rotated by r1 rotangle
    about it.  The draws come from the oracle's duni stream of the same
    seed, taken in the reference's order (gcmc.f:745-778, 871-877,
    1180-1223, 1330-1338; utility_pack.f:2545-3000): an insertion takes the
    move draw, seven proposal draws and the acceptance draw; a displacement
    the move draw, eight proposal draws and the acceptance draw only when
    estep >= 0."""
    import math
    from scipy.spatial.transform import Rotation
    from fastmc_amd.batch import Batch
    from fastmc_amd import synth
    w = synth.Workload("c1 methane", synth.c1().fw, ["ch4"], [50.0], 298.0, (8, 8, 8))
    A = w.system_arrays()
    g = A["guests"][0]
    T = np.vstack([g["x"], g["y"], g["z"]]).T
    m = np.asarray(g["mass"])
    cell = np.asarray(A["cell"], float).reshape(3, 3).T   # columns a, b, c
    ctx = fmc.Context.from_arrays(A)
    disp_only = [[0, 0, 1, 0, 0, 0, 0, 0, 0]]

    def axis(r2, r3, r4):
        b = np.array([2 * r2 - 1, 2 * r3 - 1, 2 * r4 - 1])
        return b / np.linalg.norm(b)

    n_moves = 300
    checked_ins = checked_disp = 0
    for seed in ((11, 23, 37, 5), (101, 7, 59, 130)):
        b = Batch(1)
        cid = b.add_chain(ctx, 298.0, [50.0e5], seed, freq=disp_only, delr=1.0)
        b.enable_log(cid, n_moves)
        d = oracle.Duni(*seed)
        d()                                      # the discarded first draw
        draws = [d() for _ in range(12 * n_moves)]
        off = 0                                  # the move's first draw
        cur = None
        for mv in range(n_moves):
            b.run(1)
            rec = b.get_log(cid, mv + 1)[mv]
            kind, acc, estep = int(rec[0]), int(rec[2]), rec[4]
            st = b.stats(cid)
            if kind == 1:
                assert cur is None
                if acc:
                    x, y, z = b.molecules(cid, 0)
                    cur = np.vstack([x, y, z]).T
                    ra, rb, rc, r1, r2, r3, r4 = draws[off + 1:off + 8]
                    R = Rotation.from_rotvec(2 * math.pi * r1 * axis(r2, r3, r4)).as_matrix()
                    want = T @ R.T + cell @ np.array([ra, rb, rc])
                    assert np.abs(cur - want).max() < 1e-12, (seed, mv)
                    checked_ins += 1
                off += 9
            else:
                assert kind == 3 and cur is not None
                # the step sizes adapt only every 50 / 100 displacements
                assert st["attempted"][3] < 50
                if acc:
                    x, y, z = b.molecules(cid, 0)
                    got = np.vstack([x, y, z]).T
                    moved = cur + (2 * np.array(draws[off + 1:off + 4]) - 1) * 0.5
                    com = (moved * m[:, None]).sum(axis=0) / m.sum()
                    f = np.linalg.solve(cell, com)
                    wrapped = cell @ (f - np.floor(f))
                    r1, r2, r3, r4 = draws[off + 4:off + 8]
                    R = Rotation.from_rotvec(r1 * (math.pi / 3) * axis(r2, r3, r4)).as_matrix()
                    want = (moved - com) @ R.T + wrapped
                    assert np.abs(got - want).max() < 1e-12, (seed, mv)
                    cur = got
                    checked_disp += 1
                off += 9 + (1 if estep >= 0.0 else 0)
                if st["attempted"][3] >= 45:
                    break
        b.close()
    print(f"\nchecked {checked_ins} insertions and {checked_disp} displacements")
    assert checked_ins >= 2 and checked_disp >= 20


def test_reused_device_blocks_give_the_same_trajectory(fmc):
    """A chain admitted after another system's chain left the batch takes
    that chain's released device blocks (the batch keeps them for reuse
    instead of freeing them) and still moves exactly as in a batch of its
    own: every buffer is written in full before it is read."""
    from fastmc_amd.batch import Batch
    from fastmc_amd.capi import PREC_MIXED
    A = _porous(seed=4, guests=("co2",), press=(20.0,)).system_arrays()
    X = _porous(seed=5, guests=("co2",), press=(30.0,)).system_arrays()
    n = 600

    def add(b, arrays, press, seed):
        return b.add_chain(fmc.Context.from_arrays(arrays), 298.0, [press], seed,
                           delr=arrays["delr"])

    ref_b = Batch(1, PREC_MIXED)
    ref_b.set_tile(512)
    cid = add(ref_b, A, 20.0e5, (3, 1, 4, 1))
    ref_b.enable_log(cid, n)
    ref_b.run(n)
    ref = ref_b.get_log(cid, n)
    b = Batch(2, PREC_MIXED)
    b.set_tile(512)
    for round_ in range(2):
        # a chain of another framework fills its blocks with its own state,
        # then leaves: its chain and system blocks go back for reuse
        x = add(b, X, 30.0e5, (9, 8, 7, 6 + round_))
        b.run(400)
        assert b.stats(x)["steps"] == 400
        b.remove_chain(x)
        cid = add(b, A, 20.0e5, (3, 1, 4, 1))
        b.enable_log(cid, n)
        b.run(n)
        assert np.array_equal(b.get_log(cid, n), ref), f"round {round_}"
        b.remove_chain(cid)
    b.close()
    ref_b.close()
