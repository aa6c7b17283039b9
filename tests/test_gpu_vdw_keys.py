"""Every vdW form of vdw_terms.f:189-396 (keys 1-9) on the device paths:
a porous system whose pairs take the nine forms in turn (two assignments,
so each guest-framework pair meets several forms) through

* the single-chain drop-in: random inserts, deletes and displacements with
  commits, each trial's components within 1e-10 of the oracle's;
* the FP64 batch: the oracle chain's trajectory over 1500 moves (identical
  kinds, decisions and loadings), at 256 and 512 threads per chain, with
  each energy component within 1e-10 of max(its magnitude, 100 internal
  units) and estep within 1e-10 of max(|estep|, its components' summed
  magnitude): the soft forms (Morse, the 12-10 bond) let a trial's
  components reach ~1e5 internal units while estep is their small
  difference, where summation order alone moves it by ~1e-10 relative;
* the mixed-precision batch: the FP64 decisions for the first 50 moves at
  least, and while the trajectories agree, on the moves with |estep| <
  50 kT, the bars of the batched workload: an estep error below 1e-4 kT
  and below 1e-3 relative to max(|estep|, 1e-3 kT);
* the serving path: a request on the keyed system gives the statistics of
  the same chain run alone in a batch, bit for bit."""
import numpy as np
import pytest

from keyed import keyed_arrays
from test_gpu_batch import _oracle_recs
from test_gpu_single import _random_moves

pytestmark = pytest.mark.gpu

KT = 1.9872041e-3 * 298.0


@pytest.fixture(scope="module")
def fmc():
    import fastmc_amd
    fastmc_amd.lib()
    return fastmc_amd


def _system(first=0, seed=21, press=(4.0, 8.0)):
    from fastmc_amd import synth
    rng = np.random.default_rng(seed)
    fw = synth.make_framework(rng, (26.0, 26.0, 26.0), 260)
    w = synth.Workload("keys", fw, ["co2", "n2"], list(press), 298.0, (8, 8, 8))
    A, used = keyed_arrays(w.system_arrays(), first=first)
    assert used == list(range(1, 10))
    return A


@pytest.mark.parametrize("first", [0, 4])
def test_single_chain_trials_every_key(fmc, oracle, first):
    A = _system(first)
    ctx, orc, worst = _random_moves(fmc, oracle, A, 150, seed=13, gtypes=2)
    print(f"\nkeys from {first}: worst component rel diff {worst:.3e}")
    assert sum(orc.nmol(g) for g in range(2)) > 0


@pytest.mark.parametrize("tile", [256, 512])
def test_batch_fp64_every_key(fmc, oracle, tile):
    from fastmc_amd.batch import Batch
    n = 1500
    cases = [(_system(0), (3, 5, 7, 11)), (_system(4, seed=8), (17, 19, 23, 29))]
    b = Batch(len(cases))
    b.set_tile(tile)
    ids, refs = [], []
    for A, seed in cases:
        ctx = fmc.Context.from_arrays(A)
        cid = b.add_chain(ctx, 298.0, [4.0e5, 8.0e5], seed, delr=A["delr"])
        b.enable_log(cid, n)
        ids.append(cid)
        s = oracle.OracleSystem(A["imcon"], A["cell"], A["alpha"], A["kmax"], A["rcut"],
                                A["delr"], A["ntpatm"], A["vdw"], A["fw"], A["guests"])
        recs = oracle.OracleChain(s, 298.0, [4.0e5, 8.0e5], seed, delr=A["delr"]).run(n)
        refs.append(np.array([[r.kind, r.iguest, r.accepted, r.nmol_after, r.estep,
                               r.e1, r.e2, r.vdw] for r in recs]))
    b.run(n)
    for cid, ref in zip(ids, refs):
        log = b.get_log(cid, n)
        assert np.array_equal(log[:, :4], ref[:, :4])
        comp = np.abs(log[:, 5:8] - ref[:, 5:8]) / np.maximum(np.abs(ref[:, 5:8]), 100.0)
        scale = np.abs(ref[:, 5:8]).sum(axis=1) / 418.4
        rel = np.abs(log[:, 4] - ref[:, 4]) / np.maximum(np.maximum(np.abs(ref[:, 4]), scale),
                                                         1e-3)
        print(f"chain {cid}: {int(ref[:, 2].sum())} accepted, worst component "
              f"{comp.max():.2e}, worst estep {rel.max():.2e}")
        assert comp.max() < 1e-10 and rel.max() < 1e-10
        assert ref[:, 2].sum() > 20
    b.close()


def test_batch_mixed_every_key(fmc, oracle):
    from fastmc_amd.batch import Batch
    from fastmc_amd.capi import PREC_MIXED
    n = 600
    for first, seed in ((0, (3, 5, 7, 11)), (4, (101, 23, 17, 5))):
        A = _system(first)
        ref = _oracle_recs(oracle, A, 298.0, [4.0e5, 8.0e5], seed, None, n)[0]
        b = Batch(1, PREC_MIXED)
        ctx = fmc.Context.from_arrays(A)
        cid = b.add_chain(ctx, 298.0, [4.0e5, 8.0e5], seed, delr=A["delr"])
        b.enable_log(cid, n)
        b.run(n)
        log = b.get_log(cid, n)
        b.close()
        same = np.all(log[:, :4] == ref[:, :4], axis=1)
        first_diff = np.argmin(same) if not same.all() else n
        assert first_diff > 50
        a, m = ref[:first_diff, 4], log[:first_diff, 4]
        sel = np.abs(a) < 50 * KT
        err = np.abs(m - a)[sel]
        floored = err / np.maximum(np.abs(a[sel]), 1e-3 * KT)
        raw = err / np.maximum(np.abs(a[sel]), 1e-300)
        print(f"\nkeys from {first}: identical for {first_diff} moves; {sel.sum()} with "
              f"|dE| < 50 kT: max err {err.max() / KT:.2e} kT, floored rel {floored.max():.2e}, "
              f"raw rel <= 1e-6 on {np.mean(raw <= 1e-6):.1%}")
        assert err.max() < 1e-4 * KT
        assert floored.max() < 1e-3


def test_serving_request_every_key(fmc):
    from fastmc_amd import serve
    from fastmc_amd.batch import Batch
    A = _system(4)
    req = serve.Request(0, 0.0, A, 298.0, [4.0e5, 8.0e5], (3, 5, 7, 11), 700, delr=A["delr"])
    srv = serve.Server(slots=2, slice_moves=200)
    got = srv.run([req], realtime=False)[0]
    srv.close()
    b = Batch(1, srv.precision)
    b.set_tile(srv.tile)
    ctx = fmc.Context.from_arrays(A)
    cid = b.add_chain(ctx, 298.0, [4.0e5, 8.0e5], (3, 5, 7, 11), delr=A["delr"])
    b.run(700)
    st = b.stats(cid)
    b.close()
    for key in ("steps", "nmol", "energy", "accepted", "attempted", "sum_n", "sum_e"):
        assert got.stats[key] == st[key], key
    assert sum(got.stats["accepted"]) > 20
