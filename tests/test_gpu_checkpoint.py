"""Chain checkpoints (fmc_batch_checkpoint / fmc_batch_restore_chain), the
chain's guest coordinates (fmc_batch_get_molecules) and the REVIVE record
set written from a batch chain (utility_pack.f:1800-1814)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("steps", "prod_steps", "nmol", "energy", "accepted", "attempted", "sum_n",
        "sum_n2", "sum_e", "sum_en", "sum_e2", "delrdisp", "rotangle", "e_acc",
        "e_full", "drift", "sk_dev", "recomputes", "win_len", "win_closed", "win")


@pytest.fixture(scope="module")
def fmc():
    import fastmc_amd
    fastmc_amd.lib()
    return fastmc_amd


def _porous(seed, guests=("co2",), press=(20.0,), n=260, side=26.0):
    from fastmc_amd import synth
    rng = np.random.default_rng(seed)
    fw = synth.make_framework(rng, (side, side, side), n)
    return synth.Workload("porous", fw, list(guests), list(press), 298.0, (8, 8, 8))


def _batch(n, prec=0):
    """A batch with a fixed tile: a chain's arithmetic depends on the tile,
    so batches compared bit for bit share one."""
    from fastmc_amd.batch import Batch
    b = Batch(n, prec)
    b.set_tile(256)
    return b


def _cases():
    return [
        (_porous(4).system_arrays(), [20.0e5], (101, 23, 17, 5), None),
        (_porous(21, ("co2", "n2"), (4.0, 8.0)).system_arrays(), [4.0e5, 8.0e5],
         (3, 5, 7, 11), [[1, 1, 1, 0, 0, 0, 0.5, 0.5, 0]] * 2),
    ]


@pytest.mark.parametrize("prec", [0, 1])
def test_checkpoint_resume_continues_the_trajectory(fmc, prec):
    """n1 moves, checkpoint, the batch destroyed, the chain restored into a
    new batch beside an unrelated chain, n2 more moves: every statistic and
    the move log equal those of n1 + n2 uninterrupted moves, with the
    in-launch FP64 recompute cadence kept across the split."""
    n1, n2 = 260, 340
    for A, P, seed, freq in _cases():
        print(f"prec {prec} guests {len(P)} seed {seed}", flush=True)
        ctx = fmc.Context.from_arrays(A)
        kw = dict(freq=freq, delr=A["delr"], eq_steps=50, recompute_every=200)
        full = _batch(2, prec)
        cf = full.add_chain(ctx, 298.0, P, seed, **kw)
        full.enable_log(cf, n1 + n2)
        full.run(n1 + n2)
        want = full.stats(cf)
        want_log = full.get_log(cf, n1 + n2)[n1:]
        full.close()

        part = _batch(2, prec)
        cp = part.add_chain(ctx, 298.0, P, seed, **kw)
        part.run(n1)
        blob = part.checkpoint(cp)
        part.close()

        other = fmc.Context.from_arrays(_porous(9, ("co2",), (2.0,)).system_arrays())
        res = _batch(3, prec)
        res.add_chain(other, 298.0, [2.0e5], (9, 9, 9, 9))
        cr = res.restore_chain(ctx, blob)
        assert res.stats(cr)["steps"] == n1
        res.enable_log(cr, n2)
        res.run(n2)
        got = res.stats(cr)
        got_log = res.get_log(cr, n2)
        res.close()
        other.close()
        ctx.close()
        assert got["recomputes"] == want["recomputes"] >= 3
        for k in KEYS:
            assert got[k] == want[k], (k, got[k], want[k])
        assert np.array_equal(got_log, want_log)


def test_checkpoint_preempt_and_resume_in_one_batch(fmc):
    """Serving preemption: a chain checkpointed and removed while the others
    keep moving, then restored into the same batch, ends where it would
    have ended without the break (budgeted launches)."""
    from fastmc_amd.capi import PREC_MIXED
    (A, P, seed, freq), (A2, P2, seed2, freq2) = _cases()
    ctx, ctx2 = fmc.Context.from_arrays(A), fmc.Context.from_arrays(A2)
    ref = _batch(1, PREC_MIXED)
    c = ref.add_chain(ctx, 298.0, P, seed, max_moves=900)
    ref.run(900)
    want = ref.stats(c)
    ref.close()

    b = _batch(2, PREC_MIXED)
    c = b.add_chain(ctx, 298.0, P, seed, max_moves=900)
    c2 = b.add_chain(ctx2, 298.0, P2, seed2, freq=freq2)
    b.run(300)
    blob = b.checkpoint(c)
    b.remove_chain(c)
    b.run(200)
    c = b.restore_chain(ctx, blob)
    for _ in range(400):
        if b.stats(c)["steps"] >= 900:
            break
        b.launch_budget(10 ** 12, 500.0)
        b.sync()
    got = b.stats(c)
    assert b.stats(c2)["steps"] > 500
    b.close()
    for k in KEYS:
        assert got[k] == want[k], (k, got[k], want[k])


def test_checkpoint_refuses_a_foreign_system(fmc):
    """A checkpoint restores only onto the system it was taken from, and
    only whole."""
    from fastmc_amd.capi import FmcError
    (A, P, seed, _), (A2, _, _, _) = _cases()
    ctx, ctx2 = fmc.Context.from_arrays(A), fmc.Context.from_arrays(A2)
    b = _batch(3)
    c = b.add_chain(ctx, 298.0, P, seed)
    b.run(50)
    blob = b.checkpoint(c)
    for bad_ctx, bad_blob in ((ctx2, blob), (ctx, blob[:-8]), (ctx, b"x" + blob[1:])):
        with pytest.raises(FmcError):
            b.restore_chain(bad_ctx, bad_blob)
    # a duni state the generator cannot reach: the constants cd, cm sit
    # between the carry c and the indices ir, jr
    import struct
    at = blob.find(struct.pack("<ff", 7654321.0 / 16777216.0, 16777213.0 / 16777216.0))
    assert at > 0
    ir, jr = struct.unpack_from("<ii", blob, at + 8)
    assert (ir - jr) % 97 == 64
    for off, val in ((at + 8, struct.pack("<i", ir % 97 + 1)),           # indices not 64 apart
                     (at - 4, struct.pack("<f", 0.1))):                  # carry off the 2^-24 grid
        bad = bytearray(blob)
        bad[off:off + 4] = val
        with pytest.raises(FmcError):
            b.restore_chain(ctx, bytes(bad))
    # the failed restores left the slots free
    c3 = b.restore_chain(ctx, blob)
    c4 = b.restore_chain(ctx, blob)
    assert {c3, c4} == {0, 1, 2} - {c}
    b.close()


def test_production_sums_follow_the_move_log(fmc):
    """The production sums of N, N^2, E, E*N and E^2 (gcmc.f:1830-1848)
    rebuilt on the host from the FP64 move log."""
    A, P, seed, _ = _cases()[0]
    ctx = fmc.Context.from_arrays(A)
    n, eq = 700, 100
    b = _batch(1)
    c = b.add_chain(ctx, 298.0, P, seed, eq_steps=eq)
    e = b.stats(c)["energy"][0]
    b.enable_log(c, n)
    b.run(n)
    log = b.get_log(c, n)
    st = b.stats(c)
    b.close()
    sums = np.zeros(5)
    for i in range(n):
        if log[i, 2]:
            e = e + log[i, 4]
        if i >= eq:
            k = log[i, 3]
            sums += [k, k * k, e, e * k, e * e]
    assert st["prod_steps"] == n - eq
    assert abs(st["energy"][0] - e) <= 1e-12 * abs(e)
    got = [st["sum_n"][0], st["sum_n2"][0], st["sum_e"][0], st["sum_en"][0], st["sum_e2"][0]]
    for a, w in zip(got, sums):
        assert abs(a - w) <= 1e-11 * max(1.0, abs(w)), (a, w)


def test_batch_revive_records_the_chain_state(fmc, oracle, restart_example, tmp_path):
    """From the reference's restart example (REVIVE, 90 CO2), 600 FP64
    moves; the REVIVE written from the chain reads back with the chain's
    counts, statistics and accumulated energy, and the single_point energy
    of its configuration equals the chain's own FP64 recompute of the full
    energy.  (The accumulated energy differs from single_point by the tail
    term: the move loop's elrc starts at zero, gcmc.f:138-139, so the first
    insert or delete adds the whole tail; the reference's REVIVE energy
    carries the same offset.)"""
    from fastmc_amd import inputs
    from conftest import revive_config
    ex = restart_example
    field, config, _ = revive_config(ex)
    ctx = fmc.context_from_inputs(field, config, ex["control"])
    b = _batch(1)
    c = b.add_chain(ctx, 303.15, [0.994808e5], (7, 8, 9, 10), delr=1.0, eq_steps=100)
    b.run(600)
    b.recompute()
    st = b.stats(c)
    rev = b.revive(c, field, config)
    b.close()
    p = tmp_path / "REVIVE"
    inputs.write_revive(str(p), rev)
    back = inputs.read_revive(str(p), field)
    assert back.nummols[0] == st["nmol"][0] and back.nummols[1] == 36
    assert back.energy == st["energy"] and back.is_production
    assert back.chainstats[0] == st["prod_steps"]
    assert back.chainstats[1] == st["sum_n"][0] / st["prod_steps"]
    f2, c2 = inputs.apply_revive(field, config, back)
    s, _, _ = oracle.system_from_inputs(f2, c2, ex["control"])
    e6 = s.energy_full()
    e_full = (e6[1] + e6[2] + e6[3] + e6[4] + e6[5]) / field.engunit
    print(f"revive: N {back.nummols[0]} chain energy {st['energy'][0]:.10f} "
          f"recompute {st['e_full']:.10f} single_point {e_full:.10f}")
    assert abs(e_full - st["e_full"]) < 1e-7
    assert abs(st["drift"]) < 1e-8
    # the REVCON of the same state reads back as a CONFIG
    atoms = inputs.revcon_atoms(f2, back.coords)
    inputs.write_revcon(str(tmp_path / "REVCON"), config.title, config.imcon, config.cell,
                        atoms)
    c3 = inputs.read_config(str(tmp_path / "REVCON"), f2)
    assert np.allclose(np.array(c3.coords[0])[:, 0], back.coords[0][0], rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("case", [0, 1])
def test_window_statistics_follow_the_move_log(fmc, case):
    """The kernel's end-of-run windows (production samples j*L ... j*L + L,
    L = int(samples / 5)) equal the windows of the samples rebuilt from the
    FP64 move log, across launch boundaries; the last window misses the
    sample past the end, as the reference's does; the report equals the
    restatement applied to the rebuilt samples."""
    from fastmc_amd import stats
    A, P, seed, freq = _cases()[case]
    ctx = fmc.Context.from_arrays(A)
    M, Q = 1000, 150
    b = _batch(1)
    c = b.add_chain(ctx, 298.0, P, seed, freq=freq, delr=A["delr"], eq_steps=Q,
                    max_moves=M)
    st0 = b.stats(c)
    assert st0["win_len"] == (M - Q) // 5 == 170
    b.enable_log(c, M)
    for n in (300, 300, 400):
        b.run(n)
    st = b.stats(c)
    log = b.get_log(c, M)
    rep = b.report(c, 298.0)
    b.close()
    assert st["steps"] == M and st["prod_steps"] == M - Q and st["win_closed"] == 4
    ng = len(P)
    nm = list(st0["nmol"])
    en = list(st0["energy"])
    Ns = [[] for _ in range(ng)]
    Es = [[] for _ in range(ng)]
    for i in range(M):
        g = int(log[i, 1])
        nm[g] = int(log[i, 3])
        if log[i, 2]:
            en[g] = en[g] + log[i, 4]
        if i >= Q:
            for h in range(ng):
                Ns[h].append(float(nm[h]))
                Es[h].append(en[h])
    for g in range(ng):
        sums, nwind = stats.window_sums(Ns[g], Es[g])
        assert nwind == st["win_len"]
        got = np.array([[st["win"][j][q][g] for q in range(5)] for j in range(5)])
        assert np.allclose(got, np.array(sums), rtol=1e-12, atol=1e-9), g
        want = stats.from_samples(Ns[g], Es[g], 298.0)
        for z in range(7):
            a, w = rep[g].avg[z], want.avg[z]
            assert (np.isnan(a) and np.isnan(w)) or abs(a - w) <= 1e-9 * max(1.0, abs(w))
            a, w = rep[g].std[z], want.std[z]
            assert (np.isnan(a) and np.isnan(w)) or abs(a - w) <= 1e-7 * max(1.0, abs(w))
        print(f"guest {g}: <N> {rep[g].avg[0]:.4f} +/- {rep[g].std[0]:.4f}, "
              f"Q_st {rep[g].avg[5]:.3f} +/- {rep[g].std[5]:.3f}")


@pytest.mark.parametrize("prec", [0, 1])
def test_selectivity_follows_the_move_log(fmc, prec):
    """The kernel's selectivity windows (gcmc.f:1887-1916; 120-sample
    windows, closed in the kernel across launch boundaries and the prepared
    rejections) and Batch.report's last, partial window equal the
    restatement applied to the loadings rebuilt from the move log, to 1e-12
    relative; a checkpoint carries the window state."""
    from fastmc_amd import stats
    A, P, seed, freq = _cases()[1]
    ctx = fmc.Context.from_arrays(A)
    M, Q, W = 1000, 150, 120
    b = _batch(2, prec)
    c = b.add_chain(ctx, 298.0, P, seed, freq=freq, delr=A["delr"], eq_steps=Q,
                    max_moves=M, window=W)
    st0 = b.stats(c)
    b.enable_log(c, M)
    for n in (300, 300, 400):
        b.run(n)
    st = b.stats(c)
    log = b.get_log(c, M)
    rep = b.report(c, 298.0)
    assert st["sel_window"] == W and st["sel_count"] == (M - Q) // W == 7
    nm = list(st0["nmol"])
    samples = []
    for i in range(M):
        nm[int(log[i, 1])] = int(log[i, 3])
        if i >= Q:
            samples.append([float(v) for v in nm])
    want = stats.selectivity_from_samples(samples, nwind=W)
    got = rep.selectivity
    assert sorted(got) == sorted(want) == [(0, 1)]
    g, w = got[(0, 1)], want[(0, 1)]
    print(f"\nprec {prec}: S(1/2) = {g.mean:.9f} +/- {g.var:.6f} (restatement "
          f"{w.mean:.9f} +/- {w.var:.6f}), {st['sel_count']} windows + {M - Q - 7 * W} samples")
    assert abs(g.mean - w.mean) <= 1e-12 * abs(w.mean)
    assert abs(g.var - w.var) <= 1e-10 * max(abs(w.var), 1e-6)
    assert abs(g.second - w.second) <= 1e-12 * abs(w.second)
    assert g.mean > 0
    # the state travels in a checkpoint
    blob = b.checkpoint(c)
    b2 = _batch(1, prec)
    c2 = b2.restore_chain(ctx, blob)
    st2 = b2.stats(c2)
    for k in ("sel_window", "sel_count", "sel_base", "sel_mean", "sel_second", "sel_var"):
        assert st2[k] == st[k], k
    b2.close()
    b.close()
