"""Tile shapes of the batch kernel and the autotuner that picks one.

A shape is (threads per work-group, register tile of k-vectors per thread
or the streaming form, k-group width, real-space items per pass).  The
k-group width, the items per pass and the form only change how loads are
grouped: every k still sums its trial positions in order and every thread
its real-space items in order, so the bits depend on the thread count
alone.  The autotuner times every compiled shape that can run the batch,
records the one it keeps in each chain's statistics and checkpoint, and a
batch with that shape fixed reproduces the autotuned batch bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MOVES = 500


def _systems():
    from fastmc_amd import synth
    ws = synth.c3()
    # three frameworks of the batched workload: small, mid-size and the
    # kmax-14 one, at three pressures
    return [ws[8 * 0 + 7], ws[8 * 13 + 5], ws[8 * 30 + 6]]


def _run(prec, shape=None, n=MOVES):
    import fastmc_amd
    from fastmc_amd.batch import Batch
    ws = _systems()
    b = Batch(len(ws), prec)
    if shape is not None:
        b.set_tile_shape(shape["threads"], shape["kpt"], shape["kgroup"], shape["kunroll"])
    for w in ws:
        ctx = fastmc_amd.Context.from_arrays(w.system_arrays())
        c = b.add_chain(ctx, w.temp, [p * 1e5 for p in w.press_bar], w.seed, delr=w.delr,
                        recompute_every=200)
        b.enable_log(c, n)
    b.run(n)
    logs = [b.get_log(c, n) for c in range(len(ws))]
    tiles = [b.stats(c)["tile"] for c in range(len(ws))]
    got = b.tile_shape()
    b.close()
    return logs, tiles, got


@pytest.mark.parametrize("prec", [1, 0])
def test_autotuned_batch_equals_the_recorded_fixed_shape(prec):
    logs, tiles, shape = _run(prec)
    print(f"\nprecision {prec}: autotuned shape {shape}")
    assert shape["tuned"] == 2
    for t in tiles:
        assert t == {k: shape[k] for k in ("threads", "kpt", "kgroup", "kunroll")}
    fixed = dict(shape)
    again, tiles2, shape2 = _run(prec, fixed)
    assert shape2["tuned"] == 1 and tiles2 == tiles
    for a, b in zip(logs, again):
        assert np.array_equal(a, b)
    assert sum(int(l[:, 2].sum()) for l in logs) > 20


@pytest.mark.parametrize("prec", [1, 0])
def test_group_width_unroll_and_form_leave_the_bits_alone(prec):
    """At 512 threads: every k-group width and items per pass, on the
    register tile and on the streaming form, give the same move logs."""
    base = None
    for kpt in (-1, 0):
        for kg in (2, 4, 8):
            for ku in (1, 2):
                if kpt == 0 and kg != 4:
                    continue
                logs, tiles, shape = _run(prec, dict(threads=512, kpt=kpt, kgroup=kg, kunroll=ku),
                                          n=300)
                assert tiles[0]["kgroup"] == (kg if kpt else 4) and tiles[0]["kunroll"] == ku
                if base is None:
                    base = logs
                    continue
                for a, b in zip(base, logs):
                    assert np.array_equal(a, b), (kpt, kg, ku)


def test_checkpoint_carries_the_tile():
    """A checkpoint restores into a batch without a tile (which takes the
    chain's shape) and is refused by a batch whose tile has another thread
    count."""
    import fastmc_amd
    from fastmc_amd.batch import Batch
    from fastmc_amd.capi import FmcError
    w = _systems()[0]
    ctx = fastmc_amd.Context.from_arrays(w.system_arrays())
    b = Batch(1)
    b.set_tile_shape(512, -1, 8, 2)
    c = b.add_chain(ctx, w.temp, [p * 1e5 for p in w.press_bar], w.seed, delr=w.delr)
    b.run(200)
    blob = b.checkpoint(c)
    b.close()
    other = Batch(1)
    other.set_tile(256)
    with pytest.raises(FmcError) as e:
        other.restore_chain(ctx, blob)
    assert e.value.code == -2
    other.close()
    fresh = Batch(1)
    c2 = fresh.restore_chain(ctx, blob)
    shape = fresh.tile_shape()
    assert (shape["threads"], shape["kgroup"], shape["kunroll"], shape["tuned"]) == (512, 8, 2, 1)
    fresh.run(50)
    assert fresh.stats(c2)["steps"] == 250
    fresh.close()
