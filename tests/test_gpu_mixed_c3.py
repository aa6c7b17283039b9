"""Mixed precision on the batched workload itself: one chain of every one of
the 32 synthetic frameworks (300-6000 atoms, orthorhombic 20-60 A cells,
kmax 8-14, CO2 at 0.01-10 bar, 298 K), run in mixed precision and in FP64
from the same seeds.  The FP64 batch chains follow the oracle's trajectory
(test_gpu_batch.py); here the mixed chains are compared with them, move by
move, over the prefix where the two trajectories coincide.

The bars, in kcal/mol and relative to |estep_fp64|:
  * at least 95 % of all moves have identical trajectories in the two
    precisions;
  * the absolute estep error is below 1e-4 kT on every move the acceptance
    test can take (|estep| < 50 kT);
  * the error relative to max(|estep|, 1e-3 kT) is below 1e-3 on every
    move;
  * at least 97 % of the moves have a raw relative error |err| / |estep|
    of at most 1e-6.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOVES = 600


@pytest.fixture(scope="module")
def logs():
    sys.path.insert(0, ROOT)
    import bench
    import fastmc_amd
    from fastmc_amd.capi import PREC_FP64, PREC_MIXED
    fastmc_amd.lib()
    picks = [8 * f + f % 8 for f in range(32)]
    out = []
    for prec in (PREC_FP64, PREC_MIXED):
        b, ws, _ = bench.build_batch(prec, picks=picks)
        for c in range(len(picks)):
            b.enable_log(c, MOVES)
        b.set_tile(512)
        b.run(MOVES)
        out.append([b.get_log(c, MOVES) for c in range(len(picks))])
        b.close()
    return bench, out


def test_mixed_estep_error_on_every_framework(logs):
    bench, (ref, mix) = logs
    d = bench.de_stats(ref, mix)
    kT = bench.KT_298
    print(f"\n{d['moves_compared']} moves compared, {d['identical_prefix_moves']} of "
          f"{32 * MOVES} in identical prefixes; abs max {d['abs_max_kcal']:.2e} kcal/mol "
          f"= {d['abs_max_kT']:.2e} kT, median {d['abs_median_kcal']:.2e}; raw rel max "
          f"{d['raw_rel_max']:.2e}, median {d['raw_rel_median']:.2e}; floored rel max "
          f"{d['floored_rel_max']:.2e}; raw <= 1e-6 on {d['frac_raw_le_1e-6']:.1%}, "
          f"floored <= 1e-6 on {d['frac_floored_le_1e-6']:.1%}")
    for c, (a, m) in enumerate(zip(ref, mix)):
        same = np.all(a[:, :4] == m[:, :4], axis=1)
        n = int(np.argmin(same)) if not same.all() else len(same)
        t = a[:n, 0] != 0
        e = np.abs(m[:n, 4] - a[:n, 4])[t]
        print(f"  framework {c:2d}: identical {n:4d} moves, worst {e.max() if e.size else 0:.2e}"
              f" kcal/mol")
    assert d["identical_prefix_moves"] >= 0.95 * 32 * MOVES
    assert d["abs_max_kcal"] < 1e-4 * kT
    assert d["floored_rel_max"] < 1e-3
    assert d["frac_raw_le_1e-6"] >= 0.97


def test_mixed_error_by_move_kind(logs):
    """The same comparison split by move kind -- insertions and deletions
    (S(k) changes by a whole molecule's term) and displacements (the change
    of each site's term, formed before rounding): the absolute error stays
    below 1e-4 kT on the decidable moves of each kind."""
    bench, (ref, mix) = logs
    kT = bench.KT_298
    for kinds, name in (((1,), "insert"), ((2,), "delete"), ((3, 4, 5), "displace")):
        d = bench.de_stats(ref, mix, kinds=kinds)
        print(f"\n{name}: {d['moves_compared']} moves ({d['decidable_moves']} below 50 kT), "
              f"abs max {d['abs_max_kT']:.2e} kT, median {d['abs_median_kcal'] / kT:.2e} kT, "
              f"floored max {d['floored_rel_max']:.2e}, raw <= 1e-6 on "
              f"{d['frac_raw_le_1e-6']:.1%}")
        assert d["decidable_moves"] > 100, name
        assert d["abs_max_kT"] < 1e-4, name
