"""Stage clocks of the batch kernel on the batched CO2 workload.

Runs the 256-chain workload of the benchmark (32 synthetic frameworks x 8
CO2 pressures, mixed precision, the FP64 recompute every 1000 moves of each
chain, budgeted launches) at the fixed 512x12 tile (k-group width 8 and
one real-space item per pass unless --kgroup/--kunroll say otherwise) and
prints where a move's cycles go: the mean s_memtime cycles per move of each
stage, and per wave the end of its
parallel-stage work and of its decision-stage work.

The clocks come from the profiling build of the library, in which the
512x12 tile is compiled with FMC_STAGE_CLOCKS:

    make -C fastmc-1.4.0_amd clocks
    cd fastmc-1.4.0_amd && python -m fastmc_amd.stageclock

With --plain the normal library runs the same workload and only the rate is
printed (a run to collect hardware counters on).
"""
import argparse
import ctypes as C
import os
import time

import numpy as np

from . import capi

# slot layout of the kernel's stage clocks (csrc/fmc_batch_dev.h, kClk*)
PROPOSE, TABLES, PARALLEL, DECISION, COMMIT, MOVES, FRESH, DISPLACE = range(8)
WAVE_PARALLEL, WAVE_DECISION, WAVE_RECIP, BUILD, PREP, SLOTS = 8, 16, 24, 32, 36, 40


def build(recompute_every=1000, kgroup=8, kunroll=1):
    from . import synth
    from .batch import Batch
    ws = synth.c3()
    b = Batch(len(ws), capi.PREC_MIXED)
    b.set_tile_shape(512, 12, kgroup, kunroll)
    ctxs = {}
    for w in ws:
        key = id(w.fw)
        if key not in ctxs:
            ctxs[key] = capi.Context.from_arrays(w.system_arrays())
        b.add_chain(ctxs[key], w.temp, [p * 1e5 for p in w.press_bar], w.seed,
                    delr=w.delr, recompute_every=recompute_every)
    return b, len(ws)


def table(P):
    """Markdown lines of the stage table from the clock slots P[group][slot]."""
    moves = P[:, MOVES].sum()
    out = [f"moves {moves:.0f}; fresh proposals {P[:, FRESH].sum() / moves:.3f}, "
           f"displacements {P[:, DISPLACE].sum() / moves:.3f}", "",
           "| stage | cycles per move |", "|---|---|"]
    names = ((PROPOSE, "proposal after an acceptance (wave 0)"),
             (TABLES, "tables of a fresh proposal"),
             (PARALLEL, "reciprocal + real space (+ the prepared rejection)"),
             (DECISION, "decision on thread 0; waves 1-7 build the prepared tables"),
             (COMMIT, "commit and the in-launch recompute"))
    tot = 0.0
    for s, name in names:
        v = P[:, s].sum() / moves
        tot += v
        out.append(f"| {name} | {v:.0f} |")
    out.append(f"| **total** | **{tot:.0f}** |")
    out += ["", "| wave | parallel stage ends | its reciprocal part | decision stage ends |",
            "|---|---|---|---|"]
    for w in range(8):
        out.append(f"| {w} | {P[:, WAVE_PARALLEL + w].sum() / moves:.0f} | "
                   f"{P[:, WAVE_RECIP + w].sum() / moves:.0f} | "
                   f"{P[:, WAVE_DECISION + w].sum() / moves:.0f} |")
    # the decision stage in its two phases
    out += ["", "| decision stage, mean cycles per move | |", "|---|---|"]
    for i, name in enumerate(("phase factors of the prepared move (waves 1-7), to the barrier",
                              "after it: the rows on wave 1",
                              "after it: the prepared decision inputs and prefetch (wave 7)",
                              "after it: thread 0's decision")):
        out.append(f"| {name} | {P[:, BUILD + i].sum() / moves:.0f} |")
    # the prepared rejection on the last wave, cycles from its start
    prepared = max(P[:, PREP + 2].astype(bool).sum() and moves, 1.0)
    out += ["", "| the last wave's prepared rejection, mean cycles per move from its start to the end of | |",
            "|---|---|"]
    for i, name in enumerate(("the copy of the chain state and duni table",
                              "the booking of the move as rejected", "the next proposal")):
        out.append(f"| {name} | {P[:, PREP + i].sum() / prepared:.0f} |")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--plain", action="store_true", help="the normal library, no clocks")
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--budget-ms", type=float, default=15.0)
    ap.add_argument("--kgroup", type=int, default=8, help="k-group width of the 512x12 tile")
    ap.add_argument("--kunroll", type=int, default=1, help="real-space items per pass")
    args = ap.parse_args(argv)
    if not args.plain:
        capi.LIB_PATH = os.path.join(os.path.dirname(capi.LIB_PATH), "libfastmc_amd_clocks.so")
        if not os.path.exists(capi.LIB_PATH):
            raise SystemExit(f"{capi.LIB_PATH} missing: make -C fastmc-1.4.0_amd clocks")
    L = capi.lib()
    b, n = build(kgroup=args.kgroup, kunroll=args.kunroll)
    budget = args.budget_ms * 1e3
    for _ in range(args.warmup):
        b.launch_budget(10 ** 12, budget)
        b.sync()
    clocks = None
    if not args.plain:
        clocks = L.fmc_stage_clocks_512_12
        clocks.restype = C.c_int
        clocks.argtypes = [C.c_int, C.c_int, C.c_void_p]
        if clocks(n, 1, None):
            raise SystemExit("fmc_stage_clocks_512_12 failed")
    z = sum(b.stats(c)["steps"] for c in range(n))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.launch_budget(10 ** 12, budget)
    b.sync()
    dt = time.perf_counter() - t0
    a = sum(b.stats(c)["steps"] for c in range(n))
    shape = b.tile_shape()
    print(f"{(a - z) / dt / 1e6:.3f} M moves/s over {args.steps} launches of "
          f"{args.budget_ms} ms; tile {shape}", flush=True)
    if clocks is not None:
        P = np.zeros((n, SLOTS), dtype=np.uint64)
        if clocks(n, 0, P.ctypes.data):
            raise SystemExit("fmc_stage_clocks_512_12 failed")
        print("\n".join(table(P.astype(np.float64))))
    b.close()


if __name__ == "__main__":
    main()
