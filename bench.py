"""Benchmark of the MI355X trial-move path.

Headline (``value``): whole-job GCMC trial moves per second of the batched
workload -- 256 independent chains (32 synthetic frameworks of 300-6000
atoms, orthorhombic 20-60 A cells replicated to >= 2 rcut, x 8 CO2
pressures 0.01-10 bar, 298 K, kmax 8-14), mixed precision (FP32 phase
factors and dS(k), FP64 S(k), pair terms and totals), with the FP64 full
recompute (S(k) resync, energy drift) every 1000 moves of each chain,
inside the launch.
A "step" is one persistent launch with a GPU-time budget (``--budget-ms``):
every chain moves until the budget has passed, so fast chains do not idle
behind slow ones; each chain's trajectory does not depend on where the
launches split it.  Moves are counted per chain from its own counters.
Inputs are resident in HBM before timing starts; the warmup steps
equilibrate the chains.

Reported beside it: the same workload in FP64, the plain single-chain FP64
drop-in rate (the Fortran driver over the C ABI on the 3392-atom CO2/N2
supercell), the ratio, the mixed-vs-FP64 estep error over one chain of
every framework, the dominant kernel's roofline, and the CPU baseline (the
FP64 oracle's restatement of the move loop on the host cores).
"""
import argparse
import json
import math
import multiprocessing as mp
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "fastmc-1.4.0_amd")
ORACLE = os.path.join(ROOT, "oracle")
for _p in (PKG, ORACLE):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = ("GCMC trial moves/s, batched FP32 vs plain FP64 at dE<1e-6 rel; "
          "p50/p99 isotherm latency")

# MI355X published peaks
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6


PEAK_HBM_GBS = 8000.0


def chain_work(nk, ns, n_rc, kinds, acc):
    """Algorithmic bytes and FLOPs of one trial move, averaged over the
    chain's move mix: bytes = 24 Nk (S(k) re/im + a_k read) + 16 Nk acc
    (S(k) written back on accept) + 16 P n_s N_rc (pair neighbour x/y/z +
    charge/type) + 12 P n_s (site coordinates); FLOPs = 8 P n_s Nk (phase
    products and S accumulation) + 6 Nk (|S|^2 difference * a_k) + 30 P n_s
    N_rc (distance, minimum image, two interpolations).  P = 1 for insert
    and delete, 2 for displace; N_rc = pairs per site within rcut."""
    ins, dele, disp = kinds
    tot = max(1, ins + dele + disp)
    p = (ins + dele + 2 * disp) / tot
    byt = 24 * nk + 16 * nk * acc + 16 * p * ns * n_rc + 12 * p * ns
    flo = 8 * p * ns * nk + 6 * nk + 30 * p * ns * n_rc
    return byt, flo


# ---------------------------------------------------------------------------
# CPU baseline: the oracle's move loop, one chain per core
# ---------------------------------------------------------------------------

def _cpu_worker(args):
    idx, moves = args
    import pyoracle
    from fastmc_amd import synth
    w = synth.c3()[idx]
    A = w.system_arrays()
    s = pyoracle.OracleSystem(A["imcon"], A["cell"], A["alpha"], A["kmax"],
                              A["rcut"], A["delr"], A["ntpatm"], A["vdw"],
                              A["fw"], A["guests"])
    ch = pyoracle.OracleChain(s, w.temp, [p * 1e5 for p in w.press_bar], w.seed)
    t0 = time.perf_counter()
    ch.run(moves)
    return moves, time.perf_counter() - t0


def _cpu_serial(args):
    """One core: the sample's chains one after another."""
    jobs = args
    tot, dt = 0, 0.0
    for j in jobs:
        m, t = _cpu_worker(j)
        tot += m
        dt += t
    return tot, dt


def host_cores():
    """The cores this process may run on (what nproc prints), the logical
    CPUs the machine has, and the CPU model."""
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = os.cpu_count() or 1
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return share, os.cpu_count() or share, model


# the host share a one-GPU box grants a process's worker pools
MAX_POOL = 16


def cpu_baseline(budget_s=20.0):
    """The oracle's FP64 move loop on the host: one independent chain per
    core on every core of the process's share (at most MAX_POOL), as the
    reference's MPI replicas run, plus the same chains one after another on
    one core."""
    subprocess.check_call(["make", "-s", "-C", ORACLE], stdout=subprocess.DEVNULL)
    nproc, logical, model = host_cores()
    cores = max(1, min(MAX_POOL, nproc))
    # one calibration chain on one core sizes the per-core sample
    ctxm = mp.get_context("spawn")
    with ctxm.Pool(1) as pool:
        m, dt = pool.map(_cpu_worker, [(0, 200)])[0]
    per_move = dt / m
    moves = max(200, int(budget_s / per_move / 4))
    # chain 17 i: framework (17 i) // 8, pressure (17 i) % 8 -- spans the
    # framework sizes and all 8 pressures
    picks = [(17 * i) % 256 for i in range(cores)]
    with ctxm.Pool(cores) as pool:
        t0 = time.perf_counter()
        res = pool.map(_cpu_worker, [(c, moves) for c in picks])
        wall = time.perf_counter() - t0
    total = sum(r[0] for r in res)
    # one core: every chain of the sample, moves / cores each, in turn
    m1 = max(50, moves // cores)
    with ctxm.Pool(1) as pool:
        t1, d1 = pool.map(_cpu_serial, [[(c, m1) for c in picks]])[0]
    return {"value": total / wall, "unit": "moves/s", "cores": cores,
            "kind": "port",
            "sample": f"{cores} chains of the 256-chain workload (chain 17*i: "
                      f"every other framework, all 8 pressures), one per core, "
                      f"{moves} FP64 moves each from an empty framework, oracle "
                      f"move loop compiled -O2",
            "nproc": nproc, "logical_cpus": logical, "cpu_model": model,
            "single_core": {"value": t1 / d1, "unit": "moves/s",
                            "sample": f"the same {cores} chains on one core, one "
                                      f"after another, {m1} moves each"}}


# ---------------------------------------------------------------------------
# plain FP64 drop-in: Fortran driver over the C ABI on the 3392-atom supercell
# ---------------------------------------------------------------------------

PLAIN_PARTS = ("fmc_trial set-up", "launch", "wait", "fmc_commit", "restart files",
               "the rest")


def plain_fp64_rate(moves=1600, parts=None):
    """Moves/s of the driver's run; with a dict `parts`, also its
    microseconds per move by part, as the driver's OUTPUT breaks them down
    (the host time of fmc_trial's set-up, its kernel launch and its wait
    for the sums; fmc_commit; the restart files; everything else)."""
    from fastmc_amd import synth
    exe = os.path.join(PKG, "bin", "fmc_gcmc")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} missing: run __graft_entry__.build()")
    d = tempfile.mkdtemp(prefix="fmc_plain_")
    try:
        w = synth.c2()
        w.write(d, steps=moves, eqsteps=400)
        p = os.path.join(d, "CONTROL")
        s = open(p).read().replace("finish", "seed 12 34 56 78\nfinish")
        open(p, "w").write(s)
        r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=900)
        if r.returncode:
            raise RuntimeError(r.stderr)
        out = open(os.path.join(d, "OUTPUT")).read()
        rate = float(re.search(r"trial moves per second :\s+([0-9.Ee+-]+)", out).group(1))
        m = re.search(r"microseconds per move:\s*([0-9.]+)\s+in fmc_trial: set-up\s*([0-9.]+)"
                      r"\s+launch\s*([0-9.]+)\s+wait\s*([0-9.]+)\s+fmc_commit\s*([0-9.]+)"
                      r"\s+restart files\s*([0-9.]+)\s+the rest\s*(-?[0-9.]+)", out)
        if parts is not None and m:
            parts["total"] = float(m.group(1))
            for k, v in zip(PLAIN_PARTS, m.groups()[1:]):
                parts[k] = float(v)
        return rate
    finally:
        shutil.rmtree(d, ignore_errors=True)


# ---------------------------------------------------------------------------
# batched workload
# ---------------------------------------------------------------------------

def build_batch(precision, n_chains=256, recompute_every=0, picks=None):
    """The batched workload, or the chains `picks` of it."""
    import fastmc_amd
    from fastmc_amd import synth
    from fastmc_amd.batch import Batch
    allw = synth.c3()
    ws = [allw[i] for i in picks] if picks is not None else allw[:n_chains]
    b = Batch(len(ws), precision)
    ctxs, info = {}, []
    for w in ws:
        key = id(w.fw)
        if key not in ctxs:
            ctxs[key] = fastmc_amd.Context.from_arrays(w.system_arrays())
        ctx = ctxs[key]
        b.add_chain(ctx, w.temp, [p * 1e5 for p in w.press_bar], w.seed,
                    delr=w.delr, recompute_every=recompute_every)
        inf = ctx.info()
        info.append(dict(nk=inf["nk"], ns=ctx.nsites[0], nfw=w.fw.n,
                         volm=inf["volm"], rcut=w.rcut))
    return b, ws, info


# the floor of the relative estep error: 1e-3 kT at 298 K, in kcal/mol
KT_298 = 1.9872041e-3 * 298.0
DE_FLOOR = 1e-3 * KT_298


def de_stats(ref_logs, mix_logs, kinds=None):
    """Mixed-vs-FP64 estep error over the prefix of each chain where both
    trajectories coincide (same kinds, decisions and loadings), on every
    move with a trial (no-op moves excluded; `kinds` restricts the move
    kinds).  err = |estep_mixed - estep_fp64| in kcal/mol; raw = err /
    |estep_fp64|; floored = err / max(|estep_fp64|, 1e-3 kT).  The absolute
    figures cover the moves the acceptance test can take (|estep| < 50 kT:
    a larger estep is rejected whatever its rounding); the relative figures
    cover every move."""
    import numpy as np
    errs, av = [], []
    same_moves = 0
    for a, m in zip(ref_logs, mix_logs):
        same = np.all(a[:, :4] == m[:, :4], axis=1)
        n = int(np.argmin(same)) if not same.all() else len(same)
        same_moves += n
        sel = a[:n, 0] != 0 if kinds is None else np.isin(a[:n, 0], kinds)
        errs.append(np.abs(m[:n, 4] - a[:n, 4])[sel])
        av.append(np.abs(a[:n, 4])[sel])
    err = np.concatenate(errs)
    av = np.concatenate(av)
    raw = err / np.maximum(av, 1e-300)
    floored = err / np.maximum(av, DE_FLOOR)
    dec = av < 50.0 * KT_298
    ed = err[dec] if dec.any() else np.zeros(1)
    return {"moves_compared": int(err.size), "identical_prefix_moves": same_moves,
            "decidable_moves": int(dec.sum()),
            "abs_max_kcal": float(ed.max()), "abs_median_kcal": float(np.median(ed)),
            "abs_max_kT": float(ed.max() / KT_298),
            "raw_rel_max": float(raw.max()), "raw_rel_median": float(np.median(raw)),
            "floored_rel_max": float(floored.max()),
            "frac_raw_le_1e-6": float(np.mean(raw <= 1e-6)),
            "frac_floored_le_1e-6": float(np.mean(floored <= 1e-6)),
            "floor_kcal": DE_FLOOR}


def de_error(moves=400):
    """Mixed vs FP64 estep on the same chains and seeds: one chain of every
    one of the 32 frameworks (framework f at pressure f mod 8)."""
    from fastmc_amd.capi import PREC_FP64, PREC_MIXED
    picks = [8 * f + f % 8 for f in range(32)]
    logs = []
    for prec in (PREC_FP64, PREC_MIXED):
        b, ws, _ = build_batch(prec, picks=picks)
        for c in range(len(picks)):
            b.enable_log(c, moves)
        b.run(moves)
        logs.append([b.get_log(c, moves) for c in range(len(picks))])
        b.close()
    out = de_stats(logs[0], logs[1])
    out["chains"] = f"one per framework, 32 frameworks, {moves} moves each"
    return out


def batched_fp64_rate(steps=5, warmup=3, budget_us=15000.0, recompute_every=1000):
    """The headline workload with every term in FP64 (the precision whose
    trajectories equal the oracle's): whole-job moves/s of `steps` budgeted
    launches after `warmup` of them."""
    from fastmc_amd.capi import PREC_FP64
    b, ws, _ = build_batch(PREC_FP64, recompute_every=recompute_every)
    n = len(ws)
    for _ in range(warmup):
        b.launch_budget(10 ** 12, budget_us)
        b.sync()
    z = [b.stats(c)["steps"] for c in range(n)]
    t0 = time.perf_counter()
    for _ in range(steps):
        b.launch_budget(10 ** 12, budget_us)
    b.sync()
    dt = time.perf_counter() - t0
    a = [b.stats(c)["steps"] for c in range(n)]
    b.close()
    return (sum(a) - sum(z)) / dt


def serving_leg(n_req=400, rate=200.0, moves=5000):
    """Isotherm-point serving: Poisson arrivals at `rate`/s, each request a
    framework of 300-10000 atoms with 1-3 of CO2/N2/CH4, 0.01-30 bar,
    273-348 K, `moves` trial moves, mixed precision; continuous batching
    (admission by HBM footprint on its own thread, 4 ms budgeted slices).
    Latency = arrival to result."""
    from fastmc_amd import serve
    reqs = serve.poisson_requests(n_req, rate, seed=2024, moves=moves)
    srv = serve.Server(slots=256, slice_ms=4.0, tile=512)
    t0 = time.perf_counter()
    res = srv.run(reqs, realtime=True)
    wall = time.perf_counter() - t0
    srv.close()
    lat = [r.latency for r in res]
    wait = [r.admitted - r.arrival for r in res]
    return {"requests": n_req, "rate_per_s": rate, "moves_per_request": moves,
            "p50_latency_s": serve.percentile(lat, 50),
            "p99_latency_s": serve.percentile(lat, 99),
            "p50_admission_wait_s": serve.percentile(wait, 50),
            "moves_per_s": sum(r.stats["steps"] for r in res) / wall,
            "makespan_s": wall, "slices": srv.slices}


def measured_traffic():
    """HBM bytes per launch of the batch kernel from the committed
    FETCH_SIZE/WRITE_SIZE counter runs (scaled by a plain copy kernel's
    reading), or None when they are not there."""
    p = os.path.join(ROOT, "profiles", "batch_kernel_256chains_co2_mixed_pmc.md")
    if not os.path.exists(p):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(p).read())
    return float(m.group(1)) if m else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--budget-ms", type=float, default=15.0,
                    help="GPU time per step (one launch); every chain moves "
                         "until it has passed")
    ap.add_argument("--recompute-every", type=int, default=1000,
                    help="moves between FP64 recomputes of a chain")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-fp64", action="store_true",
                    help="skip the FP64 run of the batched workload")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 256 MiB device copy (the reference "
                         "kernel for scaling FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one MI355X; chains shard by chain "
              "across GPUs with no per-move traffic, run one process per GPU",
              file=sys.stderr)
        sys.exit(2)

    from fastmc_amd.capi import PREC_MIXED

    if args.copy_calibration:
        import torch
        n = 256 << 20
        a = torch.empty(n, dtype=torch.uint8, device="cuda:0").fill_(1)
        c = torch.empty_like(a)
        for _ in range(3):
            c.copy_(a)
        torch.cuda.synchronize()
        del a, c

    cpu = None if args.no_cpu else cpu_baseline()

    b, ws, info = build_batch(PREC_MIXED, recompute_every=args.recompute_every)
    n_chains = len(ws)
    budget_us = args.budget_ms * 1e3
    never = 10 ** 12
    for _ in range(args.warmup):
        b.launch_budget(never, budget_us)
        b.sync()
    before = [b.stats(c) for c in range(n_chains)]
    # timed region: every chain resident in HBM, one budgeted launch per step
    b.timing(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.launch_budget(never, budget_us)
    b.sync()
    dt = time.perf_counter() - t0
    kernel_ms, launches = b.timing()
    after = [b.stats(c) for c in range(n_chains)]
    total_moves = sum(a["steps"] - z["steps"] for a, z in zip(after, before))
    value = total_moves / dt
    ms_per_step = dt / args.steps * 1e3

    # algorithmic bytes and flops of the timed region, chain by chain, from
    # its own Nk, density and move mix
    byt = flo = 0.0
    for c in range(n_chains):
        z, a = before[c], after[c]
        att = [a["attempted"][i] - z["attempted"][i] for i in range(6)]
        acc = [a["accepted"][i] - z["accepted"][i] for i in range(6)]
        moves = a["steps"] - z["steps"]
        kinds = (att[1], att[2], att[3] + att[4] + att[5])
        accr = sum(acc) / max(1, sum(att))
        I = info[c]
        nat = I["nfw"] + 0.5 * (a["nmol"][0] + z["nmol"][0]) * I["ns"]
        n_rc = nat / I["volm"] * 4.0 / 3.0 * math.pi * I["rcut"] ** 3
        bb, ff = chain_work(I["nk"], I["ns"], n_rc, kinds, accr)
        byt += bb * moves
        flo += ff * moves
    # per launch: algorithmic bytes / the batch kernel's average duration
    # (HIP events on the batch stream around each launch)
    k_s = kernel_ms * 1e-3
    achieved = byt / k_s / 1e9
    valu = flo / k_s / 1e12
    drift = max(abs(a["drift"]) for a in after)
    recomputes = sum(a["recomputes"] - z["recomputes"] for a, z in zip(after, before))
    nmol = [a["nmol"][0] for a in after]
    tile = b.tile_shape()
    b.close()

    plain_us = {}
    plain = None if args.no_plain else plain_fp64_rate(parts=plain_us)
    fp64 = None if args.no_fp64 else batched_fp64_rate(
        budget_us=budget_us, recompute_every=args.recompute_every)
    de = de_error()
    serving = None if args.no_serve else serving_leg()
    traffic = measured_traffic()

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "trial moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic (seeded generator, reference CONTROL/FIELD/CONFIG formats)",
        "config": {"workload": "256 batched CO2 chains: 32 frameworks 300-6000 "
                               "atoms x 8 pressures 0.01-10 bar, 298 K, kmax 8-14, "
                               "rcut 12.5 A, mixed precision, FP64 recompute "
                               "every 1000 moves of each chain",
                   "chains": n_chains, "step_budget_ms": args.budget_ms,
                   "moves_per_chain_per_step_mean": total_moves / n_chains / args.steps,
                   "recomputes_in_timed_region": recomputes,
                   "tile": {"threads": tile["threads"], "kpt": tile["kpt"],
                            "kgroup": tile["kgroup"], "kunroll": tile["kunroll"],
                            "chosen_by": {2: "autotuner", 1: "fixed", 0: "none"}[tile["tuned"]]},
                   "guest_molecules_mean": sum(nmol) / len(nmol),
                   "guest_molecules_max": max(nmol)},
        "plain_fp64_moves_per_s": plain,
        "plain_fp64_us_per_move": plain_us or None,
        "speedup_vs_plain_fp64": (value / plain) if plain else None,
        "batched_fp64_moves_per_s": fp64,
        "mixed_vs_fp64_dE": de,
        "mixed_vs_fp64_dE_rel_max": de["raw_rel_max"],
        "energy_drift_max_kcal": drift,
        "serving": serving,
        "kernel_ms_per_launch": kernel_ms / max(1, launches),
        # the model roofline: algorithmic bytes per launch over the launch's
        # duration; the kernel is limited by the latency of each chain's
        # serial move sequence, and the measured traffic's share of HBM
        # bandwidth says how far from the HBM bound it runs
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                     "traffic": traffic,
                     "limiter": "latency",
                     "traffic_frac": (traffic * max(1, launches) / k_s / 1e9 / PEAK_HBM_GBS
                                      if traffic else None)},
        # FP32 (phase products) and FP64 (S(k), pair terms) flops together,
        # against the lower (FP64) vector peak
        "valu": {"achieved": valu, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                 "frac": valu / PEAK_FP64_TFLOPS},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
