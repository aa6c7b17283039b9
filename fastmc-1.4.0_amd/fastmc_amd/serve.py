"""Continuous-batching isotherm server over the batched chains.

A request is one isotherm point: a framework, its guest species, their
partial pressures, a temperature and a number of trial moves.  Requests
are admitted as chain slots and device memory free up (admission by the
chain's projected HBM footprint, ``fmc_chain_footprint``); an admitted
request becomes one device-resident chain (coordinates, S(k), RNG and
statistics stay in HBM until it finishes).  The GPU advances every
admitted chain in slices of ``slice_moves`` moves (or of ``slice_ms`` of
GPU time, every chain moving at its own speed); requests that arrive
while a slice runs are built and admitted on the host meanwhile and join
the next slice, and finished chains leave at slice boundaries.

Each result carries the raw chain statistics and, per guest type, the
reference's end-of-run report (<N>, <E>, Q_st, C_v and their errors from
five windows of the production samples, stats.NodeStats).

A chain's arithmetic does not depend on which other chains share its
launch (one work-group per chain, a fixed tile), so a request's result is
the same alone and in the mix.  Admission runs on its own thread (the C
library guards its tables and serialises admissions), so building a new
system overlaps the slices on the GPU.
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

from .batch import Batch, footprint
from .capi import Context, PREC_MIXED


@dataclass
class Request:
    rid: int
    arrival: float                 # seconds from the start of the run
    arrays: dict                   # Context.from_arrays input
    temp: float
    press_pa: Sequence[float]
    seed: Sequence[int]
    moves: int
    delr: float = 1.0
    eq_steps: int = 0


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    done: float
    stats: dict
    footprint: int
    report: Optional[list] = None   # per guest type: stats.NodeStats, the
                                    # reference's end-of-run statistics
    selectivity: Optional[dict] = None   # {(j, i): stats.Selectivity} of each
                                         # guest pair j < i (gcmc.f:1887-1916)


    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass
class Suspended:
    """A request whose chain was checkpointed by a stopped run
    (fmc_batch_checkpoint); a later run resumes it where it stopped."""
    request: Request
    admitted: float
    blob: bytes


@dataclass
class Server:
    slots: int = 256
    precision: int = PREC_MIXED
    slice_moves: int = 250
    slice_ms: float = 0.0          # > 0: time-budgeted slices instead
    hbm_budget: int = 64 << 30     # bytes of chain state admitted at once
    tile: int = 256
    clock: Callable[[], float] = time.perf_counter
    results: Dict[int, Result] = field(default_factory=dict)

    def __post_init__(self):
        self.batch = Batch(self.slots, self.precision)
        self.batch.set_tile(self.tile)
        self._live: Dict[int, tuple] = {}   # chain id -> (request, admitted, bytes)
        self._used = 0                      # bytes admitted or being admitted
        self._inflight = 0                  # admissions under way
        self.slices = 0
        self.moves_done = 0
        self.suspended: List[Suspended] = []   # live chains of a stopped run
        self.unstarted: List[Request] = []     # requests a stopped run left

    def close(self):
        self.batch.close()

    def _retire(self, now: float, cv: threading.Condition):
        for cid in self.batch.poll():
            with cv:
                # a chain runs from its add_chain on; one that finished in its
                # first slice may not be recorded by the admission thread yet
                cv.wait_for(lambda: cid in self._live)
                req, t_adm, need = self._live[cid]
            st = self.batch.stats(cid)
            rep, sel = None, None
            if st["win_len"] > 0 and not st["capacity_fault"]:
                rep = self.batch.report(cid, req.temp)
                sel = rep.selectivity
            # the library's slot and the _live entry go together, under the
            # lock: the admission thread counts slots from _live, so it never
            # adds a chain while this one still holds a library slot, and a
            # chain it adds into the freed slot is recorded after this entry
            # is gone
            with cv:
                self.batch.remove_chain(cid)
                del self._live[cid]
                self._used -= need
                self.moves_done += st["steps"]
                self.results[req.rid] = Result(req.rid, req.arrival, t_adm, now, st, need,
                                               rep, sel)
                cv.notify_all()

    def run(self, requests: Sequence[Request], realtime: bool = True,
            resume: Sequence[Suspended] = (), stop_after_slices: Optional[int] = None
            ) -> List[Result]:
        """Serve the requests; with realtime, each arrives at its arrival
        time (seconds after the start), else all at once.  An admission
        thread builds each request's system and adds its chain while the
        GPU runs slices; this thread launches slices back to back and
        retires finished chains.

        resume: chains of an earlier stopped run, restored from their
        checkpoints before the first slice; each continues its trajectory
        exactly (its admission time is the earlier run's).
        stop_after_slices: after that many slices, stop admitting,
        checkpoint every live chain into self.suspended (their requests'
        results are not produced) and leave the requests not yet admitted in
        self.unstarted; the results returned are those of the requests
        (resumed ones first) that finished."""
        pending = sorted(requests, key=lambda r: r.arrival)
        cv = threading.Condition()
        t0 = self.clock()
        state = {"done": False, "error": None, "stop": False}
        self.suspended, self.unstarted = [], []
        for sus in resume:
            ctx = Context.from_arrays(sus.request.arrays)
            need = footprint(ctx, self.precision)
            cid = self.batch.restore_chain(ctx, sus.blob)
            ctx.close()
            self._live[cid] = (sus.request, sus.admitted, need)
            self._used += need

        def admitter():
            try:
                for i, req in enumerate(pending):
                    if realtime:
                        delay = req.arrival - (self.clock() - t0)
                        if delay > 0:
                            with cv:
                                cv.wait_for(lambda: state["stop"], timeout=delay)
                    if state["stop"]:
                        self.unstarted = list(pending[i:])
                        return
                    ctx = Context.from_arrays(req.arrays)
                    need = footprint(ctx, self.precision)
                    with cv:
                        while (not state["stop"] and
                               (len(self._live) + self._inflight >= self.slots or
                                self._used + need > self.hbm_budget)):
                            cv.wait()
                        if state["stop"]:
                            self.unstarted = list(pending[i:])
                            ctx.close()
                            return
                        self._inflight += 1
                        self._used += need
                    cid = self.batch.add_chain(ctx, req.temp, req.press_pa, req.seed,
                                               delr=req.delr, eq_steps=req.eq_steps,
                                               max_moves=req.moves)
                    ctx.close()
                    with cv:
                        self._inflight -= 1
                        self._live[cid] = (req, self.clock() - t0, need)
                        cv.notify_all()
            except BaseException as e:   # surfaced by run()
                state["error"] = e
            finally:
                with cv:
                    state["done"] = True
                    cv.notify_all()

        th = threading.Thread(target=admitter, daemon=True)
        th.start()
        n_slices = 0
        try:
            while True:
                with cv:
                    if state["error"] is not None:
                        break
                    if state["done"] and not self._live and not self._inflight:
                        break
                    if not self._live:
                        cv.wait(timeout=0.01)
                        continue
                if self.slice_ms > 0:
                    self.batch.launch_budget(10 ** 12, self.slice_ms * 1e3)
                else:
                    self.batch.launch(self.slice_moves)
                self.slices += 1
                n_slices += 1
                self.batch.sync()
                self._retire(self.clock() - t0, cv)
                if stop_after_slices is not None and n_slices >= stop_after_slices:
                    with cv:
                        state["stop"] = True
                        cv.notify_all()
                    break
        except BaseException:
            # a failed slice: the admission thread may be waiting for a slot
            # that will not come; stop it before joining
            with cv:
                state["stop"] = True
                cv.notify_all()
            raise
        finally:
            th.join()
        if state["error"] is not None:
            raise state["error"]
        if state["stop"]:
            # admissions have ended: checkpoint and release every live chain
            for cid in sorted(self._live):
                req, t_adm, need = self._live[cid]
                self.suspended.append(Suspended(req, t_adm, self.batch.checkpoint(cid)))
                self.batch.remove_chain(cid)
                self._used -= need
            self._live.clear()
        asked = [sus.request for sus in resume] + list(requests)
        return [self.results[r.rid] for r in asked if r.rid in self.results]


def percentile(v: Sequence[float], q: float) -> float:
    s = sorted(v)
    if not s:
        return float("nan")
    k = (len(s) - 1) * q / 100.0
    lo = int(k)
    hi = min(lo + 1, len(s) - 1)
    return s[lo] + (s[hi] - s[lo]) * (k - lo)


def poisson_requests(n: int, rate: float, seed: int = 7, moves: int = 5000,
                     species=None) -> List[Request]:
    """n isotherm-point requests with Poisson arrivals at `rate` per second:
    framework 300-10000 atoms (log-uniform), 1-3 of CO2/N2/CH4, total
    pressure 0.01-30 bar, 273-348 K."""
    import numpy as np
    from . import synth
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        w = synth.c4_request(rng, i)
        out.append(Request(i, t, w.system_arrays(), w.temp,
                           [p * 1e5 for p in w.press_bar], w.seed, moves, delr=w.delr))
    return out
