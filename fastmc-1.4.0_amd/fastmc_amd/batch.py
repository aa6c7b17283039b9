"""Batched device-resident chains (fmc_batch_* C ABI)."""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from .capi import (ChainParams, ChainStats, DP, FmcError, PREC_FP64, check,
                   lib)


class Report(list):
    """A chain's end-of-run report: one stats.NodeStats per guest type, and
    the selectivity of each guest pair ({(j, i): stats.Selectivity})."""
    selectivity: dict = {}


class Batch:
    """Many independent GCMC chains advanced together on the GPU, one chain
    per work-group of a persistent kernel."""

    def __init__(self, max_chains: int, precision: int = PREC_FP64):
        self.L = lib()
        h = C.c_void_p()
        check(self.L.fmc_batch_create(C.byref(h), max_chains, precision),
              "fmc_batch_create")
        self.h = h
        self.precision = precision
        self.ns = {}

    def close(self):
        if getattr(self, "h", None):
            self.L.fmc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_chain(self, ctx, temp: float, press_pa: Sequence[float],
                  seed: Sequence[int], freq: Optional[Sequence[Sequence[float]]] = None,
                  delr: float = 1.0, eq_steps: int = 0, engunit: float = 418.4,
                  ratios=(0.5, 0.5, 0.5), max_moves: int = 0,
                  recompute_every: int = 0, window: int = 0) -> int:
        p = ChainParams()
        p.temp = temp
        p.ntpguest = len(press_pa)
        for k in range(4):
            p.seed[k] = int(seed[k])
        for g, pr in enumerate(press_pa):
            p.press[g] = pr
            p.delr[g] = delr
            p.disp_ratio[g] = ratios[0]
            p.tran_ratio[g] = ratios[1]
            if freq is not None:
                for k in range(9):
                    p.freq[g][k] = freq[g][k] if k < len(freq[g]) else 0.0
        p.rota_ratio = ratios[2]
        p.engunit = engunit
        p.eq_steps = eq_steps
        p.max_moves = max_moves
        p.recompute_every = recompute_every
        p.window = window
        cid = C.c_int()
        check(self.L.fmc_batch_add_chain(self.h, ctx.h, C.byref(p), C.byref(cid)),
              "fmc_batch_add_chain")
        self.ns[cid.value] = list(ctx.nsites)
        return cid.value

    def remove_chain(self, cid: int):
        check(self.L.fmc_batch_remove_chain(self.h, cid), "fmc_batch_remove_chain")
        self.ns.pop(cid, None)

    def checkpoint(self, cid: int) -> bytes:
        """The chain's whole state (duni table, counters, statistics, step
        sizes, running S(k), guest coordinates) as opaque bytes; a chain
        restored from them continues the identical trajectory."""
        n = C.c_longlong()
        check(self.L.fmc_batch_checkpoint(self.h, cid, None, 0, C.byref(n)),
              "fmc_batch_checkpoint")
        buf = C.create_string_buffer(n.value)
        check(self.L.fmc_batch_checkpoint(self.h, cid, buf, n.value, C.byref(n)),
              "fmc_batch_checkpoint")
        return buf.raw[:n.value]

    def restore_chain(self, ctx, blob: bytes) -> int:
        """Add a chain from checkpoint bytes; ctx describes the chain's
        system (its own guest molecules are not used)."""
        cid = C.c_int()
        check(self.L.fmc_batch_restore_chain(self.h, ctx.h, blob, len(blob), C.byref(cid)),
              "fmc_batch_restore_chain")
        self.ns[cid.value] = list(ctx.nsites)
        return cid.value

    def molecules(self, cid: int, gtype: int):
        """(x, y, z) of the chain's guest molecules of type gtype, sites
        molecule-major."""
        n = C.c_int()
        check(self.L.fmc_batch_get_molecules(self.h, cid, gtype, None, None, None, 0,
                                             C.byref(n)), "fmc_batch_get_molecules")
        ns = self.ns[cid][gtype]
        m = n.value * ns
        x, y, z = (np.zeros(max(1, m)) for _ in range(3))
        check(self.L.fmc_batch_get_molecules(self.h, cid, gtype, x.ctypes.data_as(DP),
                                             y.ctypes.data_as(DP), z.ctypes.data_as(DP),
                                             m, C.byref(n)), "fmc_batch_get_molecules")
        return x[:m], y[:m], z[:m]

    def revive(self, cid: int, field, config):
        """The chain's state as the reference's REVIVE record set
        (utility_pack.f:1800-1814): guest molecule counts and coordinates
        from the chain, framework molecules from config, the rolling
        production averages <N>, <E>, <EN>, <N^2>, <E^2> (gcmc.f:1818-1848)
        in chainstats, and the accumulated energies.  Entries the chain does
        not track (<exp(-E/kT)>, the window deviations) are zero."""
        from .inputs import Revive
        st = self.stats(cid)
        prod = st["prod_steps"]
        guests = [k for k, m in enumerate(field.molecules) if m.guest]
        nummols, coords = [], []
        gi = {k: i for i, k in enumerate(guests)}
        for k, mol in enumerate(field.molecules):
            if k in gi:
                x, y, z = self.molecules(cid, gi[k])
                nummols.append(len(x) // len(mol.names))
                coords.append((x.copy(), y.copy(), z.copy()))
            else:
                c = np.asarray(config.coords[k], dtype=float).reshape(-1, 3)
                nummols.append(mol.nummols)
                coords.append((c[:, 0].copy(), c[:, 1].copy(), c[:, 2].copy()))
        stats = [float(prod)]
        for i in range(len(guests)):
            p = max(prod, 1)
            stats += [st["sum_n"][i] / p, st["sum_e"][i] / p, st["sum_en"][i] / p,
                      st["sum_n2"][i] / p, st["sum_e2"][i] / p] + [0.0] * 9
        natms = sum(len(c[0]) for c in coords)
        return Revive(-1 if prod > 0 else 0, nummols, stats, natms, coords,
                      list(st["energy"]))

    def run(self, n: int):
        check(self.L.fmc_batch_run(self.h, n), "fmc_batch_run")

    def launch(self, n: int):
        check(self.L.fmc_batch_launch(self.h, n), "fmc_batch_launch")

    def launch_budget(self, max_moves: int, budget_us: float):
        """Enqueue a launch in which every chain moves until budget_us of
        GPU time has passed (or max_moves)."""
        check(self.L.fmc_batch_launch_budget(self.h, max_moves, budget_us),
              "fmc_batch_launch_budget")

    def sync(self):
        check(self.L.fmc_batch_sync(self.h), "fmc_batch_sync")

    def done(self) -> bool:
        """True when no launch is running (does not wait)."""
        v = C.c_int()
        check(self.L.fmc_batch_query(self.h, C.byref(v)), "fmc_batch_query")
        return bool(v.value)

    def stats(self, cid: int) -> dict:
        st = ChainStats()
        r = self.L.fmc_batch_stats(self.h, cid, C.byref(st))
        if r not in (0, -4):
            raise FmcError(r, "fmc_batch_stats")
        ng = len(self.ns.get(cid, [1]))
        return dict(steps=st.steps, prod_steps=st.prod_steps,
                    accepted=list(st.accepted), attempted=list(st.attempted),
                    nmol=list(st.nmol)[:ng], energy=list(st.energy)[:ng],
                    sum_n=list(st.sum_n)[:ng], sum_n2=list(st.sum_n2)[:ng],
                    sum_e=list(st.sum_e)[:ng], sum_en=list(st.sum_en)[:ng],
                    sum_e2=list(st.sum_e2)[:ng], delrdisp=list(st.delrdisp)[:ng],
                    rotangle=st.rotangle, e_acc=st.e_acc, e_full=st.e_full,
                    drift=st.drift, sk_dev=st.sk_dev, recomputes=st.recomputes,
                    win_len=st.win_len, win_closed=st.win_closed,
                    win=[[[st.win[j][q][g] for g in range(ng)] for q in range(5)]
                         for j in range(5)],
                    sel_window=st.sel_window, sel_count=st.sel_count,
                    sel_base=list(st.sel_base)[:ng], sel_mean=list(st.sel_mean),
                    sel_second=list(st.sel_second), sel_var=list(st.sel_var),
                    tile=dict(threads=st.tile_threads, kpt=st.tile_kpt,
                              kgroup=st.tile_kgroup, kunroll=st.tile_kunroll),
                    capacity_fault=(r == -4))

    def report(self, cid: int, temp: float):
        """The reference's end-of-run statistics of a chain that has made
        its max_moves (gcmc.f:1969-2100), one stats.NodeStats per guest
        type: <N>, <E>, <EN>, <N^2>, <E^2>, Q_st, C_v and their errors from
        the five windows the kernel accumulated.  Combine replicas of one
        state point with stats.union."""
        from . import stats as S
        st = self.stats(cid)
        if st["win_len"] <= 0:
            raise ValueError("chain has no statistics windows (needs max_moves with "
                             "5 to 500000 production moves)")
        if st["win_closed"] < 4 or st["prod_steps"] < 5 * st["win_len"]:
            raise ValueError("chain has not finished its production moves")
        out = Report()
        for g in range(len(st["nmol"])):
            sums = [[st["win"][j][q][g] for q in range(5)] for j in range(5)]
            out.append(S.node_stats(sums, st["win_len"], float(st["prod_steps"]), temp))
        out.selectivity = self.selectivity(cid, st)
        return out

    def selectivity(self, cid: int, st: dict = None) -> dict:
        """The reference's selectivity of guest j over guest i, j < i
        (gcmc.f:1887-1916), for a chain at the end of its run: the kernel's
        closed windows, then the open window with weight samples / window,
        as the reference takes it at its last move.  {(j, i):
        stats.Selectivity}; empty for one guest type."""
        from . import stats as S
        st = st or self.stats(cid)
        ng = len(st["nmol"])
        if ng < 2:
            return {}
        state = S.selectivity_state(ng)
        state[0] = st["sel_count"]
        for i in range(1, ng):
            for j in range(i):
                p = i * (i - 1) // 2 + j
                state[1][i][j] = st["sel_mean"][p]
                state[2][i][j] = st["sel_second"][p]
                state[3][i][j] = st["sel_var"][p]
        nw = st["sel_window"]
        rest = st["prod_steps"] - st["sel_count"] * nw
        last_w = 1.0
        if rest > 0:
            means = [(st["sum_n"][g] - st["sel_base"][g]) / rest for g in range(ng)]
            last_w = rest / nw
            S.selectivity_update(state, means, last_w)
        return S.selectivity_report(state, last_w)

    def enable_log(self, cid: int, cap: int):
        check(self.L.fmc_batch_enable_log(self.h, cid, cap), "fmc_batch_enable_log")

    def get_log(self, cid: int, n: int) -> np.ndarray:
        out = np.zeros(8 * n)
        check(self.L.fmc_batch_get_log(self.h, cid, n, out.ctypes.data_as(DP)),
              "fmc_batch_get_log")
        return out.reshape(n, 8)

    def resync(self) -> float:
        v = C.c_double()
        check(self.L.fmc_batch_resync(self.h, C.byref(v)), "fmc_batch_resync")
        return v.value

    def recompute(self):
        """FP64 full recompute of every active chain: (largest S(k)
        deviation replaced, largest |energy drift| in kcal/mol)."""
        d = C.c_double()
        e = C.c_double()
        check(self.L.fmc_batch_recompute(self.h, C.byref(d), C.byref(e)),
              "fmc_batch_recompute")
        return d.value, e.value

    def launch_cycles(self, n_cycles: int, moves_per_cycle: int):
        """Enqueue n_cycles replays of the captured (moves + recompute)
        cycle; complete with sync()."""
        check(self.L.fmc_batch_launch_cycles(self.h, n_cycles, moves_per_cycle),
              "fmc_batch_launch_cycles")

    def poll(self, cap: int = 4096):
        """Ids of chains that reached their max_moves (or faulted)."""
        ids = (C.c_int * max(1, cap))()
        n = C.c_int()
        check(self.L.fmc_batch_poll(self.h, ids, cap, C.byref(n)), "fmc_batch_poll")
        return [ids[i] for i in range(min(n.value, cap))]

    def timing(self, reset: bool = False):
        """(kernel milliseconds, launches) since the last reset, from HIP
        events on the batch stream."""
        ms = C.c_double()
        n = C.c_longlong()
        check(self.L.fmc_batch_timing(self.h, int(reset), C.byref(ms), C.byref(n)),
              "fmc_batch_timing")
        return ms.value, n.value

    def set_tile(self, threads: int):
        check(self.L.fmc_batch_set_tile(self.h, threads), "fmc_batch_set_tile")

    def tile(self) -> int:
        v = C.c_int()
        check(self.L.fmc_batch_get_tile(self.h, C.byref(v)), "fmc_batch_get_tile")
        return v.value

    def set_tile_shape(self, threads: int, kpt: int = -1, kgroup: int = 4, kunroll: int = 1):
        """The whole tile shape (fmc_batch_set_tile_shape)."""
        check(self.L.fmc_batch_set_tile_shape(self.h, threads, kpt, kgroup, kunroll),
              "fmc_batch_set_tile_shape")

    def tile_shape(self) -> dict:
        """The tile shape in use: threads, kpt (0 streaming), kgroup, kunroll,
        tuned (0 none yet, 1 set, 2 autotuned)."""
        v = [C.c_int() for _ in range(5)]
        check(self.L.fmc_batch_get_tile_shape(self.h, *[C.byref(x) for x in v]),
              "fmc_batch_get_tile_shape")
        return dict(zip(("threads", "kpt", "kgroup", "kunroll", "tuned"),
                        (x.value for x in v)))


def footprint(ctx, precision: int = PREC_FP64) -> int:
    v = C.c_longlong()
    check(lib().fmc_chain_footprint(ctx.h, precision, C.byref(v)),
          "fmc_chain_footprint")
    return v.value
