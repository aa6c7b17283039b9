"""End-of-run statistics as the reference reports them.

The reference turns each node's production samples into five window
averages (gcmc.f:1969-2031), reports per node the window means, the
isosteric heat and heat capacity of those means and the windows' standard
deviations (gcmc.f:2051-2100), and combines nodes weighted by their sample
counts (avunion/stdunion, gcmc.f:2366-2432).  Here a node is one chain:
replicas of one isotherm point, batched on the GPU, combine the way the
reference's MPI branches do.

The window arithmetic keeps the reference's exact form:

* window x (1-based) sums the samples s = (x-1)*nwind ... (x-1)*nwind +
  nwind, both ends included (nwind + 1 samples), and divides by nwind, with
  nwind = int(samples / 5); a sample index past the last sample reads zero
  (the sample buffer is zero-filled, utility_pack.f:902);
* the window values of <N>, <E>, <EN>, <N^2>, <E^2>, Q_st and C_v are
  averaged with Welford's update and their population deviation
  sqrt(M2 / 5) is the error;
* the node's sixth statistic slot (the reference's <exp(-E/kT)> slot)
  receives the mean of the window Q_st values (gcmc.f:2008), and the union
  weighs it as the node's H entry.

Only runs of at most 500000 samples per node take the window branch.
Longer runs keep the rolling averages; their errors come from the
reference's rolling 100000-sample windows (gcmc.f:1887-1925), which are not
restated here, so they are reported as NaN.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
from typing import List, Sequence

RGAS = 8.314472 / 4184.0          # utility_pack.f:109
KBOLTZ = 1.9872041e-3             # kcal/mol/K
WINDN = 5                         # gcmc.f:119
MAXSAMP = 500000


def _div(a, b):
    """IEEE division, as the Fortran code divides (x/0 = +-inf, 0/0 = NaN)."""
    with np.errstate(all="ignore"):
        return float(np.float64(a) / np.float64(b))


def calc_qst(E2, E, N, N2, EN, temp):
    """Isosteric heat (utility_pack.f:2280-2290)."""
    return _div(-1.0 * (EN - E * N), (N2 - N * N)) + RGAS * temp


def calc_cv(E2, E, N, N2, EN, temp):
    """Heat capacity at constant volume (utility_pack.f:2305-2316)."""
    return (3.0 / 2.0 * RGAS) + _div(E2 - (E * E) - _div((EN - E * N) ** 2, (N2 - N * N)),
                                     (N * KBOLTZ * temp * temp))


def _sqrt(v):
    return math.sqrt(v) if v >= 0 else math.nan   # NaN stays NaN


@dataclass
class NodeStats:
    """One node's (chain's) report: the averages of <N>, <E>, <EN>, <N^2>,
    <E^2>, Q_st, C_v and the H slot, their errors, and the sample count."""
    weight: float
    avg: List[float]      # N, E, EN, N2, E2, Q_st, C_v, H slot
    std: List[float]      # same order


def window_sums(n_samples: Sequence[float], e_samples: Sequence[float]):
    """Sums of N, E, E*N, N^2, E^2 over the five windows of one guest's
    production samples, and nwind.  Returns (sums[5][5], nwind)."""
    n = len(n_samples)
    nwind = int(float(n) / WINDN)
    sums = [[0.0] * 5 for _ in range(WINDN)]
    for x in range(WINDN):
        y = x * nwind
        s = sums[x]
        for k in range(y, y + nwind + 1):
            a = float(n_samples[k]) if k < n else 0.0
            e = float(e_samples[k]) if k < n else 0.0
            s[0] += a
            s[1] += e
            s[2] += a * e
            s[3] += a * a
            s[4] += e * e
    return sums, nwind


def node_stats(sums, nwind: int, weight: float, temp: float,
               rolling: Sequence[float] = (0.0,) * 6) -> NodeStats:
    """A node's report from its window sums (window_sums, or the batch
    kernel's window accumulators): gcmc.f:1969-2031 then 2051-2100.
    rolling holds the rolling averages (N, E, EN, N2, E2, H) used when the
    node has more than 500000 samples (errors NaN then)."""
    if weight <= MAXSAMP:
        mean = [0.0] * 7
        m2 = [0.0] * 7
        for x in range(WINDN):
            s = sums[x]
            aN = _div(s[0], nwind)
            E = _div(s[1], nwind)
            EN = _div(s[2], nwind)
            N2 = _div(s[3], nwind)
            E2 = _div(s[4], nwind)
            w = [aN, E, EN, N2, E2, calc_qst(E2, E, aN, N2, EN, temp),
                 calc_cv(E2, E, aN, N2, EN, temp)]
            for z in range(7):
                delta = w[z] - mean[z]
                mean[z] = mean[z] + delta / (x + 1)
                m2[z] = m2[z] + delta * (w[z] - mean[z])
        cs = mean[:6]
        std = [_sqrt(m2[z] / WINDN) for z in range(7)] + [0.0]
    else:
        cs = list(rolling)
        std = [math.nan] * 8
    aN, E, EN, N2, E2 = cs[:5]
    q = calc_qst(E2, E, aN, N2, EN, temp)
    c = calc_cv(E2, E, aN, N2, EN, temp)
    return NodeStats(weight, [aN, E, EN, N2, E2, q, c, cs[5]], std)


def union(nodes: Sequence[NodeStats], temp: float) -> NodeStats:
    """Weighted combination of nodes (avunion/stdunion, gcmc.f:2366-2432):
    averages weighted by sample count, Q_st and C_v of the combined
    averages, errors from the pooled second moments."""
    sw = 0.0
    acc = [0.0] * 8
    for nd in nodes:
        sw += nd.weight
        for z in (0, 1, 2, 3, 4, 7):
            acc[z] += nd.weight * nd.avg[z]
    avg = [a / sw for a in acc]
    avg[5] = calc_qst(avg[4], avg[1], avg[0], avg[3], avg[2], temp)
    avg[6] = calc_cv(avg[4], avg[1], avg[0], avg[3], avg[2], temp)
    sq = [0.0] * 8
    for nd in nodes:
        for z in range(8):
            sq[z] += nd.weight * (nd.std[z] ** 2 + nd.avg[z] ** 2)
    std = [_sqrt(sq[z] / sw - avg[z] ** 2) for z in range(8)]
    return NodeStats(sw, avg, std)


def from_samples(n_samples, e_samples, temp: float) -> NodeStats:
    """A node's report straight from its production samples of one guest."""
    n = len(n_samples)
    if n > MAXSAMP:
        # rolling averages only (gcmc.f:1803-1848)
        roll = [0.0] * 6
        for k, (a, e) in enumerate(zip(n_samples, e_samples)):
            a, e = float(a), float(e)
            for z, v in enumerate((a, e, a * e, a * a, e * e)):
                roll[z] = roll[z] + (v - roll[z]) / (k + 1)
        return node_stats(None, 0, float(n), temp, roll)
    sums, nwind = window_sums(n_samples, e_samples)
    return node_stats(sums, nwind, float(n), temp)


# ---------------------------------------------------------------------------
# selectivity (gcmc.f:1887-1916, printed at gcmc.f:2315-2322)
# ---------------------------------------------------------------------------
NWIND = 100000                    # readinputs.f:888-889 ("averaging window")


@dataclass
class Selectivity:
    """Guest j over guest i (j < i, 0-based): the reference's running,
    window-weighted mean of <N_j>_w / <N_i>_w over the averaging windows w
    (seltot(i,j)), the quantity it prints after '+/-' (seltot(j,i): the
    weighted second moment minus the squared mean, a variance), and the
    weight sum avcount - 1 + weight of the last window."""
    mean: float
    var: float
    weight: float
    second: float = 0.0   # weighted mean of the squared ratio (selwin(i,j))


def selectivity_update(state, means, weight):
    """One window of the reference's update: avcount counts every window
    (also those skipped for <N_i> = 0); means = the window's <N_g>;
    weight = its samples / nwind (1 for a full window).  state =
    [avcount, mean[i][j], second[i][j], var[i][j]] with i > j, updated in
    place; var is recomputed, as the reference does, at each window that
    updates the pair."""
    state[0] += 1
    av = state[0]
    mean, sec, var = state[1], state[2], state[3]
    ng = len(means)
    for i in range(1, ng):
        aN = means[i]
        for j in range(i):
            if aN == 0:
                continue
            s = _div(means[j], aN)
            delta = weight * (s - mean[i][j])
            mean[i][j] = mean[i][j] + delta / av
            sec[i][j] = sec[i][j] + weight * (s * s - sec[i][j]) / (av - 1 + weight)
            w = av - 1 + weight
            var[i][j] = _div(w * sec[i][j] - (w * mean[i][j] * mean[i][j]), w)
    return state


def selectivity_state(ng: int):
    z = lambda: [[0.0] * ng for _ in range(ng)]   # noqa: E731
    return [0, z(), z(), z()]


def selectivity_report(state, last_weight: float):
    """{(j, i): Selectivity} from an update state; last_weight is the
    weight of the last window taken (the partial window's samples / nwind,
    or 1)."""
    av, mean, sec, var = state
    ng = len(mean)
    out = {}
    for i in range(1, ng):
        for j in range(i):
            out[(j, i)] = Selectivity(mean[i][j], var[i][j], av - 1 + last_weight, sec[i][j])
    return out


def selectivity_from_samples(n_samples: Sequence[Sequence[float]], nwind: int = NWIND):
    """The reference's selectivity from one node's production samples
    (n_samples[k][g] = N of guest g at sample k): windows of nwind samples,
    the last one partial (weight = its samples / nwind, taken at the last
    move as gcmc.f:1887 does when the run ends)."""
    n = len(n_samples)
    ng = len(n_samples[0]) if n else 0
    st = selectivity_state(ng)
    last_w = 1.0
    for a in range(0, n, nwind):
        blk = n_samples[a:a + nwind]
        w = len(blk) / nwind
        means = [sum(float(r[g]) for r in blk) / len(blk) for g in range(ng)]
        selectivity_update(st, means, w)
        last_w = w
    return selectivity_report(st, last_w)


def selectivity_union(nodes: Sequence[dict]) -> dict:
    """Replicas of one state point combined: means and second moments
    weighted by each replica's window weight, the variance from the pooled
    second moments (the reference prints node 0's only; identical replicas
    give the replica)."""
    out = {}
    for key in nodes[0]:
        sw = sum(nd[key].weight for nd in nodes)
        if not sw:
            out[key] = Selectivity(0.0, 0.0, 0.0, 0.0)
            continue
        m = sum(nd[key].weight * nd[key].mean for nd in nodes) / sw
        s2 = sum(nd[key].weight * nd[key].second for nd in nodes) / sw
        out[key] = Selectivity(m, s2 - m * m, sw, s2)
    return out
