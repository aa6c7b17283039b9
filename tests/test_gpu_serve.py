"""Continuous-batching server: every request completes, admission and
retirement keep the slots consistent, and each request's result is the
same alone and in the mix (bit for bit: chains do not share arithmetic)."""
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("steps", "nmol", "energy", "accepted", "attempted", "sum_n", "sum_e",
        "e_acc", "e_full", "win", "sel_count", "sel_base", "sel_mean", "sel_var")


def test_serve_same_alone_and_in_mix():
    import fastmc_amd
    from fastmc_amd import serve
    fastmc_amd.lib()
    reqs = serve.poisson_requests(6, rate=2000.0, seed=11, moves=700)
    srv = serve.Server(slots=4, slice_moves=200)
    res = srv.run(reqs, realtime=True)
    assert len(res) == 6 and not srv._live
    assert all(r.stats["steps"] == 700 for r in res)
    assert all(r.done >= r.admitted >= 0.0 and r.latency > 0 for r in res)
    # with 4 slots and 6 requests some waited for a slot
    assert max(r.admitted - r.arrival for r in res) > 0.0
    # every result carries the end-of-run report of each guest type and,
    # with several guest types, the selectivity of each pair
    multi = 0
    for r in res:
        ng = len(r.stats["nmol"])
        assert len(r.report) == ng
        assert r.report[0].weight == 700 and r.report[0].avg[0] >= 0.0
        assert sorted(r.selectivity) == [(j, i) for i in range(ng) for j in range(i)]
        multi += ng > 1
    assert multi >= 1
    srv.close()
    for k in (0, 3, 5):
        one = serve.Server(slots=1, slice_moves=300)
        alone = one.run([reqs[k]], realtime=False)[0]
        one.close()
        for key in KEYS:
            assert alone.stats[key] == res[k].stats[key], (k, key)
    print("latency ms:", [round(1e3 * r.latency, 1) for r in res])


def test_serve_budget_slices_same_results():
    """Time-budgeted slices split each request's moves at other points; the
    results equal those of fixed slices, request by request."""
    import fastmc_amd
    from fastmc_amd import serve
    fastmc_amd.lib()
    reqs = serve.poisson_requests(5, rate=2000.0, seed=12, moves=600)
    a = serve.Server(slots=8, slice_moves=200)
    ra = a.run(reqs, realtime=False)
    a.close()
    b = serve.Server(slots=8, slice_ms=1.5)
    rb = b.run(reqs, realtime=True)
    b.close()
    for x, y in zip(ra, rb):
        for key in KEYS:
            assert x.stats[key] == y.stats[key], (x.rid, key)


def test_serve_suspend_and_resume():
    """A run stopped after a few slices checkpoints its live chains; a new
    server resumes them together with the requests never admitted, and
    every request ends exactly as in one uninterrupted run."""
    import fastmc_amd
    from fastmc_amd import serve
    fastmc_amd.lib()
    reqs = serve.poisson_requests(7, rate=2000.0, seed=13, moves=600)
    whole = serve.Server(slots=3, slice_moves=150)
    want = {r.rid: r for r in whole.run(reqs, realtime=False)}
    whole.close()
    first = serve.Server(slots=3, slice_moves=150)
    done1 = first.run(reqs, realtime=False, stop_after_slices=3)
    sus, rest = first.suspended, first.unstarted
    first.close()
    assert sus and rest and not first._live
    assert len(done1) + len(sus) + len(rest) == len(reqs)
    second = serve.Server(slots=3, slice_moves=150)
    done2 = second.run(rest, realtime=False, resume=sus)
    second.close()
    got = {r.rid: r for r in done1 + done2}
    assert sorted(got) == sorted(want)
    for rid in want:
        for key in KEYS:
            assert got[rid].stats[key] == want[rid].stats[key], (rid, key)
    print(f"suspended {len(sus)} live, {len(rest)} unstarted; all {len(got)} equal")
