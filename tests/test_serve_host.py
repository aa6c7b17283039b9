"""Host-side bookkeeping of the serving loop, with a stand-in for the
device batch (no GPU)."""
import threading
import types


def test_retire_frees_the_library_slot_with_the_live_entry():
    """A finished chain stays counted in _live until the library has freed
    its slot, and both happen under the lock.  The admission thread counts
    free slots from _live under the same lock, so it can neither add a chain
    while the library is still full nor have the chain it adds into the
    freed slot (which may get the same id) forgotten by the retire."""
    from fastmc_amd import serve

    srv = serve.Server.__new__(serve.Server)
    cv = threading.Condition()

    class Batch:
        removed = []

        def poll(self):
            return [2]

        def stats(self, cid):
            return {"win_len": 0, "capacity_fault": 0, "steps": 10}

        def remove_chain(self, cid):
            assert cid in srv._live
            assert cv._is_owned()
            self.removed.append(cid)

    req = types.SimpleNamespace(rid=7, arrival=0.0, temp=298.0)
    srv.batch = Batch()
    srv._live = {2: (req, 0.5, 100), 3: (req, 0.6, 50)}
    srv._used = 150
    srv.moves_done = 0
    srv.results = {}
    srv._retire(1.0, cv)
    assert srv.batch.removed == [2]
    assert list(srv._live) == [3] and srv._used == 50 and srv.moves_done == 10
    assert srv.results[7].stats["steps"] == 10


def test_retire_waits_for_a_chain_the_admission_has_not_recorded():
    """A chain that finishes in its first slice can be polled before the
    admission thread has recorded it in _live; the retire waits for the
    record instead of failing."""
    from fastmc_amd import serve

    srv = serve.Server.__new__(serve.Server)
    cv = threading.Condition()

    class Batch:
        def poll(self):
            return [4]

        def stats(self, cid):
            return {"win_len": 0, "capacity_fault": 0, "steps": 3}

        def remove_chain(self, cid):
            pass

    req = types.SimpleNamespace(rid=9, arrival=0.0, temp=298.0)
    srv.batch = Batch()
    srv._live = {}
    srv._used = 0
    srv.moves_done = 0
    srv.results = {}

    def admit():
        with cv:
            srv._live[4] = (req, 0.1, 20)
            srv._used += 20
            cv.notify_all()

    t = threading.Timer(0.05, admit)
    t.start()
    srv._retire(1.0, cv)
    t.join()
    assert srv._live == {} and srv._used == 0 and srv.results[9].stats["steps"] == 3


def test_a_failed_slice_ends_the_run_instead_of_hanging(monkeypatch):
    """When a slice fails, run() stops the admission thread (which may be
    waiting for a slot) and raises the error."""
    from fastmc_amd import serve

    class Ctx:
        def close(self):
            pass

    monkeypatch.setattr(serve.Context, "from_arrays", staticmethod(lambda arrays: Ctx()))
    monkeypatch.setattr(serve, "footprint", lambda ctx, precision: 1)

    class Batch:
        def add_chain(self, ctx, temp, press, seed, **kw):
            return 0

        def launch(self, n):
            raise RuntimeError("slice failed")

    srv = serve.Server.__new__(serve.Server)
    srv.slots, srv.precision, srv.slice_moves, srv.slice_ms = 1, 0, 10, 0.0
    srv.hbm_budget, srv.clock, srv.results = 1 << 30, serve.time.perf_counter, {}
    srv.batch = Batch()
    srv._live, srv._used, srv._inflight = {}, 0, 0
    srv.slices = srv.moves_done = 0
    srv.suspended, srv.unstarted = [], []
    reqs = [types.SimpleNamespace(rid=i, arrival=0.0, arrays=None, temp=298.0,
                                  press_pa=[1e5], seed=(1, 2, 3, 4), delr=1.0,
                                  eq_steps=0, moves=100) for i in range(3)]
    out = {}

    def go():
        try:
            srv.run(reqs, realtime=False)
        except RuntimeError as e:
            out["error"] = str(e)

    t = threading.Thread(target=go, daemon=True)
    t.start()
    t.join(timeout=20)
    assert not t.is_alive(), "run() hung after a failed slice"
    assert out.get("error") == "slice failed"
