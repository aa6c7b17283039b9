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
