"""Batches and served streams under a small device arena.

With an arena too small for every member at once, batch members and served
requests evict one another (the youngest goes first) and are started again
later; a restarted member's queued rehash, fills and uploads are dropped and
its old hash set must outlive the rehash launch that reads it.  Every result
must still equal the plain solve of the same instance under the same arena.

Members are those whose plain solve the arena does not cut short (no
onTheFly level): a plain solve that fills the arena alone keeps as many rows
as its own largest free block holds, while a batch member evicts younger
members for the same rows, so only uncut solves are comparable.  (A per-solve
hbm_budget below the arena makes cut solves comparable too: the budget, not
the arena, then decides where they are cut.)
"""
import pytest

import vfb_amd as V

pytestmark = pytest.mark.gpu

FIELDS = ("status", "min_size", "formula", "distinct", "candidates", "level_winners",
          "otf_level", "otf_stored", "found_ordinal")
ARENA = 256 << 20


@pytest.fixture
def small_arena():
    V.set_arena_bytes(ARENA)
    try:
        yield
    finally:
        V.set_arena_bytes(0)


def test_batch_and_server_under_small_arena(small_arena):
    cands = [V.Instance.generate(4, k) for k in range(128)]
    print(f"[memory] plain solves, arena {ARENA >> 20} MiB", flush=True)
    plain_all = [V.solve(i, max_size=10) for i in cands]
    keep = [k for k, r in enumerate(plain_all) if r["otf_level"] == 0 and r["status"] in (V.OK, V.NOT_FOUND)]
    insts = [cands[k] for k in keep]
    plain = [plain_all[k] for k in keep]
    print(f"[memory] batch of {len(insts)} uncut instances", flush=True)
    assert len(insts) >= 96
    batch = V.solve_batch(insts, max_size=10)
    restarts = sum(r["restarts"] for r in batch)
    print(f"[memory] batch restarts {restarts}", flush=True)
    for k, (a, b) in enumerate(zip(plain, batch)):
        for f in FIELDS:
            assert a[f] == b[f], (k, f, a[f], b[f])
    assert restarts > 0
    print("[memory] server", flush=True)
    with V.Server(max_size=10, max_active=96) as srv:
        tickets = [srv.submit(i) for i in insts]
        served = [srv.wait(t) for t in tickets]
    print(f"[memory] server restarts {sum(r['restarts'] for r in served)}", flush=True)
    for k, (a, b) in enumerate(zip(plain, served)):
        for f in FIELDS:
            assert a[f] == b[f], (k, f, a[f], b[f])


def test_set_arena_refused_while_serving(small_arena):
    with V.Server(max_size=4, max_active=2):
        with pytest.raises(V.VfbError):
            V.set_arena_bytes(1 << 30)


def test_server_ticket_consumed_once():
    with V.Server(max_size=6, max_active=2) as srv:
        t = srv.submit(V.Instance.generate(1, 0))
        r = srv.wait(t)
        assert r["status"] == V.OK
        with pytest.raises(V.VfbError):
            srv.wait(t)
        with pytest.raises(V.VfbError):
            srv.poll(t)
        with pytest.raises(V.VfbError):
            srv.poll(12345)
