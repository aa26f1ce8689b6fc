"""vfb_solve_noalloc (no device allocation; all device memory in the
library's static scratch) against vfb_solve, through the C-ABI.

Bar: bit-exact -- status, minimal size, formula text, found ordinal, the
per-level distinct / candidate counts and the onTheFly fields -- on two calls
in a row (the second with a smaller instance, so that the scratch holds the
first call's contents), on a row-capped solve, and a solve too large for the
scratch returns VFB_CAPACITY and leaves the entry point usable.
"""
import ctypes as C

import pytest

import vfb_amd as V

pytestmark = pytest.mark.gpu

FIELDS = ("status", "min_size", "top_op", "formula", "distinct", "candidates", "level_winners",
          "level_candidates", "found_ordinal", "otf_level", "otf_stored", "levels_done")


def _noalloc(inst, **kw):
    """vfb_solve_noalloc called through the C-ABI directly."""
    o = V._opts(kw.get("max_size", V.MAX_SIZE), 3, False, 0, False, kw.get("max_rows", 0))
    r = V.Result()
    st = V.lib().vfb_solve_noalloc(inst.handle, C.byref(o), C.byref(r))
    try:
        d = V._todict(r)
    finally:
        V.lib().vfb_result_free(C.byref(r))
    assert st == d["status"]
    return d


def _same(a, b, label):
    for f in FIELDS:
        assert a[f] == b[f], (label, f, a[f], b[f])


@pytest.mark.parametrize("first,second", [((2, 1, 12), (1, 0, 10)), ((4, 3, 12), (2, 0, 10))])
def test_noalloc_equals_solve_twice(first, second):
    for cfg, k, ms in (first, second):
        print(f"[noalloc] config-{cfg} #{k} max {ms}", flush=True)
        inst = V.Instance.generate(cfg, k)
        want = V.solve(inst, max_size=ms)
        got = _noalloc(inst, max_size=ms)
        _same(want, got, (cfg, k, ms))


def test_noalloc_row_cap_equals_solve():
    inst = V.Instance.generate(2, 2)
    print("[noalloc] config-2 #2 max_rows 3000", flush=True)
    want = V.solve(inst, max_size=10, max_rows=3000)
    got = _noalloc(inst, max_size=10, max_rows=3000)
    assert want["otf_level"] > 0
    _same(want, got, "row cap")


def test_noalloc_too_large_then_usable():
    big = V.Instance.generate(3, 4)
    print("[noalloc] config-3 #4 max 18 (needs more than the scratch)", flush=True)
    r = _noalloc(big, max_size=18)
    assert r["status"] == V.CAPACITY and r["formula"] == "see_the_error"
    small = V.Instance.generate(1, 3)
    print("[noalloc] config-1 #3 max 10 after the refused solve", flush=True)
    _same(V.solve(small, max_size=10), _noalloc(small, max_size=10), "after")
