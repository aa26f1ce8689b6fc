"""The per-instance counters of large levels: launches over one instance add
their claims up per workgroup (InstDesc::defer_claims) and, on levels of
2^22 candidates and more, reserve fresh-list entries 256 at a time per
wavefront, leaving hole markers the resolve skips (InstDesc::fresh_chunk).
A batch that does not park its large members (vfb_tuning.park_min = 0)
builds the same levels in launches over several instances, which update
the counters once per wavefront: every result field must agree, field by
field, and the plain solves must agree with the OpenMP build.
"""
import pytest

import oracle_lib as O
import vfb_amd as V

pytestmark = pytest.mark.gpu

FIELDS = ("status", "min_size", "formula", "distinct", "candidates", "level_winners", "level_candidates",
          "otf_level", "otf_stored", "found_ordinal", "levels_done", "top_op", "table_slots")


@pytest.fixture
def tuning():
    old = V.get_tuning()
    yield
    V.set_tuning(**old)


def _large():
    # config-4 mix instances 219 and 456 (levels of 10^7-10^8 candidates) at
    # max size 16, config-2 instances 2 and 3 (levels of 1-5 M) at max size 14
    return [(V.Instance.generate(4, 219), 16), (V.Instance.generate(4, 456), 16),
            (V.Instance.generate(2, 2), 14), (V.Instance.generate(2, 3), 14)]


def test_large_levels_unparked_batch_equals_plain(tuning):
    print("[counters] plain (per-workgroup claims, chunked fresh lists) against an unparked batch", flush=True)
    cases = _large()
    plain = [V.solve(i, max_size=m) for i, m in cases]
    assert max(max(r["level_candidates"]) for r in plain) >= 1 << 22
    V.set_tuning(park_min=0)
    by_size = {}
    for i, m in cases:
        by_size.setdefault(m, []).append(i)
    got = {}
    for m, insts in by_size.items():
        for i, r in zip(insts, V.solve_batch(insts, max_size=m)):
            got[id(i)] = r
    for (i, m), w in zip(cases, plain):
        g = got[id(i)]
        for f in FIELDS:
            assert g[f] == w[f], (m, f, g[f], w[f])


@pytest.mark.parametrize("k", [2, 3])
def test_config2_large_levels_equal_omp(k):
    """Config-2 instances whose largest levels (1-5 M candidates) take the
    chunked fresh lists, against the OpenMP build."""
    print(f"[counters] config-2 #{k} max 14 against the OpenMP build", flush=True)
    inst = V.Instance.generate(2, k)
    names, lens, atoms, P = inst.arrays()
    want, _ = O.enumerate_omp(O.Instance(names, lens, atoms, P), 14, 3)
    got = V.solve(inst, max_size=14)
    for f in ("status", "min_size", "formula", "distinct", "candidates", "found_ordinal"):
        assert got[f] == want[f], (k, f, got[f], want[f])
