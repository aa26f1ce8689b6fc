"""Large levels: winners marked during the expansion (InstDesc::direct) and the
per-instance counters of launches over one instance.

A level of vfb_tuning.direct_min candidates and more (default 2^18) keeps
no fresh list: each fresh candidate marks its byte, a candidate that lowers
a slot's ordinal of the level marks the displaced one lost, and the count kernel packs "marked and
not lost" into the winner bitmap.  Launches over one instance add their
claims and fresh candidates up per workgroup (InstDesc::defer_claims); a
batch that does not park its large members (vfb_tuning.park_min = 0) builds
the same levels in launches over several instances, which update the
counters once per wavefront.  Every result field must agree, field by
field, with the plain solves and with the OpenMP build, also where the
level is built in a probe's passes with relaxed keys (the probe's match can
lose its key to a lower ordinal, and the level's marks then stay for the
pass that builds the rest) and where a set overflow builds a level again
(its marks cleared first).
"""
import random

import pytest

import oracle_lib as O
import vfb_amd as V

pytestmark = pytest.mark.gpu

FIELDS = ("status", "min_size", "formula", "distinct", "candidates", "level_winners", "level_candidates",
          "otf_level", "otf_stored", "found_ordinal", "levels_done", "top_op", "table_slots")
DIRECT_MIN = V.tuning_default()["direct_min"]


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
    print("[counters] plain (per-workgroup counts, direct marks) against an unparked batch", flush=True)
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
    """Config-2 instances whose largest levels (1-5 M candidates) are marked
    directly, against the OpenMP build."""
    print(f"[counters] config-2 #{k} max 14 against the OpenMP build", flush=True)
    inst = V.Instance.generate(2, k)
    names, lens, atoms, P = inst.arrays()
    want, _ = O.enumerate_omp(O.Instance(names, lens, atoms, P), 14, 3)
    got = V.solve(inst, max_size=14)
    assert max(got["level_candidates"]) >= DIRECT_MIN
    for f in ("status", "min_size", "formula", "distinct", "candidates", "found_ordinal"):
        assert got[f] == want[f], (k, f, got[f], want[f])


@pytest.mark.parametrize("rt", [1, 2, 3])
@pytest.mark.parametrize("k", [2, 3])
def test_direct_levels_in_probe_passes_equal_omp(tuning, k, rt):
    """Every level of a binary operator probed first (pass_min = 1), so that
    a found level is built up to the probe's match and, when the match lost
    its relaxed key to a lower ordinal, then above it."""
    print(f"[counters] config-2 #{k} max 12 relaxed {rt}, every level probed", flush=True)
    inst = V.Instance.generate(2, k)
    names, lens, atoms, P = inst.arrays()
    want, _ = O.enumerate_omp(O.Instance(names, lens, atoms, P), 12, rt)
    V.set_tuning(pass_min=1)
    got = V.solve(inst, max_size=12, relaxed_type=rt)
    assert max(got["level_candidates"]) >= DIRECT_MIN
    for f in ("status", "min_size", "formula", "distinct", "candidates", "found_ordinal", "top_op"):
        assert got[f] == want[f], (k, rt, f, got[f], want[f])


def test_direct_levels_built_again_after_set_overflow(tuning):
    """Sets sized for 1 % of the expected survivors overflow at the large
    levels: each such level is built again in a larger set, its marks cleared
    first, and every result but the set's size equals the default's."""
    cases = [(V.Instance.generate(2, 2), 14), (V.Instance.generate(2, 3), 14), (V.Instance.generate(4, 219), 13)]
    want = [V.solve(i, max_size=m) for i, m in cases]
    V.set_tuning(set_margin_pct=1, table_min_slots=1 << 17)
    print("[counters] set margin 1 %: overflowing large levels", flush=True)
    got = [V.solve(i, max_size=m) for i, m in cases]
    assert sum(g["retries"] for g in got) > 0
    for (i, m), w, g in zip(cases, want, got):
        assert max(g["level_candidates"]) >= DIRECT_MIN
        for f in FIELDS[:-1]:
            assert g[f] == w[f], (m, f, g[f], w[f])
    print("[counters] the same in one batch", flush=True)
    V.set_tuning(park_min=0)
    got = V.solve_batch([i for i, m in cases[:2]], max_size=14)
    for w, g in zip(want[:2], got):
        for f in FIELDS[:-1]:
            assert g[f] == w[f], (f, g[f], w[f])


@pytest.mark.parametrize("rt", [1, 2])
def test_direct_every_level_relaxed_keys_equal_oracle(tuning, rt):
    """Random instances with relaxed keys, every level marked directly and
    every level of a binary operator probed first: the probe's match loses
    its key to a lower ordinal where two CS share a relaxed key."""
    V.set_tuning(pass_min=1, direct_min=1)
    rng = random.Random(5150 + rt)
    for seed in range(12):
        T = rng.randint(2, 24)
        lens = [rng.randint(1, 64 if seed % 2 else 10) for _ in range(T)]
        A = rng.randint(1, 4)
        atoms = [[sum((rng.random() < 0.5) << j for j in range(lens[t])) for t in range(T)] for _ in range(A)]
        P = rng.randint(1, T - 1)
        names = [f"p{a}" for a in range(A)]
        print(f"[counters] relaxed {rt} random {seed}, direct everywhere", flush=True)
        want = O.enumerate_cs(O.Instance(names, lens, atoms, P), 7, rt)
        got = V.solve(V.Instance.from_bits(names, lens, atoms, P), max_size=7, relaxed_type=rt)
        for f in ("status", "min_size", "formula", "distinct", "candidates", "found_ordinal", "top_op"):
            assert got[f] == want[f], (rt, seed, f, got[f], want[f])


def test_last_pass_grows_the_set_and_rebuilds(tuning):
    """Config-4 instance 35 with relaxed keys 2 at max size 12, every level
    probed: at a level of 18.6 M candidates the probe's match loses its key,
    the set (sized for the first pass) grows for the whole level, and the
    level is built again from its first pass (counted in retries).  Results
    equal the OpenMP build's and the unprobed solve's."""
    inst = V.Instance.generate(4, 35)
    names, lens, atoms, P = inst.arrays()
    want, _ = O.enumerate_omp(O.Instance(names, lens, atoms, P), 12, 2)
    V.set_tuning(pass_min=0)
    fused = V.solve(inst, max_size=12, relaxed_type=2)
    V.set_tuning(pass_min=1)
    print("[counters] config-4 #35 relaxed 2, every level probed", flush=True)
    got = V.solve(inst, max_size=12, relaxed_type=2)
    assert got["retries"] >= 1 and fused["retries"] == 0
    for f in ("status", "min_size", "formula", "distinct", "candidates", "found_ordinal", "top_op"):
        assert got[f] == want[f], (f, got[f], want[f])
        assert fused[f] == want[f], (f, fused[f], want[f])


@pytest.mark.parametrize("setting", [dict(pass_min=1, direct_min=1), dict(pass_min=1 << 10, direct_min=1 << 10)])
def test_random_sweep_direct_and_probed_equal_oracle(tuning, setting):
    """Random instances (up to 30 traces, 5 atoms, lengths to 64, bit
    densities 0.3-0.7) at max size 7-8 with relaxed keys 1-3, every level (or
    every level of 1024 candidates and more) marked directly and probed
    first: plain and batched solves against the sequential oracle."""
    V.set_tuning(**setting)
    F = ("status", "min_size", "formula", "distinct", "candidates", "found_ordinal", "top_op")
    for rt in (1, 2, 3):
        rng = random.Random(31337 + rt)
        cases = []
        for seed in range(40):
            T = rng.randint(2, 30)
            lens = [rng.randint(1, 64 if seed % 3 else 12) for _ in range(T)]
            A = rng.randint(1, 5)
            atoms = [[sum((rng.random() < rng.choice((0.3, 0.5, 0.7))) << j for j in range(lens[t])) for t in range(T)]
                     for _ in range(A)]
            P = rng.randint(1, T - 1)
            names = [f"p{a}" for a in range(A)]
            ms = 8 if A <= 3 else 7
            want = O.enumerate_cs(O.Instance(names, lens, atoms, P), ms, rt)
            inst = V.Instance.from_bits(names, lens, atoms, P)
            got = V.solve(inst, max_size=ms, relaxed_type=rt)
            for f in F:
                assert got[f] == want[f], (setting, rt, seed, f, got[f], want[f])
            cases.append((inst, ms, want))
        print(f"[counters] sweep {setting} relaxed {rt}: plain ok, batches", flush=True)
        for ms in (7, 8):
            sub = [(i, w) for i, m, w in cases if m == ms]
            for (i, w), g in zip(sub, V.solve_batch([i for i, _ in sub], max_size=ms, relaxed_type=rt)):
                for f in F:
                    assert g[f] == w[f], (setting, rt, ms, f, g[f], w[f])
