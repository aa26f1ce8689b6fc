"""GPU parity under every kind of tuning setting: launch shapes and the pass
split change speed only, never a result.

pass_min=1 probes every level with a binary operator (target check only)
and builds a level that holds a match one canonical operator block at a time
(~, &, |, X F G, U), stopping at the block that settles it; the results must
still equal the oracle's bit for bit, including the per-level counts and the
returned formula.
"""
import random

import pytest

import oracle_lib as O
import vfb_amd as V
from ltl_semantics import separates

pytestmark = pytest.mark.gpu

SETTINGS = {
    "split_all_levels": dict(pass_min=1),
    "split_all_levels_wide": dict(pass_min=1, wide_max_waves=1 << 30),
    "never_wide_small_grids": dict(pass_min=0, wide_max_waves=0, expand_blocks_per_cu=1,
                                   resolve_blocks_per_cu=1, store_blocks_per_cu=1),
    "big_grids": dict(pass_min=1 << 16, expand_blocks_per_cu=64, resolve_blocks_per_cu=64,
                      store_blocks_per_cu=64),
    # hash sets at half or nine tenths load, one or 64 traces per store item,
    # target checks one or 16 left rows per wave
    "low_load_thin_store": dict(pass_min=1, max_load_pct=50, store_traces=1, check_rows=16,
                                wide_max_waves=0),
    "high_load_wide_store": dict(pass_min=1, max_load_pct=90, store_traces=64, check_rows=4,
                                 wide_max_waves=0, table_min_slots=1 << 20),
    "check_one_row": dict(pass_min=1, check_rows=1, wide_max_waves=0),
    # winners marked by the expansion at every level, or at none (fresh
    # lists and resolves throughout)
    "direct_everywhere": dict(pass_min=1, direct_min=1),
    "never_direct": dict(direct_min=1 << 40),
}

FIELDS = ("status", "min_size", "formula", "distinct", "candidates", "found_ordinal", "otf_level", "otf_stored")


@pytest.fixture(params=sorted(SETTINGS))
def tuned(request):
    old = V.set_tuning(**SETTINGS[request.param])
    yield request.param
    V.set_tuning(**old)


def _cases():
    out = [(1, k, 10, 0) for k in range(3)] + [(2, k, 8, 0) for k in range(4)] + [(3, 1, 6, 0)]
    out += [(2, 2, 10, 20000), (4, 14, 10, 400), (1, 0, 10, 40)]
    return out


@pytest.mark.parametrize("cfg,k,ms,rows", _cases())
def test_tuned_parity(tuned, cfg, k, ms, rows):
    inst = V.Instance.generate(cfg, k)
    names, lens, atoms, P = inst.arrays()
    print(f"[parity] tuning {tuned} cfg{cfg} #{k} max {ms} rows {rows}", flush=True)
    want = O.enumerate_cs(O.Instance(names, lens, atoms, P), ms, 3, rows)
    got = V.solve(inst, max_size=ms, max_rows=rows)
    for f in FIELDS:
        assert got[f] == want[f], (tuned, f, got[f], want[f])
    if got["status"] == V.OK:
        assert separates(got["formula"], names, lens, atoms, P)


def test_tuned_random(tuned):
    rng = random.Random(4242)
    for seed in range(6):
        T = rng.randint(2, 24)
        lens = [rng.randint(1, 64 if seed % 2 else 10) for _ in range(T)]
        A = rng.randint(1, 4)
        atoms = [[sum((rng.random() < 0.5) << j for j in range(lens[t])) for t in range(T)] for _ in range(A)]
        P = rng.randint(1, T - 1)
        inst = V.Instance.from_bits([f"p{a}" for a in range(A)], lens, atoms, P)
        print(f"[parity] tuning {tuned} random {seed}", flush=True)
        want = O.enumerate_cs(O.Instance([f"p{a}" for a in range(A)], lens, atoms, P), 7, 3)
        got = V.solve(inst, max_size=7)
        for f in FIELDS[:6]:
            assert got[f] == want[f], (tuned, seed, f, got[f], want[f])


def test_split_batch_equals_plain():
    """Instances at different passes of their levels share launches."""
    old = V.set_tuning(pass_min=1)
    try:
        insts = [V.Instance.generate(4, k) for k in range(64)]
        plain = [V.solve(i, max_size=12) for i in insts]
        batch = V.solve_batch(insts, max_size=12)
        for k, (a, b) in enumerate(zip(plain, batch)):
            for f in FIELDS + ("level_winners", "level_candidates"):
                assert a[f] == b[f], (k, f, a[f], b[f])
    finally:
        V.set_tuning(**old)


@pytest.mark.parametrize("park_min", [1, 1 << 12, 1 << 16])
def test_parked_batch_equals_plain(park_min):
    """Batch members reaching a level of park_min candidates wait, their rows
    and set kept, and go on alone after the others, oldest first; park_min=1
    parks every member at its first level, so they all run one by one from
    there.  Every result equals the plain solve's."""
    insts = [V.Instance.generate(4, k) for k in range(40)] + [V.Instance.generate(2, k) for k in range(4)]
    plain = [V.solve(i, max_size=12) for i in insts]
    old = V.set_tuning(park_min=park_min)
    try:
        print(f"[parity] parked batch park_min {park_min}", flush=True)
        batch = V.solve_batch(insts, max_size=12)
    finally:
        V.set_tuning(**old)
    for k, (a, b) in enumerate(zip(plain, batch)):
        for f in FIELDS + ("level_winners", "level_candidates"):
            assert a[f] == b[f], (park_min, k, f, a[f], b[f])


def test_split_saves_work_at_found_level():
    """With the probe and operator blocks, a level settled by its first block
    evaluates fewer candidates than the fused pass; the answer is unchanged,
    and the levels before it are evaluated in full either way."""
    inst = V.Instance.generate(2, 3)          # its formula's root is ~ (first pass)
    old = V.set_tuning(pass_min=0)
    try:
        fused = V.solve(inst, max_size=14)
        V.set_tuning(pass_min=1)
        split = V.solve(inst, max_size=14)
    finally:
        V.set_tuning(**old)
    s = fused["min_size"]
    assert split["formula"] == fused["formula"] and split["distinct"] == fused["distinct"]
    assert split["level_candidates"][s] < fused["level_candidates"][s]
    assert split["level_candidates"][:s] == fused["level_candidates"][:s]


@pytest.mark.parametrize("rt", [1, 2])
def test_probe_with_colliding_relaxed_keys(rt):
    """Relaxed keys of types 1 and 2 sample bits, so distinct CS often share a
    key: the lowest matching candidate found by the probe may then lose its
    key to a lower ordinal, and the rest of the level has to be built.  The
    results still equal the oracle's."""
    old = V.set_tuning(pass_min=1)
    try:
        for k in range(8):
            inst = V.Instance.generate(2, k)
            names, lens, atoms, P = inst.arrays()
            print(f"[parity] probe relaxed type {rt} cfg2 #{k}", flush=True)
            want = O.enumerate_cs(O.Instance(names, lens, atoms, P), 8, rt)
            got = V.solve(inst, max_size=8, relaxed_type=rt)
            for f in FIELDS[:6]:
                assert got[f] == want[f], (rt, k, f, got[f], want[f])
    finally:
        V.set_tuning(**old)
