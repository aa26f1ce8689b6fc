"""GPU parity on traces longer than one 64-bit word (k_expand_mw, the
multi-word store and atom insertion), against the oracle: per-level counts,
found ordinal and formula bit for bit, under every key type, batched with
one-word instances, with the probe / first-pass split, a row cap, the cost
function and the BS stage.
"""
import random

import pytest

import oracle_lib as O
import vfb_amd as V
from ltl_semantics import separates

pytestmark = pytest.mark.gpu

FIELDS = ("status", "min_size", "formula", "distinct", "candidates", "found_ordinal", "otf_level", "otf_stored")


def _word(rng, n):
    kind = rng.randrange(3)
    if kind == 0:
        return rng.getrandbits(n)
    if kind == 1:                        # a long run of ones somewhere
        a = rng.randrange(n)
        b = rng.randint(a, n)
        return ((1 << b) - 1) ^ ((1 << a) - 1)
    return sum(1 << k for k in range(n) if rng.random() < 0.1)


def long_instance(seed, A=None, T=None, lmax=600, short=False):
    rng = random.Random(seed)
    A = A or rng.randint(2, 4)
    T = T or rng.randint(3, 16)
    edge = [64, 65, 127, 128, 129, 192, 193]
    lens = [rng.choice(edge) if rng.random() < 0.3 else rng.randint(1 if short else 30, lmax) for _ in range(T)]
    if max(lens) <= 64:
        lens[0] = 65 + rng.randrange(200)
    atoms = [[_word(rng, n) for n in lens] for _ in range(A)]
    P = rng.randint(1, T - 1)
    names = [f"q{a}" for a in range(A)]
    return V.Instance.from_bits(names, lens, atoms, P), O.Instance(names, lens, atoms, P)


def _check(got, want, ctx):
    for f in FIELDS:
        assert got[f] == want[f], (ctx, f, got[f], want[f])


@pytest.mark.parametrize("seed", range(10))
def test_multiword_parity_hashed(seed):
    inst, oi = long_instance(500 + seed)
    print(f"[parity] multiword hashed seed {seed}", flush=True)
    want = O.enumerate_cs(oi, 7, 3)
    got = V.solve(inst, max_size=7)
    _check(got, want, seed)
    if got["status"] == V.OK:
        names, lens, atoms, P = inst.arrays()
        assert separates(got["formula"], names, lens, atoms, P)


@pytest.mark.parametrize("rt", [1, 2])
def test_multiword_parity_sampled_keys(rt):
    for seed in range(5):
        inst, oi = long_instance(600 + seed)
        print(f"[parity] multiword type {rt} seed {seed}", flush=True)
        want = O.enumerate_cs(oi, 7, rt)
        got = V.solve(inst, max_size=7, relaxed_type=rt)
        _check(got, want, (rt, seed))


def test_multiword_parity_exact_key():
    """lenSum <= 126 with a trace over 64 steps: the exact key spans words."""
    for seed in range(6):
        rng = random.Random(700 + seed)
        lens = [rng.randint(65, 100)]
        while sum(lens) < 118:
            lens.append(rng.randint(1, 126 - sum(lens)))
        A = rng.randint(2, 3)
        atoms = [[rng.getrandbits(n) for n in lens] for _ in range(A)]
        P = rng.randint(1, len(lens) - 1)
        names = [f"e{a}" for a in range(A)]
        print(f"[parity] multiword exact seed {seed}", flush=True)
        want = O.enumerate_cs(O.Instance(names, lens, atoms, P), 8, 3)
        got = V.solve(V.Instance.from_bits(names, lens, atoms, P), max_size=8)
        _check(got, want, seed)


def test_multiword_large_levels():
    """Levels large enough for one lane per candidate (the wide mode covers
    the small ones); traces up to 16 words."""
    inst, oi = long_instance(811, A=4, T=24, lmax=1000)
    print("[parity] multiword large levels", flush=True)
    want = O.enumerate_cs(oi, 8, 3)
    got = V.solve(inst, max_size=8)
    _check(got, want, "large")


def test_multiword_batch_with_one_word_instances():
    insts = []
    for k in range(12):
        if k % 3 == 0:
            insts.append(V.Instance.generate(4, k))
        else:
            insts.append(long_instance(900 + k)[0])
    plain = [V.solve(i, max_size=7) for i in insts]
    print("[parity] multiword batch", flush=True)
    batch = V.solve_batch(insts, max_size=7)
    for k, (a, b) in enumerate(zip(plain, batch)):
        for f in FIELDS + ("level_winners",):
            assert a[f] == b[f], (k, f)


def test_multiword_probe_split_and_row_cap():
    old = V.set_tuning(pass_min=1)
    try:
        for seed in range(4):
            inst, oi = long_instance(1000 + seed)
            print(f"[parity] multiword probe seed {seed}", flush=True)
            _check(V.solve(inst, max_size=7), O.enumerate_cs(oi, 7, 3), ("probe", seed))
    finally:
        V.set_tuning(**old)
    for seed, rows in ((0, 40), (1, 300)):
        inst, oi = long_instance(1100 + seed)
        print(f"[parity] multiword row cap {rows}", flush=True)
        _check(V.solve(inst, max_size=8, max_rows=rows), O.enumerate_cs(oi, 8, 3, rows), ("rows", rows))


def test_multiword_costs_and_bs_stage():
    for seed in range(3):
        inst, oi = long_instance(1200 + seed, T=8)
        costs = [1, 2, 1, 2, 1, 3, 2, 3]
        print(f"[parity] multiword costs seed {seed}", flush=True)
        want = O.enumerate_costs(oi, 8, costs, 1 + seed % 2)
        got = V.solve(inst, max_size=8, costs=costs, neg_type=1 + seed % 2)
        for f in ("status", "min_size", "formula", "distinct", "candidates", "found_ordinal"):
            assert got[f] == want[f], (seed, f)
    for seed in range(3):
        inst, oi = long_instance(1300 + seed, T=10)
        print(f"[parity] multiword bs seed {seed}", flush=True)
        want = O.enumerate_bs(oi, 3, 9)
        got = V.solve(inst, max_size=3, bs_max_size=9)
        for f in ("status", "min_size", "formula", "distinct", "candidates", "found_ordinal", "bs_stage"):
            assert got[f] == want[f], (seed, f)


def test_multiword_brute_search_rejected():
    inst, _ = long_instance(1400)
    r = V.solve(inst, max_size=4, search=V.SEARCH_BRUTE)
    assert r["status"] == V.INVALID_INPUT


def test_multiword_parked_batch_and_server():
    """Multi-word instances in a batch that parks every member (park_min=1)
    and through the server: each result equals the plain solve."""
    insts = [long_instance(1500 + k)[0] for k in range(6)] + [V.Instance.generate(4, k) for k in range(4)]
    plain = [V.solve(i, max_size=7) for i in insts]
    old = V.set_tuning(park_min=1)
    try:
        print("[parity] multiword parked batch", flush=True)
        batch = V.solve_batch(insts, max_size=7)
    finally:
        V.set_tuning(**old)
    print("[parity] multiword server", flush=True)
    with V.Server(max_size=7, max_active=4) as srv:
        got = [srv.wait(t) for t in [srv.submit(i) for i in insts]]
    for k in range(len(insts)):
        for f in FIELDS + ("level_winners",):
            assert batch[k][f] == plain[k][f], ("batch", k, f)
            assert got[k][f] == plain[k][f], ("server", k, f)
