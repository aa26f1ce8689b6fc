"""GPU parity of the BS stage (vfb_options.bs_max_size; LTL_BS.cu:808-1067,
1508-1525) against the oracle's restatement (ora_enumerate_bs): when the CS
search finds nothing up to max_size, Boolean combinations (~ & |) of the
time-0 vectors of every stored CS, level by level up to bs_max_size.  Same
status, size, formula text, per-level counts (CS levels and BS levels), and
the formula re-verified on every trace with the textbook semantics."""
import random

import pytest

import oracle_lib as O
import vfb_amd as V
from ltl_semantics import parse, separates, size

pytestmark = pytest.mark.gpu

FIELDS = ("status", "min_size", "formula", "distinct", "candidates", "found_ordinal", "bs_stage", "top_op")


def _check(inst, max_size, bs_max, label):
    names, lens, atoms, P = inst.arrays()
    print(f"[parity] bs {label} max {max_size} bs_max {bs_max}", flush=True)
    want = O.enumerate_bs(O.Instance(names, lens, atoms, P), max_size, bs_max)
    got = V.solve(inst, max_size=max_size, bs_max_size=bs_max)
    for f in FIELDS:
        assert got[f] == want[f], (label, f, got[f], want[f])
    if got["status"] == V.OK:
        assert separates(got["formula"], names, lens, atoms, P)
        assert size(parse(got["formula"], names)) == got["min_size"]
    return got


def _random(seed, A, T, lmax):
    rng = random.Random(seed)
    lens = [rng.randint(1, lmax) for _ in range(T)]
    atoms = [[rng.randrange(1 << lens[t]) for t in range(T)] for _ in range(A)]
    return V.Instance.from_bits([f"p{a}" for a in range(A)], lens, atoms, rng.randint(1, T - 1))


@pytest.mark.parametrize("seed", range(24))
def test_bs_random_tiny(seed):
    rng = random.Random(7100 + seed)
    inst = _random(8100 + seed, rng.randint(1, 2), rng.randint(3, 6), 4)
    _check(inst, 2, 7, f"tiny {seed}")


@pytest.mark.parametrize("seed", range(12))
def test_bs_random_larger(seed):
    rng = random.Random(900 + seed)
    inst = _random(9100 + seed, rng.randint(2, 3), rng.randint(8, 40), 12)
    got = _check(inst, 3, 8, f"larger {seed}")
    print("   ", got["status"], got["bs_stage"], got["min_size"], flush=True)


@pytest.mark.parametrize("k", range(4))
def test_bs_config1_low_max(k):
    _check(V.Instance.generate(1, k), 3, 9, f"cfg1 #{k}")


def test_bs_batch_equals_plain():
    insts = [_random(11000 + k, 2, 10 + k % 20, 10) for k in range(24)]
    plain = [V.solve(i, max_size=3, bs_max_size=8) for i in insts]
    batch = V.solve_batch(insts, max_size=3, bs_max_size=8)
    for k, (a, b) in enumerate(zip(plain, batch)):
        for f in FIELDS + ("level_winners",):
            assert a[f] == b[f], (k, f, a[f], b[f])
    assert any(r["bs_stage"] for r in plain)


def test_bs_off_by_default_and_for_server():
    inst = _random(12000, 2, 6, 4)
    assert V.solve(inst, max_size=2)["bs_stage"] == 0
    with pytest.raises(V.VfbError):
        V.Server(max_size=4, bs_max_size=6)
