"""GPU parity at the limits of the input space, against the oracle: the most
traces and atoms, no positive or no negative trace, a single trace, traces of
4096 steps, traces of one step, and an instance with no separating formula
searched to the largest size."""
import random

import pytest

import oracle_lib as O
import vfb_amd as V

pytestmark = pytest.mark.gpu

FIELDS = ("status", "min_size", "formula", "distinct", "candidates", "found_ordinal", "otf_level", "otf_stored")


def _inst(seed, A, lens, P, p=0.5):
    rng = random.Random(seed)
    atoms = [[sum(1 << k for k in range(n) if rng.random() < p) for n in lens] for _ in range(A)]
    names = [f"v{a}" for a in range(A)]
    return V.Instance.from_bits(names, lens, atoms, P), O.Instance(names, lens, atoms, P)


def _check(inst, oi, ms, ctx, rt=3):
    print(f"[parity] edge {ctx}", flush=True)
    want = O.enumerate_cs(oi, ms, rt)
    got = V.solve(inst, max_size=ms, relaxed_type=rt)
    for f in FIELDS:
        assert got[f] == want[f], (ctx, f, got[f], want[f])
    return got


def test_most_traces():
    rng = random.Random(1)
    lens = [rng.randint(1, 64) for _ in range(V.MAX_TRACES)]
    inst, oi = _inst(11, 3, lens, 64)
    _check(inst, oi, 6, "128 traces")


def test_most_atoms():
    # every atom holds at step 0 of every trace, so no atom separates and the
    # search builds the 64-atom levels
    rng = random.Random(12)
    lens = [5, 9, 3, 7, 2, 6]
    atoms = [[1 | sum(1 << k for k in range(1, n) if rng.random() < 0.3) for n in lens] for _ in range(V.MAX_ATOMS)]
    names = [f"v{a}" for a in range(V.MAX_ATOMS)]
    got = _check(V.Instance.from_bits(names, lens, atoms, 3), O.Instance(names, lens, atoms, 3), 3, "64 atoms")
    assert got["candidates"][1] == V.MAX_ATOMS


@pytest.mark.parametrize("P", [0, 7])
def test_no_negatives_or_no_positives(P):
    inst, oi = _inst(13 + P, 2, [6, 4, 8, 5, 3, 7, 2], P)
    _check(inst, oi, 6, f"P={P} of 7")


def test_single_trace():
    inst, oi = _inst(14, 2, [9], 1)
    _check(inst, oi, 5, "one trace")


def test_traces_of_4096_steps():
    inst, oi = _inst(15, 2, [4096, 3000, 4096], 2, p=0.02)
    _check(inst, oi, 6, "4096 steps")


def test_traces_of_one_step():
    for rt in (1, 2, 3):
        inst, oi = _inst(16, 3, [1] * 40, 20)
        _check(inst, oi, 6, f"one-step traces, type {rt}", rt)


def test_no_formula_to_the_largest_size():
    """The same trace is positive and negative: no formula separates them, and
    the search runs to size 32 (its levels stay tiny: every CS is a pair of
    equal words)."""
    w = 0b1011001
    inst = V.Instance.from_bits(["a", "b"], [7, 7], [[w, w], [w >> 1, w >> 1]], 1)
    oi = O.Instance(["a", "b"], [7, 7], [[w, w], [w >> 1, w >> 1]], 1)
    got = _check(inst, oi, 32, "no formula, size 32")
    assert got["status"] == V.NOT_FOUND and got["levels_done"] == 32
