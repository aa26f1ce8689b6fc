"""GPU parity of ltl_nocache's syntactic brute-force search
(vfb_options.search = VFB_SEARCH_BRUTE) against the oracle's restatement
(ora_bruteforce: generateMatrix / numberToFormula / evaluateRPN order,
ltl_nocache.cu:78-287, 319-446): the same first separating formula in
numbering order, the same formula text, the same number of formulas checked
per size.  Its minimal size also equals the CS enumeration's (both are
exhaustive over the same grammar)."""
import os
import random

import pytest

import oracle_lib as O
import vfb_amd as V
from ltl_semantics import separates

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ("status", "min_size", "formula", "candidates", "found_ordinal", "top_op")


def _check(inst, max_size, label):
    names, lens, atoms, P = inst.arrays()
    print(f"[parity] brute {label} max {max_size}", flush=True)
    want = O.bruteforce(O.Instance(names, lens, atoms, P), max_size)
    got = V.solve(inst, max_size=max_size, search=V.SEARCH_BRUTE)
    for f in FIELDS:
        if f == "found_ordinal" and want["min_size"] == 1:
            continue
        assert got[f] == want[f], (label, f, got[f], want[f])
    assert got["distinct"] == [0] * len(got["distinct"])
    if got["status"] == V.OK:
        assert separates(got["formula"], names, lens, atoms, P)
        cs = V.solve(inst, max_size=max_size)
        assert cs["min_size"] == got["min_size"], (label, cs["min_size"], got["min_size"])
    return got


def test_brute_example():
    inst = V.Instance.from_json(os.path.join(GOLDEN, "ExampleLTLF.json"))
    got = _check(inst, 10, "ExampleLTLF")
    assert got["min_size"] == 10


@pytest.mark.parametrize("k", range(4))
def test_brute_config1(k):
    _check(V.Instance.generate(1, k), 6, f"cfg1 #{k}")


@pytest.mark.parametrize("seed", range(16))
def test_brute_random(seed):
    rng = random.Random(3000 + seed)
    A = rng.randint(1, 3)
    T = rng.randint(2, 10)
    lens = [rng.randint(1, 64 if seed % 4 == 0 else 8) for _ in range(T)]
    atoms = [[sum((rng.random() < 0.5) << j for j in range(lens[t])) for t in range(T)] for _ in range(A)]
    inst = V.Instance.from_bits([f"p{a}" for a in range(A)], lens, atoms, rng.randint(1, T - 1))
    _check(inst, 6, f"random {seed}")


def test_brute_not_found_counts_every_formula():
    """Identical traces on both sides: nothing separates, every formula of
    every size is checked (closed-form counts)."""
    inst = V.Instance.from_bits(["a", "b"], [5, 5], [[0b10110, 0b10110], [0b01, 0b01]], 1)
    got = V.solve(inst, max_size=6, search=V.SEARCH_BRUTE)
    nf = O.num_formulas(2, 6)
    assert got["status"] == V.NOT_FOUND and got["formula"] == "Not found !"
    assert got["candidates"][1:7] == [nf[s][7] for s in range(1, 7)]
