"""GPU parity on instances whose hashed half-keys are -1 and -2.

The reference merges the two invalid half-keys (hash_set.cuh:198-204 returns
0 for both, LTL_BS.cu:262-265), so ~b is a duplicate of ~a only through that
merge (tests/fold_instance.py builds the rows by inverting SplitMix64).  The
kernels fold -1 onto -2 (key_fin); here they must agree with the oracle's
restatement on counts, formula and ordinal -- plain, batched with other
instances, and served -- and keep one distinct CS fewer at level 2 than a
full-row dedup.
"""
import pytest

import oracle_lib as O
import vfb_amd as V
from fold_instance import fold_instance
from ltl_semantics import separates

pytestmark = pytest.mark.gpu

FIELDS = ("status", "min_size", "formula", "distinct", "candidates", "found_ordinal")


def _inst(half):
    names, lens, atoms, P, _ = fold_instance(half)
    return V.Instance.from_bits(names, lens, atoms, P), O.Instance(names, lens, atoms, P)


@pytest.mark.parametrize("half", ["hi", "lo"])
@pytest.mark.parametrize("max_size", [2, 3, 8])
def test_fold_plain_equals_oracle(half, max_size):
    print(f"[fold] plain half={half} max_size={max_size}", flush=True)
    inst, oi = _inst(half)
    want = O.enumerate_cs(oi, max_size, 3)
    got = V.solve(inst, max_size=max_size)
    for f in FIELDS:
        assert got[f] == want[f], (f, got[f], want[f])
    full = O.enumerate_cs(oi, 2, O.KEY_FULLROW)
    assert got["distinct"][2] == full["distinct"][2] - 1
    if got["status"] == V.OK:
        names, lens, atoms, P, _ = fold_instance(half)
        assert separates(got["formula"], names, lens, atoms, P)


def test_fold_batched_and_served_equal_plain():
    print("[fold] batched + served", flush=True)
    folds = [_inst("hi")[0], _inst("lo")[0]]
    others = [V.Instance.generate(4, k) for k in range(6)]
    insts = [others[0], folds[0], others[1], others[2], folds[1]] + others[3:]
    plain = [V.solve(i, max_size=8) for i in insts]
    batch = V.solve_batch(insts, max_size=8)
    for k, (a, b) in enumerate(zip(plain, batch)):
        for f in FIELDS + ("level_winners",):
            assert a[f] == b[f], (k, f, a[f], b[f])
    with V.Server(max_size=8, max_active=4) as srv:
        tickets = [srv.submit(i) for i in insts]
        served = [srv.wait(t) for t in tickets]
    for k, (a, b) in enumerate(zip(plain, served)):
        for f in FIELDS + ("level_winners",):
            assert a[f] == b[f], (k, f, a[f], b[f])
