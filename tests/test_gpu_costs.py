"""GPU parity of LTL_BS's cost function and NegType (vfb_options.cost /
neg_type; LTL_BS.cu:1085-1493) against the oracle's restatement
(ora_enumerate_costs), and LTL_BS's command line end to end."""
import os
import random
import subprocess

import pytest

import oracle_lib as O
import vfb_amd as V
from test_host import _read_text_py
from ltl_semantics import separates

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ("status", "min_size", "formula", "distinct", "candidates", "found_ordinal", "top_op")


def _check(inst, max_cost, costs, neg_type, label):
    names, lens, atoms, P = inst.arrays()
    print(f"[parity] costs {label} {costs} neg {neg_type} max {max_cost}", flush=True)
    want = O.enumerate_costs(O.Instance(names, lens, atoms, P), max_cost, costs, neg_type)
    got = V.solve(inst, max_size=max_cost, costs=costs, neg_type=neg_type)
    for f in FIELDS:
        assert got[f] == want[f], (label, f, got[f], want[f])
    if got["status"] == V.OK:
        assert separates(got["formula"], names, lens, atoms, P)
    return got


@pytest.mark.parametrize("seed", range(24))
def test_costs_random(seed):
    rng = random.Random(4400 + seed)
    A = rng.randint(1, 3)
    T = rng.randint(2, 16)
    lens = [rng.randint(1, 64 if seed % 5 == 0 else 10) for _ in range(T)]
    atoms = [[rng.randrange(1 << lens[t]) for t in range(T)] for _ in range(A)]
    inst = V.Instance.from_bits([f"p{a}" for a in range(A)], lens, atoms, rng.randint(1, T - 1))
    costs = [rng.randint(1, 3) for _ in range(8)]
    _check(inst, 10, costs, 1 + seed % 2, f"random {seed}")


@pytest.mark.parametrize("k", range(4))
def test_costs_config2(k):
    rng = random.Random(k)
    costs = [1, rng.randint(1, 2), rng.randint(1, 2), rng.randint(1, 2), 1, 2, 2, rng.randint(1, 3)]
    _check(V.Instance.generate(2, k), 9, costs, 1, f"cfg2 #{k}")


def test_unit_costs_and_negtype1_equal_default():
    inst = V.Instance.generate(1, 3)
    a = V.solve(inst, max_size=9)
    b = V.solve(inst, max_size=9, costs=[1] * 8, neg_type=1)
    for f in FIELDS + ("level_winners",):
        assert a[f] == b[f], f


def test_costs_batch_equals_plain():
    rng = random.Random(77)
    costs = [1, 2, 1, 2, 1, 1, 2, 2]
    insts = [V.Instance.generate(4, k) for k in range(16)]
    plain = [V.solve(i, max_size=11, costs=costs, neg_type=2) for i in insts]
    batch = V.solve_batch(insts, max_size=11, costs=costs, neg_type=2)
    for k, (a, b) in enumerate(zip(plain, batch)):
        for f in FIELDS + ("level_winners",):
            assert a[f] == b[f], (k, f, a[f], b[f])


def test_ltlbs_cli_end_to_end():
    path = os.path.join(GOLDEN, "example_ltlbs.txt")
    env = dict(os.environ, VFB_QUIET="1")
    out = subprocess.run([V.CLI_PATH, path] + ["1"] * 8 + ["10", "12", "3", "1"], capture_output=True, text=True,
                         env=env, timeout=120)
    assert out.returncode == 0, out.stderr
    # the oracle on the traces as the Python restatement of readFile reads them
    names, lens, atoms, P = _read_text_py(path)
    want = O.enumerate_bs(O.Instance(names, lens, atoms, P), 10, 12, 3)["formula"]
    for op in "&|U":
        want = want.replace(f") {op} (", f"){op}(")
    assert f'LTL: "{want}"' in out.stdout, out.stdout
    assert "Number of Traces: 5" in out.stdout and "Cost Function: p:1" in out.stdout
