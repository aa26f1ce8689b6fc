"""GPU parity at the configurations' own sizes, against the OpenMP build of
the enumeration (oracle/cpu_omp.c, pinned to the sequential oracle in
tests/test_oracle.py).

- Config 3 (10 APs, 63 traces of 48-64 steps, hashed keys) instance 4 at
  max_formula_size 18: levels of up to ~10^8 candidates, the probe/first-pass
  split of large levels, table growth by rehash, and relaxed-key collisions
  among ~10^7 keys.  Once with the default growing set and once with the set
  started at 2^28 slots (8 GiB), the reference's capacity regime
  (LTL_BS.cu:1179-1214); results must not depend on capacity.
- Config 2 (5 APs, 40 traces of 16-32 steps) instances 0..31 at
  max_formula_size 14, the dedup-stress workload.
- Config-4 mix instances 219 and 456 at max_formula_size 16 (the batch's
  max size): levels of 10^7-10^8 candidates with probes, found levels built
  up to the probe's match only, target checks over 8 left rows per wavefront
  and diagonal splits without their repeated pairs.

Bit-exact: status, minimal size, formula text, found ordinal and the
per-level distinct / candidate counts.
"""
import pytest

import oracle_lib as O
import vfb_amd as V
from ltl_semantics import separates

pytestmark = pytest.mark.gpu

FIELDS = ("status", "min_size", "formula", "distinct", "candidates", "found_ordinal")


def _omp(inst, max_size):
    names, lens, atoms, P = inst.arrays()
    r, dt = O.enumerate_omp(O.Instance(names, lens, atoms, P), max_size, 3)
    return r, (names, lens, atoms, P)


@pytest.mark.parametrize("start_slots", [0, 1 << 28])
def test_config3_instance4_max18(start_slots):
    print(f"[scale] config-3 #4 max 18, set start {start_slots}", flush=True)
    inst = V.Instance.generate(3, 4)
    want, arrays = _omp(inst, 18)
    old = V.set_tuning(table_min_slots=start_slots) if start_slots else None
    try:
        got = V.solve(inst, max_size=18)
    finally:
        if old is not None:
            V.set_tuning(**old)
    print(f"[scale] table_slots {got['table_slots']} distinct {sum(got['distinct'])}", flush=True)
    for f in FIELDS:
        assert got[f] == want[f], (f, got[f], want[f])
    assert got["status"] == V.OK and separates(got["formula"], *arrays)
    assert sum(got["distinct"]) > 10_000_000
    if start_slots:
        assert got["table_slots"] >= 1 << 28
    else:
        assert got["table_slots"] >= 1 << 24   # grown by rehash from 2^17


def test_config2_instances_max14():
    insts = [V.Instance.generate(2, k) for k in range(32)]
    for k, inst in enumerate(insts):
        print(f"[scale] config-2 #{k} max 14", flush=True)
        want, arrays = _omp(inst, 14)
        got = V.solve(inst, max_size=14)
        for f in FIELDS:
            assert got[f] == want[f], (k, f, got[f], want[f])
        if got["status"] == V.OK:
            assert separates(got["formula"], *arrays)


@pytest.mark.parametrize("k", [219, 456])
def test_config4_large_instances_max16(k):
    print(f"[scale] config-4 #{k} max 16", flush=True)
    inst = V.Instance.generate(4, k)
    want, arrays = _omp(inst, 16)
    got = V.solve(inst, max_size=16)
    for f in FIELDS:
        assert got[f] == want[f], (f, got[f], want[f])
    assert got["status"] == V.OK and separates(got["formula"], *arrays)
    assert max(got["level_candidates"]) > 10_000_000
    assert got["probe_candidates"] > 0      # the found level was probed first
