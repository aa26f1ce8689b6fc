"""GPU parity: the HIP path (libvfb.so, through its C-ABI) against the oracle.

Bar: bit-exact -- same status, minimal size, returned formula text, and
per-level distinct / candidate counts -- on seeded instances at sizes the
oracle finishes in seconds.  Every returned formula is also re-verified on
every trace with the textbook per-position semantics.
"""
import os
import random

import pytest

import oracle_lib as O
import vfb_amd as V
from ltl_semantics import separates, parse, size

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _oracle_for(inst, max_size, relaxed_type=3):
    names, lens, atoms, P = inst.arrays()
    return O.enumerate_cs(O.Instance(names, lens, atoms, P), max_size, relaxed_type), (names, lens, atoms, P)


def _check(inst, max_size, relaxed_type=3, label=""):
    print(f"[parity] {label} max_size={max_size} relaxed={relaxed_type} shape={inst.shape()}", flush=True)
    want, (names, lens, atoms, P) = _oracle_for(inst, max_size, relaxed_type)
    got = V.solve(inst, max_size=max_size, relaxed_type=relaxed_type)
    assert got["status"] == want["status"], (label, got["status"], want["status"])
    assert got["min_size"] == want["min_size"], (label, got["min_size"], want["min_size"])
    assert got["formula"] == want["formula"], (label, got["formula"], want["formula"])
    assert got["distinct"] == want["distinct"], (label, got["distinct"], want["distinct"])
    assert got["candidates"] == want["candidates"], (label, got["candidates"], want["candidates"])
    if got["status"] == V.OK:
        assert got["found_ordinal"] == want["found_ordinal"]
        assert separates(got["formula"], names, lens, atoms, P), got["formula"]
        assert size(parse(got["formula"], names)) == got["min_size"]
    return got


def test_example_json():
    inst = V.Instance.from_json(os.path.join(GOLDEN, "ExampleLTLF.json"))
    got = _check(inst, 10, label="ExampleLTLF")
    assert got["min_size"] == 10


@pytest.mark.parametrize("k", range(6))
def test_config1(k):
    _check(V.Instance.generate(1, k), 10, label=f"cfg1 #{k}")


@pytest.mark.parametrize("k", range(6))
def test_config2_relaxed3(k):
    _check(V.Instance.generate(2, k), 7, 3, label=f"cfg2 #{k}")


@pytest.mark.parametrize("rt", [1, 2, 3])
def test_relaxed_types(rt):
    _check(V.Instance.generate(2, 2), 7, rt, label=f"cfg2 #2 type {rt}")


@pytest.mark.parametrize("k", range(4))
def test_config3(k):
    _check(V.Instance.generate(3, k), 6, 3, label=f"cfg3 #{k}")


def _random_instance(rng, A, T, lmin, lmax, p=0.5):
    lens = [rng.randint(lmin, lmax) for _ in range(T)]
    atoms = [[sum((rng.random() < p) << k for k in range(lens[t])) for t in range(T)] for _ in range(A)]
    P = rng.randint(0, T)
    return V.Instance.from_bits([f"p{a}" for a in range(A)], lens, atoms, P)


@pytest.mark.parametrize("seed", range(12))
def test_random_small(seed):
    rng = random.Random(1000 + seed)
    A = rng.randint(1, 4)
    T = rng.randint(1, 16)
    inst = _random_instance(rng, A, T, 1, 12)
    _check(inst, 7, 3, label=f"random seed {seed}")


@pytest.mark.parametrize("seed", range(4))
def test_length_64_and_many_traces(seed):
    rng = random.Random(77 + seed)
    inst = _random_instance(rng, 3, 128 if seed % 2 else 64, 60, 64, 0.4)
    _check(inst, 5, 3, label=f"long traces {seed}")


def test_batch_equals_plain():
    insts = [V.Instance.generate(4, k) for k in range(24)]
    plain = [V.solve(i, max_size=8) for i in insts]
    batch = V.solve_batch(insts, max_size=8)
    for k, (a, b) in enumerate(zip(plain, batch)):
        for f in ("status", "min_size", "formula", "distinct", "candidates", "level_winners"):
            assert a[f] == b[f], (k, f, a[f], b[f])


@pytest.mark.parametrize("cfg,k,max_rows", [(2, 2, 3000), (2, 2, 20000), (2, 3, 100000), (2, 21, 500),
                                            (1, 0, 40), (3, 4, 5000), (2, 5, 130), (1, 1, 60), (1, 2, 200),
                                            (4, 7, 90), (4, 14, 400), (4, 20, 1500), (3, 9, 300)])
def test_overflow_onthefly(cfg, k, max_rows):
    """Row cap (the reference's cache capacity): the overflowing level is stored
    up to the cap in ordinal order, the next level is check-only."""
    inst = V.Instance.generate(cfg, k)
    names, lens, atoms, P = inst.arrays()
    ms = 10
    print(f"[parity] overflow cfg{cfg} #{k} max_rows={max_rows}", flush=True)
    want = O.enumerate_cs(O.Instance(names, lens, atoms, P), ms, 3, max_rows)
    got = V.solve(inst, max_size=ms, max_rows=max_rows)
    for f in ("status", "min_size", "formula", "distinct", "candidates", "otf_level", "otf_stored"):
        assert got[f] == want[f], (f, got[f], want[f])
    if got["status"] == V.OK:
        assert separates(got["formula"], names, lens, atoms, P)


def test_batch_512_mix_equals_plain():
    insts = [V.Instance.generate(4, k) for k in range(512)]
    plain = [V.solve(i, max_size=16) for i in insts]
    batch = V.solve_batch(insts, max_size=16)
    bad = [k for k, (a, b) in enumerate(zip(plain, batch))
           if any(a[f] != b[f] for f in ("status", "min_size", "formula", "distinct", "candidates",
                                         "level_winners", "otf_level", "otf_stored"))]
    assert not bad, [(k, plain[k]["status"], batch[k]["status"]) for k in bad[:10]]


def test_server_results_equal_plain():
    """Requests arriving while others run are admitted as slots free up; each
    result equals the plain solve of the same instance."""
    import time
    insts = [V.Instance.generate(5, k) for k in range(96)]
    plain = [V.solve(i, max_size=16) for i in insts]
    with V.Server(max_size=16, max_active=16) as srv:
        tickets = []
        for k, inst in enumerate(insts):
            tickets.append(srv.submit(inst))
            if k % 8 == 7:
                time.sleep(0.002)        # arrivals spread over time
        got = [srv.wait(t) for t in tickets]
        stats = srv.stats()
    assert stats["peak_active"] <= 16
    for k, (a, b) in enumerate(zip(plain, got)):
        for f in ("status", "min_size", "formula", "distinct", "candidates", "level_winners",
                  "otf_level", "otf_stored"):
            assert a[f] == b[f], (k, f, a[f], b[f])
        assert b["ms_total"] > 0


@pytest.mark.parametrize("seed", range(16))
def test_overflow_random(seed):
    """Check-only levels over random instances of every length, incl. 64."""
    rng = random.Random(500 + seed)
    inst = _random_instance(rng, rng.randint(1, 4), rng.randint(2, 20), 1, 64 if seed % 3 == 0 else 12)
    names, lens, atoms, P = inst.arrays()
    cap = rng.randint(8, 200)
    print(f"[parity] overflow random seed {seed} cap {cap}", flush=True)
    want = O.enumerate_cs(O.Instance(names, lens, atoms, P), 8, 3, cap)
    got = V.solve(inst, max_size=8, max_rows=cap)
    for f in ("status", "min_size", "formula", "distinct", "candidates", "otf_level", "otf_stored"):
        assert got[f] == want[f], (f, got[f], want[f])
