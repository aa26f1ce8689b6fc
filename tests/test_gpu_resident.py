"""GPU tests of resident instances (vfb_instance_upload): a plain solve that
reads the atom rows from the instance's device copy returns what the solve
that uploads them returns, and what the oracle returns, bit for bit; the copy
is reused across solves, ignored by batches and the server, and freed with
the instance.
"""
import gc

import pytest

import oracle_lib as O
import vfb_amd as V
from test_gpu_multiword import long_instance

pytestmark = pytest.mark.gpu

FIELDS = ("status", "min_size", "formula", "distinct", "candidates", "level_winners", "level_candidates",
          "found_ordinal", "otf_level", "otf_stored")


def _same(a, b, ctx):
    for f in FIELDS:
        assert a[f] == b[f], (ctx, f, a[f], b[f])


@pytest.mark.parametrize("cfg,k,ms", [(1, 0, 10), (1, 3, 10), (2, 0, 14), (2, 7, 14), (3, 4, 18)])
def test_resident_equals_uploaded_and_oracle(cfg, k, ms):
    print(f"[resident] cfg{cfg} #{k} max_size={ms}", flush=True)
    plain = V.solve(V.Instance.generate(cfg, k), max_size=ms)
    inst = V.Instance.generate(cfg, k).upload()
    got = V.solve(inst, max_size=ms)
    _same(got, plain, (cfg, k))
    if cfg == 1:
        names, lens, atoms, P = inst.arrays()
        want = O.enumerate_cs(O.Instance(names, lens, atoms, P), ms, 3)
        for f in ("status", "min_size", "formula", "distinct", "candidates", "found_ordinal"):
            assert got[f] == want[f], (f, got[f], want[f])


def test_resident_reused_with_other_options():
    inst = V.Instance.generate(2, 3).upload()
    assert inst.upload() is inst                      # a second upload is a no-op
    ref = V.Instance.generate(2, 3)
    for kw in (dict(max_size=6), dict(max_size=14), dict(max_size=9, relaxed_type=1),
               dict(max_size=9, relaxed_type=2), dict(max_size=10, max_rows=3000)):
        print(f"[resident] reuse {kw}", flush=True)
        _same(V.solve(inst, **kw), V.solve(ref, **kw), kw)


def test_resident_multiword():
    for seed in range(3):
        print(f"[resident] multi-word seed {seed}", flush=True)
        inst, oi = long_instance(seed)
        want = O.enumerate_cs(oi, 7, 3)
        got = V.solve(inst.upload(), max_size=7)
        for f in ("status", "min_size", "formula", "distinct", "candidates", "found_ordinal"):
            assert got[f] == want[f], (seed, f, got[f], want[f])


def test_resident_in_batch_and_server():
    insts = [V.Instance.generate(4, k) for k in range(24)]
    plain = [V.solve(i, max_size=8) for i in insts]
    for i in insts[::2]:
        i.upload()
    print("[resident] batch of 24, half uploaded", flush=True)
    for k, (a, b) in enumerate(zip(plain, V.solve_batch(insts, max_size=8))):
        _same(a, b, ("batch", k))
    print("[resident] batch of one uploaded instance", flush=True)
    _same(V.solve_batch([insts[0]], max_size=8)[0], plain[0], "batch-1")
    print("[resident] server", flush=True)
    with V.Server(max_size=8, max_active=8) as srv:
        got = [srv.wait(t) for t in [srv.submit(i) for i in insts]]
    for k, (a, b) in enumerate(zip(plain, got)):
        for f in ("status", "min_size", "formula", "distinct", "candidates"):
            assert a[f] == b[f], ("server", k, f, a[f], b[f])


def test_resident_freed_with_instance():
    for k in range(8):
        inst = V.Instance.generate(2, k).upload()
        V.solve(inst, max_size=8)
        del inst
        gc.collect()
    fresh = V.Instance.generate(2, 9)
    _same(V.solve(fresh, max_size=10), V.solve(V.Instance.generate(2, 9).upload(), max_size=10), "after frees")
