"""The persistent level chains against the level-by-level host path.

k_chain: a solve's small levels are built inside one launch, by a team of
workgroups per instance (a single workgroup in large batches, up to 64 for a
plain solve; off by default, vfb_tuning.chain_blocks = 0).  k_chain_lds: the
smallest levels, by one workgroup per instance with the solve's whole hash
set in LDS (on by default, vfb_tuning.lds_chain = 1); every other GPU test
runs with it.  Every result must equal
the host path's field by field -- counts per level, work counters of every completed level,
formula, ordinal -- for every team size and level limit, for each key kind
(exact, sampled types 1 and 2, hashed type 3), under row caps, with NegType 2,
plain and batched; the other GPU tests compare the host path with the oracle.
"""
import pytest

import vfb_amd as V

pytestmark = pytest.mark.gpu

FIELDS = ("status", "min_size", "formula", "distinct", "candidates", "level_winners", "level_candidates",
          "otf_level", "otf_stored", "found_ordinal", "levels_done", "top_op", "table_slots")


@pytest.fixture
def tuning():
    old = V.get_tuning()
    yield
    V.set_tuning(**old)


def _same(got, want, ctx):
    for f in FIELDS:
        assert got[f] == want[f], (ctx, f, got[f], want[f])


def _insts():
    # config 1 (exact keys), config 2 (hashed keys at 40 traces), of mixed depth
    return [(1, k) for k in range(4)] + [(2, k) for k in (0, 1, 3, 4, 5, 6, 7)]


@pytest.mark.parametrize("blocks,cands", [(64, 1024), (1, 1024), (2, 64), (8, 4096), (64, 65536), (3, 100)])
def test_chain_plain_equals_host(tuning, blocks, cands):
    print(f"[chain] plain blocks={blocks} cands={cands}", flush=True)
    insts = [(c, k, V.Instance.generate(c, k)) for c, k in _insts()]
    V.set_tuning(chain_blocks=0, lds_chain=0)
    want = [V.solve(i, max_size=10) for _, _, i in insts]
    assert all(w["chain_levels"] == 0 for w in want)
    V.set_tuning(chain_blocks=blocks, chain_cands=cands, lds_chain=0)
    before = V.kernel_stats()
    got = [V.solve(i, max_size=10) for _, _, i in insts]
    after = V.kernel_stats()
    for (c, k, _), g, w in zip(insts, got, want):
        _same(g, w, (c, k))
    assert sum(g["chain_levels"] for g in got) > 0
    assert after["chain_launches"] > before["chain_launches"]
    assert after["chain_levels"] - before["chain_levels"] == sum(g["chain_levels"] for g in got)


@pytest.mark.parametrize("relaxed_type", [1, 2, 3])
def test_chain_relaxed_keys_equal_host(tuning, relaxed_type):
    print(f"[chain] relaxed type {relaxed_type}", flush=True)
    insts = [V.Instance.generate(2, k) for k in (0, 1, 4, 7)]
    V.set_tuning(chain_blocks=0, lds_chain=0)
    want = [V.solve(i, max_size=9, relaxed_type=relaxed_type) for i in insts]
    V.set_tuning(chain_blocks=64, lds_chain=0)
    got = [V.solve(i, max_size=9, relaxed_type=relaxed_type) for i in insts]
    for k, (g, w) in enumerate(zip(got, want)):
        _same(g, w, (relaxed_type, k))
    assert sum(g["chain_levels"] for g in got) > 0


@pytest.mark.parametrize("max_rows", [30, 500, 3000, 20000])
def test_chain_row_cap_equals_host(tuning, max_rows):
    """A cap inside the chain's levels: the chain leaves the capped level to
    the host (its onTheFly round), which must end alike."""
    print(f"[chain] max_rows={max_rows}", flush=True)
    insts = [V.Instance.generate(2, k) for k in (2, 3, 5)]
    V.set_tuning(chain_blocks=0, lds_chain=0)
    want = [V.solve(i, max_size=10, max_rows=max_rows) for i in insts]
    V.set_tuning(chain_blocks=64, lds_chain=0)
    got = [V.solve(i, max_size=10, max_rows=max_rows) for i in insts]
    for k, (g, w) in enumerate(zip(got, want)):
        _same(g, w, (max_rows, k))


def test_chain_neg_type_2_equals_host(tuning):
    print("[chain] NegType 2", flush=True)
    insts = [V.Instance.generate(2, k) for k in (0, 1, 4)]
    V.set_tuning(chain_blocks=0, lds_chain=0)
    want = [V.solve(i, max_size=8, neg_type=2) for i in insts]
    V.set_tuning(chain_blocks=64, lds_chain=0)
    got = [V.solve(i, max_size=8, neg_type=2) for i in insts]
    for k, (g, w) in enumerate(zip(got, want)):
        _same(g, w, ("neg2", k))
    assert sum(g["chain_levels"] for g in got) > 0


@pytest.mark.parametrize("blocks", [64, 1])
def test_chain_batch_equals_host_plain(tuning, blocks):
    """A batch (teams of several workgroups for 16 members, one each with
    blocks = 1) against plain solves without the chain."""
    print(f"[chain] batch blocks={blocks}", flush=True)
    insts = [V.Instance.generate(4, k) for k in range(12)] + [V.Instance.generate(2, k) for k in range(4)]
    V.set_tuning(chain_blocks=0, lds_chain=0)
    want = [V.solve(i, max_size=11) for i in insts]
    V.set_tuning(chain_blocks=blocks, lds_chain=0)
    got = V.solve_batch(insts, max_size=11)
    for k, (g, w) in enumerate(zip(got, want)):
        _same(g, w, ("batch", blocks, k))
    assert sum(g["chain_levels"] for g in got) > 0


def test_chain_in_second_process_equals_host(tuning):
    """While this process holds the device, the runtime refuses a second
    process's cooperative launch: its chain then runs teams of one workgroup,
    with the same results (and says so in its kernel statistics)."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    insts = [V.Instance.generate(2, k) for k in (0, 1, 4)]
    V.set_tuning(chain_blocks=0, lds_chain=0)
    want = [V.solve(i, max_size=9) for i in insts]
    code = ("import json, vfb_amd as V\n"
            "V.set_tuning(chain_blocks=64, lds_chain=0)\n"
            "got = [V.solve(V.Instance.generate(2, k), max_size=9) for k in (0, 1, 4)]\n"
            "print(json.dumps({'got': got, 'stats': V.kernel_stats()}))\n")
    print("[chain] second process", flush=True)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=root, timeout=300,
                         env=dict(os.environ, PYTHONPATH=root))
    assert out.returncode == 0, out.stderr[-2000:]
    res = json.loads(out.stdout.strip().splitlines()[-1])
    for k, (g, w) in enumerate(zip(res["got"], want)):
        _same(g, w, ("child", k))
    assert sum(g["chain_levels"] for g in res["got"]) > 0
    assert res["stats"]["chain_launches"] > 0


# ---- the level chain in LDS (k_chain_lds) ----

@pytest.mark.parametrize("chain_blocks", [0, 64])
def test_lds_chain_plain_equals_host(tuning, chain_blocks):
    """The LDS chain alone, and followed by k_chain's teams."""
    print(f"[lds chain] plain chain_blocks={chain_blocks}", flush=True)
    insts = [(c, k, V.Instance.generate(c, k)) for c, k in _insts()] + \
            [(4, k, V.Instance.generate(4, k)) for k in range(8)]
    V.set_tuning(chain_blocks=0, lds_chain=0)
    want = [V.solve(i, max_size=10) for _, _, i in insts]
    assert all(w["chain_levels"] == 0 for w in want)
    V.set_tuning(chain_blocks=chain_blocks, lds_chain=1)
    before = V.kernel_stats()
    got = [V.solve(i, max_size=10) for _, _, i in insts]
    after = V.kernel_stats()
    for (c, k, _), g, w in zip(insts, got, want):
        _same(g, w, (c, k))
    assert sum(g["chain_levels"] for g in got) > 0
    assert after["chain_levels"] - before["chain_levels"] == sum(g["chain_levels"] for g in got)


@pytest.mark.parametrize("relaxed_type", [1, 2, 3])
def test_lds_chain_relaxed_keys_equal_host(tuning, relaxed_type):
    print(f"[lds chain] relaxed type {relaxed_type}", flush=True)
    insts = [V.Instance.generate(2, k) for k in (0, 1, 4, 7)] + [V.Instance.generate(3, k) for k in (0, 1)]
    V.set_tuning(chain_blocks=0, lds_chain=0)
    want = [V.solve(i, max_size=8, relaxed_type=relaxed_type) for i in insts]
    V.set_tuning(chain_blocks=0, lds_chain=1)
    got = [V.solve(i, max_size=8, relaxed_type=relaxed_type) for i in insts]
    for k, (g, w) in enumerate(zip(got, want)):
        _same(g, w, (relaxed_type, k))
    assert sum(g["chain_levels"] for g in got) > 0


@pytest.mark.parametrize("max_rows", [10, 30, 200, 1000])
def test_lds_chain_row_cap_equals_host(tuning, max_rows):
    print(f"[lds chain] max_rows={max_rows}", flush=True)
    insts = [V.Instance.generate(2, k) for k in (2, 3, 5)] + [V.Instance.generate(1, k) for k in (0, 1)]
    V.set_tuning(chain_blocks=0, lds_chain=0)
    want = [V.solve(i, max_size=9, max_rows=max_rows) for i in insts]
    V.set_tuning(chain_blocks=0, lds_chain=1)
    got = [V.solve(i, max_size=9, max_rows=max_rows) for i in insts]
    for k, (g, w) in enumerate(zip(got, want)):
        _same(g, w, (max_rows, k))


def test_lds_chain_neg_type_2_equals_host(tuning):
    print("[lds chain] NegType 2", flush=True)
    insts = [V.Instance.generate(2, k) for k in (0, 1, 4)]
    V.set_tuning(chain_blocks=0, lds_chain=0)
    want = [V.solve(i, max_size=8, neg_type=2) for i in insts]
    V.set_tuning(chain_blocks=0, lds_chain=1)
    got = [V.solve(i, max_size=8, neg_type=2) for i in insts]
    for k, (g, w) in enumerate(zip(got, want)):
        _same(g, w, ("neg2", k))
    assert sum(g["chain_levels"] for g in got) > 0


@pytest.mark.parametrize("chain_blocks", [0, 1])
def test_lds_chain_batch_equals_host_plain(tuning, chain_blocks):
    """A batch of 40 members (one LDS workgroup each) against plain solves
    with neither chain."""
    print(f"[lds chain] batch chain_blocks={chain_blocks}", flush=True)
    insts = [V.Instance.generate(4, k) for k in range(32)] + [V.Instance.generate(2, k) for k in range(8)]
    V.set_tuning(chain_blocks=0, lds_chain=0)
    want = [V.solve(i, max_size=11) for i in insts]
    V.set_tuning(chain_blocks=chain_blocks, lds_chain=1)
    got = V.solve_batch(insts, max_size=11)
    for k, (g, w) in enumerate(zip(got, want)):
        _same(g, w, ("batch", chain_blocks, k))
    assert sum(g["chain_levels"] for g in got) > 0


def test_lds_chain_duplicate_atoms_equals_host(tuning):
    """Two atoms with equal characteristic sets: the set holds fewer keys
    than the solve's committed count, so the LDS chain leaves the levels to
    the host; the results are unchanged."""
    print("[lds chain] duplicate atoms", flush=True)
    names, lens, atoms, P = V.Instance.generate(1, 3).arrays()
    atoms = atoms + [list(atoms[0])]
    inst = V.Instance.from_bits(names + ["dup"], lens, atoms, P)
    V.set_tuning(chain_blocks=0, lds_chain=0)
    want = V.solve(inst, max_size=8)
    V.set_tuning(chain_blocks=0, lds_chain=1)
    got = V.solve(inst, max_size=8)
    _same(got, want, "dup")
