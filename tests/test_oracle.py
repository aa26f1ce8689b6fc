"""Pins the CPU oracle against things other than itself.

- operators: textbook per-position finite-trace semantics, all lengths 1..64;
- LTL identities (~~p = p, F p = true U p, G p = ~F~p, expansion law of U,
  De Morgan, idempotence of F and G);
- SplitMix64 finaliser: the published first outputs of SplitMix64 seeded 0;
- exact key: injective (decoded back bit by bit);
- relaxed keys 1 and 2: known answers worked out by hand below;
- per-level distinct counts with exact keys == semantic dedup on full rows;
- minimal size and formula == independent brute-force syntactic enumeration
  (ltl_nocache numbering), the formula re-verified per position;
- the brute-force formula counts == direct enumeration of syntax trees.
"""
import itertools
import os
import random

import pytest

import oracle_lib as O
from ltl_semantics import pos_apply, separates, parse, size

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M64 = (1 << 64) - 1


def mask(n):
    return (1 << n) - 1


@pytest.mark.parametrize("op", range(7))
def test_operators_match_position_semantics(op):
    rng = random.Random(op)
    for n in range(1, 65):
        for _ in range(12):
            a = rng.getrandbits(n)
            b = rng.getrandbits(n)
            # sparse and dense words too
            if _ % 3 == 1:
                a &= rng.getrandbits(n)
            if _ % 3 == 2:
                b |= rng.getrandbits(n)
            got = O.apply(op, a, b, n) & mask(n)
            assert got == pos_apply(op, a, b, n), (op, n, hex(a), hex(b))
            # results never carry bits at or above the trace length
            assert O.apply(op, a, b, n) >> n == 0


def test_identities():
    rng = random.Random(5)
    for n in list(range(1, 65)):
        full = mask(n)
        for _ in range(8):
            p, q = rng.getrandbits(n), rng.getrandbits(n)
            ap = lambda op, x, y=0: O.apply(op, x, y, n)
            assert ap(O.NOT, ap(O.NOT, p)) == p
            assert ap(O.FINALLY, p) == ap(O.UNTIL, full, p)
            assert ap(O.GLOBALLY, p) == ap(O.NOT, ap(O.FINALLY, ap(O.NOT, p)))
            u = ap(O.UNTIL, p, q)
            assert u == ap(O.OR, q, ap(O.AND, p, ap(O.NEXT, u)))
            assert ap(O.NOT, ap(O.AND, p, q)) == ap(O.OR, ap(O.NOT, p), ap(O.NOT, q))
            assert ap(O.FINALLY, ap(O.FINALLY, p)) == ap(O.FINALLY, p)
            assert ap(O.GLOBALLY, ap(O.GLOBALLY, p)) == ap(O.GLOBALLY, p)


def test_splitmix64_published_outputs():
    # SplitMix64 seeded with 0 yields mix(k * 0x9e3779b97f4a7c15) for k = 1, 2, 3
    g = 0x9E3779B97F4A7C15
    assert O.mix64(g) == 0xE220A8397B1DCDAF
    assert O.mix64((2 * g) & M64) == 0x6E789E6AA1B965F4
    assert O.mix64((3 * g) & M64) == 0x06C45D188009454F


def test_key_type3_known_answer():
    g = 0x9E3779B97F4A7C15
    # two traces of length 64 (lenSum 128 > 126): trace 0 feeds hi, trace 1 lo
    hi, lo = O.key([g, (2 * g) & M64], [64, 64], 3)
    assert (hi, lo) == (0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4)
    # equal words in one half cancel (XOR), a real collision of the hashed key
    w = [5, 5, 9, 1]
    assert O.key(w, [40, 40, 40, 40], 3) == O.key([7, 7, 9, 1], [40, 40, 40, 40], 3)


def test_key_exact_is_injective_and_decodes():
    rng = random.Random(11)
    for _ in range(300):
        T = rng.randint(1, 12)
        budget = rng.randint(T, 126)
        lens = []
        for t in range(T):
            room = budget - sum(lens) - (T - t - 1)
            lens.append(rng.randint(1, max(1, min(64, room))))
        words = [rng.getrandbits(n) for n in lens]
        hi, lo = O.key(words, lens, 3)
        assert hi >> 63 == 0 and lo >> 63 == 0
        bits = lo | (hi << 63)
        j = 0
        for t, n in enumerate(lens):
            assert (bits >> j) & mask(n) == words[t]
            j += n


def test_key_type1_known_answer():
    # lenSum = 2 * 64 + 1 = 129 -> stride 129 // 126 = 1: the first 126
    # trace bits in order; trace 2 (length 1) is beyond bit 126 and dropped
    lens = [64, 64, 1]
    w0, w1 = (1 << 63) | 1, (1 << 61) | (1 << 62) | 4
    hi, lo = O.key([w0, w1, 1], lens, 1)
    # global bits: trace0 bits 0..63 -> j 0..63, trace1 bits 0..61 -> j 64..125
    # j=0 -> lo bit 0; j=63 -> hi bit 0; trace1 bit2 -> j=66 -> hi bit 3;
    # trace1 bit 61 -> j=125 -> hi bit 62; trace1 bit 62 -> j=126 dropped
    assert lo == 1
    assert hi == (1 | (1 << 3) | (1 << 62))
    # stride 2: lenSum = 4 * 63 = 252 -> stride 2, samples k = 0, 2, ..., 62
    lens = [63, 63, 63, 63]
    w = [0b101, 0b10, 1 << 62, (1 << 60) | 1]
    hi, lo = O.key(w, lens, 1)
    # trace0: k=0 -> j0, k=2 -> j1; trace1: 32 samples, k=1 not sampled;
    # trace2: k=62 is sample 31 -> j = 64 + 31 = 95 -> hi bit 32;
    # trace3: k=0 -> j=96 -> hi bit 33; k=60 -> sample 30 -> j=126 dropped
    assert lo == 0b11
    assert hi == (1 << 32) | (1 << 33)


def test_key_type2_known_answer():
    lens = [64, 63]
    # lens 64, 63 -> lenSum 127: maxbits 64 and 63; j: trace0 0..63, trace1 64..126 (126 dropped)
    w0 = (1 << 63) | (1 << 62)
    w1 = (1 << 62) | (1 << 61) | 1
    hi, lo = O.key([w0, w1], lens, 2)
    # trace0 bit 62 -> j 62 -> lo bit 62; bit 63 -> j 63 -> hi bit 0
    # trace1 bit 0 -> j 64 -> hi bit 1; bit 61 -> j 125 -> hi bit 62; bit 62 -> j 126 dropped
    assert lo == 1 << 62
    assert hi == 1 | 2 | (1 << 62)


def _rand_inst(rng, A, T, lmin, lmax, p=0.5):
    lens = [rng.randint(lmin, lmax) for _ in range(T)]
    atoms = [[sum((rng.random() < p) << k for k in range(lens[t])) for t in range(T)] for _ in range(A)]
    return O.Instance([f"p{a}" for a in range(A)], lens, atoms, rng.randint(0, T))


def test_example_bruteforce_vs_cs():
    inst = O.Instance.from_json(os.path.join(GOLDEN, "ExampleLTLF.json"))
    bf = O.bruteforce(inst, 10)
    cs = O.enumerate_cs(inst, 10, 3)
    assert bf["status"] == cs["status"] == 0
    assert bf["min_size"] == cs["min_size"] == 10
    for r in (bf, cs):
        assert separates(r["formula"], inst.names, inst.lens, inst.atoms, inst.P)
        assert size(parse(r["formula"], inst.names)) == 10


@pytest.mark.parametrize("seed", range(40))
def test_min_size_matches_bruteforce(seed):
    rng = random.Random(seed)
    inst = _rand_inst(rng, rng.randint(1, 3), rng.randint(1, 6), 1, 6)
    ms = 6
    bf = O.bruteforce(inst, ms)
    for mode in (3, O.KEY_FULLROW):
        cs = O.enumerate_cs(inst, ms, mode)
        assert cs["status"] == bf["status"], (seed, mode)
        assert cs["min_size"] == bf["min_size"], (seed, mode)
        if cs["status"] == 0:
            assert separates(cs["formula"], inst.names, inst.lens, inst.atoms, inst.P)
            assert size(parse(cs["formula"], inst.names)) == cs["min_size"]


@pytest.mark.parametrize("seed", range(20))
def test_exact_counts_equal_semantic_dedup(seed):
    rng = random.Random(100 + seed)
    T = rng.randint(2, 12)
    inst = _rand_inst(rng, rng.randint(1, 4), T, 1, max(1, 126 // T))
    assert sum(inst.lens) <= 126
    a = O.enumerate_cs(inst, 7, 3)
    b = O.enumerate_cs(inst, 7, O.KEY_FULLROW)
    for f in ("status", "min_size", "distinct", "candidates", "formula", "found_ordinal"):
        assert a[f] == b[f], (seed, f)


def test_candidate_counts_closed_form():
    # candidates of level s: 4 * n(s-1) + 4 * sum_{i <= s-1-i} n(i) n(s-1-i)
    # trace 0 is repeated as the last (negative) trace: nothing separates them,
    # so every level is enumerated completely
    rng = random.Random(3)
    inst = _rand_inst(rng, 3, 5, 3, 8)
    inst = O.Instance(inst.names, inst.lens + [inst.lens[0]], [row + [row[0]] for row in inst.atoms], 1)
    r = O.enumerate_cs(inst, 6, O.KEY_FULLROW)
    assert r["status"] == 1
    n = r["distinct"]
    for s in range(2, 7):
        want = 4 * n[s - 1] + 4 * sum(n[i] * n[s - 1 - i] for i in range(1, s) if 2 * i <= s - 1)
        assert r["candidates"][s] == want


def _count_trees(A, s):
    """Direct enumeration of syntax trees in the ltl_nocache numbering's
    grammar: unary ~ X F G, & and | with left size <= right size... as the
    numbering counts them (left size j <= (s-1)//2), U with any split."""
    memo = {}

    def trees(k):
        if k in memo:
            return memo[k]
        out = []
        if k == 1:
            out = [("a", i) for i in range(A)]
        else:
            for op in "~XFG":
                out += [(op, t) for t in trees(k - 1)]
            for j in range(1, k - 1):
                for op in "&|U":
                    if op != "U" and j > (k - 1) // 2:
                        continue
                    for l, r in itertools.product(trees(j), trees(k - 1 - j)):
                        out.append((op, l, r))
        memo[k] = out
        return out

    return len(trees(s))


def test_num_formulas_closed_form():
    for A in (1, 2, 3):
        nf = O.num_formulas(A, 6)
        for s in range(1, 7):
            assert nf[s][7] == _count_trees(A, s), (A, s)


def test_onthefly_semantics():
    rng = random.Random(9)
    inst = _rand_inst(rng, 3, 8, 4, 10)
    inst = O.Instance(inst.names, inst.lens + [inst.lens[0]], [row + [row[0]] for row in inst.atoms], 1)
    full = O.enumerate_cs(inst, 7, 3)
    assert full["status"] == 1 and full["otf_level"] == 0
    # a cap above everything stored changes nothing
    big = O.enumerate_cs(inst, 7, 3, 10 ** 9)
    assert big == full
    # cap inside level 4: levels 1-4 deduped, 5 check-only, then stop
    cap = sum(full["distinct"][:4]) + 7
    r = O.enumerate_cs(inst, 7, 3, cap)
    assert r["status"] == 3 and r["otf_level"] == 4 and r["otf_stored"] == 7
    assert r["distinct"][:5] == full["distinct"][:5]
    assert r["levels_done"] == 4 and r["distinct"][5] == 0
    # the check-only level enumerates pairs over the stored rows only
    n = [0] + full["distinct"][1:4] + [7]
    want = 4 * n[4] + 4 * sum(n[i] * n[4 - i] for i in range(1, 5) if 2 * i <= 4)
    assert r["candidates"][5] == want


def _gen_instance(seed, A, T, lmax, p=0.5):
    rng = random.Random(seed)
    lens = [rng.randint(1, lmax) for _ in range(T)]
    atoms = [[sum((rng.random() < p) << k for k in range(lens[t])) for t in range(T)] for _ in range(A)]
    return O.Instance([f"p{a}" for a in range(A)], lens, atoms, rng.randint(1, T - 1))


@pytest.mark.parametrize("seed", range(24))
def test_omp_build_equals_sequential_oracle(seed):
    """The multi-core CPU baseline (cpu_omp.c) gives the sequential oracle's
    results, with and without the three-pass split of large levels."""
    rng = random.Random(9000 + seed)
    inst = _gen_instance(seed, rng.randint(1, 4), rng.randint(2, 20), 64 if seed % 3 == 0 else 10)
    key_mode = 1 + seed % 3
    want = O.enumerate_cs(inst, 7, key_mode)
    for pass_min in (0, 1, 1 << 10):
        got, _ = O.enumerate_omp(inst, 7, key_mode, pass_min)
        for f in ("status", "min_size", "formula", "distinct", "candidates", "found_ordinal", "top_op"):
            assert got[f] == want[f], (seed, pass_min, f, got[f], want[f])


def test_omp_build_on_example_and_exact_keys():
    inst = O.Instance.from_json(os.path.join(GOLDEN, "ExampleLTLF.json"))
    want = O.enumerate_cs(inst, 10, 3)
    got, _ = O.enumerate_omp(inst, 10, 3, 1)
    assert (got["formula"], got["distinct"], got["candidates"]) == (want["formula"], want["distinct"],
                                                                     want["candidates"])
    assert O.omp_threads() >= 1


# ---- BS stage ---------------------------------------------------------------

def _all_trees(A, size, memo):
    """Every syntax tree of the grammar {atoms, ~ X F G, & | U} of a size."""
    if size in memo:
        return memo[size]
    out = []
    if size == 1:
        out = [("atom", a) for a in range(A)]
    else:
        for c in _all_trees(A, size - 1, memo):
            out += [(op, c) for op in ("~", "X", "F", "G")]
        for i in range(1, size - 1):
            for l in _all_trees(A, i, memo):
                for r in _all_trees(A, size - 1 - i, memo):
                    out += [(op, l, r) for op in ("&", "|", "U")]
    memo[size] = out
    return out


def _bs_min_cost(inst, max_size, bs_max):
    """Independent minimal cost of a Boolean combination (~ & |, unit costs)
    of formulas of size <= max_size whose time-0 vector over the traces is
    the positive mask: leaves from textbook semantics, then a fixpoint over
    the 2^T vectors."""
    from ltl_semantics import eval_tree
    T, P, A = inst.T, inst.P, inst.A
    full = (1 << T) - 1
    best = {}
    memo = {}
    for s in range(1, max_size + 1):
        for tree in _all_trees(A, s, memo):
            v = 0
            for t in range(T):
                if eval_tree(tree, [inst.atoms[a][t] for a in range(A)], inst.lens[t])[0]:
                    v |= 1 << t
            best.setdefault(v, s)
    changed = True
    while changed:
        changed = False
        items = list(best.items())
        for v, c in items:
            nv, nc = ~v & full, c + 1
            if nc <= bs_max and best.get(nv, 99) > nc:
                best[nv] = nc
                changed = True
        for v, c in items:
            for w, d in items:
                for nv in (v & w, v | w):
                    if c + d + 1 <= bs_max and best.get(nv, 99) > c + d + 1:
                        best[nv] = c + d + 1
                        changed = True
    return best.get((1 << P) - 1)


@pytest.mark.parametrize("seed", range(30))
def test_bs_stage_min_cost_and_formula(seed):
    """The BS stage's size equals an independent fixpoint over the time-0
    vectors, and its formula (an LTL formula: ~ & | over LTL subformulas)
    separates the traces with exactly that size."""
    rng = random.Random(7100 + seed)
    A = rng.randint(1, 2)
    T = rng.randint(3, 6)
    lens = [rng.randint(1, 4) for _ in range(T)]
    atoms = [[rng.randrange(1 << lens[t]) for t in range(T)] for _ in range(A)]
    inst = O.Instance([f"p{a}" for a in range(A)], lens, atoms, rng.randint(1, T - 1))
    max_size, bs_max = 2, 7
    plain = O.enumerate_cs(inst, max_size, 3)
    got = O.enumerate_bs(inst, max_size, bs_max, 3)
    want = _bs_min_cost(inst, max_size, bs_max)
    if plain["status"] == 0:
        assert got["formula"] == plain["formula"] and not got["bs_stage"]
        return
    if want is None:
        assert got["status"] == 1 and got["formula"] == "Not found !"
        return
    assert got["status"] == 0 and got["bs_stage"] == 1
    assert got["min_size"] == want, (seed, got["min_size"], want, got["formula"])
    assert separates(got["formula"], inst.names, inst.lens, inst.atoms, inst.P)
    assert size(parse(got["formula"], inst.names)) == want
    # levels up to max_size are the CS stage's own
    assert got["distinct"][:max_size + 1] == plain["distinct"][:max_size + 1]


def test_bs_stage_off_when_traces_exceed_63():
    rng = random.Random(5)
    T = 64
    lens = [3] * T
    atoms = [[rng.randrange(8) for _ in range(T)]]
    inst = O.Instance(["p0"], lens, atoms, 32)
    got = O.enumerate_bs(inst, 2, 6, 3)
    assert got["bs_stage"] == 0


# ---- cost function and NegType (LTL_BS's CLI) --------------------------------

def _cost_min(inst, costs, neg_type, max_cost):
    """Independent minimal cost of a separating formula: a semantic dynamic
    program over per-position truth vectors (textbook semantics), levels by
    cost (p c1, ~ c2, & c3, | c4, X c5, F c6, G c7, U c8; NegType 2: ~ of
    atoms only), each vector kept at its lowest cost."""
    c1, c2, c3, c4, c5, c6, c7, c8 = costs
    T, A = inst.T, inst.A
    code = {"~": 0, "&": 1, "|": 2, "X": 3, "F": 4, "G": 5, "U": 6}
    seen = set()
    lv = {}

    def target(v):
        return all((v[t] & 1) == (1 if t < inst.P else 0) for t in range(T))

    def add(C, v):
        if v in seen:
            return False
        seen.add(v)
        lv.setdefault(C, []).append(v)
        return target(v)

    for a in range(A):
        if add(c1, tuple(inst.atoms[a][t] & ((1 << inst.lens[t]) - 1) for t in range(T))):
            return c1
    for C in range(c1 + 1, max_cost + 1):
        hit = False
        cand = []
        if C - c2 >= c1 and (neg_type == 1 or C - c2 == c1):
            cand += [("~", x, None) for x in lv.get(C - c2, [])]
        for op, co in (("&", c3), ("|", c4), ("U", c8)):
            for i in range(c1, C - co - c1 + 1):
                j = C - i - co
                for x in lv.get(i, []):
                    for y in lv.get(j, []):
                        cand.append((op, x, y))
        for op, co in (("X", c5), ("F", c6), ("G", c7)):
            cand += [(op, x, None) for x in lv.get(C - co, [])] if C - co >= c1 else []
        for op, x, y in cand:
            v = tuple(pos_apply(code[op], x[t], y[t] if y else 0, inst.lens[t]) for t in range(T))
            hit |= add(C, v)
        if hit:
            return C
    return None


def _formula_cost(tree, costs):
    c1, c2, c3, c4, c5, c6, c7, c8 = costs
    table = {"~": c2, "&": c3, "|": c4, "X": c5, "F": c6, "G": c7, "U": c8}
    if tree[0] == "atom":
        return c1
    return table[tree[0]] + sum(_formula_cost(c, costs) for c in tree[1:])


@pytest.mark.parametrize("seed", range(30))
def test_costs_min_and_formula(seed):
    rng = random.Random(6100 + seed)
    A = rng.randint(1, 2)
    T = rng.randint(2, 5)
    lens = [rng.randint(1, 4) for _ in range(T)]
    atoms = [[rng.randrange(1 << lens[t]) for t in range(T)] for _ in range(A)]
    inst = O.Instance([f"p{a}" for a in range(A)], lens, atoms, rng.randint(1, T - 1))
    costs = [rng.randint(1, 3) for _ in range(8)]
    neg_type = 1 + seed % 2
    max_cost = 9
    got = O.enumerate_costs(inst, max_cost, costs, neg_type)
    want = _cost_min(inst, costs, neg_type, max_cost)
    if want is None:
        assert got["status"] == 1, (seed, got["formula"])
        return
    assert got["status"] == 0 and got["min_size"] == want, (seed, costs, neg_type, got["min_size"], want)
    tree = parse(got["formula"], inst.names)
    assert separates(got["formula"], inst.names, inst.lens, inst.atoms, inst.P)
    assert _formula_cost(tree, costs) == want
    if neg_type == 2:
        def negs_on_atoms(t):
            return (t[0] != "~" or t[1][0] == "atom") and all(negs_on_atoms(c) for c in t[1:] if isinstance(c, tuple))
        assert negs_on_atoms(tree), got["formula"]


def test_unit_costs_equal_plain():
    rng = random.Random(1)
    for k in range(6):
        T = rng.randint(2, 8)
        lens = [rng.randint(1, 8) for _ in range(T)]
        atoms = [[rng.randrange(1 << lens[t]) for t in range(T)] for _ in range(2)]
        inst = O.Instance(["a", "b"], lens, atoms, rng.randint(1, T - 1))
        a = O.enumerate_cs(inst, 7, 3)
        b = O.enumerate_costs(inst, 7, [1] * 8, 1)
        assert (a["formula"], a["distinct"], a["candidates"]) == (b["formula"], b["distinct"], b["candidates"])


# ---- the -1 / -2 half-key fold (hash_set.cuh:198-204, LTL_BS.cu:262-265) ----

from fold_instance import fold_instance, mix64 as py_mix64, unmix64, type3_key  # noqa: E402


def test_unmix64_inverts_the_published_finaliser():
    rng = random.Random(9)
    g = 0x9E3779B97F4A7C15
    assert py_mix64(g) == 0xE220A8397B1DCDAF       # published SplitMix64 output
    for _ in range(200):
        x = rng.getrandbits(64)
        assert unmix64(py_mix64(x)) == x
        assert py_mix64(unmix64(x)) == x
        assert O.mix64(x) == py_mix64(x)


@pytest.mark.parametrize("half", ["hi", "lo"])
def test_fold_instance_reaches_minus_one_and_minus_two(half):
    names, lens, atoms, P, (na, nb) = fold_instance(half)
    ha, hb = type3_key(na), type3_key(nb)
    k = 0 if half == "hi" else 1
    assert ha[k] == M64 and hb[k] == M64 - 1          # raw halves -1 and -2
    assert ha[1 - k] == hb[1 - k]                      # the other halves agree
    assert na != nb                                     # different rows
    # the oracle's key (restated from the reference) merges them: -1 -> -2
    ka, kb = O.key(na, lens, 3), O.key(nb, lens, 3)
    assert ka == kb and ka[k] == M64 - 1
    # ~ of the atoms is exactly these rows
    assert [O.apply(O.NOT, w, 0, 64) for w in atoms[0]] == na
    assert [O.apply(O.NOT, w, 0, 64) for w in atoms[1]] == nb


@pytest.mark.parametrize("half", ["hi", "lo"])
def test_fold_drops_one_distinct_cs_at_level_two(half):
    names, lens, atoms, P, _ = fold_instance(half)
    inst = O.Instance(names, lens, atoms, P)
    hashed = O.enumerate_cs(inst, 2, 3)
    full = O.enumerate_cs(inst, 2, O.KEY_FULLROW)
    assert hashed["status"] == full["status"] == 1     # level 2 complete, nothing found
    assert hashed["distinct"][1] == full["distinct"][1] == 2
    # full-row dedup of the 8 unary candidates against each other and the
    # atoms, counted here on Python tuples; the fold keeps one fewer
    seen = {tuple(x) for x in atoms}
    new = 0
    for op in (O.NOT, O.NEXT, O.FINALLY, O.GLOBALLY):
        for x in atoms:
            row = tuple(O.apply(op, w, 0, 64) for w in x)
            if row not in seen:
                seen.add(row)
                new += 1
    assert full["distinct"][2] == new
    assert hashed["distinct"][2] == full["distinct"][2] - 1
    assert hashed["candidates"] == full["candidates"]


def test_omp_batch_baseline_equals_sequential_oracle():
    """The batch CPU baseline (OpenMP over instances) returns each instance's
    sequential oracle result."""
    rng = random.Random(4242)
    insts = [_rand_inst(rng, rng.randint(1, 4), rng.randint(2, 12), 1, 16) for _ in range(24)]
    got, _ = O.enumerate_many_omp(insts, 7)
    for k, inst in enumerate(insts):
        assert got[k] == O.enumerate_cs(inst, 7, 3), k
