"""The oracle on traces longer than one 64-bit word (CPU only).

A trace of n steps is ceil(n/64) words, word w holding steps 64w..64w+63.
The oracle's word-carrying operators are pinned against the position-wise
semantics, its keys against bit-position restatements of LTL_BS.cu:30-110
written over Python integers, and its whole enumeration against an
independent Python enumerator below, which evaluates every operator from its
closed form on an n-step integer (no words, no carries) and deduplicates on the
full CS row.
"""
import random

import pytest

import oracle_lib as O
from ltl_semantics import parse, pos_apply, separates, size

M64 = (1 << 64) - 1
PHI = 0x9E3779B97F4A7C15


def _rand_word(rng, n):
    """Random n-step trace, with long runs of ones and zeros mixed in."""
    kind = rng.randrange(4)
    if kind == 0:
        return rng.getrandbits(n)
    if kind == 1:                        # a run of ones ending at a random step
        a = rng.randrange(n)
        b = rng.randint(a, n)
        return ((1 << b) - 1) ^ ((1 << a) - 1)
    if kind == 2:                        # sparse
        x = 0
        for _ in range(rng.randint(0, 3)):
            x |= 1 << rng.randrange(n)
        return x
    return ((1 << n) - 1) ^ (1 << rng.randrange(n))   # all ones but one


@pytest.mark.parametrize("op", range(7))
def test_apply_trace_matches_position_semantics(op):
    rng = random.Random(900 + op)
    for n in (1, 2, 63, 64, 65, 66, 100, 127, 128, 129, 191, 192, 193, 256, 300):
        for _ in range(12):
            a, b = _rand_word(rng, n), _rand_word(rng, n)
            assert O.apply_trace(op, a, b, n) == pos_apply(op, a, b, n), (op, n, hex(a), hex(b))


def test_apply_trace_one_word_is_ora_apply():
    rng = random.Random(5)
    for op in range(7):
        for n in (1, 17, 63, 64):
            a, b = rng.getrandbits(n), rng.getrandbits(n)
            assert O.apply_trace(op, a, b, n) == O.apply(op, a, b, n)


# ---- keys: bit-position restatements of LTL_BS.cu:30-110 over integers ----

def _put(j, hl):
    if j < 63:
        hl[1] |= 1 << j
    elif j < 126:
        hl[0] |= 1 << (j - 63)


def _fold(hl):
    return tuple(M64 - 1 if h == M64 else h for h in hl)


def _key_py(cs, lens, rt):
    T, lsum = len(lens), sum(lens)
    hl = [0, 0]
    if lsum <= 126:
        j = 0
        for i in range(T):
            for k in range(lens[i]):
                if (cs[i] >> k) & 1:
                    _put(j, hl)
                j += 1
    elif rt == 1:
        stride, j = lsum // 126, 0
        for i in range(T):
            for k in range(0, lens[i], stride):
                if j >= 126:
                    break
                if (cs[i] >> k) & 1:
                    _put(j, hl)
                j += 1
    elif rt == 2:
        j = 0
        for i in range(T):
            for k in range((126 * lens[i] + lsum) // lsum):
                if j >= 126:
                    break
                if (cs[i] >> k) & 1:
                    _put(j, hl)
                j += 1
    else:
        for i in range(T):
            for w, x in enumerate(O.split_words(cs[i], lens[i])):
                h = O.mix64((x + w * PHI) & M64)
                hl[0 if i < T // 2 else 1] ^= h
    return _fold(hl)


@pytest.mark.parametrize("rt", [1, 2, 3])
def test_multiword_keys_restated(rt):
    rng = random.Random(70 + rt)
    for _ in range(150):
        T = rng.randint(1, 6)
        lens = [rng.choice([rng.randint(1, 64), rng.randint(65, 400)]) for _ in range(T)]
        cs = [_rand_word(rng, n) for n in lens]
        assert O.key(cs, lens, rt) == _key_py(cs, lens, rt), (lens, rt)


def test_type3_one_word_traces_unchanged():
    # word 0 of a trace is mixed alone, as the reference mixes each trace
    assert O.key([PHI, (2 * PHI) & M64], [64, 64], 3) == (0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4)


def test_exact_key_over_a_long_trace_is_injective():
    rng = random.Random(8)
    for _ in range(200):
        n0 = rng.randint(65, 120)
        rest = [rng.randint(1, 3) for _ in range(rng.randint(0, 2))]
        lens = [n0] + rest
        if sum(lens) > 126:
            continue
        cs = [rng.getrandbits(n) for n in lens]
        hi, lo = O.key(cs, lens, 3)
        bits, j = lo | (hi << 63), 0
        for t, n in enumerate(lens):
            assert (bits >> j) & ((1 << n) - 1) == cs[t]
            j += n


# ---- an independent enumerator: closed-form operators on n-step integers ----

def _hb(x):
    return x.bit_length() - 1


def _sem(op, x, y, n):
    full = (1 << n) - 1
    if op == O.NOT:
        return ~x & full
    if op == O.AND:
        return x & y
    if op == O.OR:
        return x | y
    if op == O.NEXT:                     # X at t = x at t+1
        return x >> 1
    if op == O.FINALLY:                  # every step up to the last true one
        return (1 << (_hb(x) + 1)) - 1 if x else 0
    if op == O.GLOBALLY:                 # every step after the last false one
        z = ~x & full
        return full & ~((1 << (_hb(z) + 1)) - 1) if z else full
    u, out = 0, 0                        # expansion law, last step first
    for t in range(n - 1, -1, -1):
        u = (y >> t) & 1 or ((x >> t) & 1 and u)
        out |= u << t
    return out


def _text(node, names):
    if node[0] == "atom":
        return names[node[1]]
    op = "~&|XFGU"[node[0]]
    if len(node) == 2:
        return f"{op}({_text(node[1], names)})"
    return f"({_text(node[1], names)}) {op} ({_text(node[2], names)})"


def sem_enumerate(inst, max_size):
    """Size-ordered enumeration in the canonical order (LTL_BS.cu:1226-1493:
    ~, & splits, | splits, X, F, G, U splits with both orientations), every AP
    stored at size 1 (LTL_BS.cu:1127-1163), only unseen CS checked and kept,
    the first matching one returned."""
    T, lens, P = inst.T, inst.lens, inst.P
    res = {"status": 1, "min_size": 0, "distinct": [0] * (O.MAX_SIZE + 1),
           "candidates": [0] * (O.MAX_SIZE + 1), "found_ordinal": 0, "formula": "Not found !"}
    want = tuple(t < P for t in range(T))

    def hit(cs):
        return tuple(bool(cs[t] & 1) for t in range(T)) == want

    seen, levels = set(), {}
    lv = []
    for a in range(inst.A):
        cs = tuple(inst.atoms[a][t] & ((1 << lens[t]) - 1) for t in range(T))
        res["candidates"][1] += 1
        if hit(cs):
            res.update(status=0, min_size=1, found_ordinal=a, formula=inst.names[a])
            return res
        seen.add(cs)
        lv.append((cs, ("atom", a)))
        res["distinct"][1] += 1
    levels[1] = lv
    for s in range(2, max_size + 1):
        cur, ordn = [], 0

        def cand(cs, node):
            nonlocal ordn
            o = ordn
            ordn += 1
            res["candidates"][s] += 1
            if cs in seen:
                return False
            seen.add(cs)
            if hit(cs):
                res.update(status=0, min_size=s, found_ordinal=o, formula=_text(node, inst.names))
                return True
            res["distinct"][s] += 1
            cur.append((cs, node))
            return False

        def un(op, src):
            for cs, nd in levels[src]:
                if cand(tuple(_sem(op, cs[t], 0, lens[t]) for t in range(T)), (op, nd)):
                    return True
            return False

        def bi(op, i):
            for l, ln in levels[i]:
                for r, rn in levels[s - 1 - i]:
                    if cand(tuple(_sem(op, l[t], r[t], lens[t]) for t in range(T)), (op, ln, rn)):
                        return True
                    if op == O.UNTIL and cand(tuple(_sem(op, r[t], l[t], lens[t]) for t in range(T)), (op, rn, ln)):
                        return True
            return False

        splits = [i for i in range(1, s) if 2 * i <= s - 1]
        done = (un(O.NOT, s - 1) or any(bi(O.AND, i) for i in splits) or any(bi(O.OR, i) for i in splits)
                or un(O.NEXT, s - 1) or un(O.FINALLY, s - 1) or un(O.GLOBALLY, s - 1)
                or any(bi(O.UNTIL, i) for i in splits))
        if done:
            return res
        levels[s] = cur
    return res


def _long_inst(rng, A, T, lmin, lmax):
    lens = [rng.randint(lmin, lmax) for _ in range(T)]
    atoms = [[_rand_word(rng, n) if rng.random() < 0.5 else rng.getrandbits(n) for n in lens] for _ in range(A)]
    return O.Instance([f"p{a}" for a in range(A)], lens, atoms, rng.randint(1, T - 1))


FIELDS = ("status", "min_size", "distinct", "candidates", "found_ordinal", "formula")


@pytest.mark.parametrize("seed", range(16))
def test_fullrow_equals_independent_enumerator(seed):
    rng = random.Random(3000 + seed)
    A, T = rng.randint(1, 2), rng.randint(2, 4)
    inst = _long_inst(rng, A, T, 50, 200)
    ms = 5 if A == 2 else 6
    want = sem_enumerate(inst, ms)
    got = O.enumerate_cs(inst, ms, O.KEY_FULLROW)
    for f in FIELDS:
        assert got[f] == want[f], (seed, f, got[f], want[f])
    if got["status"] == 0:
        assert separates(got["formula"], inst.names, inst.lens, inst.atoms, inst.P)
        assert size(parse(got["formula"], inst.names)) == got["min_size"]


@pytest.mark.parametrize("seed", range(16))
def test_hashed_and_exact_keys_equal_semantic_dedup(seed):
    """Type 3 over 128 bits does not collide at these sizes, and the exact key
    is injective, so both deduplicate exactly as the full row does."""
    rng = random.Random(4000 + seed)
    if seed % 2:
        inst = _long_inst(rng, rng.randint(1, 3), rng.randint(2, 5), 40, 300)
    else:   # lenSum <= 126 with a trace over 64 steps: exact key
        lens = [rng.randint(65, 100)]
        while sum(lens) < 120:
            lens.append(rng.randint(1, 126 - sum(lens)))
        atoms = [[rng.getrandbits(n) for n in lens] for _ in range(rng.randint(1, 3))]
        inst = O.Instance([f"p{a}" for a in range(len(atoms))], lens, atoms, rng.randint(1, len(lens) - 1))
        assert sum(inst.lens) <= 126
    a = O.enumerate_cs(inst, 6, 3)
    b = O.enumerate_cs(inst, 6, O.KEY_FULLROW)
    for f in FIELDS:
        assert a[f] == b[f], (seed, f)


def test_bruteforce_and_omp_reject_long_traces():
    inst = O.Instance(["p"], [70, 3], [[1, 1]], 1)
    assert O.bruteforce(inst, 3)["status"] == 2
    assert O.enumerate_omp(inst, 3)[0]["status"] == 2
