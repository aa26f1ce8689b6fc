"""Instances whose hashed (type-3) half-keys reach the values -1 and -2.

WarpCore's set refuses the empty key (-1) and the tombstone (-2); the
reference then uses slot id 0 for either (hash_set.cuh:198-204 and the
`H = (H > 0) ? H : -H` of LTL_BS.cu:262-265), so two CS whose hi halves are
-1 and -2 and whose lo halves are equal are one key there.  Built here by
inverting the SplitMix64 finaliser (its xorshifts and odd multipliers are
bijections on 64-bit words):

    four traces of 64 steps, two atoms a and b, P = 2 (traces 0, 1 positive);
    the type-3 key of a CS row is hi = mix(w0) ^ mix(w1), lo = mix(w2) ^ mix(w3)
    (traces i < T/2 feed hi, LTL_BS.cu:78-88);
    ~a has hi = -1, ~b has hi = -2 and both have the same lo (a and b agree on
    the traces of the other half), while ~a != ~b as rows.

So at level 2 the reference (and the fold restated in the oracle and the
kernels) keeps ~a and drops ~b, one distinct CS fewer than a full-row dedup.
`half="lo"` builds the same with the roles of the halves exchanged.
"""
import random

M64 = (1 << 64) - 1
C1 = 0xBF58476D1CE4E5B9
C2 = 0x94D049BB133111EB


def mix64(x):
    """SplitMix64 finaliser (Steele, Lea & Flood 2014), published constants."""
    x &= M64
    x = ((x ^ (x >> 30)) * C1) & M64
    x = ((x ^ (x >> 27)) * C2) & M64
    return x ^ (x >> 31)


def _unxorshift(y, s):
    x = y
    for _ in range(64 // s + 1):
        x = y ^ (x >> s)
    return x


def unmix64(z):
    """Inverse of mix64."""
    z = _unxorshift(z & M64, 31)
    z = (z * pow(C2, -1, 1 << 64)) & M64
    z = _unxorshift(z, 27)
    z = (z * pow(C1, -1, 1 << 64)) & M64
    return _unxorshift(z, 30)


def type3_key(words):
    """Unfolded type-3 key of a row of one-word traces (T/2 traces feed hi)."""
    T = len(words)
    hi = lo = 0
    for i, w in enumerate(words):
        if i < T // 2:
            hi ^= mix64(w)
        else:
            lo ^= mix64(w)
    return hi, lo


def _separates_at_bit0(pattern):
    # traces 0, 1 positive, 2, 3 negative
    return pattern == (1, 1, 0, 0)


def fold_instance(half="hi", seed=0):
    """(names, lens, atoms, P, nots): atoms[a][t] one-word traces; nots = the
    rows of ~a and ~b."""
    rng = random.Random(0xF01D + seed * 7 + (0 if half == "hi" else 1000))
    while True:
        c0 = rng.getrandbits(64)
        c1 = unmix64(mix64(c0) ^ M64)          # pair XOR = -1
        d0 = rng.getrandbits(64)
        d1 = unmix64(mix64(d0) ^ (M64 - 1))    # pair XOR = -2
        e0, e1 = rng.getrandbits(64), rng.getrandbits(64)
        if half == "hi":
            na, nb = [c0, c1, e0, e1], [d0, d1, e0, e1]
        else:
            na, nb = [e0, e1, c0, c1], [e0, e1, d0, d1]
        a = [~w & M64 for w in na]
        b = [~w & M64 for w in nb]
        # no atom and no level-2 candidate may separate the traces, so that
        # level 2 is complete: atoms and ~ (bit 0), X (bit 1); F and G of
        # random 64-step words are all 1 / all 0 at bit 0
        bad = False
        for x in (a, b):
            for sh in (0, 1):
                p = tuple((w >> sh) & 1 for w in x)
                q = tuple(1 - v for v in p)
                bad |= _separates_at_bit0(p) or _separates_at_bit0(q)
            bad |= any(w == 0 or w == M64 for w in x)
        if not bad and na != nb:
            return ["a", "b"], [64] * 4, [a, b], 2, (na, nb)
