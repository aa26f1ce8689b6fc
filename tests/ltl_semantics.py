"""Textbook finite-trace LTL semantics, evaluated position by position.

Independent of the bit-parallel operators (LTL_BS.cu:183-238) that both the
oracle and the GPU kernels use: each operator is defined here directly from
its meaning on a finite word of length `n` (strong next, strict finite
semantics), so it can pin them.
"""


def holds(w, t):
    return (w >> t) & 1 == 1


def pos_not(x, n):
    return [not x[t] for t in range(n)]


def pos_and(x, y, n):
    return [x[t] and y[t] for t in range(n)]


def pos_or(x, y, n):
    return [x[t] or y[t] for t in range(n)]


def pos_next(x, n):
    # X phi holds at t iff t+1 < n and phi holds at t+1
    return [t + 1 < n and x[t + 1] for t in range(n)]


def pos_finally(x, n):
    return [any(x[k] for k in range(t, n)) for t in range(n)]


def pos_globally(x, n):
    return [all(x[k] for k in range(t, n)) for t in range(n)]


def pos_until(x, y, n):
    # phi U psi at t: some k >= t with psi(k) and phi on [t, k)
    out = []
    for t in range(n):
        ok = False
        for k in range(t, n):
            if y[k]:
                ok = all(x[m] for m in range(t, k))
                break
            # keep scanning only while phi holds at k (strict until)
            if not x[k]:
                break
        out.append(ok)
    return out


def to_bits(w, n):
    return [holds(w, t) for t in range(n)]


def from_bits(b):
    w = 0
    for t, v in enumerate(b):
        if v:
            w |= 1 << t
    return w


OPS = ["~", "&", "|", "X", "F", "G", "U"]


def pos_apply(op, a, b, n):
    x = to_bits(a, n)
    y = to_bits(b, n)
    f = {0: lambda: pos_not(x, n), 1: lambda: pos_and(x, y, n), 2: lambda: pos_or(x, y, n),
         3: lambda: pos_next(x, n), 4: lambda: pos_finally(x, n), 5: lambda: pos_globally(x, n),
         6: lambda: pos_until(x, y, n)}[op]
    return from_bits(f())


# ---- formula text (ltl_nocache.cu:289-317 syntax) ----

def parse(s, names):
    """Parses `(l) op (r)`, `op(x)` and atom names into a tree."""
    pos = [0]

    def expect(lit):
        if not s.startswith(lit, pos[0]):
            raise ValueError(f"expected {lit!r} at {pos[0]} in {s!r}")
        pos[0] += len(lit)

    def node():
        i = pos[0]
        if s[i] == "(":
            pos[0] += 1
            left = node()
            expect(") ")
            op = s[pos[0]]
            pos[0] += 1
            expect(" (")
            right = node()
            expect(")")
            return (op, left, right)
        if s[i] in "~XFG" and i + 1 < len(s) and s[i + 1] == "(":
            op = s[i]
            pos[0] += 2
            sub = node()
            expect(")")
            return (op, sub)
        best = max((n for n in names if s.startswith(n, i)), key=len, default=None)
        if best is None:
            raise ValueError(f"bad token at {i} in {s!r}")
        pos[0] += len(best)
        return ("atom", names.index(best))

    tree = node()
    if pos[0] != len(s):
        raise ValueError(f"trailing text in {s!r}")
    return tree


def size(tree):
    if tree[0] == "atom":
        return 1
    return 1 + sum(size(c) for c in tree[1:])


def eval_tree(tree, atoms_t, n):
    """Position-wise truth list of `tree` on one trace (atoms_t[a] = word)."""
    op = tree[0]
    if op == "atom":
        return to_bits(atoms_t[tree[1]], n)
    if op == "~":
        return pos_not(eval_tree(tree[1], atoms_t, n), n)
    if op == "X":
        return pos_next(eval_tree(tree[1], atoms_t, n), n)
    if op == "F":
        return pos_finally(eval_tree(tree[1], atoms_t, n), n)
    if op == "G":
        return pos_globally(eval_tree(tree[1], atoms_t, n), n)
    x = eval_tree(tree[1], atoms_t, n)
    y = eval_tree(tree[2], atoms_t, n)
    if op == "&":
        return pos_and(x, y, n)
    if op == "|":
        return pos_or(x, y, n)
    if op == "U":
        return pos_until(x, y, n)
    raise ValueError(op)


def separates(formula, names, lens, atoms, P):
    """True iff the formula holds at time 0 on every positive trace and on no
    negative trace (LTL_BS.cu:316-319)."""
    tree = parse(formula, names)
    for t, n in enumerate(lens):
        v = eval_tree(tree, [atoms[a][t] for a in range(len(names))], n)
        if v[0] != (t < P):
            return False
    return True
