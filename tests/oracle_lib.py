"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: only tests, the smoke check and the CPU-baseline leg of
bench.py load the oracle.  The JSON reader here uses Python's json module so
that it is independent of the product's own C++ reader.
"""
import ctypes as C
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

NOT, AND, OR, NEXT, FINALLY, GLOBALLY, UNTIL, ATOM = range(8)
KEY_FULLROW = 100
MAX_SIZE = 32


class OraInstance(C.Structure):
    _fields_ = [("A", C.c_int), ("T", C.c_int), ("P", C.c_int),
                ("len", C.POINTER(C.c_int)), ("atoms", C.POINTER(C.c_uint64))]


class OraResult(C.Structure):
    _fields_ = [("status", C.c_int), ("min_size", C.c_int), ("top_op", C.c_int),
                ("levels_done", C.c_int),
                ("distinct", C.c_uint64 * (MAX_SIZE + 1)),
                ("candidates", C.c_uint64 * (MAX_SIZE + 1)),
                ("found_ordinal", C.c_uint64),
                ("otf_level", C.c_int), ("bs_stage", C.c_int), ("otf_stored", C.c_uint64),
                ("formula", C.c_char * 8192)]


LIB_OMP = os.path.join(ROOT, "oracle", "_build", "libcpu_omp.so")

_lib = None
_omp = None


def omp_lib():
    """The multi-core CPU build of the enumeration (oracle/cpu_omp.c)."""
    global _omp
    if _omp is None:
        if not os.path.exists(LIB_OMP):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _omp = C.CDLL(LIB_OMP)
        _omp.ora_enumerate_omp.restype = C.c_int
        _omp.ora_enumerate_omp.argtypes = [C.POINTER(OraInstance), C.c_int, C.c_int, C.c_uint64,
                                           C.POINTER(C.c_char_p), C.POINTER(OraResult)]
        _omp.ora_omp_threads.restype = C.c_int
        _omp.ora_enumerate_many_omp.restype = C.c_int
        _omp.ora_enumerate_many_omp.argtypes = [C.POINTER(C.POINTER(OraInstance)),
                                                C.POINTER(C.POINTER(C.c_char_p)), C.c_int, C.c_int, C.c_int,
                                                C.POINTER(OraResult)]
    return _omp


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _lib = C.CDLL(LIB)
        _lib.ora_mask.restype = C.c_uint64
        _lib.ora_mask.argtypes = [C.c_int]
        _lib.ora_apply.restype = C.c_uint64
        _lib.ora_apply.argtypes = [C.c_int, C.c_uint64, C.c_uint64, C.c_int]
        _lib.ora_mix64.restype = C.c_uint64
        _lib.ora_mix64.argtypes = [C.c_uint64]
        _lib.ora_apply_trace.restype = None
        _lib.ora_apply_trace.argtypes = [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_int,
                                         C.POINTER(C.c_uint64)]
        _lib.ora_key.restype = None
        _lib.ora_key.argtypes = [C.POINTER(C.c_uint64), C.POINTER(C.c_int), C.c_int, C.c_int,
                                 C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        _lib.ora_enumerate.restype = C.c_int
        _lib.ora_enumerate.argtypes = [C.POINTER(OraInstance), C.c_int, C.c_int, C.c_uint64,
                                       C.POINTER(C.c_char_p), C.POINTER(OraResult)]
        _lib.ora_enumerate_bs.restype = C.c_int
        _lib.ora_enumerate_bs.argtypes = [C.POINTER(OraInstance), C.c_int, C.c_int, C.c_int,
                                          C.POINTER(C.c_char_p), C.POINTER(OraResult)]
        _lib.ora_enumerate_costs.restype = C.c_int
        _lib.ora_enumerate_costs.argtypes = [C.POINTER(OraInstance), C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int,
                                             C.POINTER(C.c_char_p), C.POINTER(OraResult)]
        _lib.ora_bruteforce.restype = C.c_int
        _lib.ora_bruteforce.argtypes = [C.POINTER(OraInstance), C.c_int,
                                        C.POINTER(C.c_char_p), C.POINTER(OraResult)]
        _lib.ora_num_formulas.restype = None
        _lib.ora_num_formulas.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_uint64)]
    return _lib


def nwords(n):
    """64-bit words of a trace of n steps."""
    return (n + 63) // 64


def split_words(x, n):
    """Python int of an n-step trace -> its words, step 64w+k = bit k of word w."""
    return [(x >> (64 * w)) & 0xFFFFFFFFFFFFFFFF for w in range(nwords(n))]


def join_words(ws):
    return sum(int(w) << (64 * k) for k, w in enumerate(ws))


class Instance:
    """Plain-Python instance: names, lens, atoms[a][t] (a Python int, bit k =
    step k, any length), P positives first."""

    def __init__(self, names, lens, atoms, P):
        self.names = list(names)
        self.lens = list(lens)
        self.atoms = [list(r) for r in atoms]
        self.P = P

    @property
    def A(self):
        return len(self.names)

    @property
    def T(self):
        return len(self.lens)

    @classmethod
    def from_json(cls, path):
        with open(path) as f:
            j = json.load(f)
        names = j["atomic_propositions"]
        traces = list(j["positive_traces"]) + list(j["negative_traces"])
        lens = [len(tr[names[0]]) for tr in traces]
        atoms = []
        for v in names:
            row = []
            for tr in traces:
                w = 0
                for k, b in enumerate(tr[v]):
                    if b:
                        w |= 1 << k
                row.append(w)
            atoms.append(row)
        return cls(names, lens, atoms, len(j["positive_traces"]))

    def _c(self):
        T = self.T
        lens = (C.c_int * T)(*self.lens)
        words = [w for row in self.atoms for t, x in enumerate(row) for w in split_words(x, self.lens[t])]
        flat = (C.c_uint64 * len(words))(*words)
        names = (C.c_char_p * self.A)(*[n.encode() for n in self.names])
        inst = OraInstance(self.A, T, self.P, lens, flat)
        return inst, (lens, flat, names), names


def _result(r):
    return {
        "status": r.status, "min_size": r.min_size, "top_op": r.top_op,
        "levels_done": r.levels_done,
        "distinct": list(r.distinct), "candidates": list(r.candidates),
        "found_ordinal": r.found_ordinal,
        "otf_level": r.otf_level, "otf_stored": r.otf_stored, "bs_stage": r.bs_stage,
        "formula": r.formula.decode(),
    }


def enumerate_cs(inst, max_size, key_mode=3, max_rows=0):
    ci, keep, names = inst._c()
    r = OraResult()
    lib().ora_enumerate(C.byref(ci), max_size, key_mode, max_rows, names, C.byref(r))
    return _result(r)


def enumerate_omp(inst, max_size, key_mode=3, pass_min=1 << 22):
    """Multi-core build; returns (result, seconds of the solve alone)."""
    import time
    ci, keep, names = inst._c()
    r = OraResult()
    L = omp_lib()
    t0 = time.perf_counter()
    L.ora_enumerate_omp(C.byref(ci), max_size, key_mode, pass_min, names, C.byref(r))
    return _result(r), time.perf_counter() - t0


def enumerate_many_omp(insts, max_size, key_mode=3):
    """Batch CPU baseline: OpenMP over instances, each sequential; returns
    (results, seconds of the whole call)."""
    import time
    n = len(insts)
    conv = [i._c() for i in insts]
    cis = (C.POINTER(OraInstance) * n)(*[C.pointer(c[0]) for c in conv])
    nms = (C.POINTER(C.c_char_p) * n)(*[C.cast(c[2], C.POINTER(C.c_char_p)) for c in conv])
    out = (OraResult * n)()
    L = omp_lib()
    t0 = time.perf_counter()
    rc = L.ora_enumerate_many_omp(cis, nms, n, max_size, key_mode, out)
    dt = time.perf_counter() - t0
    if rc != 0:
        raise RuntimeError("ora_enumerate_many_omp failed")
    return [_result(out[k]) for k in range(n)], dt


def omp_threads():
    return omp_lib().ora_omp_threads()


def enumerate_bs(inst, max_size, bs_max_size, key_mode=3):
    """CS enumeration, then the BS stage up to bs_max_size if nothing matched."""
    ci, keep, names = inst._c()
    r = OraResult()
    lib().ora_enumerate_bs(C.byref(ci), max_size, key_mode, bs_max_size, names, C.byref(r))
    return _result(r)


def enumerate_costs(inst, max_cost, costs, neg_type=1, key_mode=3):
    """CS enumeration with LTL_BS's cost function (p ~ & | X F G U) and NegType."""
    ci, keep, names = inst._c()
    r = OraResult()
    lib().ora_enumerate_costs(C.byref(ci), max_cost, key_mode, (C.c_int * 8)(*costs), neg_type, names, C.byref(r))
    return _result(r)


def bruteforce(inst, max_size):
    ci, keep, names = inst._c()
    r = OraResult()
    lib().ora_bruteforce(C.byref(ci), max_size, names, C.byref(r))
    return _result(r)


def apply(op, l, r, length):
    return lib().ora_apply(op, l, r, length)


def mix64(x):
    return lib().ora_mix64(x)


def apply_trace(op, l, r, length):
    """op over one trace of `length` steps given as Python ints (any length)."""
    n = nwords(length)
    out = (C.c_uint64 * n)()
    lw = (C.c_uint64 * n)(*split_words(l, length))
    rw = (C.c_uint64 * n)(*split_words(r, length))
    lib().ora_apply_trace(op, lw, rw, length, out)
    return join_words(out)


def key(cs, lens, relaxed_type=3):
    """Key of a CS row: cs[t] is trace t's Python int (any length)."""
    T = len(cs)
    words = [w for t, x in enumerate(cs) for w in split_words(x, lens[t])]
    hi, lo = C.c_uint64(), C.c_uint64()
    lib().ora_key((C.c_uint64 * len(words))(*words), (C.c_int * T)(*lens), T, relaxed_type,
                  C.byref(hi), C.byref(lo))
    return hi.value, lo.value


def num_formulas(A, max_size):
    buf = (C.c_uint64 * ((max_size + 1) * 8))()
    lib().ora_num_formulas(A, max_size, buf)
    return [[buf[s * 8 + k] for k in range(8)] for s in range(max_size + 1)]
