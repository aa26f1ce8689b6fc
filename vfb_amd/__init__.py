"""vfb_amd -- MI355X-native LTL synthesis from positive/negative traces.

Python mirror of the C-ABI in include/vfb.h (libvfb.so, built in-tree under
vfb_amd/lib).  The hot path -- size-ordered enumeration of characteristic sets
with 128-bit keys, GPU hash-set dedup and target check -- runs in HIP kernels
for gfx950; nothing here computes on the CPU.  Loading fails loudly when the
library has not been built.

Reference interface mirrored: `LTLI(...)` (LTL_BS.cu:1069-1079) and the
`ltl_nocache <json_file> <maxLen>` process contract (ltl_nocache.cu:448-487).
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libvfb.so")
CLI_PATH = os.path.join(_HERE, "bin", "vfb")

MAX_SIZE = 32
MAX_TRACES = 128
MAX_ATOMS = 64
MAX_LEN = 4096
NOALLOC_SCRATCH_BYTES = 256 << 20
OK, NOT_FOUND, INVALID_INPUT, CAPACITY, DEVICE_ERROR, IO_ERROR = range(6)
OP_NAMES = ["~", "&", "|", "X", "F", "G", "U", "atom"]


class Options(C.Structure):
    _fields_ = [("max_size", C.c_int), ("relaxed_type", C.c_int), ("verbose", C.c_int),
                ("timing", C.c_int), ("hbm_budget", C.c_uint64), ("max_rows", C.c_uint64),
                ("search", C.c_int), ("bs_max_size", C.c_int),
                ("cost", C.c_uint16 * 8), ("neg_type", C.c_int), ("pad_opt", C.c_int)]


SEARCH_CS, SEARCH_BRUTE = 0, 1


class Result(C.Structure):
    _fields_ = [("status", C.c_int), ("min_size", C.c_int), ("top_op", C.c_int),
                ("levels_done", C.c_int),
                ("distinct", C.c_uint64 * (MAX_SIZE + 1)),
                ("candidates", C.c_uint64 * (MAX_SIZE + 1)),
                ("level_winners", C.c_uint64 * (MAX_SIZE + 1)),
                ("level_candidates", C.c_uint64 * (MAX_SIZE + 1)),
                ("found_ordinal", C.c_uint64),
                ("formula", C.c_char_p),
                ("ms_total", C.c_double), ("ms_expand", C.c_double),
                ("ms_resolve", C.c_double), ("ms_store", C.c_double),
                ("expand_launches", C.c_uint64), ("table_slots", C.c_uint64),
                ("retries", C.c_uint64), ("bytes_expand", C.c_uint64),
                ("bytes_store", C.c_uint64), ("fresh_total", C.c_uint64),
                ("otf_level", C.c_int), ("bs_stage", C.c_int), ("otf_stored", C.c_uint64),
                ("ms_probe", C.c_double), ("probe_candidates", C.c_uint64),
                ("restarts", C.c_uint32), ("parked", C.c_uint32), ("chain_levels", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("pass_min", C.c_uint64), ("wide_max_waves", C.c_uint32),
                ("expand_blocks_per_cu", C.c_uint32), ("resolve_blocks_per_cu", C.c_uint32),
                ("store_blocks_per_cu", C.c_uint32), ("park_min", C.c_uint64),
                ("table_min_slots", C.c_uint64), ("max_load_pct", C.c_uint32),
                ("store_traces", C.c_uint32), ("check_rows", C.c_uint32), ("chain_blocks", C.c_uint32),
                ("chain_cands", C.c_uint32), ("lds_chain", C.c_uint32), ("set_margin_pct", C.c_uint32),
                ("pad_tuning", C.c_uint32), ("direct_min", C.c_uint64)]


TUNING_FIELDS = [f for f, _ in Tuning._fields_ if not f.startswith("pad")]


class KernelStats(C.Structure):
    _fields_ = [("ms_expand", C.c_double), ("expand_launches", C.c_uint64), ("bytes_expand", C.c_uint64),
                ("ms_probe", C.c_double), ("probe_launches", C.c_uint64), ("bytes_probe", C.c_uint64),
                ("ms_resolve", C.c_double), ("ms_store", C.c_double), ("store_launches", C.c_uint64),
                ("bytes_store", C.c_uint64), ("ms_chain", C.c_double), ("chain_launches", C.c_uint64),
                ("chain_levels", C.c_uint64), ("bytes_chain", C.c_uint64), ("chain_solo_launches", C.c_uint64)]


# every symbol declared in include/vfb.h
EXPORTS = [
    "vfb_instance_load_json", "vfb_instance_load_text", "vfb_instance_from_bits", "vfb_instance_from_words",
    "vfb_instance_generate", "vfb_instance_write_json", "vfb_instance_shape", "vfb_instance_words", "vfb_instance_copy",
    "vfb_instance_atom_name", "vfb_instance_planted", "vfb_instance_max_size",
    "vfb_instance_upload", "vfb_instance_free", "vfb_options_default", "vfb_solve", "vfb_solve_batch",
    "vfb_result_free", "vfb_status_string", "vfb_device_info",
    "vfb_server_create", "vfb_server_submit", "vfb_server_poll", "vfb_server_wait",
    "vfb_server_stats", "vfb_server_destroy",
    "vfb_tuning_default", "vfb_get_tuning", "vfb_set_tuning", "vfb_tuning_save",
    "vfb_tuning_load", "vfb_autotune", "vfb_set_arena_bytes", "vfb_solve_noalloc", "vfb_kernel_stats_get",
]

_lib = None


def lib():
    """Loads libvfb.so; raises if it was not built (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libvfb.so not built ({LIB_PATH}); run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.vfb_instance_load_json.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.vfb_instance_load_text.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.vfb_instance_from_bits.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.c_int, C.c_int,
                                         C.POINTER(C.c_int), C.POINTER(C.c_uint64), C.POINTER(vp)]
    L.vfb_instance_from_words.argtypes = L.vfb_instance_from_bits.argtypes
    L.vfb_instance_words.argtypes = [vp]
    L.vfb_instance_words.restype = C.c_int
    L.vfb_instance_generate.argtypes = [C.c_int, C.c_uint64, C.POINTER(vp)]
    L.vfb_instance_write_json.argtypes = [vp, C.c_char_p]
    L.vfb_instance_shape.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.vfb_instance_shape.restype = None
    L.vfb_instance_copy.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_uint64)]
    L.vfb_instance_copy.restype = None
    L.vfb_instance_atom_name.argtypes = [vp, C.c_int]
    L.vfb_instance_atom_name.restype = C.c_char_p
    L.vfb_instance_planted.argtypes = [vp]
    L.vfb_instance_planted.restype = C.c_char_p
    L.vfb_instance_max_size.argtypes = [vp]
    L.vfb_instance_upload.argtypes = [vp]
    L.vfb_instance_free.argtypes = [vp]
    L.vfb_instance_free.restype = None
    L.vfb_options_default.argtypes = [C.POINTER(Options)]
    L.vfb_options_default.restype = None
    L.vfb_solve.argtypes = [vp, C.POINTER(Options), C.POINTER(Result)]
    L.vfb_solve_noalloc.argtypes = [vp, C.POINTER(Options), C.POINTER(Result)]
    L.vfb_solve_batch.argtypes = [C.POINTER(vp), C.c_size_t, C.POINTER(Options), C.POINTER(Result)]
    L.vfb_result_free.argtypes = [C.POINTER(Result)]
    L.vfb_result_free.restype = None
    L.vfb_status_string.argtypes = [C.c_int]
    L.vfb_status_string.restype = C.c_char_p
    L.vfb_device_info.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.c_int)]
    L.vfb_server_create.argtypes = [C.POINTER(Options), C.c_uint64, C.c_int, C.POINTER(vp)]
    L.vfb_server_submit.argtypes = [vp, vp, C.POINTER(C.c_uint64)]
    L.vfb_server_poll.argtypes = [vp, C.c_uint64, C.POINTER(Result), C.POINTER(C.c_int)]
    L.vfb_server_wait.argtypes = [vp, C.c_uint64, C.POINTER(Result)]
    L.vfb_server_stats.argtypes = [vp] + [C.POINTER(C.c_uint64)] * 4
    L.vfb_server_stats.restype = None
    L.vfb_server_destroy.argtypes = [vp]
    L.vfb_server_destroy.restype = None
    TP = C.POINTER(Tuning)
    L.vfb_tuning_default.argtypes = [TP]
    L.vfb_tuning_default.restype = None
    L.vfb_get_tuning.argtypes = [TP]
    L.vfb_get_tuning.restype = None
    L.vfb_set_tuning.argtypes = [TP]
    L.vfb_tuning_save.argtypes = [TP, C.c_char_p]
    L.vfb_tuning_load.argtypes = [C.c_char_p, TP]
    L.vfb_set_arena_bytes.argtypes = [C.c_uint64]
    L.vfb_kernel_stats_get.argtypes = [C.POINTER(KernelStats), C.c_int]
    L.vfb_kernel_stats_get.restype = None
    L.vfb_autotune.argtypes = [C.POINTER(vp), C.c_size_t, C.POINTER(Options), C.c_int, C.c_int,
                               C.c_char_p, TP, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    _lib = L
    return L


class VfbError(RuntimeError):
    pass


def status_string(s):
    return lib().vfb_status_string(s).decode()


class Instance:
    """Owning handle of a vfb_instance."""

    def __init__(self, handle):
        if not handle:
            raise VfbError("null instance")
        self._h = C.c_void_p(handle)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value and _lib is not None:
            _lib.vfb_instance_free(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def upload(self):
        """Keeps the atom rows resident in device memory: later solve() calls
        on this instance read them there instead of uploading them."""
        st = lib().vfb_instance_upload(self._h)
        if st != OK:
            raise VfbError(f"upload: {status_string(st)}")
        return self

    @classmethod
    def from_json(cls, path):
        h = C.c_void_p()
        st = lib().vfb_instance_load_json(os.fsencode(path), C.byref(h))
        if st != OK:
            raise VfbError(f"load_json({path}): {status_string(st)}")
        return cls(h.value)

    @classmethod
    def from_text(cls, path):
        """LTL_BS's text trace format (readFile, LTL_BS.cu:1535-1607)."""
        h = C.c_void_p()
        st = lib().vfb_instance_load_text(os.fsencode(path), C.byref(h))
        if st != OK:
            raise VfbError(f"load_text({path}): {status_string(st)}")
        return cls(h.value)

    @classmethod
    def generate(cls, config, index):
        h = C.c_void_p()
        st = lib().vfb_instance_generate(int(config), int(index), C.byref(h))
        if st != OK:
            raise VfbError(f"generate({config},{index}): {status_string(st)}")
        return cls(h.value)

    @classmethod
    def from_bits(cls, names, lens, atoms, num_pos):
        """atoms[a][t]: AP a on trace t as an integer, bit k = step k (a trace
        of more than 64 steps is several CS words); traces 0..num_pos-1 are
        positive."""
        A, T = len(names), len(lens)
        nw = [(n + 63) // 64 for n in lens]
        words = [(int(atoms[a][t]) >> (64 * w)) & 0xFFFFFFFFFFFFFFFF
                 for a in range(A) for t in range(T) for w in range(nw[t])]
        flat = (C.c_uint64 * max(1, len(words)))(*words)
        nm = (C.c_char_p * A)(*[n.encode() for n in names])
        h = C.c_void_p()
        st = lib().vfb_instance_from_words(A, nm, num_pos, T - num_pos, (C.c_int * T)(*lens), flat, C.byref(h))
        if st != OK:
            raise VfbError(f"from_bits: {status_string(st)}")
        return cls(h.value)

    def shape(self):
        a, t, p = C.c_int(), C.c_int(), C.c_int()
        lib().vfb_instance_shape(self._h, C.byref(a), C.byref(t), C.byref(p))
        return a.value, t.value, p.value

    def arrays(self):
        """(names, lens, atoms[a][t] as integers (bit k = step k), P)"""
        A, T, P = self.shape()
        W = lib().vfb_instance_words(self._h)
        lens = (C.c_int * T)()
        flat = (C.c_uint64 * (A * W))()
        lib().vfb_instance_copy(self._h, lens, flat)
        names = [lib().vfb_instance_atom_name(self._h, a).decode() for a in range(A)]
        col = [0]
        for n in lens:
            col.append(col[-1] + (n + 63) // 64)
        atoms = [[sum(flat[a * W + c] << (64 * (c - col[t])) for c in range(col[t], col[t + 1]))
                  for t in range(T)] for a in range(A)]
        return names, list(lens), atoms, P

    @property
    def planted(self):
        return lib().vfb_instance_planted(self._h).decode()

    @property
    def default_max_size(self):
        return lib().vfb_instance_max_size(self._h)

    def write_json(self, path):
        st = lib().vfb_instance_write_json(self._h, os.fsencode(path))
        if st != OK:
            raise VfbError(f"write_json: {status_string(st)}")


_DEFAULT_OPTS = None
_OPTS_CACHE = {}


def _opts(max_size, relaxed_type, timing, budget, verbose, max_rows=0, search=SEARCH_CS, bs_max_size=0,
          costs=None, neg_type=1):
    """vfb_options for these arguments (the library only reads them, so equal
    arguments share one structure)."""
    key = (max_size, relaxed_type, timing, budget, verbose, max_rows, search, bs_max_size,
           None if costs is None else tuple(costs), neg_type)
    o = _OPTS_CACHE.get(key)
    if o is None:
        o = _make_opts(*key)
        if len(_OPTS_CACHE) < 256:
            _OPTS_CACHE[key] = o
    return o


def _make_opts(max_size, relaxed_type, timing, budget, verbose, max_rows, search, bs_max_size, costs, neg_type):
    global _DEFAULT_OPTS
    if _DEFAULT_OPTS is None:
        d = Options()
        lib().vfb_options_default(C.byref(d))
        _DEFAULT_OPTS = bytes(d)
    o = Options.from_buffer_copy(_DEFAULT_OPTS)
    o.search = search
    o.bs_max_size = bs_max_size
    if costs is not None:
        for k in range(8):
            o.cost[k] = int(costs[k])
    o.neg_type = neg_type
    o.max_size = max_size
    o.relaxed_type = relaxed_type
    # True / 1: every kernel class; 2: the expand kernel only (fewer events)
    o.timing = int(timing) if not isinstance(timing, bool) else (1 if timing else 0)
    o.verbose = int(verbose) if not isinstance(verbose, bool) else (1 if verbose else 0)
    o.hbm_budget = int(budget)
    o.max_rows = int(max_rows)
    return o


def _todict(r):
    n = MAX_SIZE + 1
    return {
        "status": r.status, "min_size": r.min_size, "top_op": r.top_op,
        "levels_done": r.levels_done,
        "distinct": r.distinct[:n],
        "candidates": r.candidates[:n],
        "level_winners": r.level_winners[:n],
        "level_candidates": r.level_candidates[:n],
        "found_ordinal": r.found_ordinal,
        "formula": (r.formula or b"").decode(),
        "ms_total": r.ms_total, "ms_expand": r.ms_expand,
        "ms_resolve": r.ms_resolve, "ms_store": r.ms_store,
        "expand_launches": r.expand_launches, "table_slots": r.table_slots,
        "retries": r.retries, "bytes_expand": r.bytes_expand,
        "bytes_store": r.bytes_store, "fresh_total": r.fresh_total,
        "otf_level": r.otf_level, "otf_stored": r.otf_stored, "bs_stage": r.bs_stage,
        "ms_probe": r.ms_probe, "probe_candidates": r.probe_candidates,
        "restarts": r.restarts, "parked": r.parked, "chain_levels": r.chain_levels,
    }


def solve(inst, max_size=MAX_SIZE, relaxed_type=3, timing=False, budget=0, verbose=False, max_rows=0,
          search=SEARCH_CS, bs_max_size=0, costs=None, neg_type=1):
    """Plain single-instance solve (the drop-in for LTLI; search=SEARCH_BRUTE
    runs ltl_nocache's syntactic brute force instead; bs_max_size enables the
    BS stage after an unsuccessful search; costs = LTL_BS's c1..c8 with
    max_size as maxCost, neg_type its NegType)."""
    o = _opts(max_size, relaxed_type, timing, budget, verbose, max_rows, search, bs_max_size, costs, neg_type)
    r = Result()
    lib().vfb_solve(inst.handle, C.byref(o), C.byref(r))
    try:
        return _todict(r)
    finally:
        lib().vfb_result_free(C.byref(r))


def solve_noalloc(inst, max_size=MAX_SIZE, relaxed_type=3, budget=0, max_rows=0, bs_max_size=0, costs=None,
                  neg_type=1):
    """solve() through vfb_solve_noalloc: no device allocation, everything in
    the library's static device scratch (NOALLOC_SCRATCH_BYTES per GPU); a
    solve that needs more returns status CAPACITY."""
    o = _opts(max_size, relaxed_type, False, budget, False, max_rows, SEARCH_CS, bs_max_size, costs, neg_type)
    r = Result()
    lib().vfb_solve_noalloc(inst.handle, C.byref(o), C.byref(r))
    try:
        return _todict(r)
    finally:
        lib().vfb_result_free(C.byref(r))


def solve_batch(insts, max_size=MAX_SIZE, relaxed_type=3, timing=False, budget=0, max_rows=0, bs_max_size=0,
                costs=None, neg_type=1):
    """Many instances stepped together; each result equals solve() alone."""
    n = len(insts)
    o = _opts(max_size, relaxed_type, timing, budget, False, max_rows, SEARCH_CS, bs_max_size, costs, neg_type)
    hs = (C.c_void_p * n)(*[i.handle.value for i in insts])
    rs = (Result * n)()
    lib().vfb_solve_batch(hs, n, C.byref(o), rs)
    out = []
    for k in range(n):
        out.append(_todict(rs[k]))
        lib().vfb_result_free(C.byref(rs[k]))
    return out


def _tdict(t):
    return {f: getattr(t, f) for f in TUNING_FIELDS}


def tuning_default():
    t = Tuning()
    lib().vfb_tuning_default(C.byref(t))
    return _tdict(t)


def get_tuning():
    """Launch / pass parameters of the process (speed only, never results)."""
    t = Tuning()
    lib().vfb_get_tuning(C.byref(t))
    return _tdict(t)


def set_tuning(**kw):
    """Changes the given vfb_tuning fields; returns the previous settings."""
    old = get_tuning()
    t = Tuning(**{**old, **kw})
    st = lib().vfb_set_tuning(C.byref(t))
    if st != OK:
        raise VfbError(f"set_tuning: {status_string(st)}")
    return old


def save_tuning(tuning, path):
    st = lib().vfb_tuning_save(C.byref(Tuning(**tuning)), str(path).encode())
    if st != OK:
        raise VfbError(f"save_tuning: {status_string(st)}")


def load_tuning(path):
    t = Tuning()
    st = lib().vfb_tuning_load(str(path).encode(), C.byref(t))
    if st != OK:
        raise VfbError(f"load_tuning: {status_string(st)}")
    return _tdict(t)


def autotune(insts, max_size=MAX_SIZE, batched=True, reps=2, cache=None, relaxed_type=3):
    """Runs the library's autotuner on a sample; the best setting stays active."""
    n = len(insts)
    o = _opts(max_size, relaxed_type, False, 0, False)
    hs = (C.c_void_p * n)(*[i.handle.value for i in insts])
    best = Tuning()
    t0, t1 = C.c_double(), C.c_double()
    st = lib().vfb_autotune(hs, n, C.byref(o), 1 if batched else 0, reps,
                            str(cache).encode() if cache else None, C.byref(best), C.byref(t0), C.byref(t1))
    if st != OK:
        raise VfbError(f"autotune: {status_string(st)}")
    return {"tuning": _tdict(best), "ms_default": t0.value, "ms_best": t1.value}


def kernel_stats(reset=False):
    """Process-wide kernel time (HIP events of solves run with timing) and
    algorithmic bytes, counted once per launch."""
    s = KernelStats()
    lib().vfb_kernel_stats_get(C.byref(s), 1 if reset else 0)
    return {f: getattr(s, f) for f, _ in KernelStats._fields_}


def set_arena_bytes(nbytes):
    """Device arena of later solves (0 = 85% of the free HBM, the default)."""
    st = lib().vfb_set_arena_bytes(int(nbytes))
    if st != OK:
        raise VfbError(f"set_arena_bytes: {status_string(st)}")


def device_info():
    buf = C.create_string_buffer(256)
    cus = C.c_int()
    st = lib().vfb_device_info(buf, 256, C.byref(cus))
    if st != OK:
        return None
    return {"name": buf.value.decode(), "cus": cus.value}


class Server:
    """Continuous-batching request loop (vfb_server_*): requests are admitted
    as slots and HBM budget free up, and running requests share launches."""

    def __init__(self, max_size=MAX_SIZE, relaxed_type=3, budget=0, max_active=256, request_budget=0,
                 max_rows=0, bs_max_size=0):
        # the library refuses bs_max_size != 0 (the server steps CS levels only)
        o = _opts(max_size, relaxed_type, False, request_budget, False, max_rows, SEARCH_CS, bs_max_size)
        h = C.c_void_p()
        st = lib().vfb_server_create(C.byref(o), int(budget), int(max_active), C.byref(h))
        if st != OK:
            raise VfbError(f"server_create: {status_string(st)}")
        self._h = h

    def submit(self, inst):
        t = C.c_uint64()
        st = lib().vfb_server_submit(self._h, inst.handle, C.byref(t))
        if st != OK:
            raise VfbError(f"server_submit: {status_string(st)}")
        return t.value

    def poll(self, ticket):
        """The result if the request has finished (the ticket is then
        consumed), else None; VfbError for an unknown or consumed ticket."""
        r = Result()
        done = C.c_int()
        st = lib().vfb_server_poll(self._h, ticket, C.byref(r), C.byref(done))
        if not done.value:
            if st == INVALID_INPUT:
                raise VfbError(f"server_poll: unknown or consumed ticket {ticket}")
            return None
        try:
            return _todict(r)
        finally:
            lib().vfb_result_free(C.byref(r))

    def wait(self, ticket):
        """Blocks until the request has finished; VfbError for an unknown or
        consumed ticket, or when the server stops first."""
        r = Result()
        st = lib().vfb_server_wait(self._h, ticket, C.byref(r))
        if st == INVALID_INPUT and not r.formula:
            raise VfbError(f"server_wait: unknown or consumed ticket {ticket}, or the server stopped")
        try:
            return _todict(r)
        finally:
            lib().vfb_result_free(C.byref(r))

    def stats(self):
        v = [C.c_uint64() for _ in range(4)]
        lib().vfb_server_stats(self._h, *[C.byref(x) for x in v])
        return {"queued": v[0].value, "active": v[1].value, "bytes": v[2].value, "peak_active": v[3].value}

    def close(self):
        if self._h:
            lib().vfb_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
