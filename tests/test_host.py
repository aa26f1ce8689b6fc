"""Host-side checks of libvfb.so that need no GPU: the C-ABI exports every
symbol include/vfb.h declares, the JSON reader agrees with Python's json
module, the generator is deterministic and labels traces by its planted
formula, limits are enforced, and without a GPU a solve fails loudly
(VFB_DEVICE_ERROR) instead of falling back to the CPU."""
import ctypes as C
import json
import os
import re
import subprocess

import pytest

import vfb_amd as V
from ltl_semantics import separates, parse

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _header_functions():
    text = open(os.path.join(ROOT, "include", "vfb.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vfb_[a-z_0-9]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    L = C.CDLL(V.LIB_PATH)
    declared = _header_functions()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(V.EXPORTS) == declared


def _py_arrays(path):
    with open(path) as f:
        j = json.load(f)
    names = j["atomic_propositions"]
    traces = list(j["positive_traces"]) + list(j["negative_traces"])
    lens = [len(tr[names[0]]) for tr in traces]
    atoms = [[sum(1 << k for k, b in enumerate(tr[v]) if b) for tr in traces] for v in names]
    return names, lens, atoms, len(j["positive_traces"])


def test_json_reader_matches_python_json():
    p = os.path.join(GOLDEN, "ExampleLTLF.json")
    assert V.Instance.from_json(p).arrays() == _py_arrays(p)


@pytest.mark.parametrize("cfg,k", [(1, 0), (2, 3), (3, 1), (4, 7)])
def test_generated_roundtrip_and_labels(tmp_path, cfg, k):
    inst = V.Instance.generate(cfg, k)
    again = V.Instance.generate(cfg, k)
    assert inst.arrays() == again.arrays()          # deterministic
    path = str(tmp_path / "g.json")
    inst.write_json(path)
    assert V.Instance.from_json(path).arrays() == inst.arrays() == _py_arrays(path)
    names, lens, atoms, P = inst.arrays()
    # the planted formula labels every trace (checked per position)
    assert separates(inst.planted, names, lens, atoms, P)
    with open(path) as f:
        j = json.load(f)
    assert j["generating_seed"] == str((cfg << 32) | k)


def test_generator_shapes():
    A, T, P = V.Instance.generate(1, 0).shape()
    assert (A, T, P) == (3, 10, 5)
    A, T, P = V.Instance.generate(2, 0).shape()
    assert (A, T, P) == (5, 40, 20)
    A, T, P = V.Instance.generate(3, 0).shape()
    assert (A, T, P) == (10, 63, 32)
    for k in range(20):
        inst = V.Instance.generate(4, k)
        A, T, P = inst.shape()
        names, lens, atoms, P = inst.arrays()
        assert 2 <= A <= 8 and 8 <= T <= 63 and 1 <= P <= T - 1
        assert all(8 <= n <= 64 for n in lens)
        assert inst.default_max_size == 16
    names, lens, _, _ = V.Instance.generate(2, 5).arrays()
    assert all(16 <= n <= 32 for n in lens)
    assert "U" in V.Instance.generate(2, 5).planted
    assert V.Instance.generate(1, 3).planted == "F((a0) & (X(a1)))"


def test_limits_rejected():
    with pytest.raises(V.VfbError):
        V.Instance.from_bits(["a"], [4097], [[1]], 1)        # trace longer than VFB_MAX_LEN
    # the one-word entry point takes traces of at most 64 steps
    h = C.c_void_p()
    lib = V.lib()
    st = lib.vfb_instance_from_bits(1, (C.c_char_p * 1)(b"a"), 1, 0, (C.c_int * 1)(65), (C.c_uint64 * 1)(1),
                                    C.byref(h))
    assert st == V.INVALID_INPUT and not h.value
    with pytest.raises(V.VfbError):
        V.Instance.from_bits(["a"], [3] * 129, [[1] * 129], 1)  # more than 128 traces
    with pytest.raises(V.VfbError):
        V.Instance.from_bits(["a"], [0], [[0]], 1)           # empty trace


def test_bad_json(tmp_path):
    p = tmp_path / "bad.json"
    p.write_text('{"atomic_propositions": ["a"], "positive_traces": [{"a": [1, 0]}], '
                 '"negative_traces": [{"a": [1]}], "number_traces": 5}')
    with pytest.raises(V.VfbError):
        V.Instance.from_json(str(p))
    with pytest.raises(V.VfbError):
        V.Instance.from_json(str(tmp_path / "missing.json"))


def test_from_bits_masks_and_accessors():
    inst = V.Instance.from_bits(["x", "y"], [2, 3], [[0xFF, 0x1], [0x0, 0x7]], 1)
    names, lens, atoms, P = inst.arrays()
    assert names == ["x", "y"] and lens == [2, 3] and P == 1
    assert atoms == [[0x3, 0x1], [0x0, 0x7]]


def _has_gpu():
    return V.device_info() is not None


@pytest.mark.skipif(_has_gpu(), reason="checks the no-GPU behaviour")
def test_no_gpu_fails_loudly():
    r = V.solve(V.Instance.generate(1, 0), max_size=5)
    assert r["status"] == V.DEVICE_ERROR
    with pytest.raises(V.VfbError):
        V.Instance.generate(1, 0).upload()


def test_cli_usage_and_missing_file():
    out = subprocess.run([V.CLI_PATH], capture_output=True, text=True)
    assert out.returncode == 0 and "input_file_address" in out.stderr
    out = subprocess.run([V.CLI_PATH, "/nonexistent.json", "5"], capture_output=True, text=True)
    assert out.returncode == 0 and "Could not open" in out.stderr
    out = subprocess.run([V.CLI_PATH, os.path.join(GOLDEN, "ExampleLTLF.json"), "99"], capture_output=True, text=True)
    assert "between 1 and 32" in out.stderr


def test_parser_roundtrip():
    names = ["a0", "a1", "X1"]
    for f in ["a0", "~(a1)", "(a0) & (X(X1))", "((a0) U (G(a1))) | (F(~(X1)))"]:
        tree = parse(f, names)
        assert tree


def test_tuning_set_get_and_range_checks():
    """vfb_set_tuning accepts in-range settings, rejects others unchanged."""
    base = V.get_tuning()
    assert base == V.tuning_default() or set(base) == set(V.TUNING_FIELDS)
    try:
        V.set_tuning(pass_min=1, wide_max_waves=0, expand_blocks_per_cu=3)
        got = V.get_tuning()
        assert (got["pass_min"], got["wide_max_waves"], got["expand_blocks_per_cu"]) == (1, 0, 3)
        for bad in ({"expand_blocks_per_cu": 0}, {"resolve_blocks_per_cu": 65}, {"store_blocks_per_cu": 0},
                    {"max_load_pct": 49}, {"max_load_pct": 91}, {"store_traces": 0}, {"store_traces": 129},
                    {"check_rows": 2}, {"check_rows": 32}, {"table_min_slots": (1 << 32) + 1},
                    {"chain_blocks": 65}, {"chain_cands": 63}, {"chain_cands": 65537},
                    {"set_margin_pct": 0}, {"set_margin_pct": 1001}, {"direct_min": 0}):
            with pytest.raises(V.VfbError):
                V.set_tuning(**bad)
            assert V.get_tuning() == got
    finally:
        V.set_tuning(**base)
    assert V.get_tuning() == base


def test_tuning_file_roundtrip(tmp_path):
    t = dict(V.tuning_default(), pass_min=12345, wide_max_waves=77, store_blocks_per_cu=16)
    path = tmp_path / "tuning.txt"
    V.save_tuning(t, path)
    text = path.read_text()
    assert "pass_min 12345" in text and text.startswith("#")
    assert V.load_tuning(path) == t
    # missing keys keep their defaults; an out-of-range value is refused
    path.write_text("wide_max_waves 5\n")
    assert V.load_tuning(path) == dict(V.tuning_default(), wide_max_waves=5)
    path.write_text("expand_blocks_per_cu 0\n")
    with pytest.raises(V.VfbError):
        V.load_tuning(path)


def test_cli_brute_search_env_accepted():
    """VFB_SEARCH=brute selects ltl_nocache's search; without a GPU the CLI
    reports the device error on stderr (no CPU fallback) and exits 0."""
    env = dict(os.environ, VFB_SEARCH="brute", VFB_QUIET="1", HIP_VISIBLE_DEVICES="")
    p = subprocess.run([V.CLI_PATH, os.path.join(GOLDEN, "ExampleLTLF.json"), "4"], capture_output=True,
                       text=True, env=env, timeout=120)
    assert p.returncode == 0
    assert "device error" in p.stderr and p.stdout == ""


def test_options_have_search_field():
    o = V._opts(10, 3, False, 0, False, 0, V.SEARCH_BRUTE)
    assert o.search == V.SEARCH_BRUTE and o.max_size == 10


def _read_text_py(path):
    """Independent restatement of readFile (LTL_BS.cu:1535-1607) in Python."""
    lines = [l.rstrip("\r\n") for l in open(path)]
    while lines and not lines[-1]:
        lines.pop()
    alpha = [c for c in lines[-1] if "a" <= c <= "z"]
    letters = sorted(set(alpha))
    body = lines[:-1]
    cut = body.index("---")
    pos, rest = body[:cut], body[cut + 1:]
    neg = rest[:rest.index("---")] if "---" in rest else rest
    traces = pos + neg
    lens = [len(tr.split(";")) for tr in traces]
    atoms = [[0] * len(traces) for _ in letters]
    for t, tr in enumerate(traces):
        for k, step in enumerate(tr.split(";")):
            vals = step.split(",")
            for j, v in enumerate(vals):
                if v == "1":
                    atoms[letters.index(alpha[j])][t] |= 1 << k
    return letters, lens, atoms, len(pos)


def test_text_reader_matches_restatement_and_json():
    path = os.path.join(GOLDEN, "example_ltlbs.txt")
    inst = V.Instance.from_text(path)
    assert inst.arrays() == _read_text_py(path)
    # the same traces as ExampleLTLF.json, with a0 -> a, a1 -> b
    names, lens, atoms, P = V.Instance.from_json(os.path.join(GOLDEN, "ExampleLTLF.json")).arrays()
    assert inst.arrays() == (["a", "b"], lens, atoms, P)


def test_ltlbs_cli_argument_checks():
    path = os.path.join(GOLDEN, "example_ltlbs.txt")
    out = subprocess.run([V.CLI_PATH, path, "1", "1", "0", "1", "1", "1", "1", "1", "10", "3", "2", "1"],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "Argument number 4" in out.stdout
    out = subprocess.run([V.CLI_PATH, path] + ["1"] * 8 + ["10", "3", "4", "3"], capture_output=True, text=True,
                         timeout=60)
    assert "RlxUnqChkTyp" in out.stdout and "NegType" in out.stdout
    out = subprocess.run([V.CLI_PATH, path] + ["1"] * 8 + ["40", "3", "2", "1"], capture_output=True, text=True,
                         timeout=60)
    assert "at most 32" in out.stdout


def test_long_traces_roundtrip(tmp_path):
    """Traces over 64 steps are several CS words: from_bits / arrays keep every
    step, clear steps past the end, and the JSON writer and reader agree with
    Python's json module."""
    import random
    rng = random.Random(12)
    lens = [1, 64, 65, 128, 129, 300, 4096]
    atoms = [[rng.getrandbits(n + 7) for n in lens] for _ in range(3)]
    inst = V.Instance.from_bits(["x", "y", "z"], lens, atoms, 3)
    names, got_lens, got, P = inst.arrays()
    assert got_lens == lens and P == 3
    assert got == [[w & ((1 << n) - 1) for w, n in zip(row, lens)] for row in atoms]
    assert V.lib().vfb_instance_words(inst.handle) == sum((n + 63) // 64 for n in lens)
    path = str(tmp_path / "long.json")
    inst.write_json(path)
    assert V.Instance.from_json(path).arrays() == inst.arrays() == _py_arrays(path)


def test_text_reader_long_traces(tmp_path):
    import random
    rng = random.Random(13)
    rows = []
    for n in (3, 70, 200):
        rows.append(";".join(",".join(str(rng.randint(0, 1)) for _ in range(2)) for _ in range(n)))
    path = tmp_path / "long.txt"
    path.write_text("\n".join([rows[0], rows[1], "---", rows[2], "---", "a,b"]) + "\n")
    inst = V.Instance.from_text(str(path))
    assert inst.arrays() == _read_text_py(str(path))
    assert inst.arrays()[1] == [3, 70, 200]
