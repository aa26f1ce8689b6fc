"""The drop-in process contract end to end on the GPU: `vfb <json_file>
[max_formula_size]` (ltl_nocache.cu:448-487) prints "<file>, <formula>" on
stdout, or "<file>, Not found !" (ltl_nocache.cu:484), and exits 0.  The
printed formula must be the oracle's, bit for bit, and it must separate the
traces under the per-position semantics."""
import os
import subprocess

import pytest

import oracle_lib as O
import vfb_amd as V
from ltl_semantics import separates, parse, size

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _cli(args, **env_extra):
    env = dict(os.environ, VFB_QUIET="1", **env_extra)
    print(f"[cli] {' '.join(args)} {env_extra}", flush=True)
    out = subprocess.run([V.CLI_PATH] + args, capture_output=True, text=True, env=env, timeout=120)
    assert out.returncode == 0, out.stderr
    return out


def _oracle(inst, max_size):
    names, lens, atoms, P = inst.arrays()
    return O.enumerate_cs(O.Instance(names, lens, atoms, P), max_size, 3), (names, lens, atoms, P)


def test_example_json_formula_line():
    path = os.path.join(GOLDEN, "ExampleLTLF.json")
    inst = V.Instance.from_json(path)
    want, (names, lens, atoms, P) = _oracle(inst, 10)
    assert want["status"] == V.OK
    out = _cli([path, "10"])
    assert out.stdout == f"{path}, {want['formula']}\n", out.stdout
    assert separates(want["formula"], names, lens, atoms, P)
    assert size(parse(want["formula"], names)) == 10


def test_example_json_not_found_below_minimum():
    path = os.path.join(GOLDEN, "ExampleLTLF.json")
    out = _cli([path, "9"])
    assert out.stdout == f"{path}, Not found !\n", out.stdout


@pytest.mark.parametrize("k", range(3))
def test_config1_default_size(tmp_path, k):
    """No size argument: the bound defaults to the maximum (32); the search
    stops at the minimal size, which the oracle finds within size 10."""
    inst = V.Instance.generate(1, k)
    path = str(tmp_path / f"cfg1_{k}.json")
    inst.write_json(path)
    want, _ = _oracle(V.Instance.from_json(path), 10)
    assert want["status"] == V.OK
    out = _cli([path])
    assert out.stdout == f"{path}, {want['formula']}\n", out.stdout


def test_brute_search_formula_line(tmp_path):
    """VFB_SEARCH=brute: ltl_nocache's own syntactic search through the CLI."""
    path = str(tmp_path / "cfg1_0.json")
    V.Instance.generate(1, 0).write_json(path)
    names, lens, atoms, P = V.Instance.from_json(path).arrays()
    want = O.bruteforce(O.Instance(names, lens, atoms, P), 6)
    assert want["status"] == V.OK
    out = _cli([path, "6"], VFB_SEARCH="brute")
    assert out.stdout == f"{path}, {want['formula']}\n", out.stdout


@pytest.mark.parametrize("rt", [1, 2, 3])
def test_relaxed_type_env(tmp_path, rt):
    """VFB_RELAXED_TYPE picks the relaxed key (total length 965 > 126 bits).
    On this instance types 1 and 2 return G(F(...)) and type 3 F(... G(...)),
    so the line shows which key the GPU used."""
    path = str(tmp_path / "cfg2_1.json")
    V.Instance.generate(2, 1).write_json(path)
    names, lens, atoms, P = V.Instance.from_json(path).arrays()
    assert sum(lens) > 126
    want = O.enumerate_cs(O.Instance(names, lens, atoms, P), 14, rt)
    assert want["status"] == V.OK
    out = _cli([path, "14"], VFB_RELAXED_TYPE=str(rt))
    assert out.stdout == f"{path}, {want['formula']}\n", out.stdout
