"""GPU parity through the command lines: dataset folders in the reference's
`.bin` layout, the tools' printed results against the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import graphs as G

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "reset_amd", "bin")


def _run(args):
    exe = os.path.join(BIN, args[0])
    assert os.path.exists(exe), "command-line tools not built (python -m reset_amd.build)"
    r = subprocess.run([exe, *map(str, args[1:])], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _num(pattern, text):
    m = re.search(pattern, text)
    assert m, text
    return int(m.group(1))


@pytest.mark.parametrize("flags", [[], ["--alpha=0.5", "--bucket=2"], ["--alpha=0.25", "--bucket=3"]])
def test_tc_cli_matches_oracle(tmp_path, oracle_lib, flags):
    from reset_amd import binfmt

    rng = np.random.default_rng(5)
    n = 3000
    clique = np.array([(i, j) for i in range(60) for j in range(i + 1, 60)])
    und = G.undirected(np.concatenate([G.random_graph(n, 60000, rng), clique]), n)
    off, cols = G.orient(und, n, "degree")
    want = oracle_lib.tc_merge(off, cols)
    binfmt.write_tc_folder(tmp_path, off, cols)
    out = _run(["tc", tmp_path, *flags])
    assert _num(r"Normal hash - Triangles: (\d+)", out) == want
    assert _num(r"Hierarchical hash - Triangles: (\d+)", out) == want


def test_sss_cli_matches_float_rule(tmp_path, oracle_lib):
    """All directed edges of a symmetric graph are the candidates; a pair
    counts when r / (float)(|u|+|v|-r) >= 0.25f (sss.cu:146-149)."""
    from reset_amd import binfmt

    rng = np.random.default_rng(9)
    n = 800
    clique = np.array([(i, j) for i in range(40) for j in range(i + 1, 40)])
    und = G.undirected(np.concatenate([G.random_graph(n, 6000, rng), clique]), n)
    off, cols = G.symmetric(und, n)
    vert = np.repeat(np.arange(n, dtype=np.int32), np.diff(off))
    tabs = oracle_lib.build_tables(off, cols, 0.2, 5, 0)
    _, sizes = oracle_lib.sss_verify_hash(off, cols, tabs, vert, cols, 1, 4)
    deg = np.diff(off)
    want = sum(oracle_lib.sss_float_rule(int(r), int(deg[u]), int(deg[v]), 0.25)
               for r, u, v in zip(sizes.tolist(), vert.tolist(), cols.tolist()))
    assert want > 0
    binfmt.write_sss_folder(tmp_path, off, cols)
    out = _run(["sss", tmp_path])
    assert _num(r"greater than threshold: (\d+)", out) == want


def test_cli_competitor_methods_agree(tmp_path, oracle_lib):
    """--method=original,cuco (tc, ir) and --method=all (sss): RESET's tables
    and the competitor tables print the same counts (baseline/tc/main.cu,
    baseline/sss/main.cu, baseline/ir/main.cu)."""
    from reset_amd import binfmt

    rng = np.random.default_rng(21)
    n = 2000
    clique = np.array([(i, j) for i in range(50) for j in range(i + 1, 50)])
    und = G.undirected(np.concatenate([G.random_graph(n, 30000, rng), clique]), n)
    off, cols = G.orient(und, n, "id")
    want = oracle_lib.tc_merge(off, cols)
    binfmt.write_tc_folder(tmp_path / "tc", off, cols)
    out = _run(["tc", tmp_path / "tc", "--method=original,cuco"])
    assert _num(r"Hierarchical hash - Triangles: (\d+)", out) == want
    assert _num(r"\[cuCollections\] Triangles: (\d+)", out) == want
    soff, scols = G.symmetric(und, n)
    binfmt.write_sss_folder(tmp_path / "sss", soff, scols)
    out = _run(["sss", tmp_path / "sss", "--method=all"])
    base = _num(r"(?m)^num of pairs whose similarity is greater than threshold: (\d+)", out)
    for label in ("cuCollections", "cuckoo", "hopscotch", "roaring"):
        assert _num(r"\[" + label + r"\] num of pairs whose similarity is greater than threshold: (\d+)", out) == base
    assert "Wrong Answer" not in out
    idx_off, docs = G.inverted_index(rng, 200, 100000, 8000)
    q, qoff = G.queries(rng, idx_off, 150, 1, 5)
    _, m_cnt, _ = oracle_lib.ir_merge(idx_off, docs, q, qoff)
    binfmt.write_ir_folder(tmp_path / "ir", idx_off, docs, q, qoff)
    out = _run(["ir", tmp_path / "ir", "--method=original,cuco"])
    assert _num(r"(?m)^Total result count: (\d+)", out) == int(m_cnt.sum())
    assert _num(r"\[cuCollections\] Total result count: (\d+)", out) == int(m_cnt.sum())


def test_ir_cli_matches_chained_merge(tmp_path, oracle_lib):
    from reset_amd import binfmt

    rng = np.random.default_rng(3)
    idx_off, docs = G.inverted_index(rng, 500, 20000, 3000)
    query, qoff = G.queries(rng, idx_off, 300, 1, 5)
    _, cnt, _ = oracle_lib.ir_merge(idx_off, docs, query, qoff)
    binfmt.write_ir_folder(tmp_path, idx_off, docs, query, qoff)
    out = _run(["ir", tmp_path])
    assert _num(r"Total result count: (\d+)", out) == int(cnt.sum())


@pytest.mark.parametrize("na,nb", [(1000, 5000), (70000, 30), (0, 10)])
def test_intersection_cli_forms_agree(na, nb):
    out = _run(["intersection", na, nb, 11])
    assert _num(r"Intersection Result Count: (\d+)", out) >= 0


@pytest.mark.parametrize("flags", [[], ["--alpha=0.5", "--bucket=2"]])
def test_standalone_make_hash_table_form(tmp_path, oracle_lib, flags):
    """A TC main in the reference's free-function form (globals + make_hash_table,
    triangle-count/TC.cuh:18-38) linked with tc_hashtable.o: the published
    device globals are the handle's tables, and both layouts count the
    oracle's triangles."""
    from reset_amd import binfmt

    rng = np.random.default_rng(13)
    n = 2000
    und = G.undirected(G.random_graph(n, 30000, rng), n)
    off, cols = G.orient(und, n, "degree")
    binfmt.write_tc_folder(tmp_path, off, cols)
    out = _run(["standalone_tc", tmp_path, *flags])
    want = oracle_lib.tc_merge(off, cols)
    assert "Globals match the handle: yes" in out
    assert _num(r"Normal triangle count: (\d+)", out) == want
    assert _num(r"Hierarchical triangle count: (\d+)", out) == want
