"""GPU parity under hash collisions: every key a multiple of 4096 (1024 for
the self-join, whose token ids stay below 2^22).

The home bucket is x & (L-1) (normal layout) or (x & (M-1)) >> shift
(hierarchical), so for tables of at most 4096 buckets (M <= 4096) every key
of these inputs lands in bucket 0: builds fill a chain of full buckets and
probes walk it (hashtable.cu:53-101, TC.cu:11-31).  Each path is checked
against the oracle's merge or brute-force result on the same inputs.
"""
import itertools

import numpy as np
import pytest
import torch

import graphs as G

pytestmark = pytest.mark.gpu

STEP = 4096


@pytest.fixture(scope="module")
def api():
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from reset_amd import api as A

    A.load()
    return A


def _cuda(x, dtype=torch.int32):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dtype).cuda()


def _colliding_sets(rng, count, lo, hi, vocab=1000, step=STEP):
    """CSR of sorted sets of tokens step*i, i < vocab."""
    rows = [np.sort(rng.choice(vocab, int(rng.integers(lo, hi + 1)), replace=False)) * step
            for _ in range(count)]
    off = np.zeros(count + 1, np.int64)
    np.cumsum([len(r) for r in rows], out=off[1:])
    return off.astype(np.int32), np.concatenate(rows).astype(np.int32)


def test_tc_colliding_out_lists(api, oracle_lib):
    """K_300 on the vertices 4096*i, inside a sparse random graph: the clique
    vertices' tables hold only bucket-0 keys (M stays below 4096), in both
    layouts, staged in LDS and read in HBM."""
    rng = np.random.default_rng(3)
    n = 300 * STEP
    clique = [(STEP * i, STEP * j) for i, j in itertools.combinations(range(300), 2)]
    extra = rng.integers(0, n, size=(30000, 2))
    und = G.undirected(np.concatenate([np.array(clique), extra]), n)
    off, cols = G.orient(und, n, "degree")
    want = oracle_lib.tc_merge(off, cols)
    assert want >= 300 * 299 * 298 // 6
    t = api.HashTables(_cuda(off), _cuda(cols), 0.2, 5, 3)
    staged, units, lmax = t.tc_staged()
    assert staged > 0 and lmax <= STEP
    keys_h = t.probe_order(1)
    assert t.tc_count(0) == want
    assert t.tc_count(1) == want
    assert t.tc_count(1, probe_keys=keys_h) == want
    assert t.tc_count(0, plain=True) == want
    assert t.tc_count(1, probe_keys=keys_h, plain=True) == want
    t.close()


@pytest.mark.parametrize("p,q", [(1, 2), (4, 5)])
def test_sss_verify_colliding_sets(api, oracle_lib, p, q):
    rng = np.random.default_rng(10 + p)
    off, cols = _colliding_sets(rng, 120, 2, 250, vocab=400)
    pairs = np.array(list(itertools.combinations(range(120), 2)), np.int32)
    want = set(oracle_lib.sss_bruteforce(off, cols, p, q).tolist())
    ot = oracle_lib.build_tables(off, cols, 0.2, 5, 0)
    o_flags, o_sizes = oracle_lib.sss_verify_hash(off, cols, ot, pairs[:, 0], pairs[:, 1], p, q)
    t = api.HashTables(_cuda(off), _cuda(cols), 0.2, 5, 3)
    for layout, keys in ((0, None), (1, t.probe_order(1))):
        flags, sizes = t.sss_verify(_cuda(pairs[:, 0]), _cuda(pairs[:, 1]), p, q, layout, keys)
        assert np.array_equal(sizes.cpu().numpy(), o_sizes)
        assert np.array_equal(flags.cpu().numpy(), o_flags)
        got = {(int(i) << 32) | int(j) for (i, j), f in zip(pairs, flags.cpu().numpy()) if f}
        assert got == want
    t.close()


def test_sss_selfjoin_colliding_sets(api, oracle_lib):
    """Self-join over colliding sets, with near-duplicates so that pairs pass
    J >= 4/5.  Token ids of a self-join stay below 2^22, so the tokens are
    multiples of 1024: the grouped LDS tables of these sets (<= 512 buckets)
    still put every key in bucket 0."""
    rng = np.random.default_rng(21)
    off, cols = _colliding_sets(rng, 300, 3, 300, vocab=3000, step=1024)
    rows = [cols[off[i]:off[i + 1]].tolist() for i in range(300)]
    for i in range(60):
        r = list(rows[i])
        if len(r) > 5:
            r[int(rng.integers(0, len(r)))] = 1024 * (3000 + i)
        rows.append(sorted(set(r)))
    o = np.zeros(len(rows) + 1, np.int64)
    np.cumsum([len(r) for r in rows], out=o[1:])
    off2, cols2 = o.astype(np.int32), np.concatenate([np.array(r, np.int32) for r in rows])
    want = oracle_lib.sss_bruteforce(off2, cols2, 4, 5)
    assert len(want) >= 30
    got = api.sss_selfjoin(_cuda(off2), _cuda(cols2), 4, 5)
    assert np.array_equal(got["pairs"].cpu().numpy(), want)


def test_ir_colliding_postings(api, oracle_lib):
    rng = np.random.default_rng(31)
    idx_off, docs = _colliding_sets(rng, 60, 1, 3500, vocab=4000)
    q, qoff = G.queries(rng, idx_off, 200, 1, 5)
    m_res, m_cnt, m_roff = oracle_lib.ir_merge(idx_off, docs, q, qoff)
    assert int(m_cnt.sum()) > 0
    t = api.HashTables(_cuda(idx_off), _cuda(docs), 0.2, 5, 3)
    for layout, term0 in ((0, None), (1, t.probe_order(1))):
        res, cnt, roff = t.ir_query(_cuda(q), _cuda(qoff), layout, term0_docs=term0)
        assert np.array_equal(cnt.cpu().numpy(), m_cnt)
        got = oracle_lib.ir_result_lists(res.cpu().numpy(), cnt.cpu().numpy(), roff.cpu().numpy())
        want = oracle_lib.ir_result_lists(m_res, m_cnt, m_roff)
        assert [sorted(x) for x in got] == [sorted(x) for x in want]
    t.close()
