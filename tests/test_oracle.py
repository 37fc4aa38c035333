"""Pins the CPU oracle against closed forms, brute force and known answers.

The reference ships no golden vectors for this path, so each oracle function
is checked here against something the mathematics fixes, chosen so that a
plausible slip (dropped term, wrong shift, transposed operand, off-by-one in
the probe walk) fails at least one case.
"""
import itertools
import math

import numpy as np
import pytest

import graphs as G


# ---------------------------------------------------------------- length rule
def _closed_form_len(d):
    """alpha=0.2f, b=5: fl(d/0.2f) = 5d exactly while 5d < 2^24, so the bucket
    count is the next power of two strictly above d, at least 8."""
    if d == 0:
        return 8
    return max(8, 1 << (d.bit_length()))


def test_length_closed_form_exhaustive_small(oracle_lib):
    for d in list(range(0, 5000)) + [2**k + j for k in range(12, 21) for j in (-1, 0, 1)]:
        assert oracle_lib.calculate_length(d, 0.2, 5) == _closed_form_len(d), d


def test_length_closed_form_numpy_float32_exhaustive():
    """The float32 divide over every degree below 2^24/5 reproduces 5d, which
    is what makes the closed form exact on this range (a check of the claim,
    not of the oracle)."""
    d = np.arange(0, (1 << 24) // 5, dtype=np.int64)
    q = (d.astype(np.float32) / np.float32(0.2)).astype(np.int64)
    assert np.array_equal(q, 5 * d)


@pytest.mark.parametrize("alpha", [0.1, 0.15, 0.2, 0.25, 0.3, 0.5, 0.75, 0.95, 1.0])
@pytest.mark.parametrize("b", [1, 2, 4, 5, 8, 9])
def test_length_rule_matches_float32_semantics(oracle_lib, alpha, b):
    """Independent numpy float32 evaluation over random and boundary degrees."""
    rng = np.random.default_rng(int(alpha * 1000) * 10 + b)
    ds = np.concatenate([np.arange(0, 300), rng.integers(0, 1 << 24, 300),
                         (np.arange(1, 23)[:, None] ** 0 * (1 << np.arange(1, 23))[:, None] + np.array([-1, 0, 1])).ravel()])
    for d in ds.tolist():
        total = int(np.float32(d) / np.float32(alpha))
        ln = total // b
        want = 8 if ln == 0 else max(8, 1 << ln.bit_length())
        assert oracle_lib.calculate_length(d, alpha, b) == want, (d, alpha, b)


def test_length_rejects_overflowing_inputs(oracle_lib):
    assert oracle_lib.calculate_length(-1, 0.2, 5) == -1
    assert oracle_lib.calculate_length(1 << 30, 0.1, 1) == -1
    assert oracle_lib.calculate_length(5, 0.0, 5) == -1


# ---------------------------------------------------------------- hashes
def test_hierarchical_hash_host_form_and_monotonicity(oracle_lib):
    """utils.cu:5-7 host form x % M / (M / L), and monotone in x mod M."""
    rng = np.random.default_rng(1)
    for _ in range(2000):
        M = 1 << int(rng.integers(3, 20))
        L = 1 << int(rng.integers(3, int(math.log2(M)) + 1))
        x = int(rng.integers(0, 1 << 31))
        assert oracle_lib.hash_hierarchical(x, L, M) == x % M // (M // L)
        assert oracle_lib.hash_normal(x, L) == x % L
    M, L = 1 << 12, 1 << 5
    xs = sorted(rng.integers(0, 1 << 30, 500).tolist(), key=lambda v: v % M)
    hs = [oracle_lib.hash_hierarchical(v, L, M) for v in xs]
    assert hs == sorted(hs)


# ---------------------------------------------------------------- tables
@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("alpha,b", [(0.2, 5), (0.5, 2), (0.9, 1), (0.25, 4)])
def test_table_membership_and_invariants(oracle_lib, layout, alpha, b):
    rng = np.random.default_rng(7 + layout)
    off, cols = G.random_sets(rng, 200, 300, 1 << 20)
    t = oracle_lib.build_tables(off, cols, alpha, b, layout)
    assert int(np.count_nonzero(t.T != -1)) == len(cols)
    assert oracle_lib.table_check(off, cols, t) == 0
    for row in rng.integers(0, 200, 40).tolist():
        members = cols[off[row]:off[row + 1]]
        for key in members[:20].tolist():
            assert oracle_lib.table_search(t, row, key) == 1
        absent = np.setdiff1d(rng.integers(0, 1 << 20, 30), members)
        for key in absent.tolist():
            assert oracle_lib.table_search(t, row, key) == 0
    # the checker itself catches a planted misplacement
    bad = t.T.copy()
    nz = np.flatnonzero(bad != -1)
    bad[nz[0]] = -1
    assert oracle_lib.table_check(off, cols, t, T=bad) > 0


def test_table_check_flags_a_hole_in_a_bucket(oracle_lib):
    """An element moved behind an EMPTY slot of its own bucket breaks the
    bucket-prefix property (and its reachability): flagged."""
    off = np.array([0, 2], np.int32)
    cols = np.array([16, 3], np.int32)  # home buckets 0 and 3 of an 8-bucket table
    t = oracle_lib.build_tables(off, cols, 0.2, 5, 0)
    assert oracle_lib.table_check(off, cols, t) == 0
    bad = t.T.copy()
    BN = t.bucket_num
    assert bad[0] == 16 and bad[BN] == -1
    bad[0], bad[BN] = -1, 16  # slot 0 of bucket 0 emptied, the key moved to slot 1
    assert oracle_lib.table_check(off, cols, t, T=bad) > 0


def test_slot_plane_layout_position(oracle_lib):
    """Two sets, keys with a known home bucket: slot k of bucket j of table t
    sits at T[toff[t] + j + k*BN]."""
    off = np.array([0, 3, 5], np.int32)
    cols = np.array([16, 32, 48, 3, 11], np.int32)  # set 0: three keys with x & 7 == 0
    t = oracle_lib.build_tables(off, cols, 0.2, 5, 0)
    BN = t.bucket_num
    assert list(t.L) == [8, 8] and BN == 16
    assert [t.T[0 + 0 + k * BN] for k in range(3)] == [16, 32, 48]
    assert t.T[8 + 3] == 3 and t.T[8 + 3 + BN] == 11


def test_table_overflow_reported(oracle_lib):
    off = np.array([0, 100], np.int32)
    cols = np.arange(100, dtype=np.int32)
    with pytest.raises(oracle_lib.OracleError):
        oracle_lib.build_tables(off, cols, 4.0, 1, 0)  # 8 buckets x 1 slot < 100


def test_extract_probe_order_definition(oracle_lib):
    rng = np.random.default_rng(3)
    off, cols = G.random_sets(rng, 100, 200, 1 << 16)
    t = oracle_lib.build_tables(off, cols, 0.2, 5, 1)
    out = oracle_lib.extract_probe_order(off, t)
    BN, b = t.bucket_num, t.bucket_size
    planes = t.T.reshape(b, BN)  # planes[k, toff[t] + j]
    for row in range(100):
        L = int(t.L[row])
        grid = planes[:, t.toff[row]:t.toff[row] + L].T.ravel()  # (bucket, slot) order
        want = grid[grid != -1]
        assert np.array_equal(out[off[row]:off[row + 1]], want)
        assert sorted(want.tolist()) == cols[off[row]:off[row + 1]].tolist()


# ---------------------------------------------------------------- triangle counting
def _tc_all(o, off, cols, n):
    res = {"merge": o.tc_merge(off, cols)}
    for layout in (0, 1):
        t = o.build_tables(off, cols, 0.2, 5, layout)
        res[f"hash{layout}"] = o.tc_hash(off, cols, t)
        if layout == 1:
            res["hash1_sorted"] = o.tc_hash(off, cols, t, probe_cols=o.extract_probe_order(off, t))
    return res


@pytest.mark.parametrize("n", [3, 4, 5, 10, 17, 40])
def test_tc_complete_graph(oracle_lib, n):
    und = G.undirected(G.complete(n), n)
    for by in ("id", "degree"):
        off, cols = G.orient(und, n, by)
        assert set(_tc_all(oracle_lib, off, cols, n).values()) == {math.comb(n, 3)}


@pytest.mark.parametrize("r", [4, 5, 9, 33])
def test_tc_wheel(oracle_lib, r):
    und = G.undirected(G.wheel(r), r + 1)
    off, cols = G.orient(und, r + 1)
    assert set(_tc_all(oracle_lib, off, cols, r + 1).values()) == {r}


@pytest.mark.parametrize("edges,n", [(G.star(50), 50), (G.complete_bipartite(7, 9), 16),
                                     ([(i, i + 1) for i in range(30)], 31)])
def test_tc_triangle_free(oracle_lib, edges, n):
    off, cols = G.orient(G.undirected(edges, n), n)
    assert set(_tc_all(oracle_lib, off, cols, n).values()) == {0}


@pytest.mark.parametrize("seed", range(6))
def test_tc_bruteforce_random(oracle_lib, seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(5, 70))
    und = G.undirected(G.random_graph(n, int(rng.integers(n, 6 * n)), rng), n)
    want = G.brute_triangles(und, n)
    for by in ("id", "degree"):
        off, cols = G.orient(und, n, by)
        assert set(_tc_all(oracle_lib, off, cols, n).values()) == {want}


# ---------------------------------------------------------------- pairs
def test_pairs_known_answers(oracle_lib):
    rng = np.random.default_rng(5)
    a_off, a, b_off, b, k = G.make_pairs(rng, 300, 500, 1 << 24)
    for f in (oracle_lib.pairs_hash, oracle_lib.pairs_merge):
        assert np.array_equal(f(a_off, a, b_off, b), k)
    # A with itself -> |A|; A against a disjoint shift -> 0
    assert np.array_equal(oracle_lib.pairs_hash(a_off, a, a_off, a), np.diff(a_off))
    assert not oracle_lib.pairs_hash(a_off, a * 2, a_off, a * 2 + 1).any()


def test_pairs_against_python_sets(oracle_lib):
    rng = np.random.default_rng(6)
    a_off, a, _, _, _ = G.make_pairs(rng, 50, 400, 1000)
    b_off, b, _, _, _ = G.make_pairs(rng, 50, 400, 1000)
    want = [len(set(a[a_off[i]:a_off[i + 1]].tolist()) & set(b[b_off[i]:b_off[i + 1]].tolist()))
            for i in range(50)]
    for alpha, bs in [(0.2, 5), (0.5, 2), (0.8, 1)]:
        assert oracle_lib.pairs_hash(a_off, a, b_off, b, alpha, bs).tolist() == want


# ---------------------------------------------------------------- set similarity
@pytest.mark.parametrize("p,q,t", [(4, 5, 0.8), (1, 4, 0.25), (1, 2, 0.5)])
def test_integer_threshold_equals_float_rule(oracle_lib, p, q, t):
    """sss.cu:130-133 decides with float32; the exact rule q*r >= p*U agrees
    whenever U < 2^22 (r/U != t differs from t by >= 1/(qU)).  Exhaustive for
    U <= 4000 through numpy float32, sampled through the oracle's C rule."""
    U = np.arange(1, 4001, dtype=np.int64)
    r = np.concatenate([np.arange(0, u + 1) for u in U])
    UU = np.repeat(U, U + 1)
    flt = (r.astype(np.float32) / UU.astype(np.float32)) >= np.float32(t)
    exact = q * r >= p * UU
    assert np.array_equal(flt, exact)
    rng = np.random.default_rng(p * 10 + q)
    for _ in range(3000):
        su, sv = (int(x) for x in rng.integers(1, 2000, 2))
        rr = int(rng.integers(0, min(su, sv) + 1))
        assert oracle_lib.sss_float_rule(rr, su, sv, t) == int(q * rr >= p * (su + sv - rr))


def test_sss_verify_hash_matches_bruteforce(oracle_lib):
    rng = np.random.default_rng(11)
    off, cols = G.zipf_sets(rng, 150, 300, 2, 40)
    n = 150
    pairs = np.array(list(itertools.combinations(range(n), 2)), np.int32)
    for p, q in [(4, 5), (1, 2), (1, 4)]:
        want = set(oracle_lib.sss_bruteforce(off, cols, p, q).tolist())
        for layout in (0, 1):
            t = oracle_lib.build_tables(off, cols, 0.2, 5, layout)
            flags, sizes = oracle_lib.sss_verify_hash(off, cols, t, pairs[:, 0], pairs[:, 1], p, q)
            got = {(int(i) << 32) | int(j) for (i, j), f in zip(pairs, flags) if f}
            assert got == want
            # sizes agree with Python set intersection on a sample
            for e in rng.integers(0, len(pairs), 50).tolist():
                i, j = pairs[e]
                assert sizes[e] == len(set(cols[off[i]:off[i + 1]].tolist()) & set(cols[off[j]:off[j + 1]].tolist()))


@pytest.mark.parametrize("p,q", [(4, 5), (1, 2), (2, 3)])
def test_prefix_filter_join_equals_bruteforce(oracle_lib, p, q):
    rng = np.random.default_rng(p * 7 + q)
    off, cols = G.zipf_sets(rng, 400, 200, 2, 30)
    # plant near-duplicates so the result is non-trivial
    rows = [cols[off[i]:off[i + 1]].tolist() for i in range(400)]
    for i in range(0, 60, 2):
        r = list(rows[i])
        if len(r) > 3:
            r[-1] = 10_000 + i
        rows.append(sorted(set(r)))
    off2 = np.zeros(len(rows) + 1, np.int64)
    np.cumsum([len(r) for r in rows], out=off2[1:])
    cols2 = np.concatenate([np.array(r, np.int32) for r in rows])
    want = oracle_lib.sss_bruteforce(off2, cols2, p, q)
    got, ncand = oracle_lib.sss_prefix_join(off2, cols2, p, q)
    assert len(want) > 10
    assert np.array_equal(got, want)
    n = len(rows)
    assert ncand < n * (n - 1) // 2


# ---------------------------------------------------------------- information retrieval
def test_ir_hash_equals_chained_merge_and_sets(oracle_lib):
    rng = np.random.default_rng(21)
    idx_off, docs = G.inverted_index(rng, 120, 5000, 400)
    q, qoff = G.queries(rng, idx_off, 200, 1, 6)
    m_res, m_cnt, roff = oracle_lib.ir_merge(idx_off, docs, q, qoff)
    for i in range(200):
        terms = q[qoff[i]:qoff[i + 1]].tolist()
        want = set(docs[idx_off[terms[0]]:idx_off[terms[0] + 1]].tolist())
        for tt in terms[1:]:
            want &= set(docs[idx_off[tt]:idx_off[tt + 1]].tolist())
        assert m_res[roff[i]:roff[i] + m_cnt[i]].tolist() == sorted(want)
    for layout in (0, 1):
        t = oracle_lib.build_tables(idx_off, docs, 0.2, 5, layout)
        res, cnt, roff2 = oracle_lib.ir_hash(idx_off, docs, t, q, qoff)
        assert np.array_equal(cnt, m_cnt) and np.array_equal(roff2, roff)
        assert oracle_lib.ir_result_lists(res, cnt, roff) == oracle_lib.ir_result_lists(m_res, m_cnt, roff)
        if layout == 1:
            pd = oracle_lib.extract_probe_order(idx_off, t)
            res2, cnt2, _ = oracle_lib.ir_hash(idx_off, docs, t, q, qoff, probe_docs=pd)
            assert np.array_equal(cnt2, m_cnt)
            for i in range(200):
                got = res2[roff[i]:roff[i] + cnt2[i]].tolist()
                assert sorted(got) == m_res[roff[i]:roff[i] + m_cnt[i]].tolist()


def test_ir_special_cases(oracle_lib):
    idx_off = np.array([0, 3, 3, 6], np.int32)  # term 1 has an empty list
    docs = np.array([1, 5, 9, 5, 9, 12], np.int32)
    q = np.array([0, 0, 1, 0, 2], np.int32)
    qoff = np.array([0, 1, 3, 5], np.int32)
    t = oracle_lib.build_tables(idx_off, docs, 0.2, 5, 0)
    res, cnt, roff = oracle_lib.ir_hash(idx_off, docs, t, q, qoff)
    assert oracle_lib.ir_result_lists(res, cnt, roff) == [[1, 5, 9], [], [5, 9]]


def test_sss_lookup_bruteforce_agrees_with_selfjoin(oracle_lib):
    """Looking every row up against the collection finds itself and exactly
    the self-join partners, in both orders."""
    rng = np.random.default_rng(21)
    off, cols = G.zipf_sets(rng, 300, 200, 3, 25)
    for p, q in [(4, 5), (1, 2)]:
        pairs = oracle_lib.sss_bruteforce(off, cols, p, q)
        want = {(i, i) for i in range(len(off) - 1) if off[i + 1] > off[i]}
        for k in pairs.tolist():
            i, j = k >> 32, k & 0xFFFFFFFF
            want |= {(i, j), (j, i)}
        got = oracle_lib.sss_lookup_bruteforce(off, cols, off, cols, p, q)
        assert [(k >> 32, k & 0xFFFFFFFF) for k in got.tolist()] == sorted(want)


def test_tc_row_ranges_add_up(oracle_lib):
    """Disjoint source-row ranges partition the triangles (each is counted at
    its source vertex)."""
    rng = np.random.default_rng(12)
    n = 300
    und = G.undirected(G.random_graph(n, 5000, rng), n)
    for by in ("degree", "id"):
        off, cols = G.orient(und, n, by)
        total = oracle_lib.tc_merge(off, cols)
        cuts = [0, 37, 120, 121, 250, n]
        parts = [oracle_lib.tc_merge_rows(off, cols, a, b) for a, b in zip(cuts, cuts[1:])]
        assert sum(parts) == total == G.brute_triangles(und, n)
        # the hash count over the same row ranges: range by range equal to the
        # merge count, in both layouts
        for layout in (0, 1):
            t = oracle_lib.build_tables(off, cols, 0.2, 5, layout)
            hparts = [oracle_lib.tc_hash_rows(off, cols, t, a, b) for a, b in zip(cuts, cuts[1:])]
            assert hparts == parts
            assert oracle_lib.tc_hash_rows(off, cols, t, 0, n) == oracle_lib.tc_hash(off, cols, t) == total
    with pytest.raises(oracle_lib.OracleError):
        oracle_lib.tc_hash_rows(off, cols, t, 5, n + 1)
