"""GPU parity: set-similarity verification and IR queries / serving loop."""
import itertools
import time

import numpy as np
import pytest
import torch

import graphs as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from reset_amd import api as A

    A.load()
    return A


def _cuda(x, dtype=torch.int32):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dtype).cuda()


# ---------------------------------------------------------------- SSS verify
@pytest.mark.parametrize("p,q", [(4, 5), (1, 2), (1, 4)])
def test_sss_verify_all_pairs_matches_oracle(api, oracle_lib, p, q):
    rng = np.random.default_rng(p * 10 + q)
    off, cols = G.zipf_sets(rng, 160, 300, 2, 60)
    pairs = np.array(list(itertools.combinations(range(160), 2)), np.int32)
    t_gpu = api.HashTables(_cuda(off), _cuda(cols), 0.2, 5, 3)
    want_pairs = set(oracle_lib.sss_bruteforce(off, cols, p, q).tolist())
    ot = oracle_lib.build_tables(off, cols, 0.2, 5, 0)
    o_flags, o_sizes = oracle_lib.sss_verify_hash(off, cols, ot, pairs[:, 0], pairs[:, 1], p, q)
    for layout, keys in ((0, None), (1, t_gpu.probe_order(1))):
        flags, sizes = t_gpu.sss_verify(_cuda(pairs[:, 0]), _cuda(pairs[:, 1]), p, q, layout, keys)
        assert np.array_equal(sizes.cpu().numpy(), o_sizes)
        assert np.array_equal(flags.cpu().numpy(), o_flags)
        got = {(int(i) << 32) | int(j) for (i, j), f in zip(pairs, flags.cpu().numpy()) if f}
        assert got == want_pairs
    t_gpu.close()


def test_sss_graph_edges_reference_semantics(api, oracle_lib):
    """The reference's SSS: candidates are the CSR edges of the full
    undirected adjacency (both directions), threshold 0.25 (sss.cu:257-260)."""
    from reset_amd import inputs

    lo, hi, n = inputs.rmat_undirected(12, 16, seed=99)
    off_d, cols_d = inputs.symmetric_csr(lo, hi, n)
    t = api.HashTables(off_d, cols_d, 0.2, 5, 3)
    from reset_amd.api import expand_rows

    rows = expand_rows(off_d)
    flags, sizes = t.sss_verify(rows, cols_d, 1, 4)
    off, cols = off_d.cpu().numpy(), cols_d.cpu().numpy()
    ot = oracle_lib.build_tables(off, cols, 0.2, 5, 1)
    o_flags, o_sizes = oracle_lib.sss_verify_hash(off, cols, ot, rows.cpu().numpy(), cols, 1, 4)
    assert np.array_equal(sizes.cpu().numpy(), o_sizes)
    assert np.array_equal(flags.cpu().numpy(), o_flags)
    # the float rule of the reference gives the same decision on every pair
    su = np.diff(off)[rows.cpu().numpy()]
    sv = np.diff(off)[cols]
    r = o_sizes
    flt = (r.astype(np.float32) / (su + sv - r).astype(np.float32)) >= np.float32(0.25)
    assert np.array_equal(flt.astype(np.int32), o_flags)
    t.close()


def test_sss_large_candidates_split_across_tiles(api, oracle_lib):
    rng = np.random.default_rng(4)
    big = np.sort(rng.choice(1 << 24, 150000, replace=False))
    other = np.sort(np.concatenate([big[:90000], rng.choice(1 << 24, 30000, replace=False) + (1 << 24)]))
    other = np.unique(other)
    rows = [big, other, big[:10], big[5:20]]
    off = np.zeros(5, np.int64)
    np.cumsum([len(x) for x in rows], out=off[1:])
    cols = np.concatenate(rows).astype(np.int32)
    cand = np.array([[0, 1], [1, 0], [2, 3], [0, 2], [3, 1]], np.int32)
    t = api.HashTables(_cuda(off.astype(np.int32)), _cuda(cols), 0.2, 5, 1)
    flags, sizes = t.sss_verify(_cuda(cand[:, 0]), _cuda(cand[:, 1]), 1, 2)
    want = [len(np.intersect1d(rows[i], rows[j])) for i, j in cand]
    assert sizes.cpu().tolist() == want
    t.close()


def test_sss_rejects_out_of_range_candidates(api):
    from reset_amd.lib import ResetError

    off = np.array([0, 2, 4], np.int32)
    cols = np.array([1, 2, 2, 3], np.int32)
    t = api.HashTables(_cuda(off), _cuda(cols), 0.2, 5, 1)
    with pytest.raises(ResetError):
        t.sss_verify(_cuda(np.array([0], np.int32)), _cuda(np.array([5], np.int32)), 1, 2)
    t.close()


# ---------------------------------------------------------------- IR
def _ir_case(rng, nterms=300, ndocs=200000, maxlen=20000, nq=400, kmax=6):
    idx_off, docs = G.inverted_index(rng, nterms, ndocs, maxlen)
    q, qoff = G.queries(rng, idx_off, nq, 1, kmax)
    return idx_off, docs, q, qoff


@pytest.mark.parametrize("alpha,b", [(0.2, 5), (0.1, 1), (0.2, 2)])
def test_ir_query_matches_oracle_bit_exact(api, oracle_lib, alpha, b):
    """Both layouts against the oracle's hash path over the same tables and
    the chained merge; (0.1, 1) is the solver's pick for one lane per key,
    the IR benchmark's tables."""
    rng = np.random.default_rng(8)
    idx_off, docs, q, qoff = _ir_case(rng)
    t = api.HashTables(_cuda(idx_off), _cuda(docs), alpha, b, 3)
    ot = oracle_lib.build_tables(idx_off, docs, alpha, b, 0)
    o_res, o_cnt, o_roff = oracle_lib.ir_hash(idx_off, docs, ot, q, qoff)
    m_res, m_cnt, _ = oracle_lib.ir_merge(idx_off, docs, q, qoff)
    res, cnt, roff = t.ir_query(_cuda(q), _cuda(qoff), 0)
    assert np.array_equal(roff.cpu().numpy(), o_roff)
    assert np.array_equal(cnt.cpu().numpy(), o_cnt) and np.array_equal(o_cnt, m_cnt)
    got = oracle_lib.ir_result_lists(res.cpu().numpy(), cnt.cpu().numpy(), o_roff)
    assert got == oracle_lib.ir_result_lists(o_res, o_cnt, o_roff)
    # hierarchical tables with term0 read in probe order: same result sets, in
    # exactly the order the oracle gives for the same term0 order
    keys = t.probe_order(1)
    res_h, cnt_h, _ = t.ir_query(_cuda(q), _cuda(qoff), 1, term0_docs=keys)
    o_res_h, o_cnt_h, _ = oracle_lib.ir_hash(idx_off, docs, ot, q, qoff, probe_docs=keys.cpu().numpy())
    assert oracle_lib.ir_result_lists(res_h.cpu().numpy(), cnt_h.cpu().numpy(), o_roff) == \
        oracle_lib.ir_result_lists(o_res_h, o_cnt_h, o_roff)
    t.close()


def test_ir_long_term0_lists_multi_segment(api, oracle_lib):
    rng = np.random.default_rng(2)
    ndocs = 3_000_000
    lists = [np.sort(rng.choice(ndocs, n, replace=False)) for n in (250000, 300000, 120000, 9000)]
    idx_off = np.zeros(5, np.int64)
    np.cumsum([len(x) for x in lists], out=idx_off[1:])
    docs = np.concatenate(lists).astype(np.int32)
    idx_off = idx_off.astype(np.int32)
    q = np.array([0, 1, 0, 1, 2, 3, 0, 0], np.int32)
    qoff = np.array([0, 2, 5, 6, 7, 8], np.int32)  # {0,1} {0,1,2} {3} {0} {0}
    t = api.HashTables(_cuda(idx_off), _cuda(docs), 0.2, 5, 1)
    res, cnt, roff = t.ir_query(_cuda(q), _cuda(qoff), 0)
    m_res, m_cnt, m_roff = oracle_lib.ir_merge(idx_off, docs, q, qoff)
    assert np.array_equal(cnt.cpu().numpy(), m_cnt)
    assert oracle_lib.ir_result_lists(res.cpu().numpy(), cnt.cpu().numpy(), m_roff) == \
        oracle_lib.ir_result_lists(m_res, m_cnt, m_roff)
    t.close()


@pytest.mark.parametrize("alpha,b", [(0.2, 5), (0.1, 1)])
@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("steal_min", [0, 1, 7])
def test_ir_work_stealing_schedule_matches_merge(api, oracle_lib, layout, steal_min, alpha, b):
    """The work-stealing segment schedule (forced on small batches through
    the steal threshold) returns the chained merge's lists, ascending, for
    both layouts with term0 read from the index, on 5-slot and on one-slot
    tables."""
    rng = np.random.default_rng(30 + steal_min)
    idx_off, docs, q, qoff = _ir_case(rng, nterms=200, ndocs=400000, maxlen=40000, nq=300, kmax=8)
    t = api.HashTables(_cuda(idx_off), _cuda(docs), alpha, b, 3)
    m_res, m_cnt, m_roff = oracle_lib.ir_merge(idx_off, docs, q, qoff)
    try:
        api.ir_set_steal_threshold(steal_min)
        res, cnt, roff = t.ir_query(_cuda(q), _cuda(qoff), layout)
    finally:
        api.ir_set_steal_threshold(-1)
    assert np.array_equal(cnt.cpu().numpy(), m_cnt)
    assert oracle_lib.ir_result_lists(res.cpu().numpy(), cnt.cpu().numpy(), m_roff) == \
        oracle_lib.ir_result_lists(m_res, m_cnt, m_roff)
    t.close()


def test_ir_special_queries(api, oracle_lib):
    idx_off = np.array([0, 3, 3, 6], np.int32)
    docs = np.array([1, 5, 9, 5, 9, 12], np.int32)
    q = np.array([0, 0, 1, 0, 2], np.int32)
    qoff = np.array([0, 1, 3, 3, 5], np.int32)  # {0}, {0,1}, {}, {0,2}
    t = api.HashTables(_cuda(idx_off), _cuda(docs), 0.2, 5, 1)
    res, cnt, roff = t.ir_query(_cuda(q), _cuda(qoff), 0)
    lists = oracle_lib.ir_result_lists(res.cpu().numpy(), cnt.cpu().numpy(), roff.cpu().numpy())
    assert lists == [[1, 5, 9], [], [], [5, 9]]
    t.close()


def _serve(api, t, queries, budget=1 << 30, max_batch=64, inflight=4, gap_s=0.0):
    from reset_amd import serving

    srv = serving.IRServer(t, budget_bytes=budget, max_batch=max_batch, max_inflight=inflight)
    for i, terms in enumerate(queries):
        srv.submit(i, terms)
        if gap_s:
            time.sleep(gap_s)
    srv.drain()
    comps = srv.poll_all()
    out = {c.query_id: srv.result(c.query_id) for c in comps}
    stats = srv.stats()
    srv.close()
    return out, comps, stats


def test_ir_server_results_alone_and_in_the_mix(api, oracle_lib):
    rng = np.random.default_rng(17)
    idx_off, docs, q, qoff = _ir_case(rng, nterms=500, ndocs=500000, maxlen=60000, nq=300, kmax=8)
    queries = [q[qoff[i]:qoff[i + 1]].tolist() for i in range(300)]
    m_res, m_cnt, m_roff = oracle_lib.ir_merge(idx_off, docs, q, qoff)
    want = oracle_lib.ir_result_lists(m_res, m_cnt, m_roff)
    t = api.HashTables(_cuda(idx_off), _cuda(docs), 0.2, 5, 1)
    mixed, comps, stats = _serve(api, t, queries, budget=8 << 20, max_batch=32, inflight=4)
    assert len(comps) == 300 and all(c.status == 0 for c in comps)
    assert [mixed[i] for i in range(300)] == want
    assert stats["launches"] > 1 and stats["peak_bytes"] <= 8 << 20
    for i in rng.choice(300, 12, replace=False).tolist():
        alone, _, _ = _serve(api, t, [queries[i]])
        assert alone[0] == mixed[i] == want[i]
    t.close()


def test_ir_server_memory_admission_splits_batches(api, oracle_lib):
    """Admission by memory footprint (8 bytes per term0 entry plus offsets):
    a slot's share of the budget that holds only a few queries forces many
    launches without changing results, and a query above a slot's share is
    refused at submit (RS_ENOMEM) while the others are served."""
    from reset_amd import serving
    from reset_amd.lib import RS_ENOMEM, ResetError

    rng = np.random.default_rng(23)
    idx_off, docs, q, qoff = _ir_case(rng, nterms=200, ndocs=300000, maxlen=40000, nq=120, kmax=4)
    queries = [q[qoff[i]:qoff[i + 1]].tolist() for i in range(120)]
    m_res, m_cnt, m_roff = oracle_lib.ir_merge(idx_off, docs, q, qoff)
    want = oracle_lib.ir_result_lists(m_res, m_cnt, m_roff)
    df = np.diff(idx_off)
    fp = [8 * int(df[x[0]]) + 64 + 4 * len(x) if x else 64 for x in queries]
    t = api.HashTables(_cuda(idx_off), _cuda(docs), 0.2, 5, 1)
    slot = max(fp) + 1024
    got, comps, stats = _serve(api, t, queries, budget=2 * slot, max_batch=16, inflight=2)
    assert [got[i] for i in range(120)] == want
    assert stats["launches"] >= 20 and stats["peak_bytes"] <= 2 * slot
    # a slot share below the largest query: that query is refused, the rest served
    big = int(np.argmax(fp))
    srv = serving.IRServer(t, budget_bytes=2 * (fp[big] - 1), max_batch=16, max_inflight=2)
    with pytest.raises(ResetError) as e:
        srv.submit(big, queries[big])
    assert e.value.status == RS_ENOMEM
    small = [i for i in range(120) if fp[i] <= fp[big] - 1][:30]
    for i in small:
        srv.submit(i, queries[i])
    srv.drain()
    assert all(srv.result(i) == want[i] for i in small)
    srv.close()
    t.close()


def test_server_mixes_ir_queries_and_similarity_lookups(api, oracle_lib):
    """IR queries and similarity lookups interleaved on one serving loop:
    every result equals the request served alone and the oracle's."""
    from reset_amd import serving

    rng = np.random.default_rng(31)
    idx_off, docs, q, qoff = _ir_case(rng, nterms=300, ndocs=200000, maxlen=20000, nq=120, kmax=6)
    queries = [q[qoff[i]:qoff[i + 1]].tolist() for i in range(120)]
    m_res, m_cnt, m_roff = oracle_lib.ir_merge(idx_off, docs, q, qoff)
    want_ir = oracle_lib.ir_result_lists(m_res, m_cnt, m_roff)
    soff, scols = G.zipf_sets(rng, 3000, 500, 3, 40)
    lookups = []
    for i in rng.choice(3000, 60, replace=False).tolist():
        r = scols[soff[i]:soff[i + 1]].tolist()
        if len(r) > 3 and i % 2:
            r = sorted(set(r[1:] + [600 + i]))  # a near copy with an unknown token
        lookups.append(r)
    lookups.append([])
    lq = np.zeros(len(lookups) + 1, np.int64)
    np.cumsum([len(r) for r in lookups], out=lq[1:])
    ltok = np.concatenate([np.array(r, np.int32) for r in lookups])
    keys = oracle_lib.sss_lookup_bruteforce(soff, scols, lq.astype(np.int32), ltok, 4, 5)
    want_sss = [[k & 0xFFFFFFFF for k in keys.tolist() if k >> 32 == a] for a in range(len(lookups))]
    assert sum(len(w) for w in want_sss) >= 30
    t = api.HashTables(_cuda(idx_off), _cuda(docs), 0.2, 5, 1)
    sidx = api.SSSIndex(_cuda(soff), _cuda(scols), 4, 5)
    srv = serving.Server(t, budget_bytes=4 << 20, max_batch=16, max_inflight=3, sss_index=sidx)
    n_ir, n_s = len(queries), len(lookups)
    order = rng.permutation(n_ir + n_s).tolist()
    for k in order:
        if k < n_ir:
            srv.submit_ir(k, queries[k])
        else:
            srv.submit_sss(k, lookups[k - n_ir])
    srv.drain()
    comps = srv.poll_all()
    assert len(comps) == n_ir + n_s and all(c.status == 0 for c in comps)
    kinds = {c.query_id: c.kind for c in comps}
    got = {c.query_id: srv.result(c.query_id) for c in comps}
    assert [got[i] for i in range(n_ir)] == want_ir
    assert [got[n_ir + a] for a in range(n_s)] == want_sss
    assert all(kinds[i] == serving.REQ_IR for i in range(n_ir))
    assert all(kinds[n_ir + a] == serving.REQ_SSS for a in range(n_s))
    st = srv.stats()
    assert st["queries"] == n_ir + n_s and st["launches"] > 2
    srv.close()
    # each lookup alone, on a fresh server
    srv = serving.Server(t, sss_index=sidx)
    for a in [0, 1, 7, n_s - 1]:
        srv.submit_sss(10_000 + a, lookups[a])
        srv.drain()
        assert srv.result(10_000 + a) == want_sss[a]
    srv.close()
    with pytest.raises(Exception):
        bad = serving.Server(t)
        bad.submit_sss(1, [1, 2, 3])  # no index attached
    sidx.close()
    t.close()


def test_server_cancel_timeout_and_pause(api, oracle_lib):
    """Failure handling forced from the host.  With admission paused, queued
    requests can be cancelled (status RS_ECANCELED, never launched) or left
    to outwait a queue timeout (RS_ETIMEDOUT); the others complete exactly as
    served alone.  Cancelling an unknown or completed id is an error, and
    both request kinds are covered."""
    import time

    from reset_amd import serving
    from reset_amd.lib import RS_ECANCELED, RS_ETIMEDOUT, ResetError

    rng = np.random.default_rng(41)
    idx_off, docs, q, qoff = _ir_case(rng, nterms=200, ndocs=100000, maxlen=8000, nq=40, kmax=5)
    queries = [q[qoff[i]:qoff[i + 1]].tolist() for i in range(40)]
    m_res, m_cnt, m_roff = oracle_lib.ir_merge(idx_off, docs, q, qoff)
    want = oracle_lib.ir_result_lists(m_res, m_cnt, m_roff)
    soff, scols = G.zipf_sets(rng, 500, 300, 3, 30)
    t = api.HashTables(_cuda(idx_off), _cuda(docs), 0.2, 5, 1)
    sidx = api.SSSIndex(_cuda(soff), _cuda(scols), 4, 5)
    srv = serving.Server(t, budget_bytes=1 << 30, max_batch=8, max_inflight=2, sss_index=sidx)
    # cancellation while paused
    srv.pause()
    for i in range(40):
        srv.submit(i, queries[i])
    srv.submit_sss(1000, scols[soff[5]:soff[6]].tolist())
    for i in (3, 7, 1000):
        srv.cancel(i)
    with pytest.raises(ResetError):
        srv.cancel(5000)  # unknown
    srv.pause(False)
    srv.drain()
    comps = {c.query_id: c for c in srv.poll_all()}
    assert len(comps) == 41
    assert {i for i, c in comps.items() if c.status == RS_ECANCELED} == {3, 7, 1000}
    assert comps[1000].kind == serving.REQ_SSS and comps[3].count == -1
    for i in range(40):
        if i not in (3, 7):
            assert comps[i].status == 0 and srv.result(i) == want[i]
    with pytest.raises(ResetError):
        srv.cancel(0)  # completed
    # queue timeout: requests held past 1 ms time out unlaunched
    launches = srv.stats()["launches"]
    srv.pause()
    for i in range(100, 110):
        srv.submit(i, queries[i - 100])
    srv.submit_sss(2000, scols[soff[9]:soff[10]].tolist())
    srv.set_queue_timeout(1.0)
    time.sleep(0.05)
    srv.pause(False)
    srv.drain()
    comps = {c.query_id: c for c in srv.poll_all()}
    assert set(comps) == set(range(100, 110)) | {2000}
    assert all(c.status == RS_ETIMEDOUT and c.count == -1 for c in comps.values())
    assert srv.stats()["launches"] == launches
    # timeout off again: requests run
    srv.set_queue_timeout(0)
    srv.submit(300, queries[0])
    srv.drain()
    (c,) = srv.poll_all()
    assert c.status == 0 and srv.result(300) == want[0]
    srv.close()
