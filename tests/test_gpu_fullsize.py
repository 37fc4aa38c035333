"""GPU parity at the benchmark configurations' full sizes, against the CPU
oracle: the 1M-pair batch (every size equal to the overlap the generator
planted, and a leading sample equal to the oracle's hash count), R-MAT
scale-24 triangle counting (both layouts, against the
sorted-merge count of the whole graph), the 1M-set similarity self-join
(every planted near-duplicate found; the join restricted to a 200k-set
prefix equal to the oracle's prefix join of that prefix) and all 20k
conjunctive queries over the 10M-document Zipf index (batch, hierarchical
and served results equal to the chained sorted merge).  Same seeds as
bench.py, so the counts here are the bench's."""
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SEED = 20261203


@pytest.fixture(scope="module")
def api():
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from reset_amd import api as A

    A.load()
    return A


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def test_pairs_config2_full_batch(api, oracle_lib):
    """All 1,000,000 pairs of the headline batch (sizes log-uniform in
    [16, 65536], 2^24 universe, overlap U[0, 0.5]) in one batched call with
    the default per-bin settings: every size equals the overlap k the
    generator planted, and the first 20,000 equal the oracle's hash count."""
    from reset_amd import inputs

    P = 1_000_000
    batch = inputs.make_pair_batch(P, seed=SEED)
    _say(f"pairs: {int(batch['a_off'][-1]) + int(batch['b_off'][-1])} elements")
    plan = api.PairsPlan(P)
    sizes = plan.run(batch["a_off"], batch["a"], batch["b_off"], batch["b"])
    plan.close()
    assert torch.equal(sizes, batch["k"])
    S = 20_000
    ao = batch["a_off"][:S + 1].cpu().numpy()
    bo = batch["b_off"][:S + 1].cpu().numpy()
    a = batch["a"][:int(ao[-1])].cpu().numpy()
    b = batch["b"][:int(bo[-1])].cpu().numpy()
    assert np.array_equal(sizes[:S].cpu().numpy(), oracle_lib.pairs_hash(ao, a, bo, b))


def test_tc_rmat24_matches_oracle(api, oracle_lib):
    from reset_amd import inputs

    _say("tc scale 24: generating")
    off_d, cols_d = inputs.rmat_tc_csr(24, 16, seed=SEED + 3)
    t = api.HashTables(off_d, cols_d, 0.2, 5, 3)
    _say("tc scale 24: counting on the GPU (hierarchical, normal, plain)")
    got_h = t.tc_count(1, probe_keys=t.probe_order(1))
    got_n = t.tc_count(0)
    got_p = t.tc_count(1, plain=True)
    t.close()
    off, cols = off_d.cpu().numpy(), cols_d.cpu().numpy()
    del off_d, cols_d
    torch.cuda.empty_cache()
    _say(f"tc scale 24: oracle merge over {len(cols)} edges")
    want = oracle_lib.tc_merge(off, cols)
    assert got_h == got_n == got_p == want
    assert want == 10_286_145_463  # the bench's count (seed 20261206)


def _dense(res, cnt, roff):
    """Every query's result run, concatenated (res holds query q's survivors
    at res[roff[q] .. roff[q] + cnt[q]))."""
    cnt = np.asarray(cnt, np.int64)
    starts = np.repeat(np.asarray(roff[:-1], np.int64), cnt)
    within = np.arange(int(cnt.sum()), dtype=np.int64) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    return np.asarray(res)[starts + within]


def test_ir_config5_all_queries_match_merge(api, oracle_lib):
    from reset_amd import inputs, serving

    _say("ir: index of 10M docs / 1M terms")
    off_d, docs_d = inputs.ir_index(seed=SEED + 5)
    q, qoff = inputs.ir_queries(off_d, 20_000, seed=SEED + 6)
    off, docs = off_d.cpu().numpy(), docs_d.cpu().numpy()
    _say("ir: oracle chained merge of 20k queries")
    m_res, m_cnt, m_roff = oracle_lib.ir_merge(off, docs, q, qoff)
    want = _dense(m_res, m_cnt, m_roff)
    t = api.HashTables(off_d, docs_d, 0.2, 5, 3)
    qd, qod = torch.from_numpy(q).cuda(), torch.from_numpy(qoff).cuda()
    _say("ir: batch queries, normal layout")
    res, cnt, roff = t.ir_query(qd, qod, 0)
    cnt_h = cnt.cpu().numpy()
    assert np.array_equal(cnt_h, m_cnt) and np.array_equal(roff.cpu().numpy(), m_roff)
    assert np.array_equal(_dense(res.cpu().numpy(), cnt_h, m_roff), want)
    assert 0 < len(want) < int(m_roff[-1])
    _say("ir: batch queries, hierarchical layout, term0 in probe order")
    res_h, cnt_hh, _ = t.ir_query(qd, qod, 1, term0_docs=t.probe_order(1))
    assert np.array_equal(cnt_hh.cpu().numpy(), m_cnt)
    got_h = _dense(res_h.cpu().numpy(), m_cnt, m_roff)
    # survivors keep term0's order as read (the hierarchical table's probe
    # order); as sets per query they equal the merge's ascending lists
    qid = np.repeat(np.arange(len(m_cnt)), m_cnt)
    assert np.array_equal(got_h[np.lexsort((got_h, qid))], want)
    t.close()
    # the tables of the IR benchmark: the solver's (load factor, bucket size)
    # for one lane per key
    alpha, b, _ = api.hyper_solve(int(off[-1]) // (len(off) - 1), W=1, S_trans=128)
    _say(f"ir: batch queries, hierarchical layout, tables alpha {alpha} b {b}")
    t = api.HashTables(off_d, docs_d, alpha, b, 3)
    res_b, cnt_b, _ = t.ir_query(qd, qod, 1)
    assert np.array_equal(cnt_b.cpu().numpy(), m_cnt)
    assert np.array_equal(_dense(res_b.cpu().numpy(), m_cnt, m_roff), want)
    _say("ir: all 20k queries through the serving loop")
    srv = serving.Server(t, budget_bytes=4096 << 20, max_batch=256, max_inflight=4)
    for i in range(len(qoff) - 1):
        srv.submit(i, q[qoff[i]:qoff[i + 1]])
    srv.drain()
    comps = srv.poll_all()
    assert len(comps) == len(qoff) - 1 and all(c.status == 0 for c in comps)
    served = np.concatenate([np.asarray(srv.result(i), np.int32) for i in range(len(qoff) - 1)])
    srv.close()
    t.close()
    assert np.array_equal(served, want)


def test_sss_config4_planted_pairs_and_prefix(api, oracle_lib):
    from reset_amd import inputs

    _say("sss: 1M-set Zipf collection with 1000 planted near-duplicates")
    off_d, cols_d = inputs.zipf_set_collection(1_000_000, 1_000_000, seed=SEED + 4)
    off_d, cols_d, planted = inputs.plant_near_duplicates(off_d, cols_d, 1000, seed=SEED + 10)
    got = api.sss_selfjoin(off_d, cols_d, 4, 5)
    pairs = got["pairs"].cpu().numpy()
    # completeness: every planted (source, copy) pair has J >= 4/5 and is reported
    assert len(planted) == 1000 and np.isin(planted, pairs).all()
    off, cols = off_d.cpu().numpy(), cols_d.cpu().numpy()
    rng = np.random.default_rng(1)
    for key in rng.choice(pairs, min(2000, len(pairs)), replace=False).tolist():
        i, j = key >> 32, key & 0xFFFFFFFF
        a, b = cols[off[i]:off[i + 1]], cols[off[j]:off[j + 1]]
        r = len(np.intersect1d(a, b, assume_unique=True))
        assert 5 * r >= 4 * (len(a) + len(b) - r)
    # the join of a 200k-set prefix equals the oracle's, and equals the full
    # join's pairs inside that prefix
    S = 200_000
    _say(f"sss: oracle prefix join of the first {S} sets")
    o = off[:S + 1]
    want, ncand = oracle_lib.sss_prefix_join(o, cols[:o[-1]], 4, 5)
    sub = api.sss_selfjoin(off_d[:S + 1].contiguous(), cols_d, 4, 5)
    assert sub["candidates"] == ncand
    assert np.array_equal(sub["pairs"].cpu().numpy(), want)
    inside = pairs[(pairs >> 32) < S]
    inside = inside[(inside & 0xFFFFFFFF) < S]
    assert np.array_equal(inside, want)
