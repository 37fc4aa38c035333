"""GPU parity of the competitor tables (baseline/common/{cuco,cuckoo_hash,
hopscotch_hash,roaring_bitmap}) and their TC / SSS consumers.

Cuckoo, hopscotch and roaring builds are deterministic (serial per set in
the reference), so the device arrays must equal the oracle's slot for slot,
lost keys included; the linear-probing build places keys in parallel and is
checked through the oracle's invariant checker.  Counts are compared with the
sorted-merge and brute-force oracles (and with the oracle's own tables when
keys were lost).
"""
import itertools

import numpy as np
import pytest
import torch

import graphs as G

pytestmark = pytest.mark.gpu

KINDS = ["linear", "cuckoo", "hopscotch", "roaring"]
EMPTY = 0x7FFFFFFF


@pytest.fixture(scope="module")
def api():
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from reset_amd import api as A

    A.load()
    return A


def _cuda(x, dtype=torch.int32):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dtype).cuda()


def _csr(rows):
    off = np.zeros(len(rows) + 1, np.int64)
    np.cumsum([len(r) for r in rows], out=off[1:])
    return off.astype(np.int32), np.concatenate([np.asarray(r, np.int64) for r in rows]).astype(np.int32)


def _same_arrays(gpu: dict, ref: dict, kind):
    names = {"cuckoo": [("offsets", "toff"), ("slots", "slots"), ("stash_count", "stash_count"),
                        ("stash", "stash")],
             "hopscotch": [("offsets", "toff"), ("slots", "slots"), ("bitmap", "bitmap")],
             "roaring": [("remap", "remap"), ("node_start", "node_start"), ("word_index", "word_index"),
                         ("word_data", "word_data")]}[kind]
    for g, r in names:
        a = gpu[g].numpy()
        b = ref[r]
        if b.dtype == np.uint32:
            a = a.view(np.uint32)
        if b.dtype == np.uint64:
            a = a.view(np.uint64)
        assert np.array_equal(a, b), (kind, g)


def test_hand_cases(api, oracle_lib):
    # cuckoo: keys 2, 4 share both positions of a 2-slot table; 4 ends in the stash
    off, el = _csr([[2, 4]])
    t = api.CompetitorTables(_cuda(off), _cuda(el), "cuckoo", 1.0)
    a = t.arrays()
    assert a["slots"].tolist() == [2, EMPTY] and a["stash_count"].tolist() == [1]
    assert a["stash"][0] == 4 and t.failed == 0
    t.close()
    # hopscotch: a full neighbourhood fails the 33rd home-0 key
    keys = [257 * j for j in range(0, 65, 2)]
    off, el = _csr([keys])
    t = api.CompetitorTables(_cuda(off), _cuda(el), "hopscotch", 0.2)
    a = t.arrays()
    assert a["slots"][:32].tolist() == keys[:32] and t.failed == 1
    assert a["bitmap"].numpy().view(np.uint32)[0] == 0xFFFFFFFF
    t.close()
    # hopscotch: key 1 (home 1) is moved to slot 32 to free slot 1
    row = [1] + [257 * j for j in range(0, 63, 2)]
    off, el = _csr([row])
    t = api.CompetitorTables(_cuda(off), _cuda(el), "hopscotch", 0.2)
    a = t.arrays()
    assert a["slots"][32] == 1 and a["slots"][1] == row[-1] and t.failed == 0
    assert a["bitmap"].numpy().view(np.uint32)[1] == 0x80000000
    t.close()


@pytest.mark.parametrize("kind", KINDS)
def test_tables_and_tc_equal_oracle(api, oracle_lib, kind):
    """R-MAT scale 12 (edge factor 16, degree-oriented CSR)."""
    from reset_amd import inputs

    off_d, cols_d = inputs.rmat_tc_csr(12, 16, seed=31)
    off, cols = off_d.cpu().numpy(), cols_d.cpu().numpy()
    ref = oracle_lib.comp_build(kind, off, cols, 0.2)
    t = api.CompetitorTables(off_d, cols_d, kind, 0.2)
    assert t.failed == ref.failed == 0
    a = t.arrays()
    if kind == "linear":
        assert np.array_equal(a["offsets"].numpy(), ref.arrays["toff"])
        assert oracle_lib.comp_linear_check(off, cols, a["offsets"].numpy(), a["slots"].numpy()) == 0
    else:
        _same_arrays(a, ref.arrays, kind)
    assert t.tc_count() == oracle_lib.tc_merge(off, cols)
    t.close()


@pytest.mark.parametrize("kind,lf", [("cuckoo", 4.0), ("cuckoo", 1.0), ("hopscotch", 0.2), ("linear", 4.0)])
def test_lossy_builds_equal_oracle(api, oracle_lib, kind, lf):
    """Tables that lose keys: the cuckoo at load factors 4 and 1 overflows
    stashes; in the hopscotch, sets of 33..51 keys 257*j (j even) all share
    home s & 511 at capacity 512, so each key past the 32nd fails.  Device and
    oracle lose the same keys (the linear table the same number), and the
    counts through the lossy tables agree."""
    rng = np.random.default_rng(7)
    if kind == "hopscotch":
        rows = [np.sort(rng.choice(128, int(rng.integers(33, 52)), replace=False)) * 514 for _ in range(200)]
    else:
        rows = [np.sort(rng.choice(1 << 20, int(rng.integers(1, 300)), replace=False)) for _ in range(200)]
    off, el = _csr(rows)
    ref = oracle_lib.comp_build(kind, off, el, lf)
    t = api.CompetitorTables(_cuda(off), _cuda(el), kind, lf)
    assert t.failed == ref.failed
    if kind != "linear":
        assert ref.failed > 0
        _same_arrays(t.arrays(), ref.arrays, kind)
    pairs = np.array(list(itertools.combinations(range(0, 200, 3), 2)), np.int32)
    flags, sizes = t.sss_verify(_cuda(pairs[:, 0]), _cuda(pairs[:, 1]), 1, 4)
    if kind != "linear":
        o_flags, o_sizes = oracle_lib.comp_sss(ref, off, el, pairs[:, 0], pairs[:, 1], 1, 4)
        assert np.array_equal(sizes.cpu().numpy(), o_sizes)
        assert np.array_equal(flags.cpu().numpy(), o_flags)
    t.close()


@pytest.mark.parametrize("kind", KINDS)
def test_sss_verify_equals_bruteforce(api, oracle_lib, kind):
    rng = np.random.default_rng(12)
    off, cols = G.zipf_sets(rng, 160, 300, 2, 60)
    pairs = np.array(list(itertools.combinations(range(160), 2)), np.int32)
    want = set(oracle_lib.sss_bruteforce(off, cols, 4, 5).tolist())
    t = api.CompetitorTables(_cuda(off), _cuda(cols), kind, 0.2)
    flags, sizes = t.sss_verify(_cuda(pairs[:, 0]), _cuda(pairs[:, 1]), 4, 5)
    got = {(int(i) << 32) | int(j) for (i, j), f in zip(pairs, flags.cpu().numpy()) if f}
    assert got == want
    ot = oracle_lib.build_tables(off, cols, 0.2, 5, 0)
    _, o_sizes = oracle_lib.sss_verify_hash(off, cols, ot, pairs[:, 0], pairs[:, 1], 4, 5)
    assert np.array_equal(sizes.cpu().numpy(), o_sizes)
    t.close()


@pytest.mark.parametrize("kind", KINDS)
def test_ir_queries_equal_chained_merge(api, oracle_lib, kind):
    """Conjunctive queries through tables of an inverted index (ir_cuco.cu):
    counts and result lists (term0 order, ascending doc ids) equal the
    chained merge, single-term and empty queries included."""
    rng = np.random.default_rng(17)
    idx_off, docs = G.inverted_index(rng, 300, 200000, 20000)
    q, qoff = G.queries(rng, idx_off, 400, 1, 6)
    q = np.concatenate([q, [5]]).astype(np.int32)
    qoff = np.concatenate([qoff, [qoff[-1], qoff[-1] + 1]]).astype(np.int32)  # an empty query, then {5}
    m_res, m_cnt, m_roff = oracle_lib.ir_merge(idx_off, docs, q, qoff)
    t = api.CompetitorTables(_cuda(idx_off), _cuda(docs), kind, 0.2)
    res, cnt, roff = t.ir_query(_cuda(q), _cuda(qoff))
    assert np.array_equal(cnt.cpu().numpy(), m_cnt)
    assert oracle_lib.ir_result_lists(res.cpu().numpy(), cnt.cpu().numpy(), roff.cpu().numpy()) == \
        oracle_lib.ir_result_lists(m_res, m_cnt, m_roff)
    t.close()


def test_input_errors(api):
    from reset_amd.lib import ResetError

    off, el = _csr([[5, 3, 9]])  # descending pair inside the row
    with pytest.raises(ResetError):
        api.CompetitorTables(_cuda(off), _cuda(el), "roaring")
    off, el = _csr([[1, -4]])
    with pytest.raises(ResetError):
        api.CompetitorTables(_cuda(off), _cuda(el), "roaring")
    off, el = _csr([[1, 2], [3]])
    t = api.CompetitorTables(_cuda(off), _cuda(el), "cuckoo")
    with pytest.raises(ResetError):
        t.sss_verify(_cuda(np.array([0, 5], np.int32)), _cuda(np.array([1, 0], np.int32)))
    with pytest.raises(ResetError):  # a query term outside the index
        t.ir_query(_cuda(np.array([0, 7], np.int32)), _cuda(np.array([0, 2], np.int32)))
    t.close()
    off, el = _csr([list(range(5000))])  # 5000 / 1e-6 slots: beyond 2^30
    for kind in ("linear", "cuckoo", "hopscotch"):
        with pytest.raises(ResetError):
            api.CompetitorTables(_cuda(off), _cuda(el), kind, 1e-6)
