"""GPU parity: similarity lookups (query sets against a resident collection)
against the brute-force oracle (every query merged with every set)."""
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


def _cuda(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.int32)).cuda()


def _csr(rows):
    off = np.zeros(len(rows) + 1, np.int64)
    np.cumsum([len(r) for r in rows], out=off[1:])
    flat = np.concatenate([np.asarray(r, np.int32) for r in rows]) if off[-1] else np.zeros(0, np.int32)
    return off.astype(np.int32), flat


def _queries(rng, off, cols, vocab):
    rows = []
    n = len(off) - 1
    for i in rng.choice(n, 60, replace=False).tolist():  # exact copies
        rows.append(cols[off[i]:off[i + 1]].tolist())
    for i in rng.choice(n, 60, replace=False).tolist():  # one token swapped for an unknown one
        r = cols[off[i]:off[i + 1]].tolist()
        if r:
            r[int(rng.integers(0, len(r)))] = vocab + 17 + i
        rows.append(sorted(set(r)))
    for _ in range(40):  # random sets
        rows.append(sorted(set(rng.integers(0, vocab, int(rng.integers(1, 30))).tolist())))
    rows.append([])  # empty query
    rows.append([vocab + 1, vocab + 5, vocab + 9])  # only unknown tokens
    return _csr(rows)


@pytest.mark.parametrize("p,q", [(4, 5), (1, 2), (9, 10), (1, 4)])
def test_lookup_matches_bruteforce(api, oracle_lib, p, q):
    rng = np.random.default_rng(p * 7 + q)
    vocab = 400
    off, cols = G.zipf_sets(rng, 2000, vocab, 3, 40)
    qoff, qtok = _queries(rng, off, cols, vocab)
    want = oracle_lib.sss_lookup_bruteforce(off, cols, qoff, qtok, p, q)
    assert len(want) >= 60
    idx = api.SSSIndex(_cuda(off), _cuda(cols), p, q)
    got = idx.lookup(_cuda(qoff), _cuda(qtok))
    assert np.array_equal(got["pairs"].cpu().numpy(), want)
    # the same queries one at a time give the same answers
    for a in [0, 61, 130, len(qoff) - 3, len(qoff) - 2]:
        one = idx.lookup(_cuda(qoff[a:a + 2] - qoff[a]), _cuda(qtok[qoff[a]:qoff[a + 1]]))
        mine = [k & 0xFFFFFFFF for k in want.tolist() if k >> 32 == a]
        assert [k & 0xFFFFFFFF for k in one["pairs"].cpu().tolist()] == mine
    idx.close()


def test_lookup_config4_shape_subsample(api, oracle_lib):
    from reset_amd import inputs

    off_d, cols_d = inputs.zipf_set_collection(20000, 100000, seed=11)
    off, cols = off_d.cpu().numpy(), cols_d.cpu().numpy()
    rng = np.random.default_rng(11)
    qoff, qtok = _queries(rng, off, cols, 100000)
    want = oracle_lib.sss_lookup_bruteforce(off, cols, qoff, qtok, 4, 5)
    idx = api.SSSIndex(off_d, cols_d, 4, 5)
    got = idx.lookup(_cuda(qoff), _cuda(qtok))
    assert np.array_equal(got["pairs"].cpu().numpy(), want) and len(want) >= 60
    idx.close()


def test_lookup_rejects_unsorted_query(api):
    from reset_amd.lib import ResetError

    off = np.array([0, 3, 6], np.int32)
    cols = np.array([1, 2, 3, 2, 3, 4], np.int32)
    idx = api.SSSIndex(_cuda(off), _cuda(cols), 4, 5)
    with pytest.raises(ResetError):
        idx.lookup(_cuda(np.array([0, 3], np.int32)), _cuda(np.array([3, 1, 2], np.int32)))
    got = idx.lookup(_cuda(np.array([0, 3], np.int32)), _cuda(np.array([1, 2, 3], np.int32)))
    assert got["pairs"].cpu().tolist() == [0]
    idx.close()
