"""GPU parity: prefix-filter self-join (candidates + hash verification)."""
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


def _with_near_duplicates(rng, off, cols, count):
    rows = [cols[off[i]:off[i + 1]].tolist() for i in range(len(off) - 1)]
    for i in range(count):
        r = list(rows[2 * i])
        if len(r) > 4:
            r[int(rng.integers(0, len(r)))] = 3_000_000 + i  # one token replaced
        rows.append(sorted(set(r)))
    o = np.zeros(len(rows) + 1, np.int64)
    np.cumsum([len(r) for r in rows], out=o[1:])
    return o.astype(np.int32), np.concatenate([np.array(r, np.int32) for r in rows])


@pytest.mark.parametrize("p,q", [(4, 5), (1, 2), (2, 3), (9, 10)])
def test_selfjoin_matches_bruteforce_and_oracle(api, oracle_lib, p, q):
    rng = np.random.default_rng(p * 31 + q)
    off, cols = G.zipf_sets(rng, 500, 400, 3, 40)
    off, cols = _with_near_duplicates(rng, off, cols, 80)
    want = oracle_lib.sss_bruteforce(off, cols, p, q)
    want2, ncand = oracle_lib.sss_prefix_join(off, cols, p, q)
    assert np.array_equal(want, want2) and len(want) > 20
    got = api.sss_selfjoin(_cuda(off), _cuda(cols), p, q)
    assert np.array_equal(got["pairs"].cpu().numpy(), want)
    assert got["candidates"] == ncand


def test_selfjoin_config4_shape_subsample(api, oracle_lib):
    """A 20k-set sample of the Zipf(1.2)-size / Zipf(1.0)-token collection,
    with planted near-duplicates: pairs and candidate count equal the oracle's
    prefix join, and every reported pair meets J >= 0.8 exactly."""
    from reset_amd import inputs

    off_d, cols_d = inputs.zipf_set_collection(20000, 100000, seed=5)
    rng = np.random.default_rng(5)
    off, cols = _with_near_duplicates(rng, off_d.cpu().numpy(), cols_d.cpu().numpy(), 300)
    got = api.sss_selfjoin(_cuda(off), _cuda(cols), 4, 5)
    want, ncand = oracle_lib.sss_prefix_join(off, cols, 4, 5)
    assert got["candidates"] == ncand
    pairs = got["pairs"].cpu().numpy()
    assert np.array_equal(pairs, want) and len(pairs) >= 100
    for key in pairs[:200].tolist():
        i, j = key >> 32, key & 0xFFFFFFFF
        a, b = set(cols[off[i]:off[i + 1]].tolist()), set(cols[off[j]:off[j + 1]].tolist())
        assert 5 * len(a & b) >= 4 * len(a | b)


def test_selfjoin_edge_cases(api):
    # identical sets pair up; empty sets never do; a lone set yields nothing
    off = np.array([0, 3, 6, 6, 6, 9], np.int32)
    cols = np.array([1, 2, 3, 1, 2, 3, 4, 5, 6], np.int32)
    got = api.sss_selfjoin(_cuda(off), _cuda(cols), 4, 5)
    assert got["pairs"].cpu().tolist() == [(0 << 32) | 1]
    one = api.sss_selfjoin(_cuda(np.array([0, 2], np.int32)), _cuda(np.array([7, 9], np.int32)), 4, 5)
    assert one["pairs"].numel() == 0 and one["candidates"] == 0


def test_selfjoin_mixes_grouped_and_batched_verification(api, oracle_lib):
    """Sets above 2047 tokens cannot share the grouped LDS table: their
    candidates take the batched pair kernel while the rest are grouped; the
    joined result is the brute force's either way."""
    rng = np.random.default_rng(41)
    off, cols = G.zipf_sets(rng, 400, 300, 3, 40)
    off, cols = _with_near_duplicates(rng, off, cols, 40)
    rows = [cols[off[i]:off[i + 1]].tolist() for i in range(len(off) - 1)]
    nsmall = len(rows)
    for i in range(12):
        big = sorted(set(rng.integers(0, 200000, int(rng.integers(2100, 2600))).tolist()))
        rows.append(big)
        near = list(big)
        near[int(rng.integers(0, len(near)))] = 300000 + i
        rows.append(sorted(set(near)))
    o = np.zeros(len(rows) + 1, np.int64)
    np.cumsum([len(r) for r in rows], out=o[1:])
    off2, cols2 = o.astype(np.int32), np.concatenate([np.array(r, np.int32) for r in rows])
    want = oracle_lib.sss_bruteforce(off2, cols2, 4, 5)
    got = api.sss_selfjoin(_cuda(off2), _cuda(cols2), 4, 5)
    assert np.array_equal(got["pairs"].cpu().numpy(), want)
    big_pairs = [k for k in want.tolist() if (k >> 32) >= nsmall]
    assert len(big_pairs) >= 12 and len(want) > len(big_pairs)


@pytest.mark.parametrize("alpha,b", [(0.5, 2), (0.2, 1), (1.0, 4), (0.1, 4)])
def test_selfjoin_honours_table_hyperparameters(api, oracle_lib, alpha, b):
    """load_factor / bucket_size set the verification tables (grouped LDS
    table and the batched kernel's bins for the large build sets); the
    joined pairs equal the brute force's at every setting."""
    rng = np.random.default_rng(int(alpha * 100) + b)
    off, cols = G.zipf_sets(rng, 300, 300, 3, 40)
    off, cols = _with_near_duplicates(rng, off, cols, 40)
    rows = [cols[off[i]:off[i + 1]].tolist() for i in range(len(off) - 1)]
    nsmall = len(rows)
    for i in range(6):  # build sets above the grouped table at (0.1, 4) and (0.2, 1)
        big = sorted(set(rng.integers(0, 200000, int(rng.integers(1500, 3000))).tolist()))
        rows += [big, sorted(set(big[1:] + [400000 + i]))]
    o = np.zeros(len(rows) + 1, np.int64)
    np.cumsum([len(r) for r in rows], out=o[1:])
    off2, cols2 = o.astype(np.int32), np.concatenate([np.array(r, np.int32) for r in rows])
    want = oracle_lib.sss_bruteforce(off2, cols2, 4, 5)
    got = api.sss_selfjoin(_cuda(off2), _cuda(cols2), 4, 5, load_factor=alpha, bucket_size=b)
    assert np.array_equal(got["pairs"].cpu().numpy(), want)
    assert sum((k >> 32) >= nsmall for k in want.tolist()) >= 6 and len(want) > 12


def test_selfjoin_rejects_bad_input_and_settings(api):
    """A set that is not strictly ascending (out of order, or a repeated
    token) and verification tables without an LDS form (5-slot buckets, load
    factor above 1) are rejected with RS_EINVAL, not answered."""
    from reset_amd.lib import RS_EINVAL, ResetError

    off = np.array([0, 3, 6], np.int32)
    for cols in ([1, 3, 2, 1, 2, 3], [1, 2, 2, 1, 2, 3], [1, 2, 3, -4, 5, 6]):
        with pytest.raises(ResetError) as e:
            api.sss_selfjoin(_cuda(off), _cuda(np.array(cols, np.int32)), 4, 5)
        assert e.value.status == RS_EINVAL
    good = _cuda(np.array([1, 2, 3, 1, 2, 3], np.int32))
    for alpha, b in ((0.2, 5), (1.5, 4), (0.25, 3)):
        with pytest.raises(ResetError) as e:
            api.sss_selfjoin(_cuda(off), good, 4, 5, load_factor=alpha, bucket_size=b)
        assert e.value.status == RS_EINVAL
    assert api.sss_selfjoin(_cuda(off), good, 4, 5)["pairs"].cpu().tolist() == [1]
