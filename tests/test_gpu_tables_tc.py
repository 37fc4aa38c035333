"""GPU parity: length rule, table build, probe order and triangle counting.

Every GPU result is compared with the CPU oracle on the same inputs
(bit-exact: integer counts, lengths, offsets; schedule-dependent slot
placement is checked through the oracle's table invariants instead).
"""
import math

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


def _offsets_for(degrees):
    off = np.zeros(len(degrees) + 1, np.int64)
    np.cumsum(degrees, out=off[1:])
    assert off[-1] < 2**31
    return off.astype(np.int32)


@pytest.mark.parametrize("alpha,b", [(0.2, 5), (0.1, 1), (0.15, 9), (0.25, 4), (0.35, 3), (0.5, 2),
                                     (0.75, 7), (0.95, 1), (1.0, 8)])
def test_length_rule_device_equals_oracle(api, oracle_lib, alpha, b):
    rng = np.random.default_rng(int(alpha * 100) + b)
    chunks = [np.arange(0, 65536)]  # every degree below 2^16 (sum < 2^31)
    # degrees straddling the float32 boundaries of d/alpha near multiples of b*2^k
    edge = []
    for k in range(3, 24):
        center = int(b * (1 << k) * alpha)
        edge.extend(range(max(0, center - 3), center + 4))
    chunks.append(np.array(edge))
    for _ in range(8):
        chunks.append(rng.integers(0, 1 << 24, 120))
    for deg in chunks:
        off = _offsets_for(deg)
        L, m = api.calculate_lengths(_cuda(off), alpha, b)
        want = np.array([oracle_lib.calculate_length(int(d), alpha, b) for d in deg], np.int32)
        assert np.array_equal(L.cpu().numpy(), want)
        assert m == int(want.max())


@pytest.mark.parametrize("layout_mask", [1, 2, 3])
def test_table_geometry_and_invariants(api, oracle_lib, layout_mask):
    rng = np.random.default_rng(layout_mask)
    off, cols = G.random_sets(rng, 3000, 700, 1 << 24)
    t = api.HashTables(_cuda(off), _cuda(cols), 0.2, 5, layout_mask)
    L, toff, M = oracle_lib.geometry(off, 0.2, 5)
    d = t.describe()
    assert d.max_length == M and d.bucket_num == int(toff[-1]) and d.num_elements == len(cols)
    assert np.array_equal(t.lengths().cpu().numpy(), L)
    assert np.array_equal(t.table_offsets().cpu().numpy(), toff)
    for layout in (0, 1):
        if not layout_mask & (1 << layout):
            continue
        T = t.table(layout).cpu().numpy()
        ref = oracle_lib.Tables(L, toff, M, 5, layout, None, 0)
        assert oracle_lib.table_check(off, cols, ref, T=T) == 0
    t.close()


def test_rebuild_in_place_keeps_invariants(api, oracle_lib):
    """rs_tables_rebuild_async (fill + insert, no allocation) over the same
    CSR, then over rewritten elements with the same offsets: the tables meet
    the oracle's invariants for the elements they now hold, and a layout
    that was not built is refused."""
    from reset_amd.lib import RS_EINVAL, ResetError

    rng = np.random.default_rng(21)
    off, cols = G.random_sets(rng, 2000, 500, 1 << 24)
    cd = _cuda(cols)
    t = api.HashTables(_cuda(off), cd, 0.2, 5, 3)
    L, toff, M = oracle_lib.geometry(off, 0.2, 5)
    for new in (cols, (cols.astype(np.int64) * 7 + 3) % (1 << 24)):
        cd.copy_(_cuda(new))
        t.rebuild_async(3)
        t.rebuild_async(2)
        torch.cuda.synchronize()
        t.status()
        for layout in (0, 1):
            ref = oracle_lib.Tables(L, toff, M, 5, layout, None, 0)
            assert oracle_lib.table_check(off, new.astype(np.int32), ref, T=t.table(layout).cpu().numpy()) == 0
    t.close()
    t1 = api.HashTables(_cuda(off), cd, 0.2, 5, 1)
    with pytest.raises(ResetError) as e:
        t1.rebuild_async(2)
    assert e.value.status == RS_EINVAL
    t1.close()


@pytest.mark.parametrize("b", [1, 5, 16, 64])
def test_build_in_lds_batches_single_tables_and_hbm(api, oracle_lib, b):
    """The build takes three routes by table length: runs of small tables
    share one LDS image, a table of up to a full image is built in LDS by
    itself, and a longer one is filled and inserted in HBM.  Sets sized for
    all three (at this b), empty sets among them, keep the oracle's table
    invariants in both layouts, also after an in-place rebuild."""
    rng = np.random.default_rng(100 + b)
    sizes = np.concatenate([rng.integers(0, 200, 1500), rng.integers(1100, 2000, 12),
                            rng.integers(2100, 6000, 6), [0, 0, 4500]])
    rng.shuffle(sizes)
    rows = [np.sort(rng.choice(1 << 24, int(n), replace=False)) for n in sizes]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    cols = np.concatenate(rows).astype(np.int32)
    L, toff, M = oracle_lib.geometry(off, 0.2, b)
    t = api.HashTables(_cuda(off), _cuda(cols), 0.2, b, 3)
    assert np.array_equal(t.table_offsets().cpu().numpy(), toff)
    for rebuild in (False, True):
        if rebuild:
            t.rebuild_async(3)
            torch.cuda.synchronize()
            t.status()
        for layout in (0, 1):
            ref = oracle_lib.Tables(L, toff, M, b, layout, None, 0)
            assert oracle_lib.table_check(off, cols, ref, T=t.table(layout).cpu().numpy()) == 0
    t.close()


def test_table_overflow_is_reported(api):
    off = np.array([0, 200], np.int32)
    cols = np.arange(200, dtype=np.int32)
    from reset_amd.lib import ResetError

    with pytest.raises(ResetError) as ei:
        api.HashTables(_cuda(off), _cuda(cols), 8.0, 1, 1)
    assert ei.value.status == 3


def test_corrupted_inputs_are_rejected(api):
    """Decreasing offsets and an element equal to the EMPTY marker -1 are
    input errors (RS_EINVAL), found on the device and reported, not run."""
    from reset_amd.lib import ResetError

    with pytest.raises(ResetError) as ei:
        api.HashTables(_cuda(np.array([0, 5, 3, 8], np.int32)), _cuda(np.arange(8, dtype=np.int32)),
                       0.2, 5, 3)
    assert ei.value.status == 1 and "negative" in str(ei.value)
    with pytest.raises(ResetError) as ei:
        api.HashTables(_cuda(np.array([0, 3, 6], np.int32)), _cuda(np.array([1, -1, 7, 2, 3, 4], np.int32)),
                       0.2, 5, 3)
    assert ei.value.status == 1 and "EMPTY" in str(ei.value)


def test_probe_order_matches_oracle_on_same_table(api, oracle_lib):
    rng = np.random.default_rng(9)
    off, cols = G.random_sets(rng, 2000, 300, 1 << 20)
    t = api.HashTables(_cuda(off), _cuda(cols), 0.2, 5, 3)
    L, toff, M = oracle_lib.geometry(off, 0.2, 5)
    for layout in (0, 1):
        T = t.table(layout).cpu().numpy()
        ref = oracle_lib.Tables(L, toff, M, 5, layout, T, 0)
        want = oracle_lib.extract_probe_order(off, ref)
        got = t.probe_order(layout).cpu().numpy()
        assert np.array_equal(got, want)
    t.close()


def _tc_gpu_all(api, off, cols):
    t = api.HashTables(_cuda(off), _cuda(cols), 0.2, 5, 3)
    keys_h = t.probe_order(1)
    res = {
        "tiled_normal": t.tc_count(0),
        "tiled_hier_sorted": t.tc_count(1, probe_keys=keys_h),
        "tiled_hier_unsorted": t.tc_count(1),
        "plain_normal": t.tc_count(0, plain=True),
        "plain_hier_sorted": t.tc_count(1, probe_keys=keys_h, plain=True),
    }
    t.close()
    return res


@pytest.mark.parametrize("case", ["K40", "wheel33", "star", "bipartite", "rand1", "rand2"])
def test_tc_known_answers(api, oracle_lib, case):
    rng = np.random.default_rng(len(case))
    if case == "K40":
        n, edges, want = 40, G.complete(40), math.comb(40, 3)
    elif case == "wheel33":
        n, edges, want = 34, G.wheel(33), 33
    elif case == "star":
        n, edges, want = 300, G.star(300), 0
    elif case == "bipartite":
        n, edges, want = 50, G.complete_bipartite(20, 30), 0
    else:
        n = 400
        edges = G.random_graph(n, 4000, rng)
        want = None
    und = G.undirected(edges, n)
    if want is None:
        want = G.brute_triangles(und, n)
    for by in ("id", "degree"):
        off, cols = G.orient(und, n, by)
        assert oracle_lib.tc_merge(off, cols) == want
        assert set(_tc_gpu_all(api, off, cols).values()) == {want}


@pytest.mark.parametrize("scale", [10, 14, 16])
def test_tc_rmat_matches_oracle(api, oracle_lib, scale):
    """Config-1 shape (R-MAT a=.57 b=.19 c=.19, edge factor 16, degree-oriented
    int32 CSR); scale 16 is the config itself."""
    from reset_amd import inputs

    off_d, cols_d = inputs.rmat_tc_csr(scale, 16, seed=20261203 + scale)
    off, cols = off_d.cpu().numpy(), cols_d.cpu().numpy()
    want = oracle_lib.tc_merge(off, cols)
    tabs = oracle_lib.build_tables(off, cols, 0.2, 5, 0)
    assert oracle_lib.tc_hash(off, cols, tabs) == want
    got = _tc_gpu_all(api, off, cols)
    assert set(got.values()) == {want}, got
    # orientation invariance: the id-oriented graph has the same count
    lo, hi, n = inputs.rmat_undirected(scale, 16, seed=20261203 + scale)
    off2, cols2 = inputs.orient_csr(lo, hi, n, by="id")
    t2 = api.HashTables(off2, cols2, 0.2, 5, 1)
    assert t2.tc_count(0) == want
    t2.close()


def test_tc_hub_split_and_empty_rows(api, oracle_lib):
    """A hub whose probe work spans many tiles (K_300 core plus isolated
    vertices and empty rows at both ends)."""
    n = 2000
    core = [(100 + i, 100 + j) for i in range(300) for j in range(i + 1, 300)]
    und = G.undirected(core + [(5, 1999)], n)
    off, cols = G.orient(und, n, "degree")
    want = math.comb(300, 3)
    t = api.HashTables(_cuda(off), _cuda(cols), 0.2, 5, 3)
    probes, edges = t.tc_work()
    assert probes > 4096 * 4 and edges == len(cols)
    assert t.tc_count(0) == want and t.tc_count(1, probe_keys=t.probe_order(1)) == want
    t.close()


def test_tc_staged_and_hbm_paths_both_run(api, oracle_lib):
    """R-MAT scale 16: part of the probes go to LDS-staged tables, the rest to
    tables read in HBM; the count is the oracle's either way."""
    from reset_amd import inputs

    off_d, cols_d = inputs.rmat_tc_csr(16, 16, seed=99)
    want = oracle_lib.tc_merge(off_d.cpu().numpy(), cols_d.cpu().numpy())
    t = api.HashTables(off_d, cols_d, 0.2, 5, 3)
    probes, _ = t.tc_work()
    staged, units, lmax = t.tc_staged()
    assert 0 < staged < probes and units > 0 and lmax >= 8
    assert t.tc_count(0) == want and t.tc_count(1) == want
    assert t.tc_count(1, probe_keys=t.probe_order(1)) == want
    t.close()


def test_tc_complete_graph_all_staged_classes():
    """K_2300 oriented by id: out-degrees n-1-u give table lengths up to 4096,
    so tables of all three LDS classes (L <= 1024, 2048, 4096) are staged
    next to unstaged small tables; count = C(n, 3) exactly."""
    from reset_amd import api as A

    n = 2300
    deg = np.arange(n - 1, -1, -1, dtype=np.int64)
    off = np.zeros(n + 1, np.int64)
    np.cumsum(deg, out=off[1:])
    cols = np.concatenate([np.arange(u + 1, n, dtype=np.int32) for u in range(n)])
    t = A.HashTables(_cuda(off.astype(np.int32)), _cuda(cols), 0.2, 5, 3)
    probes, _ = t.tc_work()
    staged, units, lmax = t.tc_staged()
    assert lmax == 4096 and 0 < staged < probes
    want = math.comb(n, 3)
    assert t.tc_count(0) == want
    assert t.tc_count(1) == want
    t.close()


@pytest.mark.parametrize("alpha,b", [(0.5, 2), (0.9, 1), (0.1, 8), (0.35, 3), (0.3, 4), (0.2, 7)])
def test_tc_hyperparameters_do_not_change_the_count(api, oracle_lib, alpha, b):
    from reset_amd import inputs

    off_d, cols_d = inputs.rmat_tc_csr(12, 16, seed=7)
    want = oracle_lib.tc_merge(off_d.cpu().numpy(), cols_d.cpu().numpy())
    t = api.HashTables(off_d, cols_d, alpha, b, 3)
    assert t.tc_count(0) == want
    assert t.tc_count(1, probe_keys=t.probe_order(1)) == want
    t.close()


@pytest.mark.parametrize("b,lmax", [(9, 2048), (19, 1024)])
def test_tc_staging_at_the_lds_limit(b, lmax):
    """Bucket sizes whose largest staged class would fill the 160 KiB LDS
    exactly (b=9 at L=4096, b=19 at L=2048) without the next-unit word: those
    tables stay in HBM, the class below is staged, and K_4000 (tables up to
    L=4096) counts C(n, 3) exactly."""
    from reset_amd import api as A

    n = 4000
    deg = np.arange(n - 1, -1, -1, dtype=np.int64)
    off = np.zeros(n + 1, np.int64)
    np.cumsum(deg, out=off[1:])
    cols = np.concatenate([np.arange(u + 1, n, dtype=np.int32) for u in range(n)])
    t = A.HashTables(_cuda(off.astype(np.int32)), _cuda(cols), 0.2, b, 3)
    probes, _ = t.tc_work()
    staged, units, got_lmax = t.tc_staged()
    assert got_lmax == lmax and 0 < staged < probes
    want = math.comb(n, 3)
    assert t.tc_count(0) == want
    assert t.tc_count(1, probe_keys=t.probe_order(1)) == want
    t.close()
