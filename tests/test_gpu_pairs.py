"""GPU parity: batched and plain pair intersection, and the pair generator."""
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


def _dev(a_off, a, b_off, b):
    c = lambda x, dt: torch.from_numpy(np.ascontiguousarray(x)).to(dt).cuda()
    return c(a_off, torch.int64), c(a, torch.int32), c(b_off, torch.int64), c(b, torch.int32)


def test_generator_sets_sorted_unique_exact_overlap(api):
    from reset_amd import inputs

    batch = inputs.make_pair_batch(400, seed=3, lo=16, hi=9000)
    a_off, a = batch["a_off"].cpu().numpy(), batch["a"].cpu().numpy()
    b_off, b = batch["b_off"].cpu().numpy(), batch["b"].cpu().numpy()
    k = batch["k"].cpu().numpy()
    for i in range(400):
        A = a[a_off[i]:a_off[i + 1]]
        B = b[b_off[i]:b_off[i + 1]]
        assert np.all(np.diff(A) > 0) and np.all(np.diff(B) > 0)
        assert A.min() >= 0 and A.max() < 1 << 24 and B.max() < 1 << 24
        assert len(np.intersect1d(A, B)) == k[i]
    # deterministic: same seed, same bits
    again = inputs.make_pair_batch(400, seed=3, lo=16, hi=9000)
    assert torch.equal(again["a"], batch["a"]) and torch.equal(again["b"], batch["b"])


def test_batched_equals_oracle_and_plain_numpy_sets(api, oracle_lib):
    rng = np.random.default_rng(1)
    a_off, a, b_off, b, k = G.make_pairs(rng, 600, 3000, 1 << 24)
    d = _dev(a_off, a, b_off, b)
    plan = api.PairsPlan(600)
    got = plan.run(*d).cpu().numpy()
    assert np.array_equal(got, k)
    assert np.array_equal(got, oracle_lib.pairs_hash(a_off, a, b_off, b))
    plain = api.pairs_intersect_plain(*d).cpu().numpy()
    assert np.array_equal(plain, k)
    plan.close()


@pytest.mark.parametrize("sizes", [(0, 0), (0, 50), (1, 1), (511, 511), (512, 512), (513, 4000),
                                   (1023, 70000), (2048, 2048), (65536, 65536), (5000, 131072)])
def test_batched_chunk_boundaries(api, oracle_lib, sizes):
    """Build sides at, below and above the LDS chunk (range chunking)."""
    na, nb = sizes
    rng = np.random.default_rng(na * 7 + nb)
    universe = 1 << 24
    k = min(na, nb) // 3
    pool = rng.choice(universe, size=na + nb - k, replace=False)
    A = np.sort(pool[:na])
    B = np.sort(np.concatenate([pool[na - k:na], pool[na:]])) if nb else np.zeros(0, np.int64)
    a_off = np.array([0, na], np.int64)
    b_off = np.array([0, len(B)], np.int64)
    d = _dev(a_off, A, b_off, B)
    plan = api.PairsPlan(1)
    assert plan.run(*d).cpu().tolist() == [k]
    # swapped roles give the same size
    d2 = _dev(b_off, B, a_off, A)
    assert plan.run(*d2).cpu().tolist() == [k]
    plan.close()


@pytest.mark.parametrize("b", [4, 2, 1])
def test_colliding_keys_walk_past_full_buckets(api, oracle_lib, b):
    """Every element a multiple of 4096: in any table of <= 4096 buckets all
    keys share home bucket 0, so inserts and lookups walk a chain of full
    buckets (the displaced path), in resident and chunked pairs alike; a
    second family mixes such keys with spread ones."""
    rng = np.random.default_rng(40 + b)
    sizes = [(40, 3000), (255, 255), (256, 256), (1000, 2000), (2000, 2000), (700, 3000)]
    a_parts, b_parts, ks = [], [], []
    for i, (na, nb) in enumerate(sizes * 2):
        if i < len(sizes):
            universe = np.arange(4096, dtype=np.int64) * 4096
        else:  # half colliding, half anywhere else
            universe = np.unique(np.concatenate([np.arange(2048, dtype=np.int64) * 4096,
                                                 rng.choice(1 << 24, 6000, replace=False)]))
        k = min(na, nb) // 2
        pool = rng.choice(universe, size=na + nb - k, replace=False)
        a_parts.append(np.sort(pool[:na]))
        b_parts.append(np.sort(np.concatenate([pool[na - k:na], pool[na:]])))
        ks.append(k)
    a_off = np.concatenate([[0], np.cumsum([len(x) for x in a_parts])]).astype(np.int64)
    b_off = np.concatenate([[0], np.cumsum([len(x) for x in b_parts])]).astype(np.int64)
    A, B = np.concatenate(a_parts), np.concatenate(b_parts)
    want = oracle_lib.pairs_hash(a_off, A, b_off, B)
    assert want.tolist() == ks
    opts = api.default_pairs_opts()
    for i in range(opts.num_bins):
        opts.bins[i].bucket_size = b
    P = len(ks)
    plan = api.PairsPlan(P, opts)
    got = plan.run(*_dev(a_off, A, b_off, B)).cpu().numpy()
    plan.status()
    assert got.tolist() == ks
    got = plan.run(*_dev(b_off, B, a_off, A)).cpu().numpy()
    assert got.tolist() == ks
    plan.close()


def test_batched_config2_shape_sample(api, oracle_lib):
    """A 20k-pair sample of the 1M-pair workload (log-uniform [16, 65536],
    overlap U[0, 0.5]): batched == plain == k == oracle."""
    from reset_amd import inputs

    batch = inputs.make_pair_batch(20000, seed=20261203)
    plan = api.PairsPlan(20000)
    got = plan.run(batch["a_off"], batch["a"], batch["b_off"], batch["b"])
    assert torch.equal(got, batch["k"])
    sub = slice(0, 300)
    a_off = batch["a_off"].cpu().numpy()
    b_off = batch["b_off"].cpu().numpy()
    a = batch["a"].cpu().numpy()
    b = batch["b"].cpu().numpy()
    want = oracle_lib.pairs_hash(a_off[:301], a, b_off[:301], b)
    assert np.array_equal(got.cpu().numpy()[sub], want)
    P = 300
    plain = api.pairs_intersect_plain(batch["a_off"][:P + 1].contiguous(), batch["a"],
                                      batch["b_off"][:P + 1].contiguous(), batch["b"])
    assert np.array_equal(plain.cpu().numpy(), want)
    plan.close()


def test_graph_replays_match_eager(api):
    from reset_amd import inputs

    batch = inputs.make_pair_batch(5000, seed=11)
    plan = api.PairsPlan(5000)
    eager = plan.run(batch["a_off"], batch["a"], batch["b_off"], batch["b"]).clone()
    s = torch.cuda.Stream()
    sizes = torch.zeros(5000, dtype=torch.int32, device="cuda")
    with torch.cuda.stream(s):
        plan.capture(batch["a_off"], batch["a"], batch["b_off"], batch["b"], sizes, s)
    for rep in range(4):
        sizes.fill_(-7)
        torch.cuda.synchronize()
        plan.replay(s)
        s.synchronize()
        assert torch.equal(sizes, eager), rep
        _ = (batch["a"][:1000].float().sum())  # eager work between replays
        torch.cuda.synchronize()
    plan.status()
    plan.close()


@pytest.mark.parametrize("alpha,b,chunk,unroll", [(0.2, 4, 255, 4), (0.5, 2, 1000, 4), (0.9, 1, 1000, 2),
                                                  (0.25, 2, 511, 1), (0.12, 4, 200, 2), (0.5, 4, 1023, 2),
                                                  (0.95, 1, 64, 1)])
def test_batched_bin_parameters(api, alpha, b, chunk, unroll):
    from reset_amd import inputs

    batch = inputs.make_pair_batch(3000, seed=5, hi=20000)
    opts = api.default_pairs_opts()
    for i in range(opts.num_bins):
        opts.bins[i].load_factor = alpha
        opts.bins[i].bucket_size = b
        opts.bins[i].chunk = chunk
        opts.bins[i].unroll = unroll
    plan = api.PairsPlan(3000, opts)
    got = plan.run(batch["a_off"], batch["a"], batch["b_off"], batch["b"])
    assert torch.equal(got, batch["k"])
    plan.close()


def test_schedules(api, oracle_lib):
    """Static and work-stealing schedules give the constructed sizes.  With
    16 waves per CU (4096 waves on 256 CUs) runs are cut to leave 4 runs per
    wave: on 140k pairs runs of 1, 3 and 8 are kept and runs of 16 and 64 cut
    to 8; the small batches (an empty one included) are dealt statically;
    graph replays of a stealing batch match too (the counter is reset by a
    kernel every launch)."""
    from reset_amd import inputs
    from reset_amd.lib import RS_PAIRS_STEAL, ResetError

    P = 140000
    batch = inputs.make_pair_batch(P, seed=41, hi=2000)
    for k in (0, 1, 3, 8, 16, 64):
        opts = api.default_pairs_opts()
        opts.schedule = k
        plan = api.PairsPlan(P, opts)
        got = plan.run(batch["a_off"], batch["a"], batch["b_off"], batch["b"])
        assert torch.equal(got, batch["k"]), k
        plan.close()
    opts = api.default_pairs_opts()
    assert opts.schedule == RS_PAIRS_STEAL
    plan = api.PairsPlan(P, opts)
    for n in (0, 1, 7, 300, 20000):
        sub = plan.run(batch["a_off"][:n + 1].contiguous(), batch["a"], batch["b_off"][:n + 1].contiguous(),
                       batch["b"])
        assert torch.equal(sub, batch["k"][:n]), n
    s = torch.cuda.Stream()
    sizes = torch.zeros(P, dtype=torch.int32, device="cuda")
    with torch.cuda.stream(s):
        plan.capture(batch["a_off"], batch["a"], batch["b_off"], batch["b"], sizes, s)
    for rep in range(3):
        sizes.fill_(-7)
        torch.cuda.synchronize()
        plan.replay(s)
        s.synchronize()
        assert torch.equal(sizes, batch["k"]), rep
    plan.status()
    plan.close()
    for k in (-1, 65):
        bad = api.default_pairs_opts()
        bad.schedule = k
        with pytest.raises(ResetError):
            api.PairsPlan(10, bad)


def test_bin_parameters_rejected_when_table_exceeds_lds(api):
    from reset_amd.lib import ResetError

    opts = api.default_pairs_opts()
    opts.bins[3].chunk = 5000
    with pytest.raises(ResetError):
        api.PairsPlan(10, opts)


def test_indexed_pairs_match_oracle(api, oracle_lib):
    """Pairs given as (u, v) indices into one sorted set family (the form a
    similarity join hands over), including repeated and self pairs."""
    rng = np.random.default_rng(12)
    off, cols = G.random_sets(rng, 400, 3000, 1 << 22)
    P = 5000
    u = rng.integers(0, 400, P).astype(np.int32)
    v = rng.integers(0, 400, P).astype(np.int32)
    v[:50] = u[:50]
    plan = api.PairsPlan(P)
    sizes = torch.empty(P, dtype=torch.int32, device="cuda")
    from reset_amd.api import ptr, stream_handle
    from reset_amd.lib import check, load

    offd, colsd = torch.from_numpy(off).cuda(), torch.from_numpy(cols).cuda()
    ud, vd = torch.from_numpy(u).cuda(), torch.from_numpy(v).cuda()
    check(load().rs_pairs_intersect_indexed_async(plan._h, ptr(offd), ptr(colsd), ptr(ud), ptr(vd), P,
                                                  ptr(sizes), stream_handle()))
    plan.status()
    a_off = np.concatenate([[0], np.cumsum(np.diff(off)[u])]).astype(np.int64)
    b_off = np.concatenate([[0], np.cumsum(np.diff(off)[v])]).astype(np.int64)
    A = np.concatenate([cols[off[x]:off[x + 1]] for x in u])
    B = np.concatenate([cols[off[y]:off[y + 1]] for y in v])
    want = oracle_lib.pairs_merge(a_off, A, b_off, B)
    assert np.array_equal(sizes.cpu().numpy(), want)
    assert np.array_equal(sizes.cpu().numpy()[:50], np.diff(off)[u[:50]])
    plan.close()


def test_autotuned_opts_stay_bit_exact(api, oracle_lib):
    """The autotuner's per-bin choice keeps results exact (the tuned plan is
    checked against the constructed overlaps and the oracle)."""
    from reset_amd import inputs

    batch = inputs.make_pair_batch(3000, seed=77, lo=16, hi=20000)
    o = api.autotune_pairs(batch["a_off"], batch["a"], batch["b_off"], batch["b"], reps=1)
    table = api.pairs_opts_table(o)
    assert all(b in (1, 2, 4) and 64 <= c <= 1024 and 0 < al <= 1 and lanes in (16, 32, 64)
               for al, b, c, _, lanes in table)
    plan = api.PairsPlan(3000, o)
    sizes = torch.empty(3000, dtype=torch.int32, device="cuda")
    plan.run_async(batch["a_off"], batch["a"], batch["b_off"], batch["b"], sizes)
    torch.cuda.synchronize()
    plan.status()
    assert torch.equal(sizes, batch["k"])
    ha, hb = batch["a_off"].cpu().numpy(), batch["b_off"].cpu().numpy()
    ref = oracle_lib.pairs_merge(ha, batch["a"].cpu().numpy(), hb, batch["b"].cpu().numpy())
    assert np.array_equal(ref, sizes.cpu().numpy())
    plan.close()


def test_plain_form_reports_overflow(api):
    """The plain form returns RS_EOVERFLOW where its table cannot hold the
    larger set (load factor 8, one slot per bucket: rejected before any
    launch) and where a probe of an absent key walks a completely full table
    (load factor 2, 16 keys in 16 slots: flagged by the probe kernel)."""
    from reset_amd.lib import RS_EOVERFLOW, ResetError

    A = np.arange(1000, 1016, dtype=np.int32)          # absent from B
    B = np.arange(0, 16 * 37, 37, dtype=np.int32)
    off = np.array([0, 16], np.int64)
    d = _dev(off, A, off, B)
    for alpha, b in ((8.0, 1), (2.0, 1)):
        with pytest.raises(ResetError) as e:
            api.pairs_intersect_plain(*d, load_factor=alpha, bucket_size=b)
        assert e.value.status == RS_EOVERFLOW
    # same pair at the default hyperparameters: exact, and the error word was cleared
    assert api.pairs_intersect_plain(*d).cpu().tolist() == [0]
    d2 = _dev(off, B, off, B)
    assert api.pairs_intersect_plain(*d2, load_factor=2.0, bucket_size=1).cpu().tolist() == [16]


def test_batched_flags_unsorted_sets_on_the_chunked_path(api, oracle_lib):
    """A pair whose smaller set exceeds the bin's chunk is built chunk by
    chunk in value order: shuffled sets there are reported (RS_EINVAL), not
    counted.  Pairs whose smaller set fits one table (the resident path) need
    no order and stay exact when shuffled."""
    from reset_amd.lib import RS_EINVAL, ResetError

    rng = np.random.default_rng(8)
    a_off, a, b_off, b, k = G.make_pairs(rng, 200, 3000, 1 << 24)
    small = (np.diff(a_off) <= 255) | (np.diff(b_off) <= 255)
    for off, el in ((a_off, a), (b_off, b)):
        for i in range(len(off) - 1):
            rng.shuffle(el[off[i]:off[i + 1]])
    plan = api.PairsPlan(200)
    with pytest.raises(ResetError) as e:
        plan.run(*_dev(a_off, a, b_off, b))
    assert e.value.status == RS_EINVAL
    # the resident-path pairs alone: exact although shuffled
    idx = np.flatnonzero(small)
    assert len(idx) >= 10
    A = [a[a_off[i]:a_off[i + 1]] for i in idx]
    B = [b[b_off[i]:b_off[i + 1]] for i in idx]
    ao = np.concatenate([[0], np.cumsum([len(x) for x in A])]).astype(np.int64)
    bo = np.concatenate([[0], np.cumsum([len(x) for x in B])]).astype(np.int64)
    got = plan.run(*_dev(ao, np.concatenate(A), bo, np.concatenate(B))).cpu().numpy()
    assert np.array_equal(got, k[idx])
    plan.close()


def _clustered_set(rng, n, lo, hi):
    """n distinct ascending values in [lo, hi]: runs of consecutive integers
    at a few random places, so a 256-key chunk's values are bunched into a
    sliver of its range."""
    vals = set()
    while len(vals) < n:
        start = int(rng.integers(lo, hi))
        run = int(rng.integers(1, 200))
        vals.update(range(start, min(hi + 1, start + run)))
    return np.array(sorted(vals)[:n], np.int64)


def test_ordered_chunk_tables_clustered_and_extreme_values(api, oracle_lib):
    """Pairs whose sets are within 8x of each other and larger than the
    chunk take ordered chunk tables.  Values bunched into runs of consecutive
    integers put many keys into few buckets (long displacements, the walk
    past the home bucket); values 0 and 2^31 - 1 sit at the ends of the
    order-preserving hash; sets start at every offset mod 4 within their
    arrays (the staging groups' leading inactive lanes).  Each size equals
    the oracle's, with the roles of the two sets swapped as well."""
    rng = np.random.default_rng(77)
    a_parts, b_parts = [], []
    for i in range(48):
        ns = int(rng.integers(300, 3000))
        ng = int(ns * rng.uniform(1.0, 7.5))
        if i % 3 == 0:
            S = _clustered_set(rng, ns, 0, 1 << 20)
            G_ = np.union1d(S[rng.random(ns) < 0.4], _clustered_set(rng, ng, 0, 1 << 20))[:ng]
        elif i % 3 == 1:
            S = np.unique(np.concatenate([[0, (1 << 31) - 1], rng.choice(1 << 31, ns, replace=False)]))
            G_ = np.unique(np.concatenate([[0, (1 << 31) - 1], S[rng.random(len(S)) < 0.3],
                                           rng.choice(1 << 31, ng, replace=False)]))
        else:  # consecutive integers: the chunk's range is exactly its size
            base = int(rng.integers(0, 1 << 24))
            S = np.arange(base, base + ns, dtype=np.int64)
            G_ = np.unique(np.concatenate([S[::3], np.arange(base + ns, base + ns + ng // 2)]))
        a_parts.append(S)
        b_parts.append(np.sort(G_))
    # pad so that the sets start at every offset mod 4 within their arrays
    pad_a = [np.array([7] * (i % 4), np.int64) for i in range(len(a_parts))]
    A = np.concatenate([x for pa, s_ in zip(pad_a, a_parts) for x in (pa, s_)])
    starts = np.cumsum([0] + [len(pa) + len(s_) for pa, s_ in zip(pad_a, a_parts)])[:-1] + \
        np.array([len(pa) for pa in pad_a])
    a_off_lo = starts.astype(np.int64)
    a_off_hi = (starts + np.array([len(s_) for s_ in a_parts])).astype(np.int64)
    # per-pair CSR over a copy with the pads: pair i = A[a_off_lo[i]:a_off_hi[i]]
    Ac = np.concatenate([A[lo_:hi_] for lo_, hi_ in zip(a_off_lo, a_off_hi)])
    assert np.array_equal(Ac, np.concatenate(a_parts))
    b_off = np.concatenate([[0], np.cumsum([len(x) for x in b_parts])]).astype(np.int64)
    Bc = np.concatenate(b_parts)
    a_off = np.concatenate([[0], np.cumsum([len(x) for x in a_parts])]).astype(np.int64)
    want = np.array([len(np.intersect1d(x, y)) for x, y in zip(a_parts, b_parts)], np.int32)
    assert np.array_equal(oracle_lib.pairs_hash(a_off, Ac, b_off, Bc), want)
    plan = api.PairsPlan(len(want))
    # unaligned starts: the A family read in place from the padded array, one
    # pair per call (its offsets index the padded array)
    dA = torch.from_numpy(A.astype(np.int32)).cuda()
    for i in range(len(want)):
        ao = torch.tensor([a_off_lo[i], a_off_hi[i]], dtype=torch.int64, device="cuda")
        bo = torch.tensor([b_off[i], b_off[i + 1]], dtype=torch.int64, device="cuda")
        dB = torch.from_numpy(Bc.astype(np.int32)).cuda()
        got = plan.run(ao, dA, bo, dB).cpu().tolist()
        assert got == [int(want[i])], (i, got, int(want[i]))
    # the whole batch, both role orders
    got = plan.run(*_dev(a_off, Ac, b_off, Bc)).cpu().numpy()
    assert np.array_equal(got, want)
    got = plan.run(*_dev(b_off, Bc, a_off, Ac)).cpu().numpy()
    assert np.array_equal(got, want)
    plan.close()


def test_ordered_chunk_flags_unsorted_or_duplicated_keys(api):
    """Two swapped neighbours or a repeated key inside one chunk of an
    ordered-path pair are reported (RS_EINVAL), not counted."""
    from reset_amd.lib import RS_EINVAL, ResetError

    rng = np.random.default_rng(91)
    S = np.sort(rng.choice(1 << 24, 2000, replace=False))
    G_ = np.sort(rng.choice(1 << 24, 5000, replace=False))
    off_s = np.array([0, len(S)], np.int64)
    off_g = np.array([0, len(G_)], np.int64)
    plan = api.PairsPlan(1)
    assert plan.run(*_dev(off_s, S, off_g, G_)).cpu().tolist() == [len(np.intersect1d(S, G_))]
    for bad in ("swap", "dup"):
        T = S.copy()
        if bad == "swap":
            T[1000], T[1001] = T[1001], T[1000]
        else:
            T[1001] = T[1000]
        with pytest.raises(ResetError) as e:
            plan.run(*_dev(off_s, T, off_g, G_))
        assert e.value.status == RS_EINVAL
    plan.close()


def _mixed_opts(api):
    """Bins with different bucket sizes and lanes per pair: bins 0-3 on 16
    lanes (b 1), 4-5 on 32 lanes (b 2), 6 on 16 lanes (b 4), 7-9 a wave with
    b 2, 10-12 a wave with b 1, the rest the defaults (b 4) -- six launches."""
    opts = api.default_pairs_opts()
    def put(bins, alpha, b, chunk, lanes):
        for i in bins:
            opts.bins[i].load_factor, opts.bins[i].bucket_size = alpha, b
            opts.bins[i].chunk, opts.bins[i].lanes = chunk, lanes
    put(range(0, 4), 0.125, 1, 255, 16)
    put(range(4, 6), 0.25, 2, 255, 32)
    put([6], 0.125, 4, 255, 16)
    put(range(7, 10), 0.125, 2, 127, 64)
    put(range(10, 13), 0.0625, 1, 127, 64)
    return opts


def test_mixed_bin_settings_stay_bit_exact(api, oracle_lib):
    """Per-bin bucket size and probe width (lanes per pair), launched as one
    kernel per run of equal settings: sizes equal the constructed overlaps and
    the oracle, eagerly and from graph replays."""
    from reset_amd import inputs

    batch = inputs.make_pair_batch(6000, seed=19, lo=1, hi=40000)
    opts = _mixed_opts(api)
    plan = api.PairsPlan(6000, opts)
    got = plan.run(batch["a_off"], batch["a"], batch["b_off"], batch["b"])
    assert torch.equal(got, batch["k"])
    ha, hb = batch["a_off"].cpu().numpy(), batch["b_off"].cpu().numpy()
    ref = oracle_lib.pairs_merge(ha, batch["a"].cpu().numpy(), hb, batch["b"].cpu().numpy())
    assert np.array_equal(ref, got.cpu().numpy())
    # bins 0..6 hold pairs (sizes from 1), so the part-wave kernels ran
    tot = np.diff(ha) + np.diff(hb)
    assert (tot < 128).sum() > 50
    s = torch.cuda.Stream()
    sizes = torch.zeros(6000, dtype=torch.int32, device="cuda")
    with torch.cuda.stream(s):
        plan.capture(batch["a_off"], batch["a"], batch["b_off"], batch["b"], sizes, s)
    for rep in range(3):
        sizes.fill_(-7)
        torch.cuda.synchronize()
        plan.replay(s)
        s.synchronize()
        assert torch.equal(sizes, batch["k"]), rep
        torch.cuda.synchronize()
    plan.status()
    plan.close()


def test_part_wave_lanes_rejected_where_they_do_not_fit(api):
    """16 lanes per pair in the last bin, or where a table of 2^bin - 1 keys
    exceeds a quarter of the wave's region, is refused at plan creation."""
    from reset_amd.lib import ResetError

    for bin_, lanes in ((19, 16), (12, 16), (9, 32)):
        opts = api.default_pairs_opts()
        opts.bins[bin_].lanes = lanes
        with pytest.raises(ResetError):
            api.PairsPlan(10, opts)
    opts = api.default_pairs_opts()
    opts.bins[3].lanes = 17
    with pytest.raises(ResetError):
        api.PairsPlan(10, opts)
