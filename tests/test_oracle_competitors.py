"""CPU pinning of the oracle's competitor tables (oracle/competitors.c).

Known answers worked out by hand from the reference's rules, invariants that
any correct table of the kind satisfies, and the consumers' counts against
the sorted-merge and brute-force oracles.
"""
import itertools

import numpy as np
import pytest

import graphs as G
import oracle

EMPTY = oracle.COMP_EMPTY


def _csr(rows):
    off = np.zeros(len(rows) + 1, np.int64)
    np.cumsum([len(r) for r in rows], out=off[1:])
    el = np.concatenate([np.asarray(r, np.int64) for r in rows]) if rows else np.zeros(0)
    return off.astype(np.int32), el.astype(np.int32)


def _unfmix32(h):
    """Inverse of MurmurHash3's finalizer: xorshift inverses and the modular
    inverses of its two multipliers (0x85ebca6b^-1 = 0xa5cb9243,
    0xc2b2ae35^-1 = 0x7ed1b41d mod 2^32)."""
    m = 0xFFFFFFFF
    h ^= h >> 16
    h = (h * 0x7ED1B41D) & m
    h ^= (h >> 13) ^ (h >> 26)
    h = (h * 0xA5CB9243) & m
    h ^= h >> 16
    return h


def test_fmix32_known_answers_and_bijection():
    assert oracle.fmix32(0) == 0
    assert oracle.fmix32(1) == 0x514E28B7  # MurmurHash3 fmix32(1), published value
    assert (0x85EBCA6B * 0xA5CB9243) & 0xFFFFFFFF == 1 and (0xC2B2AE35 * 0x7ED1B41D) & 0xFFFFFFFF == 1
    rng = np.random.default_rng(0)
    xs = rng.integers(0, 1 << 32, 2000, dtype=np.uint64).tolist()
    hs = [oracle.fmix32(int(x)) for x in xs]
    assert [_unfmix32(h) for h in hs] == xs
    assert len(set(hs)) == len(set(xs))


def test_cuckoo_two_keys_sharing_both_positions_end_in_the_stash():
    """Set 0, keys 2 and 4, load factor 1: capacity 2, and both keys have
    both positions at slot 0 ((k ^ k>>8 ^ 0) & 1 = 0, and 0xAAAAAAAA is even).
    4 evicts 2, 2 evicts 4, ... 6000 evictions: the last (the 6000th, odd
    index 5999) writes 2 back, leaving 4 to the stash."""
    off, el = _csr([[2, 4]])
    t = oracle.comp_build("cuckoo", off, el, 1.0)
    a = t.arrays
    assert a["toff"].tolist() == [0, 2]
    assert a["slots"].tolist() == [2, EMPTY]
    assert a["stash_count"].tolist() == [1]
    assert a["stash"][0] == 4 and set(a["stash"][1:].tolist()) == {EMPTY}
    assert t.failed == 0
    assert oracle.comp_contains(t, 0, 2) and oracle.comp_contains(t, 0, 4)
    assert not oracle.comp_contains(t, 0, 6)


def test_cuckoo_stash_overflow_counts_failures():
    """200 keys at load factor 8: capacity ceil(200/8) = 25 -> 32 slots, so
    the table and the 16-entry stash hold 48 keys at most and every other
    key is lost and counted."""
    rows = [list(range(0, 400, 2))]
    off, el = _csr(rows)
    t = oracle.comp_build("cuckoo", off, el, 8.0)
    cap = int(t.arrays["toff"][1])
    placed = int(np.sum(t.arrays["slots"] != EMPTY)) + int(t.arrays["stash_count"][0])
    assert cap == 32 and t.arrays["stash_count"][0] == 16
    assert placed + t.failed == 200
    kept = [k for k in rows[0] if oracle.comp_contains(t, 0, k)]
    assert len(kept) == placed


def _hop_home(k, s, cap):
    return (k ^ (k >> 8) ^ s) & (cap - 1)


def test_hopscotch_full_neighbourhood_fails_the_33rd_key():
    """33 keys 257*j (j even, j <= 64) of set 0 all have home 0 under the
    capacity 512 (33 / 0.1 -> 512): the first 32 fill slots 0..31, the 33rd
    finds slot 32 free but no home in 1..32 holds an element to swap back."""
    keys = [257 * j for j in range(0, 65, 2)]
    assert len(keys) == 33 and {_hop_home(k, 0, 512) for k in keys} == {0}
    off, el = _csr([keys])
    t = oracle.comp_build("hopscotch", off, el, 0.2)
    a = t.arrays
    assert a["toff"].tolist() == [0, 512 + 32]
    assert a["slots"][:32].tolist() == keys[:32] and set(a["slots"][32:].tolist()) == {EMPTY}
    assert int(a["bitmap"][0]) == 0xFFFFFFFF and not a["bitmap"][1:].any()
    assert t.failed == 1
    assert not oracle.comp_contains(t, 0, keys[32])


def test_hopscotch_displacement_moves_a_neighbour_back():
    """Row [1, 0, 514, ..., 257*60, 257*62]: key 1 has home 1 (slot 1); the 31
    home-0 keys before the last fill slots 0 and 2..31; the last home-0 key
    finds slot 32 free, the latest home with an element before slot 32 is 1
    (bit 0), so key 1 moves to slot 32 (bit 31 of home 1) and the new key
    takes slot 1 (bit 1 of home 0)."""
    zero = [257 * j for j in range(0, 63, 2)]
    assert _hop_home(1, 0, 512) == 1 and {_hop_home(k, 0, 512) for k in zero} == {0}
    row = [1] + zero
    off, el = _csr([row])
    t = oracle.comp_build("hopscotch", off, el, 0.2)
    a = t.arrays
    s = a["slots"].tolist()
    assert s[0] == 0 and s[1] == zero[-1] and s[2:32] == zero[1:31] and s[32] == 1
    assert int(a["bitmap"][0]) == 0xFFFFFFFF and int(a["bitmap"][1]) == 0x80000000
    assert t.failed == 0
    assert all(oracle.comp_contains(t, 0, k) for k in row)


def _random_sets(seed, count=150, hi=400, universe=1 << 20):
    rng = np.random.default_rng(seed)
    return G.random_sets(rng, count, hi, universe)


@pytest.mark.parametrize("kind", ["linear", "cuckoo", "hopscotch", "roaring"])
def test_membership_matches_the_sets(kind):
    off, el = _random_sets(3)
    t = oracle.comp_build(kind, off, el, 0.2)
    assert t.failed == 0
    rng = np.random.default_rng(4)
    for s in range(0, len(off) - 1, 7):
        row = set(el[off[s]:off[s + 1]].tolist())
        probes = list(row)[:40] + rng.integers(0, 1 << 20, 40).tolist()
        for k in probes:
            assert oracle.comp_contains(t, s, int(k)) == (k in row)


def test_cuckoo_keys_sit_at_their_positions_or_in_the_stash():
    off, el = _random_sets(5)
    t = oracle.comp_build("cuckoo", off, el, 0.5)
    a = t.arrays
    for s in range(len(off) - 1):
        cap = int(a["toff"][s + 1] - a["toff"][s])
        region = a["slots"][a["toff"][s]:a["toff"][s + 1]]
        for pos, k in enumerate(region.tolist()):
            if k != EMPTY:
                h1 = (k ^ (k >> 8) ^ s) & (cap - 1)
                h2 = (k ^ (k >> 16) ^ (s << 1) ^ 0xAAAAAAAA) & (cap - 1)
                assert pos in (h1, h2)
        stored = [k for k in region.tolist() if k != EMPTY]
        stored += a["stash"][16 * s:16 * s + a["stash_count"][s]].tolist()
        assert sorted(stored) == sorted(el[off[s]:off[s + 1]].tolist())


def test_hopscotch_maps_point_at_their_homes_keys():
    off, el = _random_sets(6)
    t = oracle.comp_build("hopscotch", off, el, 0.2)
    a = t.arrays
    for s in range(len(off) - 1):
        base, total = int(a["toff"][s]), int(a["toff"][s + 1] - a["toff"][s])
        cap = total - oracle.HOP_H
        found = []
        for h in range(total):
            b = int(a["bitmap"][base + h])
            for j in range(32):
                if b >> j & 1:
                    k = int(a["slots"][base + h + j])
                    assert _hop_home(k, s, cap) == h
                    found.append(k)
        assert sorted(found) == sorted(el[off[s]:off[s + 1]].tolist())


def test_roaring_words_decode_to_the_sets_and_remap_is_the_rank():
    off, el = _random_sets(7, universe=5000)
    n = len(off) - 1
    t = oracle.comp_build("roaring", off, el)
    a = t.arrays
    ids = sorted(set(range(n)) | set(el.tolist()))
    rank = {v: i for i, v in enumerate(ids)}
    assert len(a["remap"]) == max(n, int(el.max()) + 1)
    assert all(a["remap"][v] == rank[v] for v in ids)
    assert all(a["remap"][v] == -1 for v in range(len(a["remap"])) if v not in rank)
    inv = {i: v for v, i in rank.items()}
    for s in range(n):
        lo, hi = a["node_start"][s], a["node_start"][s + 1]
        wi = a["word_index"][lo:hi].tolist()
        assert wi == sorted(set(wi))
        got = [inv[64 * w + b] for w, d in zip(wi, a["word_data"][lo:hi].tolist()) for b in range(64)
               if int(d) >> b & 1]
        assert sorted(got) == sorted(el[off[s]:off[s + 1]].tolist())


def test_linear_checker_catches_planted_mistakes():
    off, el = _random_sets(8, count=40)
    t = oracle.comp_build("linear", off, el, 0.5)
    toff, slots = t.arrays["toff"], t.arrays["slots"].copy()
    assert oracle.comp_linear_check(off, el, toff, slots) == 0
    # a dropped key
    k = int(np.flatnonzero(slots != EMPTY)[3])
    bad = slots.copy()
    bad[k] = EMPTY
    assert oracle.comp_linear_check(off, el, toff, bad) > 0
    # a key moved past an EMPTY slot of its probe sequence
    bad = slots.copy()
    empties = np.flatnonzero(bad[:toff[1]] == EMPTY)
    keys = np.flatnonzero(bad[:toff[1]] != EMPTY)
    if len(empties) and len(keys):
        bad[empties[0]], bad[keys[0]] = bad[keys[0]], EMPTY
        if oracle.fmix32(int(bad[empties[0]])) & (toff[1] - 1) != empties[0]:
            assert oracle.comp_linear_check(off, el, toff, bad) > 0


@pytest.mark.parametrize("kind", ["linear", "cuckoo", "hopscotch", "roaring"])
def test_consumers_equal_merge_and_bruteforce(kind):
    n = 300
    und = G.undirected(G.random_graph(n, 5000, np.random.default_rng(9)), n)
    for by in ("id", "degree"):
        off, cols = G.orient(und, n, by)
        t = oracle.comp_build(kind, off, cols, 0.2)
        assert oracle.comp_tc(t, off, cols) == oracle.tc_merge(off, cols)
    off, cols = G.zipf_sets(np.random.default_rng(10), 120, 300, 2, 50)
    pairs = np.array(list(itertools.combinations(range(120), 2)), np.int32)
    t = oracle.comp_build(kind, off, cols, 0.2)
    flags, sizes = oracle.comp_sss(t, off, cols, pairs[:, 0], pairs[:, 1], 1, 2)
    want = set(oracle.sss_bruteforce(off, cols, 1, 2).tolist())
    assert {(int(i) << 32) | int(j) for (i, j), f in zip(pairs, flags) if f} == want
    merged = oracle.pairs_merge(
        np.concatenate([[0], np.cumsum(np.diff(off)[pairs[:, 0]])]).astype(np.int64),
        np.concatenate([cols[off[i]:off[i + 1]] for i in pairs[:, 0]]),
        np.concatenate([[0], np.cumsum(np.diff(off)[pairs[:, 1]])]).astype(np.int64),
        np.concatenate([cols[off[j]:off[j + 1]] for j in pairs[:, 1]]))
    assert np.array_equal(sizes, merged)
