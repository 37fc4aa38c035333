"""Synthetic workload inputs built on the device (seeded, reproducible).

R-MAT graphs: HIP edge generator, then dedup / self-loop removal / orientation
/ CSR with torch device ops.  Pair batches: per-pair sizes and overlaps drawn
on the host with numpy, sets generated on the device (sorted, unique, exact
overlap).  IR / set-similarity collections: Zipf inputs.
These are input plumbing, not the measured path.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import api

RMAT_A, RMAT_B, RMAT_C = 0.57, 0.19, 0.19


def rmat_undirected(scale, edge_factor=16, seed=20261203, device="cuda"):
    """Unique undirected edges (u < v, no self-loops) of an R-MAT graph:
    int64 tensors (lo, hi), n = 2^scale."""
    n = 1 << scale
    m = edge_factor * n
    e = api.gen_rmat_edges(scale, m, RMAT_A, RMAT_B, RMAT_C, seed=seed, device=device).long()
    u, v = e[:, 0], e[:, 1]
    del e
    keep = u != v
    u, v = u[keep], v[keep]
    lo = torch.minimum(u, v)
    hi = torch.maximum(u, v)
    del u, v, keep
    key = torch.unique(lo * n + hi)
    return key // n, key % n, n


def orient_csr(lo, hi, n, by="degree"):
    """Acyclic orientation low-degree -> high-degree (ties by id), or by id
    (u -> v for u < v, the reference preprocessing's rule); int32 CSR with
    rows sorted ascending."""
    if by == "degree":
        deg = torch.bincount(torch.cat([lo, hi]), minlength=n)
        fwd = (deg[lo] < deg[hi]) | ((deg[lo] == deg[hi]) & (lo < hi))
        src = torch.where(fwd, lo, hi)
        dst = torch.where(fwd, hi, lo)
        del deg, fwd
    else:
        src, dst = lo, hi
    key, _ = torch.sort(src * n + dst)
    src = key // n
    dst = (key % n).to(torch.int32)
    off = torch.zeros(n + 1, dtype=torch.int64, device=lo.device)
    off[1:] = torch.cumsum(torch.bincount(src, minlength=n), 0)
    return off.to(torch.int32), dst


def symmetric_csr(lo, hi, n):
    """Both directions (the set-similarity CSR of preprocessor.py:161-184)."""
    src = torch.cat([lo, hi])
    dst = torch.cat([hi, lo])
    key, _ = torch.sort(src * n + dst)
    off = torch.zeros(n + 1, dtype=torch.int64, device=lo.device)
    off[1:] = torch.cumsum(torch.bincount(key // n, minlength=n), 0)
    return off.to(torch.int32), (key % n).to(torch.int32)


def rmat_tc_csr(scale, edge_factor=16, seed=20261203, by="degree", device="cuda"):
    lo, hi, n = rmat_undirected(scale, edge_factor, seed, device)
    return orient_csr(lo, hi, n, by)


# ---------------------------------------------------------------- pairs
def pair_meta(num_pairs, seed=20261203, lo=16, hi=65536, overlap_max=0.5):
    """Sizes log-uniform in [lo, hi] (floor(exp(U[ln lo, ln(hi+1))))) and
    overlap k = floor(f * min(|A|, |B|)), f ~ U[0, overlap_max]."""
    rng = np.random.default_rng(seed)
    ln_lo, ln_hi = math.log(lo), math.log(hi + 1)
    na = np.floor(np.exp(rng.uniform(ln_lo, ln_hi, num_pairs))).astype(np.int64)
    nb = np.floor(np.exp(rng.uniform(ln_lo, ln_hi, num_pairs))).astype(np.int64)
    na = np.clip(na, lo, hi)
    nb = np.clip(nb, lo, hi)
    f = rng.uniform(0.0, overlap_max, num_pairs)
    k = np.floor(f * np.minimum(na, nb)).astype(np.int64)
    return na, nb, k


def make_pair_batch(num_pairs, seed=20261203, device="cuda", **kw):
    """-> dict(a_off, a, b_off, b, k) on the device (int64 offsets, int32 sets)."""
    na, nb, k = pair_meta(num_pairs, seed, **kw)
    a_off = np.zeros(num_pairs + 1, np.int64)
    b_off = np.zeros(num_pairs + 1, np.int64)
    np.cumsum(na, out=a_off[1:])
    np.cumsum(nb, out=b_off[1:])
    a_off_d = torch.from_numpy(a_off).to(device)
    b_off_d = torch.from_numpy(b_off).to(device)
    k_d = torch.from_numpy(k.astype(np.int32)).to(device)
    a, b = api.gen_pair_sets(a_off_d, b_off_d, k_d, seed)
    return dict(a_off=a_off_d, a=a, b_off=b_off_d, b=b, k=k_d)


# ---------------------------------------------------------------- set collections
def _distinct_rows(counts, universe, draw, gen, device):
    """CSR of rows with exactly counts[i] distinct values each, sorted: draw the
    missing number of values per row (draw(n, gen) -> int64 values), dedup the
    (row, value) keys, repeat until every row is full."""
    n = counts.numel()
    keys = torch.empty(0, dtype=torch.int64, device=device)
    have = torch.zeros(n, dtype=torch.int64, device=device)
    for _ in range(64):
        deficit = counts - have
        tot = int(deficit.sum().item())
        if tot == 0:
            break
        rows = torch.repeat_interleave(torch.arange(n, device=device), deficit)
        vals = draw(tot, gen)
        keys = torch.unique(torch.cat([keys, rows * universe + vals]))
        have = torch.bincount(keys // universe, minlength=n)
    else:
        raise RuntimeError("distinct sampling did not converge")
    off = torch.zeros(n + 1, dtype=torch.int64, device=device)
    off[1:] = torch.cumsum(have, 0)
    return off.to(torch.int32), (keys % universe).to(torch.int32)


def zipf_set_collection(num_sets=1_000_000, vocab=1_000_000, s_min=10, s_max=2000, z_size=1.2, z_tok=1.0,
                        seed=20261203, device="cuda"):
    """Set-similarity input: sizes Zipf(z_size) on [s_min, s_max] (inverse CDF),
    tokens of rank r drawn with probability proportional to r^-z_tok (token id =
    r - 1), each set distinct and sorted."""
    rng = np.random.default_rng(seed)
    ks = np.arange(s_min, s_max + 1)
    ps = ks.astype(np.float64) ** -z_size
    sizes = rng.choice(ks, size=num_sets, p=ps / ps.sum())
    sizes = np.minimum(sizes, vocab)
    w = torch.arange(1, vocab + 1, dtype=torch.float64, device=device) ** -z_tok
    cdf = torch.cumsum(w, 0)
    cdf /= cdf[-1].clone()
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)

    def draw(k, g):
        u = torch.rand(k, generator=g, dtype=torch.float64, device=device)
        return torch.clamp(torch.searchsorted(cdf, u), max=vocab - 1)

    return _distinct_rows(torch.from_numpy(sizes).to(device), vocab, draw, gen, device)


def ir_index(num_docs=10_000_000, num_terms=1_000_000, z=1.1, cap=2_000_000, seed=20261203, device="cuda"):
    """Inverted index: term of rank r has L_r = clamp(floor(D / r^z), 1, cap)
    postings, doc ids uniform without replacement in [0, D), sorted."""
    r = np.arange(1, num_terms + 1, dtype=np.float64)
    lens = np.clip(np.floor(num_docs / r ** z), 1, cap).astype(np.int64)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)

    def draw(k, g):
        return torch.randint(0, num_docs, (k,), generator=g, device=device, dtype=torch.int64)

    return _distinct_rows(torch.from_numpy(lens).to(device), num_docs, draw, gen, device)


def ir_queries(index_offsets, num_queries, kmin=2, kmax=8, seed=20261203):
    """Conjunctive queries: k ~ U{kmin..kmax} distinct terms drawn with
    probability proportional to their posting-list length, ordered by
    (document frequency, id) as the reference preprocessing orders them
    (preprocessor.py:345-349).  -> (terms int32, offsets int32) on the host."""
    off = np.asarray(index_offsets.cpu().numpy() if hasattr(index_offsets, "cpu") else index_offsets, np.int64)
    df = np.diff(off)
    cdf = np.cumsum(df, dtype=np.float64)
    cdf /= cdf[-1]
    rng = np.random.default_rng(seed)
    ks = rng.integers(kmin, kmax + 1, size=num_queries)
    terms, qoff = [], [0]
    for k in ks.tolist():
        got = []
        while len(got) < k:  # draws proportional to df, repeats rejected
            t = int(min(np.searchsorted(cdf, rng.random()), len(df) - 1))
            if t not in got:
                got.append(t)
        got.sort(key=lambda t: (df[t], t))
        terms.extend(got)
        qoff.append(len(terms))
    return np.array(terms, np.int32), np.array(qoff, np.int32)


def plant_near_duplicates(off, cols, count, seed=20261213, vocab=1_000_000):
    """Replace the last `count` sets of a collection by copies of distinct
    earlier sets with one token swapped for a vocabulary token the set lacks:
    J(copy, source) = (s - 1) / (s + 1) >= 4/5 for sets of s >= 9 tokens.
    -> (offsets, elements, planted) with planted the int64 keys
    source << 32 | copy (source < copy), ascending."""
    rng = np.random.default_rng(seed)
    offh = off.cpu().numpy().astype(np.int64)
    n = len(offh) - 1
    count = min(count, n // 2)
    keep = int(offh[n - count])
    tail, planted = [], []
    for j, src in enumerate(rng.choice(n - count, count, replace=False).tolist()):
        r = cols[int(offh[src]):int(offh[src + 1])].cpu().numpy().copy()
        have = set(r.tolist())
        x = int(rng.integers(0, vocab))
        while x in have:
            x = int(rng.integers(0, vocab))
        r[int(rng.integers(0, len(r)))] = x
        tail.append(np.sort(r))
        planted.append((src << 32) | (n - count + j))
    lens = np.array([len(r) for r in tail], np.int64)
    new_off = offh.copy()
    new_off[n - count + 1:] = keep + np.cumsum(lens)
    new_cols = torch.cat([cols[:keep], torch.from_numpy(np.concatenate(tail).astype(np.int32)).to(cols.device)])
    return (torch.from_numpy(new_off.astype(np.int32)).to(off.device), new_cols,
            np.sort(np.array(planted, np.int64)))
