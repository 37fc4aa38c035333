"""Small numpy graph / set builders shared by the CPU and GPU tests."""
from __future__ import annotations

import itertools

import numpy as np


def undirected(edges, n):
    """Dedup, drop self-loops -> array of (u, v) with u < v."""
    e = np.asarray(list(edges), np.int64).reshape(-1, 2)
    e = e[e[:, 0] != e[:, 1]]
    lo = np.minimum(e[:, 0], e[:, 1])
    hi = np.maximum(e[:, 0], e[:, 1])
    key = np.unique(lo * n + hi)
    return np.stack([key // n, key % n], 1)


def orient(und, n, by="degree"):
    """Acyclic orientation: by id (reference, preprocessor.py v > u) or by
    (degree, id) (low degree -> high degree)."""
    u, v = und[:, 0], und[:, 1]
    if by == "id":
        src, dst = u, v
    else:
        deg = np.bincount(np.concatenate([u, v]), minlength=n)
        fwd = (deg[u] < deg[v]) | ((deg[u] == deg[v]) & (u < v))
        src = np.where(fwd, u, v)
        dst = np.where(fwd, v, u)
    return csr(src, dst, n)


def symmetric(und, n):
    """Full undirected adjacency (both directions), as the SSS inputs."""
    return csr(np.concatenate([und[:, 0], und[:, 1]]), np.concatenate([und[:, 1], und[:, 0]]), n)


def csr(src, dst, n):
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    off = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(src, minlength=n), out=off[1:])
    return off.astype(np.int32), dst.astype(np.int32)


def complete(n):
    return [(i, j) for i, j in itertools.combinations(range(n), 2)]


def wheel(r):
    """hub 0 plus an r-cycle on 1..r: exactly r triangles for r >= 4."""
    rim = [(1 + i, 1 + (i + 1) % r) for i in range(r)]
    return [(0, 1 + i) for i in range(r)] + rim


def star(n):
    return [(0, i) for i in range(1, n)]


def complete_bipartite(a, b):
    return [(i, a + j) for i in range(a) for j in range(b)]


def random_graph(n, m, rng):
    return rng.integers(0, n, size=(m, 2))


def brute_triangles(und, n):
    A = np.zeros((n, n), np.int64)
    A[und[:, 0], und[:, 1]] = 1
    A[und[:, 1], und[:, 0]] = 1
    return int(np.trace(A @ A @ A) // 6)


def random_sets(rng, count, max_size, universe, min_size=0):
    """CSR of sorted unique random sets."""
    sizes = rng.integers(min_size, max_size + 1, size=count)
    rows = [np.unique(rng.integers(0, universe, size=s)) for s in sizes]
    off = np.zeros(count + 1, np.int64)
    np.cumsum([len(r) for r in rows], out=off[1:])
    cols = np.concatenate(rows) if rows else np.zeros(0, np.int32)
    return off.astype(np.int32), cols.astype(np.int32)


def zipf_sets(rng, count, vocab, min_size, max_size, s_size=1.2, s_tok=1.0):
    """Sets with Zipf sizes and Zipf token frequencies (sorted, unique)."""
    ks = np.arange(min_size, max_size + 1)
    ps = ks.astype(float) ** -s_size
    sizes = rng.choice(ks, size=count, p=ps / ps.sum())
    ranks = np.arange(1, vocab + 1, dtype=float) ** -s_tok
    ranks /= ranks.sum()
    rows = []
    for s in sizes:
        s = min(int(s), vocab)
        got = set()
        while len(got) < s:
            got.update(rng.choice(vocab, size=s - len(got), p=ranks).tolist())
        rows.append(np.array(sorted(got), np.int32))
    off = np.zeros(count + 1, np.int64)
    np.cumsum([len(r) for r in rows], out=off[1:])
    return off.astype(np.int32), np.concatenate(rows).astype(np.int32)


def make_pairs(rng, count, max_size, universe, overlap_frac=0.5):
    """Pairs of sorted unique sets with a chosen overlap -> (a_off, a, b_off, b, k)."""
    A, B, K = [], [], []
    for _ in range(count):
        na, nb = rng.integers(0, max_size + 1, size=2)
        k = int(rng.integers(0, int(min(na, nb) * overlap_frac) + 1))
        pool = rng.choice(universe, size=na + nb - k, replace=False)
        a = np.sort(pool[:na])
        b = np.sort(np.concatenate([pool[na - k:na], pool[na:]]))
        A.append(a)
        B.append(b)
        K.append(k)
    a_off = np.zeros(count + 1, np.int64)
    b_off = np.zeros(count + 1, np.int64)
    np.cumsum([len(x) for x in A], out=a_off[1:])
    np.cumsum([len(x) for x in B], out=b_off[1:])
    cat = lambda xs: (np.concatenate(xs) if xs else np.zeros(0)).astype(np.int32)
    return a_off, cat(A), b_off, cat(B), np.array(K, np.int32)


def inverted_index(rng, num_terms, num_docs, max_len):
    lens = rng.integers(0, max_len + 1, size=num_terms)
    rows = [np.sort(rng.choice(num_docs, size=int(min(l, num_docs)), replace=False)) for l in lens]
    off = np.zeros(num_terms + 1, np.int64)
    np.cumsum([len(r) for r in rows], out=off[1:])
    return off.astype(np.int32), np.concatenate(rows).astype(np.int32)


def queries(rng, idx_off, count, kmin=1, kmax=5):
    """Queries of distinct terms sorted by (df, id), as ir preprocessing does."""
    nterms = len(idx_off) - 1
    df = np.diff(idx_off)
    q, qoff = [], [0]
    for _ in range(count):
        k = int(rng.integers(kmin, kmax + 1))
        ts = rng.choice(nterms, size=min(k, nterms), replace=False)
        ts = sorted(ts.tolist(), key=lambda t: (df[t], t))
        q.extend(ts)
        qoff.append(len(q))
    return np.array(q, np.int32), np.array(qoff, np.int32)
