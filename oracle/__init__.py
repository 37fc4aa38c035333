"""ctypes front end of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import this module, and only
as the checker or as the timed CPU baseline.  Nothing under ``reset_amd/``
imports it.

All arrays are numpy; functions return fresh numpy outputs.  Each wrapper
names the reference code it restates (see oracle.c for line citations).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

NORMAL = 0
HIERARCHICAL = 1


def build() -> None:
    """Compile oracle.c into liboracle.so (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_p = C.c_void_p
_i32, _i64, _f32 = C.c_int32, C.c_int64, C.c_float


def _declare(L):
    L.orc_calculate_length.restype = _i32
    L.orc_calculate_length.argtypes = [_i32, _f32, _i32]
    L.orc_hash_normal.restype = C.c_uint32
    L.orc_hash_normal.argtypes = [C.c_uint32, C.c_uint32]
    L.orc_hash_hierarchical.restype = C.c_uint32
    L.orc_hash_hierarchical.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
    L.orc_table_geometry.restype = C.c_int
    L.orc_table_geometry.argtypes = [_p, _i64, _f32, _i32, _p, _p, _p]
    L.orc_table_build.restype = C.c_int
    L.orc_table_build.argtypes = [_p, _p, _i64, _p, _p, _i32, _i32, _i32, _p, _p]
    L.orc_table_search.restype = C.c_int
    L.orc_table_search.argtypes = [_p, _i64, C.c_uint32, _i32, _i32, _i32]
    L.orc_table_check.restype = _i64
    L.orc_table_check.argtypes = [_p, _p, _i64, _p, _p, _i32, _i32, _i32, _p]
    L.orc_extract_probe_order.restype = C.c_int
    L.orc_extract_probe_order.argtypes = [_p, _i64, _p, _i32, _p, _p]
    L.orc_tc_hash.restype = C.c_int
    L.orc_tc_hash.argtypes = [_p, _p, _i64, _p, _p, _i32, _i32, _i32, _p, C.c_int, _p]
    L.orc_tc_merge.restype = C.c_uint64
    L.orc_tc_merge.argtypes = [_p, _p, _i64, C.c_int]
    L.orc_tc_hash_rows.restype = C.c_int
    L.orc_tc_hash_rows.argtypes = [_p, _p, _i64, _i64, _i64, _p, _p, _i32, _i32, _i32, _p, C.c_int, _p]
    L.orc_tc_merge_rows.restype = C.c_uint64
    L.orc_tc_merge_rows.argtypes = [_p, _p, _i64, _i64, C.c_int]
    L.orc_pairs_hash.restype = C.c_int
    L.orc_pairs_hash.argtypes = [_p, _p, _p, _p, _i64, _f32, _i32, C.c_int, _p]
    L.orc_pairs_merge.restype = None
    L.orc_pairs_merge.argtypes = [_p, _p, _p, _p, _i64, C.c_int, _p]
    L.orc_sss_verify_hash.restype = C.c_int
    L.orc_sss_verify_hash.argtypes = [_p, _p, _i64, _p, _p, _i32, _i32, _i32, _p,
                                      _p, _p, _i64, _i32, _i32, C.c_int, _p, _p]
    L.orc_sss_float_rule.restype = C.c_int
    L.orc_sss_float_rule.argtypes = [_i32, _i32, _i32, _f32]
    L.orc_sss_lookup_bruteforce.restype = _i64
    L.orc_sss_lookup_bruteforce.argtypes = [_p, _p, _i64, _p, _p, _i64, _i32, _i32, _p, _i64]
    L.orc_sss_bruteforce.restype = _i64
    L.orc_sss_bruteforce.argtypes = [_p, _p, _i64, _i32, _i32, _p, _i64]
    L.orc_sss_prefix_join.restype = _i64
    L.orc_sss_prefix_join.argtypes = [_p, _p, _i64, _i32, _i32, C.c_int, _p, _i64, _p]
    L.orc_ir_hash.restype = C.c_int
    L.orc_ir_hash.argtypes = [_p, _p, _i64, _p, _i32, _i32, _i32, _p, _p, _p, _i64,
                              _p, C.c_int, _p, _p]
    L.orc_ir_merge.restype = None
    L.orc_ir_merge.argtypes = [_p, _p, _p, _p, _i64, _p, C.c_int, _p, _p]
    L.orc_max_threads.restype = C.c_int
    L.orc_max_threads.argtypes = []


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class OracleError(RuntimeError):
    pass


def max_threads() -> int:
    return int(lib().orc_max_threads())


def calculate_length(degree: int, load_factor: float, bucket_size: int) -> int:
    """triangle-count/hashtable.cu:20-36"""
    return int(lib().orc_calculate_length(int(degree), float(load_factor), int(bucket_size)))


def hash_normal(x: int, L: int) -> int:
    return int(lib().orc_hash_normal(x & 0xFFFFFFFF, L))


def hash_hierarchical(x: int, L: int, M: int) -> int:
    return int(lib().orc_hash_hierarchical(x & 0xFFFFFFFF, L, M))


class Tables:
    """A family of per-set tables in the reference's slot-plane layout."""

    def __init__(self, L, toff, max_len, bucket_size, layout, T, conflicts):
        self.L, self.toff, self.max_len = L, toff, max_len
        self.bucket_size, self.layout, self.T = bucket_size, layout, T
        self.conflicts = conflicts

    @property
    def bucket_num(self) -> int:
        return int(self.toff[-1])


def geometry(off, load_factor=0.2, bucket_size=5):
    """hashtable.cu:103-130 -> (L, toff, max_len)"""
    off = _c(off, np.int32)
    n = len(off) - 1
    L = np.empty(n, np.int32)
    toff = np.empty(n + 1, np.int64)
    m = C.c_int32(0)
    rc = lib().orc_table_geometry(_ptr(off), n, load_factor, bucket_size, _ptr(L), _ptr(toff), C.byref(m))
    if rc:
        raise OracleError(f"geometry rc={rc}")
    return L, toff, int(m.value)


def build_tables(off, elems, load_factor=0.2, bucket_size=5, layout=NORMAL) -> Tables:
    """hashtable.cu:53-101 (sequential schedule)"""
    off = _c(off, np.int32)
    elems = _c(elems, np.int32)
    n = len(off) - 1
    L, toff, M = geometry(off, load_factor, bucket_size)
    T = np.empty(int(toff[-1]) * bucket_size, np.int32)
    conf = C.c_int64(0)
    rc = lib().orc_table_build(_ptr(off), _ptr(elems), n, _ptr(L), _ptr(toff), bucket_size, M, layout,
                               _ptr(T), C.byref(conf))
    if rc:
        raise OracleError(f"build rc={rc}")
    return Tables(L, toff, M, bucket_size, layout, T, int(conf.value))


def table_check(off, elems, tables: Tables, T=None) -> int:
    """Invariant violations of a table T (default: tables.T) for this CSR."""
    off = _c(off, np.int32)
    elems = _c(elems, np.int32)
    T = tables.T if T is None else _c(T, np.int32)
    return int(lib().orc_table_check(_ptr(off), _ptr(elems), len(off) - 1, _ptr(tables.L), _ptr(tables.toff),
                                      tables.bucket_size, tables.max_len, tables.layout, _ptr(T)))


def extract_probe_order(off, tables: Tables, T=None):
    """baseline/common/utils.cu:76-146"""
    off = _c(off, np.int32)
    T = tables.T if T is None else _c(T, np.int32)
    out = np.empty(int(off[-1]), np.int32)
    rc = lib().orc_extract_probe_order(_ptr(off), len(off) - 1, _ptr(tables.toff), tables.bucket_size, _ptr(T),
                                       _ptr(out))
    if rc:
        raise OracleError(f"extract rc={rc}")
    return out


def table_search(tables: Tables, t: int, key: int) -> int:
    L = int(tables.L[t])
    bucket = hash_hierarchical(key, L, tables.max_len) if tables.layout else hash_normal(key, L)
    base = tables.T[int(tables.toff[t]):]
    return int(lib().orc_table_search(_ptr(base), tables.bucket_num, bucket, L, tables.bucket_size, key))


def tc_hash(off, cols, tables: Tables, probe_cols=None, nthreads=0) -> int:
    """TC.cu:128-190: probe keys come from probe_cols (default cols)."""
    off = _c(off, np.int32)
    keys = _c(cols if probe_cols is None else probe_cols, np.int32)
    out = C.c_uint64(0)
    rc = lib().orc_tc_hash(_ptr(off), _ptr(keys), len(off) - 1, _ptr(tables.L), _ptr(tables.toff),
                           tables.bucket_size, tables.max_len, tables.layout, _ptr(tables.T), nthreads,
                           C.byref(out))
    if rc:
        raise OracleError(f"tc rc={rc}")
    return int(out.value)


def tc_hash_rows(off, cols, tables: Tables, row_begin, row_end, nthreads=0) -> int:
    """tc_hash over the edges whose source row lies in [row_begin, row_end)."""
    off, cols = _c(off, np.int32), _c(cols, np.int32)
    out = C.c_uint64(0)
    rc = lib().orc_tc_hash_rows(_ptr(off), _ptr(cols), len(off) - 1, int(row_begin), int(row_end), _ptr(tables.L),
                                _ptr(tables.toff), tables.bucket_size, tables.max_len, tables.layout, _ptr(tables.T),
                                nthreads, C.byref(out))
    if rc:
        raise OracleError(f"tc rows rc={rc}")
    return int(out.value)


def tc_merge(off, cols, nthreads=0) -> int:
    off = _c(off, np.int32)
    cols = _c(cols, np.int32)
    return int(lib().orc_tc_merge(_ptr(off), _ptr(cols), len(off) - 1, nthreads))


def tc_merge_rows(off, cols, row_begin, row_end, nthreads=0) -> int:
    """Triangles whose source vertex lies in [row_begin, row_end)."""
    off, cols = _c(off, np.int32), _c(cols, np.int32)
    return int(lib().orc_tc_merge_rows(_ptr(off), _ptr(cols), int(row_begin), int(row_end), nthreads))


def pairs_hash(a_off, a, b_off, b, load_factor=0.2, bucket_size=5, nthreads=0):
    """set_size/intersection.cu:128-181 applied to every pair"""
    a_off, b_off = _c(a_off, np.int64), _c(b_off, np.int64)
    a, b = _c(a, np.int32), _c(b, np.int32)
    P = len(a_off) - 1
    out = np.empty(P, np.int32)
    rc = lib().orc_pairs_hash(_ptr(a_off), _ptr(a), _ptr(b_off), _ptr(b), P, load_factor, bucket_size,
                              nthreads, _ptr(out))
    if rc:
        raise OracleError(f"pairs rc={rc}")
    return out


def pairs_merge(a_off, a, b_off, b, nthreads=0):
    a_off, b_off = _c(a_off, np.int64), _c(b_off, np.int64)
    a, b = _c(a, np.int32), _c(b, np.int32)
    P = len(a_off) - 1
    out = np.empty(P, np.int32)
    lib().orc_pairs_merge(_ptr(a_off), _ptr(a), _ptr(b_off), _ptr(b), P, nthreads, _ptr(out))
    return out


def sss_verify_hash(off, cols, tables: Tables, cand_u, cand_v, p, q, probe_cols=None, nthreads=0):
    """sss.cu:93-154 with the exact integer threshold -> (flags, sizes)"""
    off = _c(off, np.int32)
    keys = _c(cols if probe_cols is None else probe_cols, np.int32)
    cu, cv = _c(cand_u, np.int32), _c(cand_v, np.int32)
    P = len(cu)
    flags = np.empty(P, np.int32)
    sizes = np.empty(P, np.int32)
    rc = lib().orc_sss_verify_hash(_ptr(off), _ptr(keys), len(off) - 1, _ptr(tables.L), _ptr(tables.toff),
                                   tables.bucket_size, tables.max_len, tables.layout, _ptr(tables.T),
                                   _ptr(cu), _ptr(cv), P, p, q, nthreads, _ptr(flags), _ptr(sizes))
    if rc:
        raise OracleError(f"sss rc={rc}")
    return flags, sizes


def sss_float_rule(r, su, sv, threshold) -> int:
    return int(lib().orc_sss_float_rule(r, su, sv, threshold))


def sss_bruteforce(off, cols, p, q, cap=1 << 22):
    off, cols = _c(off, np.int32), _c(cols, np.int32)
    out = np.empty(cap, np.int64)
    n = lib().orc_sss_bruteforce(_ptr(off), _ptr(cols), len(off) - 1, p, q, _ptr(out), cap)
    if n < 0:
        raise OracleError("bruteforce cap too small")
    return out[:n].copy()


def sss_lookup_bruteforce(off, cols, qoff, qtok, p, q, cap=1 << 22):
    """Every (query, row) with J >= p/q -> keys (a << 32 | j), ascending."""
    off, cols = _c(off, np.int32), _c(cols, np.int32)
    qoff, qtok = _c(qoff, np.int32), _c(qtok, np.int32)
    out = np.empty(cap, np.int64)
    n = lib().orc_sss_lookup_bruteforce(_ptr(off), _ptr(cols), len(off) - 1, _ptr(qoff), _ptr(qtok),
                                        len(qoff) - 1, p, q, _ptr(out), cap)
    if n < 0:
        raise OracleError("lookup cap too small")
    return out[:n].copy()


def sss_prefix_join(off, cols, p, q, cap=1 << 24, nthreads=0):
    """-> (sorted pair keys i<<32|j, number of deduplicated candidates)"""
    off, cols = _c(off, np.int32), _c(cols, np.int32)
    out = np.empty(cap, np.int64)
    nc = C.c_int64(0)
    n = lib().orc_sss_prefix_join(_ptr(off), _ptr(cols), len(off) - 1, p, q, nthreads, _ptr(out), cap,
                                  C.byref(nc))
    if n < 0:
        raise OracleError("prefix join cap too small")
    return out[:n].copy(), int(nc.value)


def ir_result_offsets(idx_off, query, qoff):
    """ir.cu:324-330: result region of a query = |term0's list|"""
    idx_off = np.asarray(idx_off, np.int64)
    query = np.asarray(query, np.int64)
    qoff = np.asarray(qoff, np.int64)
    Q = len(qoff) - 1
    n0 = np.zeros(Q, np.int64)
    nz = qoff[1:] > qoff[:-1]
    t0 = query[qoff[:-1][nz]]
    n0[nz] = idx_off[t0 + 1] - idx_off[t0]
    roff = np.zeros(Q + 1, np.int64)
    np.cumsum(n0, out=roff[1:])
    return roff


def ir_hash(idx_off, docs, tables: Tables, query, qoff, probe_docs=None, nthreads=0):
    """ir.cu:94-176 -> (result, count, roff)"""
    idx_off = _c(idx_off, np.int32)
    d = _c(docs if probe_docs is None else probe_docs, np.int32)
    query, qoff = _c(query, np.int32), _c(qoff, np.int32)
    roff = ir_result_offsets(idx_off, query, qoff)
    Q = len(qoff) - 1
    res = np.full(max(int(roff[-1]), 1), -1, np.int32)
    cnt = np.empty(Q, np.int32)
    rc = lib().orc_ir_hash(_ptr(idx_off), _ptr(d), len(idx_off) - 1, _ptr(tables.toff), tables.bucket_size,
                           tables.max_len, tables.layout, _ptr(tables.T), _ptr(query), _ptr(qoff), Q,
                           _ptr(roff), nthreads, _ptr(res), _ptr(cnt))
    if rc:
        raise OracleError(f"ir rc={rc}")
    return res, cnt, roff


def ir_merge(idx_off, docs, query, qoff, nthreads=0):
    idx_off, docs = _c(idx_off, np.int32), _c(docs, np.int32)
    query, qoff = _c(query, np.int32), _c(qoff, np.int32)
    roff = ir_result_offsets(idx_off, query, qoff)
    Q = len(qoff) - 1
    res = np.full(max(int(roff[-1]), 1), -1, np.int32)
    cnt = np.empty(Q, np.int32)
    lib().orc_ir_merge(_ptr(idx_off), _ptr(docs), _ptr(query), _ptr(qoff), Q, _ptr(roff), nthreads, _ptr(res),
                       _ptr(cnt))
    return res, cnt, roff


def ir_result_lists(res, cnt, roff):
    """Per-query result lists as Python lists of ints."""
    return [res[roff[i]:roff[i] + cnt[i]].tolist() for i in range(len(cnt))]


# ------------------------------------------------------- competitor tables
# baseline/common/{cuco,cuckoo_hash,hopscotch_hash,roaring_bitmap}.cu(h),
# restated in competitors.c.
COMP_KINDS = {"linear": 0, "cuckoo": 1, "hopscotch": 2, "roaring": 3}
COMP_EMPTY = 0x7FFFFFFF
CUCKOO_STASH = 16
HOP_H = 32


class _OrcComp(C.Structure):
    _fields_ = [("kind", C.c_int32), ("toff", C.c_void_p), ("slots", C.c_void_p),
                ("stash_count", C.c_void_p), ("stash", C.c_void_p), ("bitmap", C.c_void_p),
                ("remap", C.c_void_p), ("remap_len", C.c_int64), ("node_start", C.c_void_p),
                ("word_index", C.c_void_p), ("word_data", C.c_void_p)]


def _declare_comp(L):
    if getattr(L, "_comp_declared", False):
        return
    L.orc_fmix32.restype = C.c_uint32
    L.orc_fmix32.argtypes = [C.c_uint32]
    for g in ("orc_comp_linear_geometry", "orc_comp_cuckoo_geometry", "orc_comp_hopscotch_geometry"):
        getattr(L, g).restype = _i64
        getattr(L, g).argtypes = [_p, _i64, _f32, _p]
    L.orc_comp_linear_build.restype = _i64
    L.orc_comp_linear_build.argtypes = [_p, _p, _i64, _p, _p]
    L.orc_comp_linear_check.restype = _i64
    L.orc_comp_linear_check.argtypes = [_p, _p, _i64, _p, _p]
    L.orc_comp_cuckoo_build.restype = _i64
    L.orc_comp_cuckoo_build.argtypes = [_p, _p, _i64, _p, _p, _p, _p]
    L.orc_comp_hopscotch_build.restype = _i64
    L.orc_comp_hopscotch_build.argtypes = [_p, _p, _i64, _p, _p, _p]
    L.orc_comp_roaring_remap.restype = _i64
    L.orc_comp_roaring_remap.argtypes = [_p, _p, _i64, _p, _i64]
    L.orc_comp_roaring_words.restype = _i64
    L.orc_comp_roaring_words.argtypes = [_p, _p, _i64, _p]
    L.orc_comp_roaring_build.restype = _i64
    L.orc_comp_roaring_build.argtypes = [_p, _p, _i64, _p, _p, _p, _p]
    L.orc_comp_contains.restype = C.c_int
    L.orc_comp_contains.argtypes = [C.POINTER(_OrcComp), _i32, _i32]
    L.orc_comp_tc.restype = C.c_uint64
    L.orc_comp_tc.argtypes = [C.POINTER(_OrcComp), _p, _p, _i64]
    L.orc_comp_sss.restype = None
    L.orc_comp_sss.argtypes = [C.POINTER(_OrcComp), _p, _p, _p, _p, _i64, _i32, _i32, _p, _p]
    L._comp_declared = True


class CompTable:
    """One competitor table family built on the CPU: ``arrays`` holds the
    kind's arrays by name (toff/slots, stash_count/stash, bitmap, remap/
    node_start/word_index/word_data) and ``failed`` the lost inserts."""

    def __init__(self, kind: str, arrays: dict, failed: int):
        self.kind, self.arrays, self.failed = kind, arrays, failed

    def struct(self) -> _OrcComp:
        a = self.arrays
        s = _OrcComp()
        s.kind = COMP_KINDS[self.kind]
        for name in ("toff", "slots", "stash_count", "stash", "bitmap", "remap", "node_start",
                     "word_index", "word_data"):
            if name in a:
                setattr(s, name, a[name].ctypes.data)
        s.remap_len = len(a["remap"]) if "remap" in a else 0
        return s


def comp_build(kind: str, off, el, load_factor: float = 0.2) -> CompTable:
    L = lib()
    _declare_comp(L)
    off, el = _c(off, np.int32), _c(el, np.int32)
    n = len(off) - 1
    if kind == "roaring":
        remap_len = max(n, int(el.max()) + 1 if len(el) else 0)
        remap = np.empty(max(remap_len, 1), np.int32)
        if L.orc_comp_roaring_remap(_ptr(off), _ptr(el), n, _ptr(remap), remap_len) < 0:
            raise OracleError("roaring: element out of range")
        remap = remap[:remap_len]
        W = int(L.orc_comp_roaring_words(_ptr(off), _ptr(el), n, _ptr(remap)))
        node_start = np.empty(n + 1, np.int32)
        word_index = np.empty(max(W, 1), np.int32)
        word_data = np.empty(max(W, 1), np.uint64)
        L.orc_comp_roaring_build(_ptr(off), _ptr(el), n, _ptr(remap), _ptr(node_start),
                                 _ptr(word_index), _ptr(word_data))
        return CompTable(kind, dict(remap=remap, node_start=node_start, word_index=word_index[:W],
                                    word_data=word_data[:W]), 0)
    toff = np.empty(n + 1, np.int64)
    geo = {"linear": L.orc_comp_linear_geometry, "cuckoo": L.orc_comp_cuckoo_geometry,
           "hopscotch": L.orc_comp_hopscotch_geometry}[kind]
    total = int(geo(_ptr(off), n, float(load_factor), _ptr(toff)))
    slots = np.empty(max(total, 1), np.int32)
    if kind == "linear":
        failed = L.orc_comp_linear_build(_ptr(off), _ptr(el), n, _ptr(toff), _ptr(slots))
        return CompTable(kind, dict(toff=toff, slots=slots[:total]), int(failed))
    if kind == "cuckoo":
        stash_count = np.empty(max(n, 1), np.int32)
        stash = np.empty(max(n * CUCKOO_STASH, 1), np.int32)
        failed = L.orc_comp_cuckoo_build(_ptr(off), _ptr(el), n, _ptr(toff), _ptr(slots),
                                         _ptr(stash_count), _ptr(stash))
        return CompTable(kind, dict(toff=toff, slots=slots[:total], stash_count=stash_count[:n],
                                    stash=stash[:n * CUCKOO_STASH]), int(failed))
    bitmap = np.empty(max(total, 1), np.uint32)
    failed = L.orc_comp_hopscotch_build(_ptr(off), _ptr(el), n, _ptr(toff), _ptr(slots), _ptr(bitmap))
    return CompTable(kind, dict(toff=toff, slots=slots[:total], bitmap=bitmap[:total]), int(failed))


def fmix32(x: int) -> int:
    L = lib()
    _declare_comp(L)
    return int(L.orc_fmix32(x & 0xFFFFFFFF))


def comp_linear_check(off, el, toff, slots) -> int:
    """Violations of the linear-probing invariants (0 = a valid table)."""
    L = lib()
    _declare_comp(L)
    off, el = _c(off, np.int32), _c(el, np.int32)
    toff, slots = _c(toff, np.int64), _c(slots, np.int32)
    return int(L.orc_comp_linear_check(_ptr(off), _ptr(el), len(off) - 1, _ptr(toff), _ptr(slots)))


def comp_contains(t: CompTable, s: int, key: int) -> bool:
    L = lib()
    _declare_comp(L)
    st = t.struct()
    return bool(L.orc_comp_contains(C.byref(st), int(s), int(key)))


def comp_tc(t: CompTable, off, cols) -> int:
    """Sum over CSR entries (u, v) of |N(s) n N(g)| probed through g's table."""
    L = lib()
    _declare_comp(L)
    off, cols = _c(off, np.int32), _c(cols, np.int32)
    st = t.struct()
    return int(L.orc_comp_tc(C.byref(st), _ptr(off), _ptr(cols), len(off) - 1))


def comp_sss(t: CompTable, off, el, cand_u, cand_v, p, q):
    L = lib()
    _declare_comp(L)
    off, el = _c(off, np.int32), _c(el, np.int32)
    cu, cv = _c(cand_u, np.int32), _c(cand_v, np.int32)
    P = len(cu)
    flags, sizes = np.empty(P, np.int32), np.empty(P, np.int32)
    st = t.struct()
    L.orc_comp_sss(C.byref(st), _ptr(off), _ptr(el), _ptr(cu), _ptr(cv), P, int(p), int(q),
                   _ptr(flags), _ptr(sizes))
    return flags, sizes
