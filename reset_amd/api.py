"""Python host API over the C ABI, using torch only for device memory/streams.

Mirrors the reference's object interface (baseline/common/hash_table.cuh,
baseline/{tc,sss,ir}/*.cuh): ``HashTables`` is HashTableBuilder, and the
``tc_*``, ``sss_*``, ``ir_*``, ``pairs_*`` functions are the run_* methods'
kernels.  All inputs are int32/int64 CUDA tensors; nothing falls back to CPU.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import lib as _L
from .lib import BUILD_BOTH, HIERARCHICAL, NORMAL, check, load


_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        _HIP = C.CDLL("libamdhip64.so")
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _HIP.hipMemcpy.restype = C.c_int
    return _HIP


def _need_cuda(*ts):
    for t in ts:
        if t is not None and (not t.is_cuda or not t.is_contiguous()):
            raise ValueError("expected contiguous CUDA tensors")


def ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def stream_handle(stream=None):
    s = torch.cuda.current_stream() if stream is None else stream
    return C.c_void_p(s.cuda_stream)


def calculate_length(degree, load_factor=0.2, bucket_size=5):
    """triangle-count/hashtable.cu:20-36"""
    return int(load().rs_calculate_length(int(degree), float(load_factor), int(bucket_size)))


def calculate_lengths(offsets, load_factor=0.2, bucket_size=5, stream=None):
    """hashtable.cu:39-51 on the device -> (lengths int32[n], max length)"""
    _need_cuda(offsets)
    n = offsets.numel() - 1
    out = torch.empty(n, dtype=torch.int32, device=offsets.device)
    m = C.c_int32(0)
    check(load().rs_calculate_lengths(ptr(offsets), n, float(load_factor), int(bucket_size), ptr(out),
                                      C.byref(m), stream_handle(stream)))
    return out, int(m.value)


def expand_rows(offsets, num_elements=None, stream=None):
    _need_cuda(offsets)
    n = offsets.numel() - 1
    E = int(offsets[-1].item()) if num_elements is None else int(num_elements)
    rows = torch.empty(max(E, 1), dtype=torch.int32, device=offsets.device)[:E]
    check(load().rs_expand_rows(ptr(offsets), n, E, ptr(rows) if E else None, stream_handle(stream)))
    return rows


class HashTables:
    """Per-set bucketed hash tables (HashTableBuilder, hash_table.cuh:12-95)."""

    def __init__(self, offsets, elements, load_factor=0.2, bucket_size=5, layouts=BUILD_BOTH,
                 stream=None):
        _need_cuda(offsets, elements)
        assert offsets.dtype == torch.int32 and elements.dtype == torch.int32
        self.offsets, self.elements = offsets, elements  # borrowed by the handle
        self._h = C.c_void_p()
        n = offsets.numel() - 1
        check(load().rs_tables_build(ptr(offsets), n, ptr(elements), float(load_factor), int(bucket_size),
                                     int(layouts), stream_handle(stream), C.byref(self._h)))
        self._plan = None

    @property
    def handle(self):
        return self._h

    def describe(self) -> _L.rs_tables_desc:
        d = _L.rs_tables_desc()
        check(load().rs_tables_describe(self._h, C.byref(d)))
        return d

    def _dev(self, addr, n, dtype):
        """Copy of a device array owned by the handle (hipMemcpy D2D)."""
        out = torch.empty(n, dtype=dtype, device=self.offsets.device)
        if not addr or n == 0:
            return out
        torch.cuda.synchronize()
        nbytes = n * out.element_size()
        rc = _hip().hipMemcpy(C.c_void_p(out.data_ptr()), C.c_void_p(addr), C.c_size_t(nbytes), 3)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed: {rc}")
        return out

    def lengths(self):
        d = self.describe()
        return self._dev(d.d_lengths, d.num_sets, torch.int32)

    def table_offsets(self):
        d = self.describe()
        return self._dev(d.d_table_offsets, d.num_sets + 1, torch.int64)

    def table(self, layout=NORMAL):
        d = self.describe()
        addr = d.d_table_hierarchical if layout == HIERARCHICAL else d.d_table_normal
        return self._dev(addr, d.bucket_num * d.bucket_size, torch.int32)

    def status(self, stream=None):
        check(load().rs_tables_status(self._h, stream_handle(stream)))

    def set_timing(self, enable=True):
        check(load().rs_tables_timing(self._h, int(bool(enable))))

    def kernel_ms(self):
        """Duration of the last TC count / IR query kernels (HIP events on their stream)."""
        ms = C.c_float(0)
        check(load().rs_tables_kernel_ms(self._h, C.byref(ms)))
        return float(ms.value)

    def rebuild_async(self, layouts=BUILD_BOTH, stream=None):
        """Fill + insert again into the existing tables (no allocation; the
        build phases the reference times, hashtable.cu:164-177)."""
        check(load().rs_tables_rebuild_async(self._h, int(layouts), stream_handle(stream)))

    def probe_order(self, layout=HIERARCHICAL, stream=None):
        """baseline/common/utils.cu:148-175 (extract_hashtable_to_csr)"""
        out = torch.empty_like(self.elements)
        check(load().rs_tables_probe_order(self._h, int(layout), ptr(out), stream_handle(stream)))
        return out

    # ------------------------------------------------------------ TC
    def tc_plan(self, stream=None):
        if self._plan is None:
            h = C.c_void_p()
            check(load().rs_tc_plan_create(self._h, stream_handle(stream), C.byref(h)))
            self._plan = h
        return self._plan

    def tc_work(self):
        probes, edges = C.c_int64(0), C.c_int64(0)
        check(load().rs_tc_plan_work(self.tc_plan(), C.byref(probes), C.byref(edges)))
        return int(probes.value), int(edges.value)

    def tc_staged(self):
        """(probes served from LDS-staged tables, units, largest staged length)."""
        probes, units, lmax = C.c_int64(0), C.c_int64(0), C.c_int64(0)
        check(load().rs_tc_plan_staged(self.tc_plan(), C.byref(probes), C.byref(units), C.byref(lmax)))
        return int(probes.value), int(units.value), int(lmax.value)

    def tc_count_async(self, out, layout=NORMAL, probe_keys=None, plain=False, stream=None):
        fn = load().rs_tc_count_plain_async if plain else load().rs_tc_count_async
        check(fn(self._h, self.tc_plan(stream), int(layout), ptr(probe_keys), ptr(out), stream_handle(stream)))
        return out

    def tc_count(self, layout=NORMAL, probe_keys=None, plain=False, stream=None):
        out = torch.empty(1, dtype=torch.int64, device=self.offsets.device)
        self.tc_count_async(out, layout, probe_keys, plain, stream)
        self.status(stream)
        return int(out.item())

    # ------------------------------------------------------------ SSS
    def sss_verify(self, cand_u, cand_v, p, q, layout=NORMAL, probe_keys=None, stream=None):
        _need_cuda(cand_u, cand_v)
        P = cand_u.numel()
        flags = torch.empty(P, dtype=torch.int32, device=cand_u.device)
        sizes = torch.empty(P, dtype=torch.int32, device=cand_u.device)
        check(load().rs_sss_verify_async(self._h, int(layout), ptr(probe_keys), ptr(cand_u), ptr(cand_v), P,
                                         int(p), int(q), ptr(flags), ptr(sizes), stream_handle(stream)))
        self.status(stream)
        return flags, sizes

    # ------------------------------------------------------------ IR
    def ir_query(self, query, query_offsets, layout=NORMAL, term0_docs=None, stream=None):
        """-> (result int32[roff[Q]], count int32[Q], roff int64[Q+1])"""
        _need_cuda(query, query_offsets)
        Q = query_offsets.numel() - 1
        roff = ir_result_offsets(self.offsets, query, query_offsets)
        total = int(roff[-1].item()) if Q else 0
        res = torch.full((max(total, 1),), -1, dtype=torch.int32, device=query.device)
        ws = torch.empty(max(total, 1), dtype=torch.int32, device=query.device)
        cnt = torch.empty(max(Q, 1), dtype=torch.int32, device=query.device)
        check(load().rs_ir_query_async(self._h, int(layout), ptr(term0_docs), ptr(query), ptr(query_offsets), Q,
                                       ptr(roff), ptr(res), ptr(cnt), ptr(ws), stream_handle(stream)))
        self.status(stream)
        return res[:total], cnt[:Q], roff

    def close(self):
        if self._plan is not None:
            load().rs_tc_plan_free(self._plan)
            self._plan = None
        if self._h:
            load().rs_tables_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ir_result_offsets(index_offsets, query, query_offsets):
    """ir.cu:324-330: a query's result region holds |list(term0)| entries."""
    qo = query_offsets.long()
    Q = qo.numel() - 1
    n0 = torch.zeros(Q, dtype=torch.int64, device=query.device)
    nz = qo[1:] > qo[:-1]
    if bool(nz.any()):
        t0 = query.long()[qo[:-1][nz]]
        io = index_offsets.long()
        n0[nz] = io[t0 + 1] - io[t0]
    roff = torch.zeros(Q + 1, dtype=torch.int64, device=query.device)
    roff[1:] = torch.cumsum(n0, 0)
    return roff


# ---------------------------------------------------------------- set similarity
def _sss_result(h, device):
    try:
        npairs, ncand, dptr = C.c_int64(0), C.c_int64(0), C.c_void_p()
        check(load().rs_sss_result_info(h, C.byref(npairs), C.byref(ncand), C.byref(dptr)))
        probes, elems = C.c_int64(0), C.c_int64(0)
        check(load().rs_sss_result_work(h, C.byref(probes), C.byref(elems)))
        ms = (C.c_float * 3)()
        check(load().rs_sss_result_timings(h, ms))
        pairs = torch.empty(max(npairs.value, 1), dtype=torch.int64, device=device)
        if npairs.value:
            torch.cuda.synchronize()
            rc = _hip().hipMemcpy(C.c_void_p(pairs.data_ptr()), dptr, C.c_size_t(8 * npairs.value), 3)
            if rc != 0:
                raise RuntimeError(f"hipMemcpy failed: {rc}")
        return dict(pairs=pairs[:npairs.value], candidates=int(ncand.value), verify_probes=int(probes.value),
                    verify_elements=int(elems.value), ms=[float(x) for x in ms])
    finally:
        load().rs_sss_result_free(h)


class CompetitorTables:
    """The baseline directory's competitor tables over one CSR family
    (baseline/common/{cuco,cuckoo_hash,hopscotch_hash,roaring_bitmap}): kind
    is "linear" (cuCollections linear probing), "cuckoo", "hopscotch" or
    "roaring".  The CSR tensors must outlive the object (it borrows them)."""

    _ARRAYS = {"offsets": (0, torch.int64), "slots": (1, torch.int32), "stash_count": (2, torch.int32),
               "stash": (3, torch.int32), "bitmap": (4, torch.int32), "remap": (5, torch.int32),
               "node_start": (6, torch.int32), "word_index": (7, torch.int32),
               "word_data": (8, torch.int64)}

    def __init__(self, offsets, elements, kind="cuckoo", load_factor=0.2, stream=None):
        _need_cuda(offsets, elements)
        self.offsets, self.elements = offsets, elements
        self.kind = kind
        self.n = offsets.numel() - 1
        h = C.c_void_p()
        check(load().rs_comp_build(ptr(offsets), self.n, ptr(elements), _L.COMP_KINDS[kind],
                                   float(load_factor), stream_handle(stream), C.byref(h)))
        self._h = h
        slots, failed = C.c_int64(0), C.c_int64(0)
        check(load().rs_comp_info(self._h, C.byref(slots), C.byref(failed)))
        self.slots, self.failed = int(slots.value), int(failed.value)

    def arrays(self) -> dict:
        """The table's arrays as host tensors (names as in the C header)."""
        out = {}
        for name, (code, dt) in self._ARRAYS.items():
            nb = C.c_int64(0)
            check(load().rs_comp_array_bytes(self._h, code, C.byref(nb)))
            if nb.value == 0:
                continue
            t = torch.empty(nb.value // torch.empty(0, dtype=dt).element_size(), dtype=dt)
            check(load().rs_comp_copy_out(self._h, code, C.c_void_p(t.data_ptr()), nb.value))
            out[name] = t
        return out

    def tc_count(self, stream=None) -> int:
        c = C.c_uint64(0)
        check(load().rs_comp_tc_count(self._h, C.byref(c), stream_handle(stream)))
        return int(c.value)

    def sss_verify(self, cand_u, cand_v, p=4, q=5, stream=None):
        _need_cuda(cand_u, cand_v)
        P = cand_u.numel()
        flags = torch.empty(max(P, 1), dtype=torch.int32, device=cand_u.device)[:P]
        sizes = torch.empty(max(P, 1), dtype=torch.int32, device=cand_u.device)[:P]
        check(load().rs_comp_sss_verify_async(self._h, ptr(cand_u), ptr(cand_v), P, int(p), int(q),
                                              ptr(flags), ptr(sizes), stream_handle(stream)))
        check(load().rs_comp_status(self._h, stream_handle(stream)))
        return flags, sizes

    def ir_query(self, query, query_offsets, stream=None):
        """-> (result int32[roff[Q]], count int32[Q], roff int64[Q+1]); the
        tables must be built over an inverted index (sets = terms)."""
        _need_cuda(query, query_offsets)
        Q = query_offsets.numel() - 1
        roff = ir_result_offsets(self.offsets, query, query_offsets)
        total = int(roff[-1].item()) if Q else 0
        res = torch.full((max(total, 1),), -1, dtype=torch.int32, device=query.device)
        cnt = torch.empty(max(Q, 1), dtype=torch.int32, device=query.device)
        check(load().rs_comp_ir_query_async(self._h, ptr(query), ptr(query_offsets), Q, ptr(roff), ptr(res),
                                            ptr(cnt), stream_handle(stream)))
        check(load().rs_comp_status(self._h, stream_handle(stream)))
        return res[:total], cnt[:Q], roff

    def close(self):
        if getattr(self, "_h", None):
            load().rs_comp_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sss_selfjoin(offsets, elements, p=4, q=5, load_factor=0.25, bucket_size=4, stream=None):
    """Prefix-filter self-join at Jaccard >= p/q with hash verification;
    load_factor / bucket_size (1, 2 or 4) set the verification tables.
    -> dict(pairs: int64 tensor of i<<32|j ascending (device), candidates,
            verify_probes, ms: [order, candidates, verify])"""
    _need_cuda(offsets, elements)
    h = C.c_void_p()
    check(load().rs_sss_selfjoin(ptr(offsets), offsets.numel() - 1, ptr(elements), int(p), int(q),
                                 float(load_factor), int(bucket_size), stream_handle(stream), C.byref(h)))
    return _sss_result(h, offsets.device)


class SSSIndex:
    """Resident similarity-lookup index over a set collection (borrowed CSR)."""

    def __init__(self, offsets, elements, p=4, q=5, stream=None):
        _need_cuda(offsets, elements)
        assert offsets.dtype == torch.int32 and elements.dtype == torch.int32
        self.offsets, self.elements, self.p, self.q = offsets, elements, int(p), int(q)
        self._h = C.c_void_p()
        check(load().rs_sss_index_build(ptr(offsets), offsets.numel() - 1, ptr(elements), self.p, self.q,
                                        stream_handle(stream), C.byref(self._h)))

    def lookup(self, query_offsets, query_tokens, stream=None):
        """-> dict(pairs: int64 (query << 32 | set) ascending, candidates, ...)"""
        _need_cuda(query_offsets, query_tokens)
        assert query_offsets.dtype == torch.int32 and query_tokens.dtype == torch.int32
        h = C.c_void_p()
        check(load().rs_sss_lookup(self._h, ptr(query_offsets), query_offsets.numel() - 1, ptr(query_tokens),
                                   stream_handle(stream), C.byref(h)))
        return _sss_result(h, query_offsets.device)

    def close(self):
        if self._h:
            load().rs_sss_index_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- pairs
def default_pairs_opts():
    o = _L.rs_pairs_opts()
    load().rs_pairs_default_opts(C.byref(o))
    return o


def hyper_cost(arr_size, alpha, b, W=32, S_trans=32, S_slot=4):
    """Analytic cost of (alpha, b) for a table of arr_size entries."""
    return float(load().rs_hyper_cost(int(arr_size), float(alpha), int(b), int(W), int(S_trans), int(S_slot)))


def hyper_solve(arr_size, W=32, S_trans=128, S_slot=4):
    """Least-cost (alpha, b, cost) over the reference solver's grid."""
    al, bb, co = C.c_double(0), C.c_int32(0), C.c_double(0)
    check(load().rs_hyper_solve(int(arr_size), int(W), int(S_trans), int(S_slot), C.byref(al), C.byref(bb),
                                C.byref(co)))
    return float(al.value), int(bb.value), float(co.value)


def autotune_pairs(a_off, a, b_off, b, opts=None, reps=2, stream=None):
    """Per-size-bin (load factor, bucket size, chunk, lanes per pair) picked by
    timing candidates on these pairs (a sample of the workload); returns the
    tuned rs_pairs_opts."""
    _need_cuda(a_off, a, b_off, b)
    o = opts if opts is not None else default_pairs_opts()
    check(load().rs_pairs_autotune(ptr(a_off), ptr(a), ptr(b_off), ptr(b), a_off.numel() - 1, int(reps),
                                   C.byref(o), stream_handle(stream)))
    return o


def pairs_opts_table(o):
    """The bins of an rs_pairs_opts as plain tuples (alpha, b, chunk, unroll, lanes)."""
    return [(round(o.bins[i].load_factor, 4), o.bins[i].bucket_size, o.bins[i].chunk, o.bins[i].unroll,
             o.bins[i].lanes) for i in range(o.num_bins)]


class PairsPlan:
    """Batched persistent pair-intersection launch (size-binned plan)."""

    def __init__(self, max_pairs, opts=None):
        self._h = C.c_void_p()
        self.opts = opts if opts is not None else default_pairs_opts()
        check(load().rs_pairs_plan_create(int(max_pairs), C.byref(self.opts), C.byref(self._h)))

    def run_async(self, a_off, a, b_off, b, sizes, stream=None):
        P = a_off.numel() - 1
        check(load().rs_pairs_intersect_batched_async(self._h, ptr(a_off), ptr(a), ptr(b_off), ptr(b), P,
                                                      ptr(sizes), stream_handle(stream)))
        return sizes

    def run(self, a_off, a, b_off, b, stream=None):
        _need_cuda(a_off, a, b_off, b)
        sizes = torch.empty(max(a_off.numel() - 1, 1), dtype=torch.int32, device=a.device)
        self.run_async(a_off, a, b_off, b, sizes, stream)
        self.status(stream)
        return sizes[:a_off.numel() - 1]

    def capture(self, a_off, a, b_off, b, sizes, stream):
        check(load().rs_pairs_graph_capture(self._h, ptr(a_off), ptr(a), ptr(b_off), ptr(b), a_off.numel() - 1,
                                            ptr(sizes), stream_handle(stream)))

    def replay(self, stream=None):
        check(load().rs_pairs_graph_launch(self._h, stream_handle(stream)))

    def status(self, stream=None):
        check(load().rs_pairs_plan_status(self._h, stream_handle(stream)))

    def set_timing(self, enable=True):
        check(load().rs_pairs_plan_timing(self._h, int(bool(enable))))

    def kernel_ms(self):
        """Duration of the last batched kernel (HIP events on its stream)."""
        ms = C.c_float(0)
        check(load().rs_pairs_plan_kernel_ms(self._h, C.byref(ms)))
        return float(ms.value)

    def close(self):
        if self._h:
            load().rs_pairs_plan_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ir_set_steal_threshold(segments=-1):
    """IR batches of at least `segments` segments use work stealing (-1: default)."""
    check(load().rs_ir_set_steal_threshold(int(segments)))


def pairs_intersect_plain(a_off, a, b_off, b, load_factor=0.2, bucket_size=5, stream=None):
    """set_size/intersection.cu shape: per-pair table fill + build + probe launches."""
    _need_cuda(a_off, a, b_off, b)
    h_a, h_b = a_off.cpu().contiguous(), b_off.cpu().contiguous()
    P = a_off.numel() - 1
    maxb = int(torch.maximum(h_a[1:] - h_a[:-1], h_b[1:] - h_b[:-1]).max().item()) if P else 0
    L = calculate_length(maxb, load_factor, bucket_size)
    ws = torch.empty(max(L * bucket_size, 1), dtype=torch.int32, device=a.device)
    sizes = torch.empty(max(P, 1), dtype=torch.int32, device=a.device)
    check(load().rs_pairs_intersect_plain(ptr(a_off), ptr(a), ptr(b_off), ptr(b),
                                          C.c_void_p(h_a.data_ptr()), C.c_void_p(h_b.data_ptr()), P,
                                          float(load_factor), int(bucket_size), ptr(ws), ptr(sizes),
                                          stream_handle(stream)))
    return sizes[:P]


# ---------------------------------------------------------------- generators
def gen_pair_sets(a_off, b_off, k, seed, stream=None):
    """Sorted unique sets with |A_i cap B_i| = k_i exactly (device)."""
    _need_cuda(a_off, b_off, k)
    a = torch.empty(max(int(a_off[-1].item()), 1), dtype=torch.int32, device=a_off.device)
    b = torch.empty(max(int(b_off[-1].item()), 1), dtype=torch.int32, device=a_off.device)
    check(load().rs_gen_pair_sets_async(ptr(a_off), ptr(b_off), ptr(k), a_off.numel() - 1, int(seed), ptr(a),
                                        ptr(b), stream_handle(stream)))
    return a, b


def gen_rmat_edges(scale, num_edges, a=0.57, b=0.19, c=0.19, seed=1, device="cuda", stream=None):
    """R-MAT edge list int32[num_edges, 2] (labels scrambled)."""
    ta, tb, tc = (int(round(x * 2**32)) for x in (a, b, c))
    out = torch.empty((max(num_edges, 1), 2), dtype=torch.int32, device=device)
    check(load().rs_gen_rmat_async(int(scale), int(num_edges), ta, tb, tc, int(seed), ptr(out),
                                   stream_handle(stream)))
    return out[:num_edges]
