"""Python handle of the native serving loop (rs_server_* in the C ABI).

The loop itself -- admission by memory footprint, batching, overlapping
launches on their own streams, completion -- runs in the library's own host
threads; this class only submits requests and reads completions/results.
Two request kinds share one loop: IR conjunctive queries against the
resident index, and (with an attached SSSIndex) similarity lookups against a
resident set collection.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as _L
from .lib import NORMAL, REQ_IR, REQ_SSS, check, load  # noqa: F401  (kinds re-exported)


class Server:
    def __init__(self, tables, layout=NORMAL, budget_bytes=1 << 30, max_batch=256, max_inflight=4,
                 sss_index=None):
        self.tables = tables  # keeps the index and its tables alive
        self._h = C.c_void_p()
        check(load().rs_server_create(tables.handle, int(layout), int(budget_bytes), int(max_batch),
                                      int(max_inflight), C.byref(self._h)))
        self._buf = (_L.rs_completion * 4096)()
        self.sss_index = None
        if sss_index is not None:
            self.attach_sss(sss_index)

    def attach_sss(self, index):
        self.sss_index = index  # keeps the collection and its index alive
        check(load().rs_server_attach_sss(self._h, index._h))

    def submit(self, query_id, terms):
        """An IR query: term ids, term0 (shortest list) first."""
        arr = np.ascontiguousarray(terms, dtype=np.int32)
        check(load().rs_server_submit_ir(self._h, int(query_id), arr.ctypes.data_as(C.c_void_p), len(arr)))

    submit_ir = submit

    def submit_sss(self, query_id, tokens):
        """A similarity lookup: the query set's tokens, strictly ascending."""
        arr = np.ascontiguousarray(tokens, dtype=np.int32)
        check(load().rs_server_submit_sss(self._h, int(query_id), arr.ctypes.data_as(C.c_void_p), len(arr)))

    def poll(self):
        n = C.c_int32(0)
        check(load().rs_server_poll(self._h, self._buf, len(self._buf), C.byref(n)))
        return [self._buf[i] for i in range(n.value)]

    def poll_all(self):
        out = []
        while True:
            got = self.poll()
            if not got:
                return out
            out.extend(_copy(c) for c in got)

    def result(self, query_id):
        """The request's result ids (releases the stored result)."""
        cap = 1 << 14
        while True:
            buf = np.empty(cap, np.int32)
            cnt = C.c_int64(0)
            check(load().rs_server_result(self._h, int(query_id), buf.ctypes.data_as(C.c_void_p), cap,
                                          C.byref(cnt)))
            if cnt.value <= cap:
                return buf[:cnt.value].tolist()
            cap = int(cnt.value)

    def drain(self):
        check(load().rs_server_drain(self._h))

    def cancel(self, query_id):
        """Cancel a still-queued request (it completes with status
        RS_ECANCELED); raises ResetError if it is not queued."""
        check(load().rs_server_cancel(self._h, int(query_id)))

    def set_queue_timeout(self, ms):
        """Requests that wait longer than `ms` before admission complete with
        RS_ETIMEDOUT, unlaunched (0 disables)."""
        check(load().rs_server_set_queue_timeout(self._h, float(ms)))

    def pause(self, paused=True):
        """Hold (or resume) admission; launched batches still complete."""
        check(load().rs_server_pause(self._h, 1 if paused else 0))

    def stats(self):
        a, b, c = C.c_int64(0), C.c_int64(0), C.c_int64(0)
        check(load().rs_server_stats(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return {"launches": a.value, "queries": b.value, "peak_bytes": c.value}

    def close(self):
        if self._h:
            load().rs_server_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


IRServer = Server


def _copy(c):
    out = _L.rs_completion()
    C.pointer(out)[0] = c
    return out
