"""GPU: rs_ir_query_noalloc (scratch in a fixed device array, no allocation)
against rs_ir_query_async (per-call scratch), through the C ABI: two calls in
a row, the second with a smaller batch, bit-exact results and counts, for
both layouts; the per-call limit is enforced."""
import sys

import numpy as np
import pytest
import torch

import graphs as G

pytestmark = pytest.mark.gpu


def _run(lib, fn, t, layout, q, qoff, roff, total):
    from reset_amd.api import ptr, stream_handle

    Q = qoff.numel() - 1
    res = torch.full((max(total, 1),), -1, dtype=torch.int32, device="cuda")
    ws = torch.empty(max(total, 1), dtype=torch.int32, device="cuda")
    cnt = torch.empty(max(Q, 1), dtype=torch.int32, device="cuda")
    st = getattr(lib, fn)(t._h, int(layout), None, ptr(q), ptr(qoff), Q, ptr(roff), ptr(res), ptr(cnt),
                          ptr(ws), stream_handle())
    torch.cuda.synchronize()
    return st, res, cnt


def test_noalloc_matches_allocating_entry_point(oracle_lib):
    assert torch.cuda.is_available(), "GPU tests need the MI355X"
    from reset_amd import api
    from reset_amd.lib import RS_EINVAL, RS_OK, load

    lib = load()
    rng = np.random.default_rng(77)
    idx_off, docs = G.inverted_index(rng, 400, 600000, 60000)
    t = api.HashTables(torch.from_numpy(idx_off).cuda(), torch.from_numpy(docs).cuda(), 0.2, 5, 3)
    for call, nq in enumerate((1500, 240)):  # the second call is smaller
        q, qoff = G.queries(rng, idx_off, nq, 1, 6)
        qd, qod = torch.from_numpy(q).cuda(), torch.from_numpy(qoff).cuda()
        roff = api.ir_result_offsets(t.offsets, qd, qod)
        total = int(roff[-1].item())
        for layout in (0, 1):
            print(f"noalloc: call {call}, {nq} queries, layout {layout}", file=sys.stderr, flush=True)
            s0, r0, c0 = _run(lib, "rs_ir_query_async", t, layout, qd, qod, roff, total)
            s1, r1, c1 = _run(lib, "rs_ir_query_noalloc", t, layout, qd, qod, roff, total)
            assert s0 == RS_OK and s1 == RS_OK
            t.status()
            c0h, c1h = c0[:nq].cpu().numpy(), c1[:nq].cpu().numpy()
            assert np.array_equal(c0h, c1h)
            lists0 = oracle_lib.ir_result_lists(r0.cpu().numpy(), c0h, roff.cpu().numpy())
            lists1 = oracle_lib.ir_result_lists(r1.cpu().numpy(), c1h, roff.cpu().numpy())
            assert lists0 == lists1
            m_res, m_cnt, m_roff = oracle_lib.ir_merge(idx_off, docs, q, qoff)
            assert lists1 == oracle_lib.ir_result_lists(m_res, m_cnt, m_roff)
    # above the per-call query limit: refused before any work
    over = 65536 + 1
    qd = torch.zeros(over, dtype=torch.int32, device="cuda")
    qod = torch.arange(over + 1, dtype=torch.int32, device="cuda")
    roff = api.ir_result_offsets(t.offsets, qd, qod)
    st, _, _ = _run(lib, "rs_ir_query_noalloc", t, 0, qd, qod, roff, int(roff[-1].item()))
    assert st == RS_EINVAL
    t.close()
