"""The analytic (load factor, bucket size) model against the reference's own
cached solver results (hyperparameter_solver/cache/recommended_values.json,
written by hyperparameter_solver/src/solver.py + c_api.py:288-338), copied to
tests/golden/solver_recommended.json by tests/golden/make_solver_golden.py.

The cache holds 216 entries over 72 distinct (size, W, S_trans, S_slot)
configurations (the tc/sss/ir modes share each configuration's answer).  A
solve evaluates 171 grid points; this suite checks one configuration per
(S_trans, S_slot) pair, spread over the table sizes, to stay within seconds.
"""
import json
import os
import re

import pytest

from reset_amd import api

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "solver_recommended.json")


def _entries():
    with open(GOLDEN) as f:
        return json.load(f)["entries"]


def _configs():
    out = {}
    for k, v in _entries().items():
        m = re.match(r"(\w+)_size(\d+)_W(\d+)_T(\d+)_S(\d+)$", k)
        assert m, k
        cfg = tuple(int(g) for g in m.groups()[1:])
        out.setdefault(cfg, []).append((m.group(1), v))
    return out


def test_fixture_shape():
    cfgs = _configs()
    assert len(_entries()) == 216 and len(cfgs) == 72
    for cfg, rows in cfgs.items():
        assert sorted(r[0] for r in rows) == ["ir", "sss", "tc"]
        assert len({tuple(r[1]) for r in rows}) == 1, cfg


def _sample():
    cfgs = sorted(_configs())
    sizes = sorted({c[0] for c in cfgs})
    picked = []
    for i, (t, s) in enumerate(sorted({(c[2], c[3]) for c in cfgs})):
        n = sizes[i % len(sizes)]
        picked.append(next(c for c in cfgs if c[0] == n and c[2] == t and c[3] == s))
    return picked


@pytest.mark.parametrize("cfg", _sample(), ids=lambda c: "n%d_W%d_T%d_S%d" % c)
def test_solve_matches_reference_cache(cfg):
    alpha, b, cost = _configs()[cfg][0][1]
    a2, b2, c2 = api.hyper_solve(*cfg)
    assert (a2, b2) == (pytest.approx(alpha, abs=1e-12), b)
    assert c2 == pytest.approx(cost, rel=1e-9)
    # the recommended point is the model's cost there
    assert api.hyper_cost(cfg[0], alpha, b, cfg[1], cfg[2], cfg[3]) == pytest.approx(cost, rel=1e-9)


def test_cost_special_cases():
    n, W, T, S = 100000, 32, 64, 8
    # no keys (alpha 0): the expected occupancy is 0
    assert api.hyper_cost(n, 0.0, 4, W, T, S) == 1e10
    # single-slot buckets at alpha 1 saturate: almost every bucket overflows
    assert api.hyper_cost(n, 1.0, 1, W, T, S) >= 1e9
    assert 0 < api.hyper_cost(n, 0.2, 8, W, T, S) < 1e9
    with pytest.raises(Exception):
        api.hyper_solve(0, W, T, S)


def test_cost_sparse_closed_form():
    """At alpha b = 0.09 with b = 9 a bucket practically never overflows
    (P(m >= 9) ~ 1e-15), so pi = delta_0, psi = u (Binomial, mean n p =
    alpha b): E_S = 0.09, E_L = E[m] + 1 = 1.09 and
    T_step = ceil(W S_slot / (E_S S_trans)) + 1 = ceil(128 / 2.88) + 1 = 46."""
    c = api.hyper_cost(1_000_000, 0.01, 9, 32, 32, 4)
    assert c == pytest.approx(1.09 * 46, rel=1e-9)
    # swapping S_trans and S_slot changes T_step: ceil(1024 / 0.36) + 1 = 2846
    c = api.hyper_cost(1_000_000, 0.01, 9, 32, 4, 32)
    assert c == pytest.approx(1.09 * 2846, rel=1e-9)


# res/sc18.csv, sc19.csv and sc20.csv hold the full grid of costs of the
# reference's solver for the R-MAT scale-18/19/20 graphs at W = 32,
# S_trans = 128, S_slot = 4, with the table size taken as the graph's
# degree-weighted mean degree (hyperparameter_solver/src/util.py:131-146),
# which the files do not record.  One row, (0.9, 4), fixes the size (its cost
# rises monotonically with the size, one step at a time); the other 170 rows
# of each grid must then match to the 6 printed decimals.
GRID_SIZES = {"sc18": 1254, "sc19": 1683, "sc20": 2123}


@pytest.mark.parametrize("sc", sorted(GRID_SIZES))
def test_full_grid_matches_reference_csv(sc):
    from concurrent.futures import ThreadPoolExecutor

    with open(os.path.join(os.path.dirname(__file__), "golden", "solver_grid.json")) as f:
        rows = json.load(f)["grids"][sc]
    assert len(rows) == 19 * 9
    assert sorted((a, b) for a, b, _ in rows) == sorted(
        (round(0.1 + 0.05 * i, 2), b) for i in range(19) for b in range(1, 10))
    n = GRID_SIZES[sc]
    with ThreadPoolExecutor(8) as ex:
        costs = list(ex.map(lambda r: api.hyper_cost(n, r[0], r[1], 32, 128, 4), rows))
    for (a, b, ref), c in zip(rows, costs):
        if ref >= 1e8:  # pruned points: the penalty constants themselves
            assert c == ref, (a, b)
        else:
            assert abs(c - ref) <= 5.01e-7, (a, b, c, ref)
    # the neighbouring sizes do not fit: the grid pins the size, and so the
    # binomial (not Poisson) occupancy model
    for m in (n - 1, n + 1):
        assert abs(api.hyper_cost(m, 0.9, 4, 32, 128, 4) - dict(((a, b), c) for a, b, c in rows)[(0.9, 4)]) > 2e-6
    # the recommendation of the reference's summary (res/*_summary.json)
    assert api.hyper_solve(n, 32, 128, 4)[:2] == (0.2, 5)


def test_one_lane_probes_pick_one_slot_buckets():
    """The IR query kernels probe with one lane per key (W = 1). At 128-byte
    transactions the model's least-cost pair is then load factor 0.1 with
    one-slot buckets at every posting length of the IR index: the pair the
    IR benchmark builds its tables with."""
    for n in (2, 68, 4096, 2_000_000):
        assert api.hyper_solve(n, 1, 128, 4)[:2] == (0.1, 1)
