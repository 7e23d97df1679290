"""Autotuner host logic: buckets, configurations within each kernel's geometry,
table round trip, and a choice that depends on the LP's own shape only."""

import numpy as np
import pytest

from sovlp.lp.autotune import CANDIDATES, NNZ_EDGES, Autotuner, bucket_of, bulk_rate, config_of
from sovlp.lp.batch import batch_geometry
from sovlp.workloads import known_optimum_lp, uneven_batch_member


def test_buckets_follow_the_nonzero_edges():
    small = known_optimum_lp(10, 30, 3, seed=1).problem          # 90 nnz
    mid = known_optimum_lp(200, 600, 5, seed=2).problem          # ~3000 nnz
    big = known_optimum_lp(800, 3000, 8, seed=3).problem         # ~24000 nnz
    assert bucket_of(small) == 0
    assert NNZ_EDGES[bucket_of(mid)] <= mid.A.nnz < NNZ_EDGES[bucket_of(mid) + 1]
    assert bucket_of(big) == len(NNZ_EDGES) - 2


@pytest.mark.parametrize("precision", [32, 64])
def test_candidates_fit_their_kernels(precision):
    for cand in CANDIDATES:
        b, r, t = config_of(precision, *cand)
        rmax, tmax = batch_geometry(precision, b)
        assert b in (64, 128, 256, 512) and 1 <= r <= rmax and 64 <= t <= tmax


def test_table_round_trip_and_choice_by_shape():
    t = Autotuner(32)
    t.table = {0: config_of(32, 64, 4, 1.0), 1: config_of(32, 128, 1, 1.0)}
    t.measured = {0: {config_of(32, 64, 4, 1.0): 100.0}}
    u = Autotuner.from_json(t.to_json())
    assert u.table == t.table and u.measured == t.measured
    p = uneven_batch_member(1234).problem
    # the choice is a function of the LP alone
    assert u.choose(p) == u.choose(p.copy())
    b = bucket_of(p)
    assert u.choose(p) == (u.table[b] if b in u.table else u.table[max(k for k in u.table if k < b)])


def test_empty_buckets_take_the_nearest_tuned_configuration():
    from sovlp.core.problem import Problem
    from sovlp.core.sparse import SparseMatrix
    t = Autotuner(32)
    t.table = {1: config_of(32, 256, 4, 1.0), 2: config_of(32, 512, 4, 1.0)}

    def lp_with(nnz):
        rows = np.arange(nnz) % 50
        cols = np.arange(nnz) // 50
        n = int(cols.max()) + 1
        A = SparseMatrix.from_triplets(rows, cols, np.ones(nnz), 50, n)
        return Problem(A=A, c=np.ones(n), row_lb=np.zeros(50), row_ub=np.full(50, 1e30),
                       col_lb=np.zeros(n), col_ub=np.ones(n))

    assert t.choose(lp_with(20000)) == t.table[2]       # bucket 3: the largest tuned below
    assert t.choose(lp_with(500)) == t.table[1]         # bucket 0: none below, nearest above
    assert t.choose(lp_with(2000)) == t.table[1]
    assert Autotuner(32).choose(lp_with(500)) == Autotuner(32).default()


def test_table_with_an_oversized_tile_is_refused():
    d = Autotuner(32).to_json()
    d["table"] = {"0": [64, 64 * 8, 512]}
    with pytest.raises(ValueError):
        Autotuner.from_json(d)


def test_bulk_rate_uses_the_95th_percentile_finish():
    class I:
        def __init__(self, s, d):
            self.start_tick, self.device_ticks = s, d
    infos = [I(0, 10)] * 95 + [I(0, 1000)] * 5
    assert bulk_rate(infos, hz=100.0) == pytest.approx(95 / 0.10)
