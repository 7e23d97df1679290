"""Batched PDLP on the GPU: every LP of an uneven batch reaches its closed-form
optimum, and an LP's result is identical alone and in the mix."""

import numpy as np
import pytest

from oracle import oracle
from sovlp.core.problem import Status
from sovlp.lp.batch import solve_batch
from sovlp.lp.pdlp import PDLPParams, solve_pdlp
from sovlp.workloads import uneven_batch_member

pytestmark = pytest.mark.gpu


def _batch(count, base=2000):
    return [uneven_batch_member(base + i) for i in range(count)]


def test_fp64_batch_matches_closed_form_and_oracle():
    print("[gpu] fp64 batch of 12 uneven LPs", flush=True)
    kos = _batch(12)
    par = PDLPParams(eps_abs=1e-6, eps_rel=1e-6)
    res = solve_batch([k.problem for k in kos], par, precision=64, block=256)
    print(f"[gpu] batch done in {res.seconds:.3f}s, {res.graphs} graph replays", flush=True)
    for ko, s in zip(kos, res.solutions):
        assert s.status == Status.OPTIMAL, (ko.problem.name, s)
        assert abs(s.objective - ko.f_star) <= 1e-5 * max(1.0, abs(ko.f_star))
        rv, bv, _ = ko.problem.violation(s.x)
        assert rv <= 1e-6 and bv <= 1e-12
    # the oracle on two of them
    for ko, s in list(zip(kos, res.solutions))[:2]:
        r = oracle.solve(ko.problem, eps_abs=1e-6, eps_rel=1e-6)
        assert r.status == 0
        assert abs(s.objective - r.objective) <= 1e-5 * max(1.0, abs(r.objective))


def test_result_is_identical_alone_and_in_the_mix():
    print("[gpu] alone vs mixed", flush=True)
    kos = _batch(6, base=3000)
    par = PDLPParams(eps_abs=1e-6, eps_rel=1e-6)
    mixed = solve_batch([k.problem for k in kos], par, precision=64, block=256)
    for i in (0, 3):
        alone = solve_batch([kos[i].problem], par, precision=64, block=256)
        a, b = alone.solutions[0], mixed.solutions[i]
        assert a.iterations == b.iterations
        assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
        assert a.objective == b.objective
    # the single-LP engine with one workgroup runs the same arithmetic
    s = solve_pdlp(kos[0].problem, PDLPParams(device="gpu", team="block", eps_abs=1e-6, eps_rel=1e-6))
    assert s.iterations == mixed.solutions[0].iterations
    assert np.array_equal(s.x, mixed.solutions[0].x)


def test_fp32_refined_batch_reaches_closed_form_to_1e8():
    """FP32 iterates + FP64 reductions + iterative refinement: every LP meets
    _converged at 1e-6 (feas_cap 1e-6) and its objective is within 1e-8 of the
    closed-form optimum; each refinement round's record is reported."""
    print("[gpu] fp32 + refinement batch of 10", flush=True)
    kos = _batch(10, base=4000)
    par = PDLPParams(eps_abs=1e-6, eps_rel=1e-6, feas_cap=1e-6)
    res = solve_batch([k.problem for k in kos], par, precision=32, block=256)
    print(f"[gpu] fp32 batch done in {res.seconds:.3f}s", flush=True)
    for ko, s in zip(kos, res.solutions):
        assert s.status == Status.OPTIMAL, (ko.problem.name, s.status, s.info.refinement)
        assert abs(s.objective - ko.f_star) <= 1e-8 * max(1.0, abs(ko.f_star)), \
            (ko.problem.name, s.objective, ko.f_star, s.info.refinement)
        rv, bv, _ = ko.problem.violation(s.x)
        assert rv <= 1e-6 and bv <= 1e-6
        ref = s.info.refinement
        # the reported point is one of the evaluated ones (the best): its row has
        # the reported objective and meets the feasibility cap
        row = min(ref, key=lambda r: abs(r[4] - s.objective))
        assert abs(row[4] - s.objective) <= 1e-9 * max(1.0, abs(s.objective))
        assert row[1] <= 1e-6
        # textbook duals on output, as in FP64
        np.testing.assert_allclose(s.reduced_costs, ko.problem.c - ko.problem.A.rmatvec(s.y),
                                   rtol=1e-9, atol=1e-9)


def test_fp32_result_is_identical_alone_and_in_the_mix():
    print("[gpu] fp32 alone vs mixed", flush=True)
    kos = _batch(5, base=6000)
    par = PDLPParams(eps_abs=1e-6, eps_rel=1e-6, feas_cap=1e-6)
    mixed = solve_batch([k.problem for k in kos], par, precision=32, block=256)
    alone = solve_batch([kos[2].problem], par, precision=32, block=256)
    a, b = alone.solutions[0], mixed.solutions[2]
    assert a.iterations == b.iterations
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)


def test_small_workgroups_agree_with_large():
    print("[gpu] block 64 vs 256", flush=True)
    kos = _batch(4, base=5000)
    par = PDLPParams(eps_abs=1e-6, eps_rel=1e-6)
    a = solve_batch([k.problem for k in kos], par, precision=64, block=64)
    for ko, s in zip(kos, a.solutions):
        assert s.status == Status.OPTIMAL
        assert abs(s.objective - ko.f_star) <= 1e-5 * max(1.0, abs(ko.f_star))


def test_autotuned_mixed_configurations_in_one_batch():
    """Per-LP kernel configurations from an autotuner table (64, 128 and 512
    threads, different tile caps) in one batch: every LP reaches its closed
    form, and an LP's result is identical alone and in the mix."""
    from sovlp.lp.autotune import Autotuner, bucket_of, config_of
    print("[gpu] mixed configurations", flush=True)
    t = Autotuner(32)
    t.table = {0: config_of(32, 64, 4, 1.0), 1: config_of(32, 128, 1, 1.0),
               2: config_of(32, 512, 4, 0.5), 3: config_of(32, 128, 4, 1.0)}
    kos = _batch(12, base=9000)
    assert len({bucket_of(k.problem) for k in kos}) >= 2
    par = PDLPParams(eps_abs=1e-6, eps_rel=1e-6, feas_cap=1e-6)
    mixed = solve_batch([k.problem for k in kos], par, precision=32, tuner=t)
    for ko, s in zip(kos, mixed.solutions):
        assert s.status == Status.OPTIMAL, (ko.problem.name, s.status)
        assert abs(s.objective - ko.f_star) <= 1e-6 * max(1.0, abs(ko.f_star))
    for i in (0, 5):
        alone = solve_batch([kos[i].problem], par, precision=32, tuner=t).solutions[0]
        assert alone.iterations == mixed.solutions[i].iterations
        assert np.array_equal(alone.x, mixed.solutions[i].x)
        assert np.array_equal(alone.y, mixed.solutions[i].y)


@pytest.mark.parametrize("precision", [32, 64])
def test_lds_resident_matrix_is_bit_for_bit_the_streamed_one(precision):
    """The run kernel holds an LP's scaled matrix and step vectors in LDS when
    they fit (the PDHG loop then reads no matrix from memory), and otherwise its
    step vectors alone (the products gather from LDS): every LP of a mixed
    batch, in either mode or neither, ends with the same iterations, x and y as
    with everything streamed from memory (PDLPParams(lds_matrix=False)); the
    single-LP engine on one workgroup does the same."""
    from sovlp.lp.autotune import Autotuner, packaged_table
    from sovlp.lp.batch import BatchEngine
    print(f"[gpu] LDS-resident vs streamed matrix, FP{precision}", flush=True)
    kos = _batch(48, base=7000)
    probs = [k.problem for k in kos]
    tuner = Autotuner.load(packaged_table(32)) if precision == 32 else None
    out, modes = {}, None
    for flag in (True, False):
        par = PDLPParams(eps_abs=1e-6, eps_rel=1e-6, feas_cap=1e-6, device="gpu",
                         lds_matrix=flag)
        eng = BatchEngine(probs, par, precision=precision, block=256, tuner=tuner)
        try:
            eng.run()
            out[flag], _ = eng.solutions()
            if flag:
                modes = eng.lds_modes()
        finally:
            eng.close()
    resident = modes == 1
    print(f"[gpu] {int(resident.sum())} of {len(probs)} LPs resident, "
          f"{int((modes == 2).sum())} with only their step vectors in LDS", flush=True)
    assert 0 < resident.sum() < len(probs) and (modes == 2).sum() > 0
    for i, (a, b) in enumerate(zip(out[True], out[False])):
        assert a.status == b.status and a.iterations == b.iterations, i
        assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y), i
        assert np.array_equal(a.reduced_costs, b.reduced_costs), i
    i = int(np.flatnonzero(resident)[0])
    single = [solve_pdlp(probs[i], PDLPParams(device="gpu", team="block", eps_abs=1e-6,
                                              eps_rel=1e-6, precision=precision,
                                              lds_matrix=flag)) for flag in (True, False)]
    assert single[0].iterations == single[1].iterations
    assert np.array_equal(single[0].x, single[1].x) and np.array_equal(single[0].y, single[1].y)
    if precision == 64:      # same workgroup size as the batch: the batch's arithmetic
        assert single[0].iterations == out[False][i].iterations
        assert np.array_equal(single[0].x, out[False][i].x)
