"""Host logic of the serving loop: HBM admission, slot reuse, continuous admission,
warm-start routing and latency accounting, driven by a stand-in device on a
virtual clock (no GPU: the device side is covered by test_gpu_serve.py)."""

import numpy as np
import pytest

from sovlp.core.problem import Solution, Status
from sovlp.lp.serve import HbmAdmission, Request, ServeLoop, hbm_budget, summarize
from sovlp.workloads import blending_lp, blending_stream, perturbed, small_lp


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 1e-4)


class FakeDevice:
    """Each LP finishes after ``steps_for(problem)`` steps; a step takes 1 ms of the
    virtual clock. Records what the loop asks of it."""

    def __init__(self, slots, clock, steps_for=lambda p: 3, nbytes=lambda p: 100):
        self.slots = slots
        self.clock = clock
        self.steps_for = steps_for
        self.nbytes = nbytes
        self.left = {}
        self.occupant = {}
        self.admits = []          # (time, slot, problem name, warm)
        self.max_active = 0

    def request_bytes(self, p, warm=False):
        return self.nbytes(p)

    def admit(self, slot, p, warm=None):
        assert slot not in self.left, "slot admitted twice"
        self.left[slot] = self.steps_for(p)
        self.occupant[slot] = (p, warm)
        self.admits.append((self.clock.t, slot, p.name, warm is not None))
        self.max_active = max(self.max_active, len(self.left))

    def step(self, launches=1):
        self.clock.t += 1e-3 * launches
        done = []
        for s in list(self.left):
            self.left[s] -= launches
            if self.left[s] <= 0:
                done.append(s)
        return done

    def result(self, slot):
        p, warm = self.occupant[slot]
        assert self.left[slot] <= 0, "result of a running slot"
        sol = Solution(Status.OPTIMAL, x=np.full(p.n, 1.0), objective=float(p.n),
                       y=np.full(p.m, 2.0), iterations=10 if warm else 100)

        class I:
            primal_weight = 3.0
        sol.info = I()
        return sol

    def release(self, slot):
        assert self.left.pop(slot) <= 0
        del self.occupant[slot]


def _req(key, n=10, m=4, arrival=0.0, model=None, resolve=False):
    from sovlp.core.problem import Problem
    from sovlp.core.sparse import SparseMatrix
    A = SparseMatrix.from_triplets(np.arange(m), np.arange(m), np.ones(m), m, n)
    p = Problem(A=A, c=np.ones(n), row_lb=np.zeros(m), row_ub=np.ones(m), col_lb=np.zeros(n),
                col_ub=np.ones(n), name=f"r{key}")
    return Request(key, p, arrival, model, resolve)


def test_admission_never_exceeds_the_hbm_budget():
    clk = Clock()
    dev = FakeDevice(slots=8, clock=clk)
    loop = ServeLoop(dev, HbmAdmission(250), clock=clk, sleep=clk.sleep)
    out = loop.run([_req(k) for k in range(5)])
    assert len(out) == 5 and all(s.solution is not None for s in out)
    assert dev.max_active == 2                 # 2 x 100 B fit in 250 B, a third does not
    assert loop.admission.used == 0 and loop.admission.peak == 200
    # arrival order is kept: request k is admitted no later than request k+1
    admitted = {s.key: s.admitted for s in out}
    assert all(admitted[k] <= admitted[k + 1] for k in range(4))


def test_request_larger_than_the_budget_is_refused_and_others_proceed():
    clk = Clock()
    dev = FakeDevice(slots=4, clock=clk, nbytes=lambda p: 1000 if p.name == "r1" else 100)
    out = ServeLoop(dev, HbmAdmission(500), clock=clk, sleep=clk.sleep).run(
        [_req(k) for k in range(3)])
    by = {s.key: s for s in out}
    assert by[1].solution is None and "HBM" in by[1].refused
    assert by[0].solution is not None and by[2].solution is not None


def test_slots_are_reused_and_requests_overlap():
    clk = Clock()
    # request 0 is long (50 steps); requests 1, 2 arrive later and are short
    dev = FakeDevice(slots=2, clock=clk, steps_for=lambda p: 50 if p.name == "r0" else 3)
    reqs = [_req(0), _req(1, arrival=0.005), _req(2, arrival=0.006)]
    out = ServeLoop(dev, HbmAdmission(10_000), clock=clk, sleep=clk.sleep).run(reqs)
    by = {s.key: s for s in out}
    # 1 joins while 0 runs (continuous batching), 2 reuses 1's slot when it frees
    assert by[1].admitted < by[0].finished
    assert by[2].slot == by[1].slot and by[2].admitted >= by[1].finished
    assert by[1].finished < by[0].finished
    assert all(s.latency >= 0 for s in out)


def test_resolves_warm_start_from_the_models_last_solution():
    clk = Clock()
    dev = FakeDevice(slots=4, clock=clk)
    reqs = [_req(0, model=7), _req(1, arrival=0.001, model=7, resolve=True),
            _req(2, arrival=0.5, model=7, resolve=True), _req(3, arrival=0.6, model=8,
                                                             resolve=True)]
    out = ServeLoop(dev, HbmAdmission(10_000), clock=clk, sleep=clk.sleep).run(reqs)
    by = {s.key: s for s in out}
    assert not by[0].warm
    assert not by[1].warm          # model 7 was still running when 1 arrived: cold
    assert by[2].warm              # model 7 has a solution by then: warm
    assert not by[3].warm          # no earlier solution of model 8
    warm_admits = [a for a in dev.admits if a[3]]
    assert [a[2] for a in warm_admits] == ["r2"]


def test_warm_start_of_a_different_shape_is_dropped():
    clk = Clock()
    dev = FakeDevice(slots=2, clock=clk)
    reqs = [_req(0, n=10, model=1), _req(1, n=12, arrival=0.5, model=1, resolve=True)]
    out = ServeLoop(dev, HbmAdmission(10_000), clock=clk, sleep=clk.sleep).run(reqs)
    assert not {s.key: s for s in out}[1].warm


def test_hbm_budget_and_summary():
    assert hbm_budget(0.5, probe=lambda: (1000, 4000)) == 500
    with pytest.raises(ValueError):
        HbmAdmission(0)
    clk = Clock()
    dev = FakeDevice(slots=1, clock=clk, steps_for=lambda p: 10)
    out = ServeLoop(dev, HbmAdmission(10_000), clock=clk, sleep=clk.sleep).run(
        [_req(k) for k in range(4)])
    s = summarize(out)
    assert s["finished"] == 4 and s["optimal"] == 4
    # one slot, 10 steps of 1 ms each, all arriving at 0: latencies 10, 20, 30, 40 ms
    lat = sorted(x.latency for x in out)
    np.testing.assert_allclose(lat, [0.010, 0.020, 0.030, 0.040], atol=1e-9)
    assert s["latency_p50_s"] == pytest.approx(0.025)
    assert s["lps_per_s"] == pytest.approx(4 / 0.040)


def test_blending_generator_shape_and_stream():
    p = blending_lp(3, K=4, P=2, T=5)
    K, P, T = 4, 2, 5
    assert p.n == K * P * T + K * T
    assert p.m == T * (2 * K + 4 * P)
    eq, _, le, _, _ = p.row_types()
    assert eq == T * (K + P) and le == T * (K + 3 * P)
    ev = list(blending_stream(rate=50.0, duration=2.0, seed=1))
    assert all(b[0] <= 2.0 for b in ev)
    assert 50 <= len(ev) <= 150
    kinds = [e[2] for e in ev]
    assert kinds[0] == "new" and 0.15 < kinds.count("resolve") / len(ev) < 0.45
    q = perturbed(p, 5, 0.01)
    assert q.m == p.m and q.n == p.n
    fin = p.row_ub < 1e30
    assert np.all(np.abs(q.row_ub[fin] - p.row_ub[fin]) <= 0.01 * np.abs(p.row_ub[fin]) + 1e-12)
    assert np.all(np.abs(q.c - p.c) <= 0.01 * np.abs(p.c) + 1e-12)


@pytest.mark.parametrize("warm", [False, True])
def test_non_optimal_results_escalate_to_the_fallback(warm):
    clk = Clock()

    class Failing(FakeDevice):
        def result(self, slot):
            sol = super().result(slot)
            p, _ = self.occupant[slot]
            if p.name in ("r1", "r3"):
                sol.status = Status.NUMERICAL
            return sol

    first = Failing(slots=2, clock=clk)
    second = FakeDevice(slots=1, clock=clk, steps_for=lambda p: 5)
    adm = HbmAdmission(1000)
    out = ServeLoop(first, adm, clock=clk, sleep=clk.sleep, fallback=second,
                    escalate_warm=warm).run([_req(k, arrival=0.001 * k) for k in range(5)])
    by = {s.key: s for s in out}
    assert len(out) == 5 and all(int(s.solution.status) == 0 for s in out)
    assert {k for k, s in by.items() if s.escalated} == {1, 3}
    assert by[1].first_status == int(Status.NUMERICAL)
    # the fallback got exactly the escalated requests, cold by default
    assert [(a[2], a[3]) for a in second.admits] == [("r1", warm), ("r3", warm)]
    assert by[1].finished > by[0].finished and adm.used == 0
