"""Host logic of the serving loop (no GPU): admission by slot count and memory budget in
arrival order, every request finished once, accounting of the budget."""
import numpy as np
import pytest

from kitamr_hip import cases, serve
from kitamr_hip.solver import Residual, Status


class FakeCtx:
    def __init__(self, log, rid, held=None):
        self.log, self.rid, self.held = log, rid, held

    def device_bytes(self):
        return self.held

    def step(self, dt, step, residual=True):
        self.log.append(("step", self.rid))
        return np.zeros(4), np.zeros(4)

    def sync(self):
        pass

    def get_state(self, df=True):
        return {"rid": self.rid}

    def close(self):
        self.log.append(("close", self.rid))


class FakeKA:
    def __init__(self, case, log, held=None):
        self.case = case
        self.config = case.config
        self.ctx = FakeCtx(log, case.rid, held if held is not None else case.nbytes)
        self.status = Status(1e-3, 1e-3, residual=Residual(np.zeros(4), np.zeros(4), np.ones(4)))

    def close(self):
        self.ctx.close()


def _requests(sizes, steps, arrivals):
    base = cases.riemann_2d(n=4, nv=8)
    out = []
    for i, (nb, st, t) in enumerate(zip(sizes, steps, arrivals)):
        c = type("C", (), {})()
        c.config, c.topo, c.phase_cells, c.n_cells, c.rid = (base.config, base.topo,
                                                             base.phase_cells, base.n_cells, i)
        c.nbytes = nb
        out.append(serve.Request(rid=i, case=c, arrival=t, max_steps=st, nbytes=nb))
    return out


def test_admission_respects_slots_and_budget():
    log = []
    state = {"active": 0, "max_active": 0}
    srv = serve.Server(budget_bytes=100, max_slots=2, factory=lambda c: FakeKA(c, log))
    reqs = _requests([60, 30, 50, 20, 40], [3, 5, 2, 4, 1], [0.0] * 5)
    rep = srv.run(reqs)
    assert rep["requests"] == 5
    assert srv.peak_active <= 2 and srv.peak_used <= 100
    assert srv.admit_order == [0, 1, 2, 3, 4]          # FIFO
    for r in reqs:
        assert r.steps == r.max_steps and r.result == {"rid": r.rid}
        assert r.finished >= r.admitted >= r.arrival
    assert [e for e in log if e[0] == "close"].__len__() == 5
    assert srv.used == 0


def test_head_of_line_waits_for_memory():
    log = []
    srv = serve.Server(budget_bytes=100, max_slots=4, factory=lambda c: FakeKA(c, log))
    reqs = _requests([70, 60, 10], [4, 1, 1], [0.0, 0.0, 0.0])
    srv.run(reqs)
    # request 1 (60 B) cannot join request 0 (70 B); request 2 queues behind it (FIFO)
    assert srv.admit_order == [0, 1, 2]
    assert reqs[1].admitted >= reqs[0].finished


def test_charged_the_bytes_the_context_holds():
    """Admission reserves the estimate, then the budget is charged what the created context
    holds: two 60 B requests whose contexts hold 30 B each run together in a 100 B budget."""
    log = []
    srv = serve.Server(budget_bytes=100, max_slots=4, factory=lambda c: FakeKA(c, log, held=30))
    reqs = _requests([60, 60], [5, 5], [0.0, 0.0])
    srv.run(reqs)
    assert srv.peak_active == 2 and srv.peak_used == 60
    assert [r.charged for r in reqs] == [30, 30] and srv.used == 0


def test_c5_specs_follow_the_stated_distributions():
    sp = serve.c5_specs(400, 7)
    n = np.array([s["n"] for s in sp])
    m = np.array([s["nv"] for s in sp])
    kn = np.array([s["Kn"] for s in sp])
    ma = np.array([s["Ma"] for s in sp])
    assert np.all(n % 2 == 0) and np.all(n * n >= 4000) and np.all(n * n <= 2.1e5)
    assert set(m.tolist()) <= set(range(16, 65, 8)) and len(set(m.tolist())) == 7
    assert np.all((kn >= 1e-3) & (kn <= 1.0)) and np.all((ma >= 0.5) & (ma <= 5.0))
    kinds = [s["kind"] for s in sp]
    assert 150 < kinds.count("riemann") < 250
    assert np.all([s["vmax"] >= max(5.0, s["Ma"] * (5 / 6) ** 0.5 + 5.0) - 1e-12 for s in sp])
    assert serve.c5_specs(5, 3) == serve.c5_specs(5, 3)


def test_request_over_budget_is_rejected():
    srv = serve.Server(budget_bytes=10, max_slots=1, factory=lambda c: None)
    with pytest.raises(ValueError):
        srv.run(_requests([11], [1], [0.0]))


def test_device_bytes_estimate_covers_state():
    c = cases.riemann_2d(n=8, nv=16)
    est = serve.device_bytes_estimate(c)
    assert est >= c.phase_cells * 2 * 8 * 4     # f ping-pong + 2D slopes, FP64


def test_poisson_arrivals_seeded():
    a = serve.poisson_arrivals(50, 10.0, 3)
    b = serve.poisson_arrivals(50, 10.0, 3)
    assert np.array_equal(a, b) and np.all(np.diff(a) > 0)
    assert 2.0 < a[-1] < 10.0
