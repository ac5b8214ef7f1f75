"""The serving loop on the device: cases of different sizes, forms and parameters arrive
while others run, are admitted as slots and memory free up (charged the device bytes their
contexts hold, admitted on an upper bound of them), and each result is bit-for-bit the result
of the same case run alone (contexts are independent, the kernels deterministic)."""
import numpy as np
import pytest

from _util import say
from kitamr_hip import cases, serve
from kitamr_hip import solver as S


def _mix():
    return [
        ("riemann8", lambda: cases.riemann_2d(n=8, nv=16), 0.0, 12),
        ("riemann16_amr", lambda: cases.riemann_2d(n=16, nv=8, ps_levels=1, halo=(2.0, 1.5),
                                                   Kn=0.1), 0.0, 9),
        ("periodic_fp32f", lambda: cases.periodic_box_2d(n=8, nv=16, precision="fp32f"), 0.01,
         10),
        ("box3d", lambda: cases.box_3d(n=4, nv=8), 0.02, 7),
        ("riemann8_vsamr", lambda: cases.riemann_2d(n=8, nv=8, vs_levels=1, Kn=1e-3), 0.02, 8),
        ("wall", lambda: cases.wall_channel(n=8, nv=8), 0.03, 11),
        ("cylinder", lambda: cases.cylinder_2d(n=16, nv=16), 0.01, 9),
        ("cylinder_vsamr_fp32f", lambda: cases.cylinder_2d(n=16, nv=8, vs_levels=1,
                                                           precision="fp32f"), 0.02, 6),
    ]


def _alone(case, steps):
    ka = S.KA(case)
    for _ in range(steps):
        S.limit_dt(ka)
        S.step(ka)
    ka.ctx.sync()
    st = ka.ctx.get_state()
    ka.close()
    return st


@pytest.mark.gpu
def test_results_alone_equal_results_in_the_mix():
    mix = _mix()
    reqs = []
    for i, (tag, mk, t, steps) in enumerate(mix):
        reqs.append(serve.Request(rid=i, case=mk(), arrival=t, max_steps=steps, tag=tag))
    # budget and slots small enough that some requests wait for others to finish
    budget = sum(sorted(serve.device_bytes_estimate(r.case) for r in reqs)[-3:])
    srv = serve.Server(budget_bytes=budget, max_slots=3, log=say)
    rep = srv.run(reqs)
    say(f"[serving] {rep}")
    assert srv.peak_active <= 3 and srv.peak_used <= budget
    for r, (tag, mk, t, steps) in zip(reqs, mix):
        assert r.steps == steps, tag
        assert 0 < r.charged <= r.nbytes, tag       # admitted on an upper bound
        ref = _alone(mk(), steps)
        for k in ("w", "prim", "qf", "df"):
            assert np.array_equal(r.result[k], ref[k]), (tag, k)


@pytest.mark.gpu
def test_serving_workload_requests_alone_equal_mix():
    """Requests drawn from the serving workload's distributions (Riemann or cylinder, Kn
    log-uniform in [1e-3, 1], Ma in [0.5, 5], 4k-200k cells, 16^2-64^2 velocity grids, FP64 or
    the FP32 form; serve.c5_specs), a few steps each, arriving while others run under a budget
    and four slots: each result equals the same request run alone, bit for bit."""
    specs = serve.c5_specs(6, 20261205)
    steps = [6, 4, 5, 3, 6, 4]
    reqs = []
    for i, sp in enumerate(specs):
        case = serve.c5_case(sp)
        say(f"[serving c5] {i}: {serve.c5_tag(sp)} P={case.phase_cells}")
        reqs.append(serve.Request(rid=i, case=case, arrival=0.002 * i, max_steps=steps[i],
                                  tag=serve.c5_tag(sp)))
    budget = sum(sorted(serve.device_bytes_estimate(r.case) for r in reqs)[-3:])
    srv = serve.Server(budget_bytes=budget, max_slots=4, log=say)
    rep = srv.run(reqs)
    say(f"[serving c5] {rep}")
    assert srv.peak_active <= 4 and srv.peak_used <= budget
    for r, sp, st in zip(reqs, specs, steps):
        assert r.steps == st and 0 < r.charged <= r.nbytes, r.tag
        ref = _alone(serve.c5_case(sp), st)
        for k in ("w", "prim", "qf", "df"):
            assert np.array_equal(r.result[k], ref[k]), (r.tag, k)
