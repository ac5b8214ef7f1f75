"""The allocation-free entry point kah_vs_resolution_noalloc against kah_vs_resolution.

Two calls in a row on one GPU, the second on a context with fewer cells (the fixed device
scratch keeps the first call's contents), each equal to the allocating entry point's result
on the same state; in FP64 and the FP32 form.
"""
import numpy as np
import pytest

from _util import say
from kitamr_hip import abi, cases
from kitamr_hip.solver import kah_params


def _ctx(case):
    d = abi.HipContext(kah_params(case.config))
    d.set_topology(case.topo)
    d.set_state(case.w, case.prim, case.df)
    d.step(case.dt_xi, 0)
    return d


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32f"])
def test_vs_resolution_noalloc_equals_allocating(precision):
    big = cases.riemann_2d(n=16, nv=8, vs_levels=2, precision=precision)
    small = cases.riemann_2d(n=8, nv=12, vs_levels=1, precision=precision)
    assert small.n_cells < big.n_cells
    for name, case in (("big", big), ("small", small)):
        say(f"[noalloc] {precision} {name}: cells={case.n_cells} P={case.phase_cells}")
        d = _ctx(case)
        got = d.vs_resolution_noalloc()
        want = d.vs_resolution()
        say(f"[noalloc] {name}: {got} vs {want}")
        assert np.array_equal(got, want), (name, got, want)
        assert np.all(got > 0)
        d.close()
