"""Size-independent properties at the benchmark sizes (the oracle agreement itself is tested
on small cases; at these sizes the oracle would take minutes per step):

- conservation: on a periodic, 2-level refined box of 75k leaves with a 48x48 grid (173 M
  phase-space cells), the totals of mass, momentum and energy stay within round-off over 3
  steps, in FP64 and in the FP32 form (every face flux enters its two cells with opposite
  signs, whatever the size);
- the two forms agree: on the 49k-leaf Riemann bench workload the FP32 form's moments stay
  within 1e-5 (relative, per cell) of the FP64 form's after 3 steps;
- the velocity-adaptive cylinder bench workload (200k leaves, adapted grids, 112 IB faces):
  every SolidNeighbor carries zero net mass through the wall (in the FP32 form to the float
  rounding of its stored values: the wall density is formed in FP64, the result stored as
  float), and the state stays finite with positive density away from the wall.
"""
import numpy as np
import pytest

from _util import rel_err_rows, say
from kitamr_hip import abi, cases
from kitamr_hip.solver import kah_params


def _dev(case):
    d = abi.HipContext(kah_params(case.config))
    d.set_topology(case.topo)
    d.set_state(case.w, case.prim, case.df)
    return d


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32f"])
def test_conservation_periodic_refined_box(precision):
    case = cases.periodic_box_2d(n=128, nv=48, ps_levels=2, precision=precision)
    say(f"[full] periodic {precision}: cells={case.n_cells} P={case.phase_cells}")
    assert case.phase_cells > 1e8
    dev = _dev(case)
    n = case.n_cells
    V = np.prod(case.topo["cell_ds"][:n].reshape(n, 2), axis=1)
    tot0 = (np.asarray(case.w).reshape(n, 4) * V[:, None]).sum(0)
    for s in range(3):
        dev.step(case.dt_xi, s, residual=False)
    w = dev.get_state(df=False)["w"].reshape(n, 4)
    assert np.isfinite(w).all()
    tot = (w * V[:, None]).sum(0)
    drift = np.abs(tot - tot0) / np.maximum(np.abs(tot0), 1.0)
    say(f"[full] periodic {precision}: relative drift {drift}")
    assert np.all(drift <= 1e-12), drift
    dev.close()


@pytest.mark.gpu
def test_fp32_form_tracks_fp64_on_the_bench_workload():
    c64 = cases.riemann_2d(n=64, nv=48, ps_levels=3)
    c32 = cases.riemann_2d(n=64, nv=48, ps_levels=3, precision="fp32f")
    say(f"[full] riemann fp32 vs fp64: cells={c64.n_cells} P={c64.phase_cells}")
    out = []
    for c in (c64, c32):
        d = _dev(c)
        for s in range(3):
            d.step(c.dt_xi, s, residual=False)
        out.append(d.get_state(df=False))
        d.close()
    for k in ("w", "prim"):
        e = rel_err_rows(out[1][k], out[0][k], 4)
        say(f"[full] {k}: {e:.3g}")
        assert e <= 1e-5, k


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp32f", 1e-6), ("fp64", 1e-12)])
def test_adaptive_cylinder_wall_and_state(precision, tol):
    case = cases.cylinder_2d(n=448, nv=20, vmax=15.0, L=8.0, radius=0.5, Ma=5.0, Kn=0.1,
                             omega=0.81, vs_levels=2, vs_variants=6, precision=precision)
    say(f"[full] adaptive cylinder {precision}: cells={case.n_cells} P={case.phase_cells}")
    dev = _dev(case)
    for s in range(3):
        dev.step(case.dt_xi, s, residual=False)
    aux = dev.get_ib()
    t = case.topo
    off = np.asarray(t["vs_offset"])
    nr = t["ib_normal"].reshape(-1, 2)
    for k, p in enumerate(t["ib_face_pair"]):
        h = int(t["pair_here"][p])
        a, b = int(off[h]), int(off[h + 1])
        vm = np.asarray(t["vs_mid"])[2 * a:2 * b].reshape(2, b - a).T
        wt = np.asarray(t["vs_weight"])[a:b]
        fl = wt * (vm @ nr[k]) * aux[k][0]
        assert abs(fl.sum()) < tol * np.abs(fl).sum(), k
    w = dev.get_state(df=False)["w"].reshape(-1, 4)
    fluid = t["bound_enc"] == 0
    assert np.isfinite(w).all() and (w[fluid, 0] > 0).all()
    dev.close()
