"""The FP32 form (f and its slopes stored and advected in FP32; moment sums, w, prim, tau,
q and the Maxwellian parameters in FP64) against the FP64 oracle: moments rho, rho*U, rho*E
and the heat flux q within 1e-5 relative after several steps (CAIDVM_Marching, and
CIP_Marching whose projection stays FP64), and mass, momentum and energy conserved to
round-off on a periodic box (the macro face fluxes stay antisymmetric FP64 sums)."""
import numpy as np
import pytest

from _util import pair, rel_err, rel_err_rows, say
from kitamr_hip import cases

MOMENT_TOL = 1e-5      # rho, rho*U, rho*E and q against the FP64 oracle
CONSERVE_TOL = 1e-12   # relative change of the domain totals of w (periodic box)

CASES = {
    "riemann8_amr1_nv16": lambda: cases.riemann_2d(n=8, nv=16, ps_levels=1, halo=(2.0, 1.5),
                                                   precision="fp32f"),
    "riemann16_amr2": lambda: cases.riemann_2d(n=16, nv=12, ps_levels=2, halo=(3.0, 2.0),
                                               precision="fp32f"),
    "riemann8_vsamr": lambda: cases.riemann_2d(n=8, nv=8, vs_levels=2, precision="fp32f"),
    "riemann8_cip": lambda: cases.riemann_2d(n=8, nv=16, marching="CIP_Marching",
                                             precision="fp32f"),
    "box3d_cip": lambda: cases.box_3d(n=4, nv=12, marching="CIP_Marching", precision="fp32f"),
    "wall_channel": lambda: cases.wall_channel(n=8, nv=8),
    "box3d_nv8": lambda: cases.box_3d(n=4, nv=8),
    "box3d_amr_vsamr": lambda: cases.box_3d(n=4, nv=6, ps_levels=1, vs_levels=1),
}


def _fp32_pair(case):
    dev, orc = pair(case, precision=1)   # KAH_PREC_FP32F on the device; the oracle is FP64
    return dev, orc


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_fp32f_moments_against_fp64_oracle(name):
    case = CASES[name]()
    say(f"[fp32f] {name}: cells={case.n_cells} P={case.phase_cells}")
    dev, orc = _fp32_pair(case)
    dt = case.dt_xi
    dim = case.config.solver.DIM
    NW = dim + 2
    for s in range(5):
        dev.step(dt, s)
        orc.step(dt, s)
    g, o = dev.get_state(), orc.get_state()
    ew = rel_err_rows(g["w"], o["w"], NW)
    eq = rel_err(g["qf"], o["qf"])
    edf = rel_err(g["df"], o["df"])
    per = [rel_err(np.asarray(g["w"]).reshape(-1, NW)[:, m], np.asarray(o["w"]).reshape(-1, NW)[:, m])
           for m in range(NW)]
    say(f"[fp32f] {name}: w {ew:.2e} (rho {per[0]:.2e}, rhoU {max(per[1:1 + dim]):.2e}, "
        f"rhoE {per[-1]:.2e}) q {eq:.2e} df {edf:.2e}")
    assert ew <= MOMENT_TOL, "w"
    assert eq <= MOMENT_TOL, "q"
    assert edf <= 1e-4, "df"
    dev.close()


@pytest.mark.gpu
@pytest.mark.parametrize("vs_levels", [0, 1])
def test_fp32f_conservation_periodic(vs_levels):
    case = cases.periodic_box_2d(n=8, nv=16, ps_levels=1, vs_levels=vs_levels,
                                 precision="fp32f")
    say(f"[fp32f] periodic vs_levels={vs_levels}: cells={case.n_cells} P={case.phase_cells}")
    dev, orc = _fp32_pair(case)
    n = case.n_cells
    V = np.prod(case.topo["cell_ds"][:n].reshape(n, 2), axis=1)
    tot0 = (np.asarray(case.w).reshape(n, 4) * V[:, None]).sum(0)
    for s in range(10):
        dev.step(case.dt_xi, s)
    tot = (dev.get_state()["w"].reshape(n, 4) * V[:, None]).sum(0)
    scale = np.maximum(np.abs(tot0), 1.0)
    drift = np.abs(tot - tot0) / scale
    say(f"[fp32f] periodic vs_levels={vs_levels}: relative drift of the totals {drift}")
    assert np.all(drift <= CONSERVE_TOL), drift
    dev.close()
