"""Writes the golden fixtures of tests/test_golden.py from the oracle.

The reference ships no golden vectors and Julia is not available, so these fixtures are the
oracle's outputs on small seeded cases: they pin the HIP path (and any later change of the
oracle) to fixed numbers. Run from the repository root:  python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "kitamr.jl_amd"))
sys.path.insert(0, str(ROOT / "oracle"))

import oracle as O  # noqa: E402
from kitamr_hip import cases  # noqa: E402
from kitamr_hip.solver import kah_params  # noqa: E402

CASES = {
    "riemann8_caidvm": lambda: cases.riemann_2d(n=8, nv=8),
    "riemann8_amr2_vs1": lambda: cases.riemann_2d(n=4, nv=6, ps_levels=2, halo=(1.5, 1.0),
                                                  vs_levels=1),
    "wall_channel8": lambda: cases.wall_channel(n=8, nv=8),
    "riemann8_euler_dvm": lambda: cases.riemann_2d(n=8, nv=8, marching="Euler", flux="DVM"),
    "box3d": lambda: cases.box_3d(n=3, nv=6, bcs=("SuperSonicInflow", "UniformOutflow",
                                                  "Period", "Period", "Maxwellian",
                                                  "UniformOutflow")),
    "riemann8_cip": lambda: cases.riemann_2d(n=8, nv=16, marching="CIP_Marching"),
}
STEPS = 3
STRIDE = 7


def run(name):
    case = CASES[name]()
    o = O.Oracle(kah_params(case.config), case.topo)
    o.set_state(case.w, case.prim, case.df)
    res = None
    for s in range(STEPS):
        r, a = o.step(case.dt_xi, s)
        res = (r, a) if s == 0 else res
    st = o.get_state()
    df = st["df"]
    return dict(w=st["w"], prim=st["prim"], qf=st["qf"], df_sub=df[::STRIDE],
                df_sum=np.array([df.sum(), np.abs(df).sum(), (df * df).sum()]),
                res0=res[0], avg0=res[1], steps=np.array([STEPS]), stride=np.array([STRIDE]))


if __name__ == "__main__":
    out = Path(__file__).resolve().parent
    for name in sys.argv[1:] or CASES:
        np.savez_compressed(out / f"{name}.npz", **run(name))
        print("wrote", name)
