"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the
oracle): the oracle must keep reproducing them (CPU), and the HIP path must match them within
the FP64 parity tolerance (GPU)."""
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from _util import rel_err, rel_err_rows, say
from kitamr_hip import abi
from kitamr_hip.solver import kah_params

GOLD = Path(__file__).resolve().parent / "golden"
import importlib.util  # noqa: E402

_spec = importlib.util.spec_from_file_location("make_golden", GOLD / "make_golden.py")
MG = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(MG)


@pytest.mark.parametrize("name", list(MG.CASES))
def test_oracle_reproduces_golden(name):
    g = np.load(GOLD / f"{name}.npz")
    out = MG.run(name)
    for k in ("w", "prim", "qf", "df_sub", "df_sum", "res0", "avg0"):
        assert np.allclose(out[k], g[k], rtol=1e-13, atol=1e-300), k


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(MG.CASES))
def test_hip_matches_golden(name):
    g = np.load(GOLD / f"{name}.npz")
    case = MG.CASES[name]()
    say(f"[golden] {name}")
    dev = abi.HipContext(kah_params(case.config))
    dev.set_topology(case.topo)
    dev.set_state(case.w, case.prim, case.df)
    r0 = a0 = None
    for s in range(int(g["steps"][0])):
        r, a = dev.step(case.dt_xi, s)
        if s == 0:
            r0, a0 = r, a
    st = dev.get_state()
    dim = case.config.solver.DIM
    assert rel_err_rows(st["w"], g["w"], dim + 2) <= 1e-12
    assert rel_err_rows(st["prim"], g["prim"], dim + 2) <= 1e-12
    assert rel_err(st["qf"], g["qf"]) <= 1e-10
    assert rel_err(st["df"][::int(g["stride"][0])], g["df_sub"]) <= 1e-12
    assert np.allclose(r0, g["res0"], rtol=1e-9, atol=1e-30)
    assert np.allclose(a0, g["avg0"], rtol=1e-12)
    dev.close()
