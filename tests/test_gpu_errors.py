"""Error behaviour of the C ABI on the device: every failure is a status code with a message
(the reference raises `error(...)`), nothing aborts, and the context stays usable."""
import numpy as np
import pytest

from _util import say
from kitamr_hip import abi, cases
from kitamr_hip.solver import kah_params


def _ctx(case, **kw):
    dev = abi.HipContext(kah_params(case.config))
    return dev


@pytest.mark.gpu
def test_step_before_state_and_bad_dt():
    case = cases.riemann_2d(n=4, nv=8)
    say("[errors] state/dt")
    dev = _ctx(case)
    with pytest.raises(abi.KahError, match="no topology"):
        dev.step(case.dt_xi, 0)
    dev.set_topology(case.topo)
    with pytest.raises(abi.KahError, match="no state"):
        dev.step(case.dt_xi, 0)
    dev.set_state(case.w, case.prim, case.df)
    for bad in (0.0, -1.0, float("nan"), float("inf")):
        with pytest.raises(abi.KahError, match="dt must be positive"):
            dev.step(bad, 0)
    dev.step(case.dt_xi, 0)          # still usable
    assert np.all(np.isfinite(dev.get_state()["df"]))
    dev.close()


@pytest.mark.gpu
def test_invalid_topology_is_rejected_with_a_message():
    case = cases.riemann_2d(n=4, nv=8)
    say("[errors] topology")
    dev = _ctx(case)
    bad = dict(case.topo)
    bad["nb_cell"] = np.array(case.topo["nb_cell"], copy=True)
    bad["nb_cell"][0] = 10 ** 6
    with pytest.raises(abi.KahError, match="neighbour id out of range"):
        dev.set_topology(bad)
    bad = dict(case.topo)
    bad["image_src"] = np.array(case.topo["image_src"], copy=True)
    if len(bad["image_src"]):
        bad["image_src"][0] = -5
        with pytest.raises(abi.KahError, match="image source out of range"):
            dev.set_topology(bad)
    dev.set_topology(case.topo)      # a valid topology afterwards works
    dev.set_state(case.w, case.prim, case.df)
    dev.step(case.dt_xi, 0)
    dev.close()


@pytest.mark.gpu
def test_invalid_options():
    case = cases.riemann_2d(n=4, nv=8)
    say("[errors] options")
    dev = _ctx(case)
    for name, value in (("flux_waves", 6), ("flux_waves", 1), ("slope_vpt", 0), ("slope_vpt", 7), ("vel_tiles", 2),
                        ("overlap_maps", -1), ("graphs", 3), ("branches", 0), ("branches", 5),
                        ("regular", 2)):
        with pytest.raises(abi.KahError):
            dev.set_option(name, value)
    with pytest.raises(abi.KahError, match="unknown option"):
        dev._check(dev.lib.kah_set_option(dev.h, 99, 1), "kah_set_option")
    dev.close()


@pytest.mark.gpu
def test_dvm_with_maxwellian_wall_raises_like_the_reference():
    """src/Flux/DVM.jl:12 reads undefined variables: the device rejects the configuration with
    the reference's UndefVarError message (the oracle raises it in flux!)."""
    case = cases.riemann_2d(n=4, nv=8, flux="DVM", bc="Maxwellian")
    say("[errors] DVM + wall")
    dev = _ctx(case)
    with pytest.raises(abi.KahError, match="here_weight"):
        dev.set_topology(case.topo)
    dev.close()
