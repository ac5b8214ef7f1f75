"""Restart arrays gathered from and scattered to HBM (kah_save_restart / kah_load_restart)
against the layout of save_for_restart (src/IO/Restart.jl:80-117) and the loader
_restart_build_ps_list (:306-321).

The expected arrays are assembled on the host from kah_get_state and the topology: a
transposed ws or vs_df, a solid cell that keeps its velocity cells, or a shifted row fails the
bit-for-bit comparison. A checkpoint loaded into a fresh context (one file, or two rank files
of another partition) gives back the state bit for bit, prim = get_prim(w), and the same next
step.
"""
import numpy as np
import pytest

import oracle as O
from _util import pair, rel_err, say
from kitamr_hip import abi, cases
from kitamr_hip import restart as R
from kitamr_hip.solver import kah_params
from test_gpu_mirrors import _with_solid

CASES = {
    "riemann8": lambda: cases.riemann_2d(n=8, nv=8),
    "riemann16_amr2": lambda: cases.riemann_2d(n=16, nv=12, ps_levels=2, halo=(3.0, 2.0)),
    "riemann8_vsamr": lambda: cases.riemann_2d(n=8, nv=8, vs_levels=2),
    "box3d_amr_vsamr": lambda: cases.box_3d(n=4, nv=6, ps_levels=1, vs_levels=1),
}


def host_restart(topo, st, dim, ndf):
    """The six arrays of save_for_restart from host copies of the state and the topology."""
    nw = dim + 2
    off = np.asarray(topo["vs_offset"])
    benc = np.asarray(topo["bound_enc"])
    uni = int(topo.get("vs_uniform", 0))
    vmid = np.asarray(topo["vs_mid"])
    vlev = np.asarray(topo["vs_level"])
    nc = len(off) - 1
    vs_nums = np.where(benc < 0, 0, np.diff(off)).astype(np.int64)
    ws = st["w"].reshape(nc, nw).copy()
    ws[benc < 0] = 0.0
    lev, mid, df = [], [], []
    for c in range(nc):
        if benc[c] < 0:
            continue
        a, b = int(off[c]), int(off[c + 1])
        n = b - a
        g = 0 if uni else a
        lev.append(vlev[g:g + n])
        mid.append(vmid[g * dim:(g + n) * dim].reshape(dim, n).T)
        df.append(st["df"][a * ndf:b * ndf].reshape(ndf, n).T)
    return dict(vs_nums=vs_nums, bound_encs=benc.astype(np.int64), ws=ws,
                vs_levels=np.concatenate(lev).astype(np.int8), vs_midpoints=np.concatenate(mid),
                vs_df=np.concatenate(df))


def _ctx(case, precision="fp64"):
    params = kah_params(case.config)
    params.precision = abi.PREC[precision]
    dev = abi.HipContext(params)
    dev.set_topology(case.topo)
    return dev


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32f"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_save_matches_reference_layout(name, precision):
    case = CASES[name]()
    dim, ndf = case.config.solver.DIM, case.config.solver.NDF
    say(f"[restart] save {name} {precision}")
    dev = _ctx(case, precision)
    dev.set_state(case.w, case.prim, case.df)
    dev.step(case.dt_xi, 0)
    got = dev.save_restart()
    exp = host_restart(case.topo, dev.get_state(), dim, ndf)
    for k in R.ARRAYS:
        assert got[k].shape == exp[k].shape, k
        assert np.array_equal(got[k], exp[k]), k
    dev.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32f"])
@pytest.mark.parametrize("name", ["riemann16_amr2", "riemann8_vsamr", "box3d_amr_vsamr"])
def test_restart_resumes_bit_for_bit(name, precision, tmp_path):
    case = CASES[name]()
    say(f"[restart] resume {name} {precision}")
    a = _ctx(case, precision)
    a.set_state(case.w, case.prim, case.df)
    a.step(case.dt_xi, 0)
    a.step(case.dt_xi, 1)
    R.save_for_restart(a, tmp_path, {"step": 2, "sim_time": 2 * case.dt_xi})
    b = _ctx(case, precision)
    st = R.restart(b, tmp_path)
    assert st["step"] == 2
    sa, sb = a.get_state(), b.get_state()
    assert np.array_equal(sa["w"], sb["w"]) and np.array_equal(sa["df"], sb["df"])
    # prim = get_prim(w) (src/IO/Restart.jl:319)
    nw = a.dim + 2
    g = case.config.gas.gamma
    pexp = np.concatenate([O.get_prim(a.dim, sb["w"][nw * c:nw * (c + 1)], g)
                           for c in range(case.n_cells)])
    assert rel_err(sb["prim"], pexp) < 1e-14
    a.step(case.dt_xi, 2)
    b.step(case.dt_xi, 2)
    sa, sb = a.get_state(), b.get_state()
    assert np.array_equal(sa["w"], sb["w"]) and np.array_equal(sa["df"], sb["df"])
    a.close()
    b.close()


@pytest.mark.gpu
def test_restart_from_another_partition(tmp_path):
    """Two rank files (cells [0, k) and [k, N)) load into one context, and one file loads
    into two contexts holding the halves' cells."""
    case = CASES["riemann16_amr2"]()
    say("[restart] partitions")
    a = _ctx(case)
    a.set_state(case.w, case.prim, case.df)
    a.step(case.dt_xi, 0)
    arr = a.save_restart()
    n = case.n_cells
    k = n // 3
    cum = np.concatenate([[0], np.cumsum(arr["vs_nums"])])
    for r, (c0, c1) in enumerate([(0, k), (k, n)]):
        v0, v1 = int(cum[c0]), int(cum[c1])
        R.write_rank_file(tmp_path, r, {"vs_nums": arr["vs_nums"][c0:c1],
                                        "bound_encs": arr["bound_encs"][c0:c1],
                                        "ws": arr["ws"][c0:c1],
                                        "vs_levels": arr["vs_levels"][v0:v1],
                                        "vs_midpoints": arr["vs_midpoints"][v0:v1],
                                        "vs_df": arr["vs_df"][v0:v1]})
    R.write_manifest(tmp_path, 2, 2, [k, n - k], [int(cum[k]), int(cum[n] - cum[k])])
    R.write_status(tmp_path)
    b = _ctx(case)
    R.restart(b, tmp_path)
    assert np.array_equal(a.get_state()["df"], b.get_state()["df"])
    back = b.save_restart()
    for key in R.ARRAYS:
        assert np.array_equal(back[key], arr[key])
    a.close()
    b.close()


@pytest.mark.gpu
def test_load_rejects_another_topology():
    say("[restart] rejects")
    u = CASES["riemann8"]()
    v = CASES["riemann8_vsamr"]()
    a = _ctx(v)
    a.set_state(v.w, v.prim, v.df)
    arr = a.save_restart()
    b = _ctx(u)
    b.set_state(u.w, u.prim, u.df)
    before = b.get_state()
    with pytest.raises(abi.KahError, match="do not match the topology"):
        b.load_restart(arr)
    after = b.get_state()
    assert np.array_equal(before["df"], after["df"])
    a.close()
    b.close()


@pytest.mark.gpu
def test_solid_cells_in_restart_arrays():
    case = cases.riemann_2d(n=8, nv=8)
    say("[restart] solid")
    topo = _with_solid(case, 27)
    dev = abi.HipContext(kah_params(case.config))
    dev.set_topology(topo)
    dev.set_state(case.w, case.prim, case.df)
    got = dev.save_restart()
    assert got["vs_nums"][27] == 0 and got["bound_encs"][27] == -1
    assert np.all(got["ws"][27] == 0.0)
    exp = host_restart(topo, dev.get_state(), 2, 2)
    for k in R.ARRAYS:
        assert np.array_equal(got[k], exp[k]), k
    dev.close()
