"""Velocity-space adaptation on the device (kah_vs_resolution / kah_vs_flags / kah_vs_remap /
kah_vs_project) against the oracle.

With the oracle holding exactly the device's state (w, prim, f and slopes), the resolution
and every velocity cell's refine / coarsen flag agree bit for bit (both criteria modes, 2D
and 3D, adapted grids, FP64 and the FP32 form). A whole adaptation (refine pass, coarsen
pass, the vs_balance! rounds where asked, conservative I-projection of the changed cells)
equals the same sequence done with the oracle's flags, numpy remaps of the downloaded f and
the oracle's I-projection: unprojected
cells bit for bit, projected ones within 1e-12 (FP64). The adapted state then steps in parity
with the oracle on the new grids.
"""
import numpy as np
import pytest

import oracle as O
from _util import pair, rel_err, rel_err_rows, say
from kitamr_hip import abi, cases, vsadapt
from kitamr_hip.solver import kah_params

CASES = {
    "riemann8_vsamr": lambda: cases.riemann_2d(n=8, nv=8, vs_levels=2),
    "box3d_vsamr": lambda: cases.box_3d(n=4, nv=6, vs_levels=1),
}
# grids whose neighbours break the 2:1 velocity-level balance as built
BALANCE_CASES = {
    "riemann8_vsamr3": lambda: cases.riemann_2d(n=8, nv=8, vs_levels=3),
    # faces across the periodic boundaries name image records of their cells
    "periodic8_vsamr3": lambda: cases.periodic_box_2d(n=8, nv=8, vs_levels=3),
    "box3d_vsamr2": lambda: cases.box_3d(n=4, nv=6, vs_levels=2),
}


def _sync_oracle(dev, orc):
    st = dev.get_state(sdf=True)
    orc.set_state(st["w"], st["prim"], st["df"])
    orc.set_sdf(st["sdf"])
    return st


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32f"])
@pytest.mark.parametrize("mode", ["lohner", "contribution"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_flags_equal_oracle(name, mode, precision):
    case = CASES[name]()
    say(f"[vs] flags {name} {mode} {precision}")
    dev, orc = pair(case, precision=abi.PREC[precision])
    for s in range(2):
        dev.step(case.dt_xi, s)
    _sync_oracle(dev, orc)
    vr = dev.vs_resolution()
    assert np.array_equal(vr, orc.vs_resolution())
    p = abi.vs_adapt_params(mode, coeff_local=1e-2 if mode == "lohner" else 1e-3)
    got, exp = dev.vs_flags(p, vr), orc.vs_flags(p, vr)
    assert np.array_equal(got, exp), np.flatnonzero(got != exp)[:10]
    assert (got & 2).any()
    dev.close()
    orc.close()


def _remap(f, off_old, off_new, src, ndf, nch):
    out = np.empty(off_new[-1] * ndf)
    for c in range(len(off_new) - 1):
        no, nn = off_old[c + 1] - off_old[c], off_new[c + 1] - off_new[c]
        fo = f[off_old[c] * ndf:off_old[c + 1] * ndf].reshape(ndf, no)
        s = src[off_new[c]:off_new[c + 1]]
        fn = np.empty((ndf, nn))
        for i, j in enumerate(s):
            if j >= 0:
                fn[:, i] = fo[:, j]
            else:
                acc = np.zeros(ndf)
                for g in range(nch):
                    acc = acc + fo[:, -j - 1 + g]
                fn[:, i] = acc / nch
        out[off_new[c] * ndf:off_new[c + 1] * ndf] = fn.reshape(-1)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["lohner", "contribution"])
@pytest.mark.parametrize("name,balance", [(n, False) for n in sorted(CASES)] +
                         [("riemann8_vsamr", True)])
def test_adaptation_equals_oracle_sequence(name, mode, balance):
    case = CASES[name]()
    dim, ndf = case.config.solver.DIM, case.config.solver.NDF
    nch = 1 << dim
    say(f"[vs] adapt {name} {mode} balance={balance}")
    params = kah_params(case.config)
    dev, orc = pair(case)
    for s in range(2):
        dev.step(case.dt_xi, s)
    st0 = _sync_oracle(dev, orc)
    p = abi.vs_adapt_params(mode, coeff_local=1e-2 if mode == "lohner" else 1e-3)
    # the device path
    topo2, va = vsadapt.vs_adapt(dev, case.topo, p, dim, balance=balance)
    got = dev.get_state(sdf=True)
    # the same sequence with the oracle's criteria, numpy remaps and the oracle's projection
    vr = orc.vs_resolution()
    f1 = orc.vs_flags(p, vr)
    t1, src1, ch1 = vsadapt.plan_refine(case.topo, f1, dim)
    df1 = _remap(st0["df"], np.asarray(case.topo["vs_offset"]), np.asarray(t1["vs_offset"]),
                 src1, ndf, nch)
    o1 = O.Oracle(params, t1)
    o1.set_state(st0["w"], st0["prim"], df1)
    off0, off1 = np.asarray(case.topo["vs_offset"]), np.asarray(t1["vs_offset"])
    sdf1 = np.zeros(off1[-1] * ndf * dim)
    for c in range(case.n_cells):
        if not ch1[c]:
            sdf1[off1[c] * ndf * dim:off1[c + 1] * ndf * dim] = \
                st0["sdf"][off0[c] * ndf * dim:off0[c + 1] * ndf * dim]
    o1.set_sdf(sdf1)
    f2 = o1.vs_flags(p, vr)
    t2, src2, ch2 = vsadapt.plan_coarsen(t1, f2, dim)
    df2 = _remap(df1, off1, np.asarray(t2["vs_offset"]), src2, ndf, nch)
    rounds = 0
    while balance:   # vs_balance!: injection onto the balanced grids, before the projection
        t3, src3, ch3 = vsadapt.plan_balance(t2, dim)
        if not ch3.any():
            break
        df2 = _remap(df2, np.asarray(t2["vs_offset"]), np.asarray(t3["vs_offset"]), src3, ndf,
                     nch)
        t2, rounds = t3, rounds + 1
    assert rounds > 0 or not balance, "the case does not exercise the balance"
    assert np.array_equal(np.asarray(t2["vs_offset"]), np.asarray(topo2["vs_offset"]))
    assert np.array_equal(np.asarray(t2["vs_level"]), np.asarray(topo2["vs_level"]))
    va_ref = (ch1 | ch2).astype(bool)
    assert np.array_equal(va, va_ref)
    assert va.any(), "no cell changed: the case does not exercise the adaptation"
    off2 = np.asarray(t2["vs_offset"])
    mid = np.asarray(t2["vs_mid"])
    wt = np.asarray(t2["vs_weight"])
    nw = dim + 2
    for c in np.flatnonzero(va):
        a, b = off2[c], off2[c + 1]
        blk = df2[a * ndf:b * ndf]
        df2[a * ndf:b * ndf] = O.i_projection(dim, ndf, mid[a * dim:b * dim], wt[a:b], blk,
                                              st0["w"][c * nw:(c + 1) * nw])
    # unprojected cells bit for bit, projected ones to the parity tolerance
    for c in range(case.n_cells):
        a, b = off2[c] * ndf, off2[c + 1] * ndf
        if va[c]:
            assert rel_err_rows(got["df"][a:b], df2[a:b], b - a) < 1e-12, c
        else:
            assert np.array_equal(got["df"][a:b], df2[a:b]), c
    assert not got["sdf"].any()
    # the adapted state steps in parity with the oracle on the new grids
    o2 = O.Oracle(params, t2)
    o2.set_state(got["w"], got["prim"], got["df"])
    for s in range(2):
        dev.step(case.dt_xi, 2 + s)
        o2.step(case.dt_xi, 2 + s)
    g, o = dev.get_state(), o2.get_state()
    assert rel_err_rows(g["w"], o["w"], nw) < 1e-12
    assert rel_err(g["df"], o["df"]) < 1e-12
    for x in (dev, orc, o1, o2):
        x.close()


@pytest.mark.gpu
def test_adaptation_fp32_form_moments():
    """The FP32 form adapts on the device too; the projection restores each changed cell's
    moments to its w within the float rounding of f."""
    case = cases.riemann_2d(n=8, nv=8, vs_levels=2, precision="fp32f")
    dim, ndf = 2, 2
    say("[vs] adapt fp32f")
    dev, _ = pair(case, precision=1)
    for s in range(2):
        dev.step(case.dt_xi, s)
    w0 = dev.get_state()["w"]
    topo2, va = vsadapt.vs_adapt(dev, case.topo, abi.vs_adapt_params("lohner"), dim)
    assert va.any()
    st = dev.get_state()
    off = np.asarray(topo2["vs_offset"])
    mid = np.asarray(topo2["vs_mid"])
    wt = np.asarray(topo2["vs_weight"])
    for c in np.flatnonzero(va)[:20]:
        a, b = off[c], off[c + 1]
        m = O.micro_to_macro(dim, ndf, mid[a * dim:b * dim], st["df"][a * ndf:b * ndf], wt[a:b])
        wc = w0[4 * c:4 * c + 4]
        assert np.max(np.abs(m - wc)) <= 1e-6 * np.max(np.abs(wc)), (c, m, wc)
    dev.step(case.dt_xi, 2)
    assert np.all(np.isfinite(dev.get_state()["df"]))
    dev.close()


@pytest.mark.gpu
def test_remap_rejects_bad_sources():
    case = CASES["riemann8_vsamr"]()
    say("[vs] remap errors")
    dev, _ = pair(case)
    off = np.asarray(case.topo["vs_offset"])
    src = np.concatenate([np.arange(off[c + 1] - off[c]) for c in range(case.n_cells)]).astype(np.int32)
    bad = src.copy()
    bad[3] = 10 ** 6
    with pytest.raises(abi.KahError, match="out of range"):
        dev.vs_remap(case.topo, bad, np.ones(case.n_cells, np.uint8))
    bad = src.copy()
    bad[0] = -(int(off[1]) + 1)          # a mean that would run past the grid
    with pytest.raises(abi.KahError, match="out of range"):
        dev.vs_remap(case.topo, bad, np.ones(case.n_cells, np.uint8))
    bad = src.copy()
    bad[[0, 1]] = bad[[1, 0]]
    with pytest.raises(abi.KahError, match="unchanged but remapped"):
        dev.vs_remap(case.topo, bad, np.zeros(case.n_cells, np.uint8))
    # identity remap keeps the state and the context steps on
    before = dev.get_state()
    dev.vs_remap(case.topo, src, np.zeros(case.n_cells, np.uint8))
    after = dev.get_state()
    assert np.array_equal(before["df"], after["df"]) and np.array_equal(before["w"], after["w"])
    dev.step(case.dt_xi, 0)
    dev.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32f"])
@pytest.mark.parametrize("name", sorted(BALANCE_CASES))
def test_balance_rounds_inject_on_the_device(name, precision):
    """vs_balance! alone on grids that break the balance as built (2D: three rounds, 3D: one):
    the device's f after the rounds equals the numpy injection of the downloaded f onto the
    same host-planned grids, bit for bit, and the balanced state steps in parity with the
    oracle."""
    case = BALANCE_CASES[name]()
    dim, ndf = case.config.solver.DIM, case.config.solver.NDF
    say(f"[vs] balance {name} {precision}")
    dev, _ = pair(case, precision=abi.PREC[precision])
    for s in range(2):
        dev.step(case.dt_xi, s)
    st0 = dev.get_state()
    topo, rounds = vsadapt.balance_rounds(dev, case.topo, dim)
    assert rounds >= 1
    got = dev.get_state()
    t, df, r = case.topo, st0["df"], 0
    while True:
        t1, src, ch = vsadapt.plan_balance(t, dim)
        if not ch.any():
            break
        df = _remap(df, np.asarray(t["vs_offset"]), np.asarray(t1["vs_offset"]), src, ndf,
                    1 << dim)
        t, r = t1, r + 1
    assert r == rounds
    assert np.array_equal(np.asarray(t["vs_offset"]), np.asarray(topo["vs_offset"]))
    assert np.array_equal(got["df"], df)
    assert np.array_equal(got["w"], st0["w"])
    if precision == "fp64":
        o = O.Oracle(kah_params(case.config), topo)
        o.set_state(got["w"], got["prim"], got["df"])
        for s in range(2):
            dev.step(case.dt_xi, 2 + s)
            o.step(case.dt_xi, 2 + s)
        g, oo = dev.get_state(), o.get_state()
        assert rel_err_rows(g["w"], oo["w"], dim + 2) < 1e-12
        assert rel_err(g["df"], oo["df"]) < 1e-12
        o.close()
    dev.close()
