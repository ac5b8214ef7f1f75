"""Immersed boundary on the device (update_solid_cell!, update_solid_neighbor! and the IB
faces, src/Boundary/Immersed_boundary.jl and src/Flux/Flux.jl:458-487) against the oracle:
the split phases and the fused step in FP64 (cylinders in 2D, spheres in 3D), graph replay,
the FP32 form, positivity_preserving_ib! under CIP_Marching, the zero wall mass flux of every
SolidNeighbor and the rejected configurations.

Cells inside the solid (no faces, never updated) are left out of the state comparisons: the
device does not carry their f across the two state buffers.
"""
import numpy as np
import pytest

import oracle as O
from _util import pair, rel_err, rel_err_cols, rel_err_rows, say
from kitamr_hip import abi, cases
from kitamr_hip.solver import kah_params

TOL = 1e-12

CASES = {
    "cyl16": lambda: cases.cylinder_2d(n=16, nv=12),
    "cyl32_nv16": lambda: cases.cylinder_2d(n=32, nv=16),
    "cyl16_amr1": lambda: cases.cylinder_2d(n=16, nv=8, ps_levels=1),
    "cyl16_euler": lambda: cases.cylinder_2d(n=16, nv=8, marching="Euler"),
    "sph8": lambda: cases.sphere_3d(n=8, nv=6, L=2.0),
    "sph12_nv8": lambda: cases.sphere_3d(n=12, nv=8),
    # per-cell adapted velocity grids (six variants: neighbours differ)
    "cyl16_vsamr": lambda: cases.cylinder_2d(n=16, nv=8, vs_levels=2),
    "cyl16_vsamr_amr1": lambda: cases.cylinder_2d(n=16, nv=8, vs_levels=1, ps_levels=1),
    # CIP_Marching: positivity_preserving_ib! on the donors
    "cyl16_cip": lambda: cases.cylinder_2d(n=16, nv=16, marching="CIP_Marching"),
    "cyl16_vsamr_cip": lambda: cases.cylinder_2d(n=16, nv=8, vs_levels=1,
                                                 marching="CIP_Marching"),
    "sph8_cip": lambda: cases.sphere_3d(n=8, nv=12, L=2.0, marching="CIP_Marching"),
    # the DVM flux: the wall side reconstructed with the SolidNeighbor's boundary slope
    "cyl16_dvm": lambda: cases.cylinder_2d(n=16, nv=12, flux="DVM"),
    "cyl16_vsamr_dvm": lambda: cases.cylinder_2d(n=16, nv=8, vs_levels=1, flux="DVM"),
    "sph8_dvm": lambda: cases.sphere_3d(n=8, nv=6, L=2.0, flux="DVM"),
    "sph8_vsamr_dvm": lambda: cases.sphere_3d(n=8, nv=6, L=2.0, vs_levels=1, flux="DVM"),
}


def live_cells(case):
    """Fluid cells and the solid cells that update_solid_cell! rebuilds."""
    benc = case.topo["bound_enc"]
    keep = benc >= 0
    keep[case.topo["ib_solid"]] = True
    return keep


def live_rows(case):
    """The f entries of the live cells (cell-major, ndf*n per cell)."""
    keep = live_cells(case)
    off = np.asarray(case.topo["vs_offset"]) * case.config.solver.NDF
    m = np.zeros(int(off[-1]), bool)
    for c in np.flatnonzero(keep):
        m[off[c]:off[c + 1]] = True
    return m


def check_state(case, g, o, tol, what):
    keep = live_cells(case)
    nw = case.config.solver.DIM + 2
    rows = live_rows(case)
    assert rel_err(g["df"][rows], o["df"][rows]) <= tol, f"{what}: df"
    for k in ("w", "prim"):
        assert rel_err_rows(g[k].reshape(-1, nw)[keep], o[k].reshape(-1, nw)[keep], nw) <= tol, \
            f"{what}: {k}"


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_split_phases(name):
    case = CASES[name]()
    say(f"[ib split] {name}: cells={case.n_cells} faces={len(case.topo['ib_face_pair'])}")
    dev, orc = pair(case)
    dt = case.dt_xi
    for s in range(2):
        dev.slope(dt)
        orc.slope()
        dev.flux(dt)
        orc.flux(dt)
        assert rel_err(dev.get_ib(flat=True), orc.get_ib(flat=True)) <= TOL, "SolidNeighbor df"
        gf, gv = dev.get_flux()
        of, ov = orc.get_flux()
        assert rel_err(gv, ov) <= TOL, "vs flux"
        assert rel_err_cols(gf, of, case.config.solver.DIM + 2) <= 1e-11, "macro flux"
        rg = dev.iterate(dt, s)
        ro = orc.iterate(dt, s)
        check_state(case, dev.get_state(), orc.get_state(), TOL, f"step {s}")
        assert rel_err(rg[1], ro[1]) <= 1e-11
    dev.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_fused_steps(name):
    case = CASES[name]()
    say(f"[ib fused] {name}: cells={case.n_cells} P={case.phase_cells}")
    dev, orc = pair(case)
    for s in range(4):
        dev.step(case.dt_xi, s)
        orc.step(case.dt_xi, s)
    check_state(case, dev.get_state(), orc.get_state(), TOL, "4 steps")
    assert rel_err(dev.get_ib(flat=True), orc.get_ib(flat=True)) <= TOL
    dev.close()


@pytest.mark.gpu
def test_graph_replay_equals_launches():
    case = CASES["cyl32_nv16"]()
    say("[ib graphs]")
    a, _ = pair(case)
    b, _ = pair(case, options={"graphs": 1})
    for s in range(5):
        a.step(case.dt_xi, s)
        b.step(case.dt_xi, s)
        ga, gb = a.get_state(), b.get_state()
        for k in ("df", "w", "prim"):
            assert np.array_equal(ga[k], gb[k]), (s, k)
        assert np.array_equal(a.get_ib(), b.get_ib()), s
    a.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["cyl16_vsamr", "cyl16_vsamr_amr1"])
def test_overlap_maps_equal_search(name):
    """Adapted grids: the IB faces read through precomputed overlap maps or through the
    in-kernel key search give bit-identical steps."""
    case = CASES[name]()
    say(f"[ib maps] {name}")
    a, _ = pair(case)
    b, _ = pair(case, options={"overlap_maps": 0})
    for s in range(3):
        a.step(case.dt_xi, s)
        b.step(case.dt_xi, s)
    ga, gb = a.get_state(), b.get_state()
    for k in ("df", "w", "prim"):
        assert np.array_equal(ga[k], gb[k]), k
    assert np.array_equal(a.get_ib(flat=True), b.get_ib(flat=True))
    a.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["cyl16", "cyl32_nv16", "sph12_nv8", "cyl16_vsamr"])
def test_wall_carries_no_mass(name):
    case = CASES[name]()
    say(f"[ib wall] {name}")
    dev = abi.HipContext(kah_params(case.config))
    dev.set_topology(case.topo)
    dev.set_state(case.w, case.prim, case.df)
    for s in range(3):
        dev.step(case.dt_xi, s)
    aux = dev.get_ib()
    t = case.topo
    dim = case.config.solver.DIM
    nr = t["ib_normal"].reshape(-1, dim)
    off = np.asarray(t["vs_offset"])
    uni = int(t.get("vs_uniform", 0))
    for k, p in enumerate(t["ib_face_pair"]):
        h = int(t["pair_here"][p])
        g = 0 if uni else int(off[h])
        n = int(off[h + 1] - off[h])
        vm = np.asarray(t["vs_mid"])[g * dim:(g + n) * dim].reshape(dim, n).T
        wt = np.asarray(t["vs_weight"])[g:g + n]
        flux = wt * (vm @ nr[k]) * aux[k][0]
        assert abs(flux.sum()) < 1e-13 * np.abs(flux).sum(), k
    dev.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["cyl16", "cyl32_nv16", "sph12_nv8", "cyl16_vsamr", "cyl16_dvm",
                                  "sph8_vsamr_dvm", "cyl16_vsamr_cip"])
def test_fp32_form(name):
    """FP32 form (f stored and advected in FP32) against the FP64 oracle after 3 steps:
    moments within 1e-5 of each cell's scale, the SolidNeighbor df within 1e-5."""
    case = CASES[name]()
    say(f"[ib fp32f] {name}")
    dev, orc = pair(case, precision=abi.PREC["fp32f"])
    for s in range(3):
        dev.step(case.dt_xi, s)
        orc.step(case.dt_xi, s)
    g, o = dev.get_state(), orc.get_state()
    keep = live_cells(case)
    nw = case.config.solver.DIM + 2
    for k in ("w", "prim"):
        assert rel_err_rows(g[k].reshape(-1, nw)[keep], o[k].reshape(-1, nw)[keep], nw) <= 1e-5, k
    assert rel_err(dev.get_ib(flat=True), orc.get_ib(flat=True)) <= 1e-5
    dev.close()


@pytest.mark.gpu
def test_rejections():
    say("[ib rejects]")
    case = CASES["cyl16"]()
    p = kah_params(case.config)
    dev = abi.HipContext(p)
    with pytest.raises(abi.KahError, match="no immersed-boundary faces"):
        dev.get_ib()
    t = dict(case.topo)
    img = np.array(t["ib_image_cell"], copy=True)
    img[0] = int(t["ib_solid"][0])
    t["ib_image_cell"] = img
    with pytest.raises(abi.KahError, match="is not fluid"):
        dev.set_topology(t)
    dev.set_topology(case.topo)
    dev.set_state(case.w, case.prim, case.df)
    with pytest.raises(abi.KahError, match="no immersed-boundary faces"):
        dev.get_ib()
    dev.close()


@pytest.mark.gpu
def test_velocity_adaptation_with_immersed_boundary():
    """vs_adaptive_mesh_refinement! on a cylinder with adapted grids: the SolidNeighbors and
    cut cells follow the donors' new grids, and the adapted state steps in parity with the
    oracle built on the new topology."""
    from kitamr_hip import ib, vsadapt
    case = CASES["cyl16_vsamr"]()
    say("[ib vs adapt]")
    params = kah_params(case.config)
    dev, _ = pair(case)
    for s in range(2):
        dev.step(case.dt_xi, s)
    p = abi.vs_adapt_params("lohner", coeff_local=1e-2)
    topo2, va = vsadapt.vs_adapt(dev, case.topo, p, 2)
    assert va.any()
    cut = ib.cut_lists(topo2, 2)
    for key, arr in zip(("ib_cut_off", "ib_cut_index", "ib_cut_gas", "ib_cut_solid"), cut):
        assert np.array_equal(topo2[key], arr), key
    got = dev.get_state()
    o2 = O.Oracle(params, topo2)
    o2.set_state(got["w"], got["prim"], got["df"])
    case2 = type(case)(case.config, topo2, got["w"], got["prim"], got["df"], case.dt_xi)
    for s in range(2):
        dev.step(case.dt_xi, 2 + s)
        o2.step(case.dt_xi, 2 + s)
    check_state(case2, dev.get_state(), o2.get_state(), TOL, "after adaptation")
    assert rel_err(dev.get_ib(flat=True), o2.get_ib(flat=True)) <= TOL
    dev.close()
    o2.close()


@pytest.mark.gpu
def test_balance_with_immersed_boundary():
    """vs_balance! next to the immersed circle (grids of three levels that break the balance
    as built): the rounds skip solid and inside-solid cells, whose grids stay as they were;
    the cut cells follow the donors' refined grids; and the balanced state steps in parity
    with the oracle built on the new topology."""
    from kitamr_hip import ib, vsadapt
    case = cases.cylinder_2d(n=16, nv=8, vs_levels=3)
    say("[ib balance]")
    params = kah_params(case.config)
    dev, _ = pair(case)
    for s in range(2):
        dev.step(case.dt_xi, s)
    topo2, rounds = vsadapt.balance_rounds(dev, case.topo, 2)
    assert rounds >= 1
    g0, g1 = vsadapt.cell_grids(case.topo, 2), vsadapt.cell_grids(topo2, 2)
    benc = np.asarray(case.topo["bound_enc"])
    for c in np.flatnonzero(benc < 0):
        assert np.array_equal(g0[c][0], g1[c][0]) and np.array_equal(g0[c][1], g1[c][1]), c
    assert any(len(g1[c][0]) > len(g0[c][0]) for c in np.flatnonzero(benc >= 0))
    cut = ib.cut_lists(topo2, 2)
    for key, arr in zip(("ib_cut_off", "ib_cut_index", "ib_cut_gas", "ib_cut_solid"), cut):
        assert np.array_equal(topo2[key], arr), key
    got = dev.get_state()
    o2 = O.Oracle(params, topo2)
    o2.set_state(got["w"], got["prim"], got["df"])
    case2 = type(case)(case.config, topo2, got["w"], got["prim"], got["df"], case.dt_xi)
    for s in range(2):
        dev.step(case.dt_xi, 2 + s)
        o2.step(case.dt_xi, 2 + s)
    check_state(case2, dev.get_state(), o2.get_state(), TOL, "after balance")
    assert rel_err(dev.get_ib(flat=True), o2.get_ib(flat=True)) <= TOL
    dev.close()
    o2.close()
