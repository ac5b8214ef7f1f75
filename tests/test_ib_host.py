"""Immersed boundary on the host: the circle geometry (kitamr_hip.ib) and the oracle's
update_solid_cell! / update_solid_neighbor! restatement (src/Boundary/Immersed_boundary.jl),
pinned without the reference's outputs:

- cut cells: the two parts fill the velocity cell, a line (plane) through the cell's centre
  halves it, a line through a corner cuts off its closed-form triangle, a plane cuts off the
  corner tetrahedron of volume l^3/6, a normal with one zero component gives the prism over
  the 2D cut, and the solid part is the one on the side u.n > 0 (the side the wall
  Maxwellian fills);
- roles and faces: no face joins two solid cells, every solid/fluid face is owned by the
  fluid cell with its rot and area sign flipped, every donor has a solid face neighbour;
- calc_intersect (circle and sphere): the point lies on the boundary between donor and solid
  cell, the normal is the unit outward radius;
- update_solid_cell! / image_df: for f linear in x with exact slopes the direction-weighted
  extrapolation is exact (weights sum to one; a wrong dx sign or midpoint breaks it), so the
  solid cell holds f(x_solid) and the gas side of every SolidNeighbor holds f(aux point)
  (boundary_slope! then equals the exact gradient, and dxL carries f from the image point
  to the wall);
- the DVM flux next to the wall: on the same linear fields every fluid face carries the
  exact value a + b.(x_face - u dt), and an immersed face's incoming rows the
  SolidNeighbor's f plus its boundary slope times (x_face - u dt - x_solid)[dir];
- the wall density: after the Maxwellian fill and the cut-cell mix the SolidNeighbor carries
  no net mass through the wall, sum_i weight_i (u_i.n) f_i = 0, and the filled cells hold
  rho_w times the unit-density wall Maxwellian.
"""
import math

import numpy as np
import pytest

import oracle as O
from kitamr_hip import cases, ib
from kitamr_hip.solver import discrete_maxwell, kah_params


def test_cut_areas_fill_the_cell_and_halve_on_the_diagonal():
    n = np.array([1.0, 1.0]) / math.sqrt(2.0)
    mids = np.array([[0.0, 0.0], [0.5, 0.5], [-0.5, 0.5], [3.0, 3.0]])
    idx, gas, sol = ib.cut_cells_2d(n, mids, np.zeros(4, np.int8), np.array([1.0, 1.0]))
    # [0.5, 0.5]: the line u.n = 0 passes through its corner (0, 0) only; [3, 3] is far away
    assert idx.tolist() == [0, 2]
    assert np.allclose(gas + sol, 1.0, rtol=0, atol=1e-15)
    assert gas[0] == pytest.approx(0.5, abs=1e-15)       # through the centre
    assert gas[1] == pytest.approx(0.5, abs=1e-15)       # [-0.5, 0.5]: the diagonal of the cell
    # a steeper line through the corner (0, 0) of the cell [0, 1] x [0, 1]
    n2 = np.array([2.0, -1.0]) / math.sqrt(5.0)          # u.n = 0 on y = 2x
    idx, gas, sol = ib.cut_cells_2d(n2, np.array([[0.5, 0.5]]), np.zeros(1, np.int8),
                                    np.array([1.0, 1.0]))
    # gas side 2x - y < 0: the triangle above y = 2x within the square, area 1/4
    assert idx.tolist() == [0]
    assert gas[0] == pytest.approx(0.25, abs=1e-15) and sol[0] == pytest.approx(0.75, abs=1e-15)


def test_cut_solid_part_is_on_the_normal_side():
    n = np.array([0.6, 0.8])
    # a cell straddling the line, shifted toward +n: its solid (u.n > 0) part is the larger
    m = np.array([[0.1 * 0.6, 0.1 * 0.8]])
    idx, gas, sol = ib.cut_cells_2d(n, m, np.zeros(1, np.int8), np.array([1.0, 1.0]))
    assert idx.tolist() == [0] and sol[0] > gas[0]
    idx, gas, sol = ib.cut_cells_2d(n, -m, np.zeros(1, np.int8), np.array([1.0, 1.0]))
    assert sol[0] < gas[0]
    # axis-aligned normals: no cut cells (the line runs along cell edges)
    idx, _, _ = ib.cut_cells_2d(np.array([0.0, 1.0]), m, np.zeros(1, np.int8), np.ones(2))
    assert len(idx) == 0


def test_cut_areas_of_a_refined_cell_scale_with_its_level():
    n = np.array([1.0, 2.0]) / math.sqrt(5.0)
    m = np.array([[0.05, -0.02]])
    for lev in range(3):
        idx, g, s = ib.cut_cells_2d(n, m, np.full(1, lev, np.int8), np.array([0.4, 0.4]))
        assert idx.tolist() == [0]
        assert g[0] + s[0] == pytest.approx((0.4 / 2 ** lev) ** 2, rel=1e-14)


def test_cut_volumes_of_cubes():
    lvl = np.zeros(1, np.int8)
    # a plane through the centre halves the cube, for any normal
    for n in ([1.0, 1.0, 1.0], [1.0, -2.0, 0.5], [0.3, 0.0, 1.0]):
        n = np.asarray(n) / np.linalg.norm(n)
        idx, g, s_ = ib.cut_cells_3d(n, np.zeros((1, 3)), lvl, np.ones(3))
        assert idx.tolist() == [0] and g[0] == pytest.approx(0.5, abs=1e-15)
    # the corner tetrahedron x + y + z < 0 of the cube [-0.25, 0.75]^3: legs 0.75
    n = np.ones(3) / math.sqrt(3.0)
    idx, g, s_ = ib.cut_cells_3d(n, np.full((1, 3), 0.25), lvl, np.ones(3))
    assert g[0] == pytest.approx(0.75 ** 3 / 6.0, rel=1e-14)
    assert s_[0] == pytest.approx(1.0 - 0.75 ** 3 / 6.0, rel=1e-14)
    # the solid part lies on u.n > 0
    idx, g, s_ = ib.cut_cells_3d(-n, np.full((1, 3), 0.25), lvl, np.ones(3))
    assert s_[0] == pytest.approx(0.75 ** 3 / 6.0, rel=1e-14)
    # one zero component: the cube is a prism over the 2D cut, volume = extent x area
    n3 = np.array([0.6, 0.8, 0.0])
    m = np.array([[0.1, -0.05, 0.3]])
    idx, g, s_ = ib.cut_cells_3d(n3, m, lvl, np.array([1.0, 1.0, 0.5]))
    i2, g2, s2 = ib.cut_cells_2d(n3[:2], m[:, :2], lvl, np.array([1.0, 1.0]))
    assert g[0] == pytest.approx(0.5 * g2[0], rel=1e-14)
    # touching a corner only, and normals along an axis: no cut cell
    assert len(ib.cut_cells_3d(n, np.full((1, 3), 0.5), lvl, np.ones(3))[0]) == 0
    assert len(ib.cut_cells_3d(np.array([0.0, 0.0, 1.0]), m, lvl, np.ones(3))[0]) == 0


def test_sphere_roles_and_intersections():
    case = cases.sphere_3d(n=8, nv=6, L=2.0)
    t = case.topo
    n = case.n_cells
    mid = t["cell_mid"].reshape(-1, 3)[:n]
    inside = np.linalg.norm(mid, axis=1) < 1.0
    assert np.array_equal(t["bound_enc"] < 0, inside)
    ap = t["ib_aux_point"].reshape(-1, 3)
    nr = t["ib_normal"].reshape(-1, 3)
    assert np.allclose(np.linalg.norm(ap, axis=1), 1.0, atol=1e-14)
    assert np.allclose(nr, ap, atol=1e-15)
    ph, pt = t["pair_here"], t["pair_there"]
    for k, p in enumerate(t["ib_face_pair"]):
        d = int(t["pair_dir"][p])
        # the intersection lies between donor and solid along the face direction
        lo, hi = sorted((mid[ph[p], d], mid[pt[p], d]))
        assert lo <= ap[k, d] <= hi
        other = [e for e in range(3) if e != d]
        assert np.array_equal(ap[k, other], mid[ph[p], other])


@pytest.mark.parametrize("ps_levels", [0, 1])
def test_roles_faces_and_intersections(ps_levels):
    case = cases.cylinder_2d(n=16, nv=8, ps_levels=ps_levels)
    t = case.topo
    n = case.n_cells
    benc = t["bound_enc"]
    mid = t["cell_mid"].reshape(-1, 2)[:n]
    inside = np.hypot(mid[:, 0], mid[:, 1]) < 1.0
    assert np.array_equal(benc < 0, inside)
    ph, pt = t["pair_here"], t["pair_there"]
    assert np.all(benc[ph] >= 0), "a pair is owned by a solid cell"
    real = pt >= 0
    assert not np.any(inside[ph[real]] & inside[pt[real]])
    # the faces towards solid cells are exactly the IB faces, one per donor/solid neighbour
    sf = np.flatnonzero(real & (benc[np.where(real, pt, 0)] < 0))
    assert sorted(sf.tolist()) == sorted(t["ib_face_pair"].tolist())
    assert set(ph[sf].tolist()) == set(np.flatnonzero(benc > 0).tolist())
    # flipped faces keep the geometric face: midpoint between here and there along dir
    for p in sf:
        d = int(t["pair_dir"][p])
        xm = t["pair_mid"].reshape(-1, 2)[p, d]
        xh, xt = mid[ph[p], d], mid[pt[p], d]
        assert xm == pytest.approx(0.5 * (xh + xt), abs=1e-14)
        # rot = +1 on the here cell's low face (the solid lies below), -1 on its high face
        assert t["pair_rot"][p] == (1.0 if xt < xh else -1.0)
        assert np.sign(t["pair_area"][p]) == t["pair_rot"][p]
    ap = t["ib_aux_point"].reshape(-1, 2)
    nr = t["ib_normal"].reshape(-1, 2)
    assert np.allclose(np.hypot(ap[:, 0], ap[:, 1]), 1.0, atol=1e-14)
    assert np.allclose(nr, ap, atol=1e-15)
    # the image stencil ends with the donor, the solid cells' neighbours are fluid
    ioff = t["ib_image_off"]
    for k, p in enumerate(t["ib_face_pair"]):
        assert t["ib_image_cell"][ioff[k + 1] - 1] == ph[p]
        assert np.all(benc[t["ib_image_cell"][ioff[k]:ioff[k + 1]]] >= 0)
    assert np.all(benc[t["ib_solid_nb"]] >= 0)
    assert np.all(benc[t["ib_solid"]] < 0)


def test_geometry_rejects_unsupported_meshes():
    case = cases.riemann_2d(n=8, nv=8)
    with pytest.raises(ValueError, match="domain boundary"):
        ib.immerse_circle(case.topo, (0.0, 0.0), 0.3)


def test_cut_cells_follow_each_donors_grid():
    """Adapted grids: every IB face's cut cells index its donor's grid, and their two parts
    add up to the weight of that donor's velocity cell (the Newton-Cotes weight is the cell
    area); refresh_cuts after the grids change gives the cuts of the new grids."""
    case = cases.cylinder_2d(n=16, nv=8, vs_levels=2)
    t = case.topo
    off = np.asarray(t["vs_offset"])
    here = t["pair_here"][t["ib_face_pair"]]
    co = t["ib_cut_off"]
    for k, h in enumerate(here):
        idx = t["ib_cut_index"][co[k]:co[k + 1]]
        assert len(idx) and idx.max() < off[h + 1] - off[h]
        wt = np.asarray(t["vs_weight"])[off[h] + idx]
        tot = t["ib_cut_gas"][co[k]:co[k + 1]] + t["ib_cut_solid"][co[k]:co[k + 1]]
        assert np.allclose(tot, wt, rtol=1e-13, atol=0)
    same = ib.refresh_cuts(t)
    for key in ("ib_cut_off", "ib_cut_index", "ib_cut_gas", "ib_cut_solid"):
        assert np.array_equal(same[key], t[key])


def _linear_ib_state(case, seed=3):
    """f(x, u) = a(u) + b(u).x on every cell (a > 0, small b), with the exact slopes b."""
    rng = np.random.default_rng(seed)
    t = case.topo
    n = case.n_cells
    dim, ndf = case.config.solver.DIM, case.config.solver.NDF
    n0 = int(t["vs_offset"][1])
    mid = t["cell_mid"].reshape(-1, dim)[:n]
    a = rng.uniform(0.5, 1.5, (ndf, n0))
    g = rng.uniform(-0.02, 0.02, (dim, ndf, n0))          # [d, k, v]
    df = a[None] + np.einsum("cd,dkv->ckv", mid, g)
    sdf = np.broadcast_to(g[None], (n, dim, ndf, n0))   # [c][d][k][v]
    return df, a, g, np.ascontiguousarray(sdf).reshape(-1)


IB_CASES = {
    "cylinder": lambda: cases.cylinder_2d(n=16, nv=8),
    "sphere": lambda: cases.sphere_3d(n=8, nv=6, L=2.0),
}


@pytest.mark.parametrize("name", list(IB_CASES))
def test_oracle_ib_is_exact_on_linear_fields(name):
    case = IB_CASES[name]()
    t = case.topo
    n = case.n_cells
    dim, ndf = case.config.solver.DIM, case.config.solver.NDF
    df, a, g, sdf = _linear_ib_state(case)
    o = O.Oracle(kah_params(case.config), t)
    o.set_state(case.w, case.prim, df.reshape(-1))
    o.set_sdf(sdf)
    o.flux(case.dt_xi)
    st = o.get_state()
    got = st["df"].reshape(n, ndf, -1)
    mid = t["cell_mid"].reshape(-1, dim)[:n]
    for s in t["ib_solid"]:
        exact = a + np.einsum("d,dkv->kv", mid[s], g)
        assert np.max(np.abs(got[s] - exact)) < 1e-14
    aux = o.get_ib()
    vm = np.asarray(t["vs_mid"]).reshape(dim, -1).T
    ap = t["ib_aux_point"].reshape(-1, dim)
    nr = t["ib_normal"].reshape(-1, dim)
    coff = t["ib_cut_off"]
    checked = 0
    for k in range(len(t["ib_face_pair"])):
        vn = vm @ nr[k]
        cut = np.zeros(len(vn), bool)
        cut[t["ib_cut_index"][coff[k]:coff[k + 1]]] = True
        gas = (vn < 0) & ~cut
        exact = a + np.einsum("d,dkv->kv", ap[k], g)
        assert np.max(np.abs(aux[k][:, gas] - exact[:, gas])) < 1e-13
        checked += gas.sum()
    assert checked > 0


@pytest.mark.parametrize("name", ["cylinder", "sphere"])
def test_oracle_wall_carries_no_mass_and_fills_a_maxwellian(name):
    case = (cases.cylinder_2d(n=16, nv=12) if name == "cylinder"
            else cases.sphere_3d(n=8, nv=6, L=2.0))
    t = case.topo
    dim = case.config.solver.DIM
    o = O.Oracle(kah_params(case.config), t)
    o.set_state(case.w, case.prim, case.df)
    for s in range(3):
        o.slope()
        o.flux(case.dt_xi)
        o.iterate(case.dt_xi, s)
    aux = o.get_ib()
    vm = np.asarray(t["vs_mid"]).reshape(dim, -1).T
    wt = np.asarray(t["vs_weight"])
    nr = t["ib_normal"].reshape(-1, dim)
    coff = t["ib_cut_off"]
    bc = t["ib_bc"].reshape(-1, dim + 2)
    ncut = 0
    for k in range(len(t["ib_face_pair"])):
        vn = vm @ nr[k]
        flux = wt * vn * aux[k, 0]
        assert abs(flux.sum()) < 1e-13 * np.abs(flux).sum()
        cut = np.zeros(len(vn), bool)
        cut[t["ib_cut_index"][coff[k]:coff[k + 1]]] = True
        ncut += cut.sum()
        fill = (vn >= 0) & ~cut
        pw = bc[k].copy()
        pw[0] = 1.0
        M = discrete_maxwell(vm, pw, case.config.gas.K)     # [k, v]
        ratio = aux[k][:, fill] / M[:, fill]
        # one density rho_w per face (negative where the extrapolated gas side carries mass
        # toward the wall out of an expansion, as the reference's formula allows)
        assert np.ptp(ratio) < 1e-12 * np.abs(ratio).max()
    assert ncut > 0


def _brute_extrapolate(case, sdf, cells, x, tgt):
    """update_solid_cell!'s extrapolation by geometric overlap of velocity cells: for each
    target cell, the source cells whose boxes overlap it (a finer source counts 2^-D per level,
    a coarser one is injected)."""
    t = case.topo
    dim, ndf = case.config.solver.DIM, case.config.solver.NDF
    off = np.asarray(t["vs_offset"])
    du = ib._root_du(t, dim)
    mid = t["cell_mid"].reshape(-1, dim)
    lt, mt = ib._cell_grid(t, tgt, dim)
    n = len(lt)
    u = mt
    nu = np.sqrt((u * u).sum(axis=1))
    w = []
    for c in cells:
        l = x - mid[c]
        l = l / np.sqrt((l * l).sum())
        r = (u @ l) / nu
        w.append(np.maximum(0.0, r) ** 2)
    w = np.array(w)
    ws = w.sum(axis=0)
    out = np.zeros((ndf, n))
    for j, c in enumerate(cells):
        wn = np.where(ws == 0.0, 1.0 / len(cells), w[j] / np.where(ws == 0.0, 1.0, ws))
        ls, ms = ib._cell_grid(t, c, dim)
        ns = len(ls)
        f = case.df_now[off[c] * ndf:off[c + 1] * ndf].reshape(ndf, ns)
        s = sdf[off[c] * ndf * dim:off[c + 1] * ndf * dim].reshape(dim, ndf, ns)
        val = f + np.einsum("d,dkv->kv", x - mid[c], s)
        hs = du[None, :] / 2.0 ** ls[:, None].astype(float)
        for i in range(n):
            hi = du / 2.0 ** int(lt[i])
            ov = np.all(np.abs(ms - mt[i][None, :]) < 0.5 * (hs + hi[None, :]) - 1e-12, axis=1)
            acc = np.zeros(ndf)
            for jj in np.flatnonzero(ov):
                fac = 2.0 ** (-dim * (int(ls[jj]) - int(lt[i]))) if ls[jj] > lt[i] else 1.0
                acc += val[:, jj] * fac
            out[:, i] += acc * wn[i]
    return out


def test_oracle_solid_cells_on_adapted_grids_match_overlap():
    case = cases.cylinder_2d(n=16, nv=8, vs_levels=2)
    t = case.topo
    dim, ndf = 2, 2
    rng = np.random.default_rng(11)
    P = int(t["vs_offset"][-1])
    sdf = rng.uniform(-0.05, 0.05, P * ndf * dim)
    o = O.Oracle(kah_params(case.config), t)
    o.set_state(case.w, case.prim, case.df)
    o.set_sdf(sdf)
    o.flux(case.dt_xi)
    got = o.get_state()["df"]
    case.df_now = np.asarray(case.df)
    off = np.asarray(t["vs_offset"])
    mid = t["cell_mid"].reshape(-1, dim)
    so = t["ib_solid_nb_off"]
    for q, s in enumerate(t["ib_solid"]):
        nbs = t["ib_solid_nb"][so[q]:so[q + 1]]
        exp = _brute_extrapolate(case, sdf, nbs, mid[s], s)
        g = got[off[s] * ndf:off[s + 1] * ndf].reshape(ndf, -1)
        assert np.max(np.abs(g - exp)) <= 1e-13 * np.max(np.abs(exp)), s
    # and the wall still carries no mass on the donors' adapted grids
    aux = o.get_ib()
    nr = t["ib_normal"].reshape(-1, dim)
    for k, p in enumerate(t["ib_face_pair"]):
        h = t["pair_here"][p]
        lev, mv = ib._cell_grid(t, h, dim)
        wt = np.asarray(t["vs_weight"])[off[h]:off[h + 1]]
        fl = wt * (mv @ nr[k]) * aux[k][0]
        assert abs(fl.sum()) < 1e-13 * np.abs(fl).sum()


def test_oracle_ib_positivity_closed_form():
    """positivity_preserving_ib! (CIP_Marching) on f = M + b.x with exact slopes b: the
    boundary slope is b[dir], so the perturbance plus the slope term telescope to
    b[dir] (x_face - aux - u dt)[dir] per velocity cell; its moments times dt/V are added to
    each donor's w (theta = 1 for this small perturbation), and nothing else changes w."""
    case = cases.cylinder_2d(n=16, nv=8, marching="CIP_Marching")
    t = case.topo
    n, dim, ndf = case.n_cells, 2, 2
    n0 = int(t["vs_offset"][1])
    mid = t["cell_mid"].reshape(-1, dim)[:n]
    rng = np.random.default_rng(5)
    M = np.asarray(case.df).reshape(n, ndf, n0)[0]
    g = 1e-3 * M[None] * rng.uniform(-1.0, 1.0, (dim, ndf, n0))     # [d, k, v]
    df = M[None] + np.einsum("cd,dkv->ckv", mid, g)
    sdf = np.ascontiguousarray(np.broadcast_to(g[None], (n, dim, ndf, n0))).reshape(-1)
    o = O.Oracle(kah_params(case.config), t)
    o.set_state(case.w, case.prim, df.reshape(-1))
    o.set_sdf(sdf)
    dt = case.dt_xi
    o.flux(dt)
    F, _ = o.get_flux()
    w0 = np.asarray(case.w).reshape(n, 4)
    o.iterate(dt, 0)
    w1 = o.get_state()["w"].reshape(n, 4)
    ds = t["cell_ds"].reshape(-1, dim)[:n]
    V = ds[:, 0] * ds[:, 1]
    expect = w0 + F.reshape(n, 4) * dt / V[:, None]
    vm = np.asarray(t["vs_mid"]).reshape(dim, -1).T
    wt = np.asarray(t["vs_weight"])
    ap = t["ib_aux_point"].reshape(-1, dim)
    fm = t["pair_mid"].reshape(-1, dim)
    checked = set()
    for k, p in enumerate(t["ib_face_pair"]):
        c = int(t["pair_here"][p])
        d = int(t["pair_dir"][p])
        rot, area = t["pair_rot"][p], t["pair_area"][p]
        up = rot * vm[:, d] > 0
        mi = np.where(up[None], g[d] * (fm[p, d] - ap[k, d] - vm[:, d] * dt) * vm[:, d] * area, 0.0)
        we = np.array([(wt * mi[0]).sum(), (wt * vm[:, 0] * mi[0]).sum(),
                       (wt * vm[:, 1] * mi[0]).sum(),
                       0.5 * (wt * ((vm ** 2).sum(axis=1) * mi[0] + mi[1])).sum()])
        expect[c] += we * (dt / V[c])
        checked.add(c)
    fluid = t["bound_enc"] >= 0
    assert len(checked) > 0
    err = np.abs(w1[fluid] - expect[fluid]).max(axis=1) / np.abs(expect[fluid]).max(axis=1)
    assert err.max() < 1e-13
    # the IB term is visible: without it the donors would be off by more than round-off
    donors = np.array(sorted(checked))
    base = w0[donors] + F.reshape(n, 4)[donors] * dt / V[donors, None]
    assert np.abs(w1[donors] - base).max() > 1e3 * 1e-13 * np.abs(base).max()


@pytest.mark.parametrize("name", ["cylinder", "sphere"])
def test_oracle_dvm_flux_is_exact_on_linear_fields(name):
    """The DVM flux on f = a + b.x with exact slopes b: across a fluid face both sides
    reconstruct to a + b.(x_face - u dt), whichever is upwind, so every fluid cell away from
    the domain walls receives sum(+-area u_dir (a + b.(x_face - u dt))). At an immersed face
    the incoming rows (rot u_dir > 0) take the SolidNeighbor's f plus its boundary slope
    (b[dir] on a linear field) times (x_face - u dt - x_solid)[dir], and the solid cell
    receives nothing."""
    case = (cases.cylinder_2d(n=16, nv=8, flux="DVM") if name == "cylinder"
            else cases.sphere_3d(n=8, nv=6, L=2.0, flux="DVM"))
    t = case.topo
    n = case.n_cells
    dim, ndf = case.config.solver.DIM, case.config.solver.NDF
    n0 = int(t["vs_offset"][1])
    df, a, g, sdf = _linear_ib_state(case)
    o = O.Oracle(kah_params(case.config), t)
    o.set_state(case.w, case.prim, df.reshape(-1))
    o.set_sdf(sdf)
    dt = case.dt_xi
    o.flux(dt)
    _, vf = o.get_flux()
    vf = vf.reshape(n, ndf, n0)
    aux = o.get_ib()
    vm = np.asarray(t["vs_mid"]).reshape(dim, -1).T
    mid = t["cell_mid"].reshape(-1, dim)[:n]
    fmid = t["pair_mid"].reshape(-1, dim)
    ibk = {int(p): k for k, p in enumerate(t["ib_face_pair"])}
    expect = np.zeros((n, ndf, n0))
    plain = np.zeros((n, ndf, n0))   # the same without the wall side's slope term
    walled = np.zeros(n, bool)
    n_inc = 0
    for p in range(int(t["n_pairs"])):
        h, th = int(t["pair_here"][p]), int(t["pair_there"][p])
        if th < 0:
            walled[h] = True
            continue
        d, area = int(t["pair_dir"][p]), float(t["pair_area"][p])
        ff = a + np.einsum("vd,dkv->kv", fmid[p][None, :] - vm * dt, g)
        if p in ibk:
            k = ibk[p]
            inc = t["pair_rot"][p] * vm[:, d] > 0
            wall = aux[k] + g[d] * (fmid[p, d] - vm[:, d] * dt - mid[th, d])
            ff = np.where(inc[None], wall, ff)
            n_inc += int(inc.sum())
            expect[h] += area * vm[:, d] * ff
            plain[h] += area * vm[:, d] * np.where(inc[None], aux[k], ff)
            continue
        expect[h] += area * vm[:, d] * ff
        expect[th] -= area * vm[:, d] * ff
        plain[h] += area * vm[:, d] * ff
        plain[th] -= area * vm[:, d] * ff
    fluid = (t["bound_enc"][:n] >= 0) & ~walled
    donors = np.zeros(n, bool)
    donors[t["pair_here"][t["ib_face_pair"]]] = True
    assert n_inc > 0 and (fluid & donors).sum() > 0
    scale = np.abs(expect[fluid]).max()
    assert np.abs(vf[fluid] - expect[fluid]).max() < 1e-13 * scale
    # the wall side's slope term is visible: the plain SolidNeighbor f (the CAIDVM there
    # side) misses the donors' flux by far more than round-off
    sel = fluid & donors
    assert np.abs(vf[sel] - plain[sel]).max() > 1e4 * 1e-13 * scale
