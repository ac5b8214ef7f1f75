"""Host side of velocity-space adaptation (kitamr_hip.vsadapt): the grid streams of
src/Velocity_space/Rebuild.jl on levels, midpoints and weights, and the sources that let the
device move f. Pinned by: refine equals VelocityGrid.refine (an independent loop form of
refine_grid_stream!); coarsening a refined grid with every cell willing restores it exactly;
a group is only merged when all 2^dim siblings are willing and aligned; mass and energy
moments of a remapped f are preserved by refine (injection keeps sum(weight*f)) and by coarsen
(the mean times 2^dim weight); the oracle's criteria agree with a brute-force restatement
whose face neighbours come from a geometric search; the vs_balance! rounds leave balanced,
forced refinements of the original grids (checked on velocity-cell boxes)."""
import numpy as np
import pytest

import oracle as O
from kitamr_hip import abi, cases, vsadapt
from kitamr_hip.solver import kah_params
from kitamr_hip.vspace import VelocityGrid


def _grid(dim=2, n=4, maxlevel=2, seed=0):
    g = VelocityGrid(dim, [-4.0, 4.0] * dim, [n] * dim, maxlevel=maxlevel)
    rng = np.random.default_rng(seed)
    for _ in range(maxlevel):
        g.refine(rng.random(g.n) < 0.4)
    return g


@pytest.mark.parametrize("dim", [2, 3])
def test_refine_stream_equals_grid_refine(dim):
    g = _grid(dim, n=4, maxlevel=2)
    rng = np.random.default_rng(5)
    flags = (rng.random(g.n) < 0.3) & (g.level < 2)
    ch, lv, md, wt, src = vsadapt.refine_stream(g.level, g.midpoint, g.weight, flags, g.ds, dim)
    h = g.copy().refine(flags)
    assert ch
    assert np.array_equal(lv, h.level) and np.array_equal(md, h.midpoint)
    assert np.array_equal(wt, h.weight)
    # injection: every child takes its parent's value
    f = rng.random(g.n)
    assert np.array_equal(f[src][np.repeat(~flags, np.where(flags, 1 << dim, 1))],
                          f[~flags])


@pytest.mark.parametrize("dim", [2, 3])
def test_coarsen_all_willing_restores_the_refined_parent_grid(dim):
    base = VelocityGrid(dim, [-4.0, 4.0] * dim, [4] * dim, maxlevel=1)
    flags = np.zeros(base.n, bool)
    flags[::3] = True
    fine = base.copy().refine(flags)
    ch, lv, md, wt, src = vsadapt.coarsen_stream(fine.level, fine.midpoint, fine.weight,
                                                 np.ones(fine.n, bool), fine.ds, 1, dim)
    assert ch
    # every level-1 group merges: the result is the root grid
    root = VelocityGrid(dim, [-4.0, 4.0] * dim, [4] * dim, maxlevel=1)
    assert np.array_equal(lv, root.level)
    assert np.array_equal(md, root.midpoint)
    assert np.allclose(wt, root.weight, rtol=0, atol=0)
    # mass of f is kept by the mean: sum(w_new*mean) = sum over children
    f = np.random.default_rng(1).random(fine.n)
    fn = np.array([f[s] if s >= 0 else f[-s - 1:-s - 1 + (1 << dim)].sum() / (1 << dim)
                   for s in src])
    assert np.isclose((fn * wt).sum(), (f * fine.weight).sum(), rtol=1e-14)


def test_coarsen_needs_every_sibling():
    g = VelocityGrid(2, [-4.0, 4.0] * 2, [2, 2], maxlevel=1)
    fine = g.copy().refine(np.ones(g.n, bool))          # 16 cells, 4 groups
    ok = np.ones(fine.n, bool)
    ok[5] = False                                        # second group not willing
    ch, lv, md, wt, src = vsadapt.coarsen_stream(fine.level, fine.midpoint, fine.weight, ok,
                                                 fine.ds, 1, 2)
    assert ch and len(lv) == 3 + 4
    assert list(src[:1]) == [-1] and list(src[1:5]) == [4, 5, 6, 7]


def _brute_flags(orc, case, params, vr):
    """Lohner / contribution flags with face neighbours found geometrically."""
    topo, dim, ndf = case.topo, case.config.solver.DIM, case.config.solver.NDF
    st = orc.get_state()
    off = np.asarray(topo["vs_offset"])
    L = int(topo["vs_maxlevel"])
    hf = np.array([(topo["vs_hi"][d] - topo["vs_lo"][d]) / topo["vs_trees_num"][d]
                   for d in range(dim)]) / 2.0 ** L
    out = np.zeros(off[-1], np.uint8)
    nw = dim + 2
    for c, (lv, md, wt) in enumerate(vsadapt.cell_grids(topo, dim)):
        n = len(lv)
        f = st["df"][off[c] * ndf:off[c + 1] * ndf].reshape(ndf, n)
        s = st["sdf"][off[c] * ndf * dim:off[c + 1] * ndf * dim].reshape(dim, ndf, n)
        w = st["w"][c * nw:(c + 1) * nw]
        U = st["prim"][c * nw + 1:c * nw + 1 + dim]
        ds = np.asarray(topo["cell_ds"]).reshape(-1, dim)[c]
        size = hf[None, :] * (2.0 ** (L - lv.astype(np.int64)))[:, None]
        lo, hi = md - 0.5 * size, md + 0.5 * size
        s1 = np.abs(f[0]).max(initial=0.0)
        s2 = np.abs(f[1]).max(initial=0.0) if ndf == 2 else 0.0

        def locate(p):
            inside = np.all((p[None, :] >= lo) & (p[None, :] < hi), axis=1)
            k = np.flatnonzero(inside)
            return int(k[0]) if len(k) else -1

        def ratio(a, b, cc, dl, dr, sc):
            num = abs(dr * a - (dl + dr) * b + dl * cc)
            den = (dr * abs(a - b) + dl * abs(cc - b) +
                   1e-2 * (dr * abs(a) + (dl + dr) * abs(b) + dl * abs(cc)) + 1e-3 * sc * (dl + dr))
            return num / den if den > 0 else 0.0

        SU = U[0] * U[0]
        for d in range(1, dim):
            SU = SU + U[d] * U[d]
        den = w[-1] - 0.5 * w[0] * SU
        for i in range(n):
            cdf = [f[k, i] + max([abs(s[d, k, i] * ds[d]) for d in range(dim)] + [0.0])
                   for k in range(ndf)]
            S = (U[0] - md[i, 0]) ** 2
            for d in range(1, dim):
                S = S + (U[d] - md[i, d]) ** 2
            c1w = cdf[0] * wt[i]
            if ndf == 2:
                e = 0.5 * (S * cdf[0] + cdf[1]) * wt[i]
                lr = max(abs(e) / den, c1w / w[0])
            else:
                lr = max(abs(0.5 * S * cdf[0] * wt[i]) / den, c1w / w[0])
                e = 0.5 * (S * cdf[0]) * wt[i]
            lc = max(e / den, c1w / w[0])
            P2 = 2.0 ** dim
            loc_r, loc_c = lr > params.coeff_local, lc < params.coeff_local / P2
            cg = params.coeff_global
            glob_r = c1w > cg * vr[0] or e > vr[1] * cg
            glob_c = c1w < cg * vr[0] / P2 and e < vr[1] * cg / P2
            eta = 0.0
            if params.mode == 0:
                for d in range(dim):
                    nb = []
                    for sgn in (-1, 1):
                        p = md[i].copy()
                        p[d] += sgn * (0.5 * size[i, d] + 0.5 * hf[d])
                        nb.append(locate(p))
                    dl = size[i, d] if nb[0] < 0 else 0.5 * (size[i, d] + size[nb[0], d])
                    dr = size[i, d] if nb[1] < 0 else 0.5 * (size[i, d] + size[nb[1], d])
                    for k, sc in ((0, s1), (1, s2))[:ndf]:
                        a = 0.0 if nb[0] < 0 else f[k, nb[0]]
                        b = 0.0 if nb[1] < 0 else f[k, nb[1]]
                        eta = max(eta, ratio(a, f[k, i], b, dl, dr, sc))
                base = eta > params.coeff_lohner or loc_r
                coarse = eta < 0.3 * params.coeff_lohner and loc_c
            else:
                base = loc_r or glob_r
                coarse = loc_c and glob_c
            out[off[c] + i] = (1 if (lv[i] < L and base) else 0) | (2 if coarse else 0)
    return out


@pytest.mark.parametrize("mode", ["lohner", "contribution"])
@pytest.mark.parametrize("dim", [2, 3])
def test_oracle_flags_equal_brute_force(mode, dim):
    case = (cases.riemann_2d(n=4, nv=8, vs_levels=2) if dim == 2 else
            cases.box_3d(n=2, nv=4, vs_levels=1))
    o = O.Oracle(kah_params(case.config), case.topo)
    o.set_state(case.w, case.prim, case.df)
    o.step(case.dt_xi, 0)
    vr = o.vs_resolution()
    # a lower local threshold in contribution mode so that both decisions occur
    p = abi.vs_adapt_params(mode, coeff_local=1e-2 if mode == "lohner" else 1e-3)
    got = o.vs_flags(p, vr)
    exp = _brute_flags(o, case, p, vr)
    assert np.array_equal(got, exp), np.flatnonzero(got != exp)[:10]
    assert (got & 2).any()
    if dim == 2:    # the 3D grid is already at its maximum level where it matters
        assert (got & 1).any()
    o.close()


def test_oracle_resolution_is_the_peak_contribution():
    case = cases.riemann_2d(n=4, nv=8, vs_levels=2)
    o = O.Oracle(kah_params(case.config), case.topo)
    o.set_state(case.w, case.prim, case.df)
    vr = o.vs_resolution()
    st = o.get_state()
    L = int(case.topo["vs_maxlevel"])
    wf = 100.0 / 64 / 4 ** L
    off = np.asarray(case.topo["vs_offset"])
    dmax = max(st["df"][off[c] * 2:off[c + 1] * 2].max() for c in range(case.n_cells))
    assert vr[0] == dmax * wf
    o.close()


def _boxes(topo, c, dim):
    lev, mid, _ = vsadapt.cell_grids(topo, dim)[c]
    h = vsadapt.root_ds(topo, dim)[None, :] / 2.0 ** lev.astype(np.float64)[:, None]
    return lev.astype(np.int64), mid - 0.5 * h, mid + 0.5 * h


def _fluid_pairs(topo):
    """Face pairs of fluid cells; a periodic face's image record is read as its source."""
    benc = np.asarray(topo["bound_enc"])
    n = len(benc)
    img = np.asarray(topo.get("image_src", []), dtype=np.int64)
    src = lambda r: int(r) if r < n else int(img[r - n])
    out = []
    for h, t in zip(topo["pair_here"], topo["pair_there"]):
        if t < 0:
            continue
        h, t = src(h), src(t)
        if benc[h] >= 0 and benc[t] >= 0:
            out.append((h, t))
    return out


def _balance_case(kind, levels):
    if kind == "periodic":
        return cases.periodic_box_2d(n=8, nv=8, vs_levels=levels)
    return cases.riemann_2d(n=8, nv=8, vs_levels=levels)


@pytest.mark.parametrize("kind,levels", [("riemann", 2), ("riemann", 3), ("periodic", 3)])
def test_balance_is_a_forced_two_to_one_refinement(kind, levels):
    """vs_balance! (src/Velocity_space/Balance.jl): pinned by geometry, with overlaps found by
    brute force over velocity-cell boxes. After the rounds (a) no two overlapping velocity
    cells of face-neighbouring fluid cells differ by more than one level; (b) every grid
    refines its original (each original cell is tiled by cells at least as fine, weights
    adding up) and each new cell's source is the original cell holding it (injection);
    (c) every refinement was forced: the parent box of each cell made by balance overlaps a
    neighbour's cell at least two levels finer than that parent."""
    dim = 2
    case = _balance_case(kind, levels)
    if kind == "periodic":   # faces across the periodic boundaries name image records
        assert np.asarray(case.topo["pair_there"]).max() >= case.n_cells
    t0 = case.topo
    t, rounds = t0, 0
    while True:
        t1, src, ch = vsadapt.plan_balance(t, dim)
        if not ch.any():
            break
        off0, off1 = np.asarray(t["vs_offset"]), np.asarray(t1["vs_offset"])
        for c in np.flatnonzero(ch):
            _, lo0, hi0 = _boxes(t, c, dim)
            _, lo1, hi1 = _boxes(t1, c, dim)
            s = src[off1[c]:off1[c + 1]]
            assert np.all(lo1 >= lo0[s] - 1e-12) and np.all(hi1 <= hi0[s] + 1e-12)
        t, rounds = t1, rounds + 1
    assert rounds >= 1
    pairs = _fluid_pairs(t)
    assert pairs
    # (a) balanced
    for h, n in pairs:
        la, loa, hia = _boxes(t, h, dim)
        lb, lob, hib = _boxes(t, n, dim)
        ov = np.all((loa[:, None, :] < hib[None, :, :] - 1e-12) &
                    (lob[None, :, :] < hia[:, None, :] - 1e-12), axis=2)
        ia, ib = np.nonzero(ov)
        assert np.all(np.abs(la[ia] - lb[ib]) <= 1), (h, n)
    # (b) a refinement of the original grids
    for c in range(case.n_cells):
        l0, lo0, hi0 = _boxes(t0, c, dim)
        l1, lo1, hi1 = _boxes(t, c, dim)
        w0 = vsadapt.cell_grids(t0, dim)[c][2]
        w1 = vsadapt.cell_grids(t, dim)[c][2]
        inside = np.all((lo1[:, None, :] >= lo0[None, :, :] - 1e-12) &
                        (hi1[:, None, :] <= hi0[None, :, :] + 1e-12), axis=2)
        assert np.all(inside.sum(axis=1) == 1), c
        home = inside.argmax(axis=1)
        assert np.all(l1 >= l0[home]), c
        assert np.allclose(np.bincount(home, weights=w1, minlength=len(w0)), w0, rtol=1e-13)
    # (c) forced
    nbrs = {}
    for h, n in pairs:
        nbrs.setdefault(h, set()).add(n)
        nbrs.setdefault(n, set()).add(h)
    made = 0
    lo_v = np.asarray(t0["vs_lo"][:dim], dtype=np.float64)
    ds = vsadapt.root_ds(t0, dim)
    for c in range(case.n_cells):
        l0, lo0, hi0 = _boxes(t0, c, dim)
        l1, lo1, hi1 = _boxes(t, c, dim)
        inside = np.all((lo1[:, None, :] >= lo0[None, :, :] - 1e-12) &
                        (hi1[:, None, :] <= hi0[None, :, :] + 1e-12), axis=2)
        new = np.flatnonzero(l1 > l0[inside.argmax(axis=1)])
        for i in new:
            hp = ds / 2.0 ** (l1[i] - 1)
            plo = lo_v + np.floor((lo1[i] - lo_v) / hp + 1e-9) * hp
            phi = plo + hp
            forced = False
            for n in nbrs.get(c, ()):
                lb, lob, hib = _boxes(t, n, dim)
                ov = np.all((plo[None, :] < hib - 1e-12) & (lob < phi[None, :] - 1e-12), axis=1)
                forced |= bool(np.any(lb[ov] >= l1[i] + 1))
            assert forced, (c, i)
            made += 1
    assert made > 0


def _balance_flags_every_pair(topo, dim):
    """The balance flags by the direct loop over every fluid face pair and both directions
    (no grouping by grid): the finest neighbour cell inside each velocity cell."""
    off = np.asarray(topo["vs_offset"])
    L = int(topo["vs_maxlevel"])
    grids = vsadapt.cell_grids(topo, dim)
    keys = [vsadapt.grid_keys(topo, g[0], g[1], dim) for g in grids]
    need = [np.full(len(g[0]), -1, dtype=np.int64) for g in grids]
    for h, t in _fluid_pairs(topo):
        for a, b in ((h, t), (t, h)):
            ka, la = keys[a], grids[a][0].astype(np.int64)
            kb, lb = keys[b], grids[b][0].astype(np.int64)
            for j in range(len(kb)):
                i = np.searchsorted(ka, kb[j], side="right") - 1
                if i >= 0 and kb[j] < ka[i] + (1 << (dim * (L - la[i]))):
                    need[a][i] = max(need[a][i], lb[j])
    flags = np.zeros(int(off[-1]), dtype=np.uint8)
    benc = np.asarray(topo["bound_enc"])
    for c, g in enumerate(grids):
        if benc[c] >= 0:
            flags[off[c]:off[c + 1]] = need[c] > g[0].astype(np.int64) + 1
    return flags


@pytest.mark.parametrize("make", [
    lambda: cases.riemann_2d(n=8, nv=8, vs_levels=3),
    lambda: cases.periodic_box_2d(n=8, nv=8, vs_levels=3),
    lambda: cases.cylinder_2d(n=16, nv=8, vs_levels=2),
])
def test_balance_flags_grouped_by_grid_equal_every_pair(make):
    """The balance planning computes each distinct ordered pair of grids once and skips faces
    between cells on one grid; its flags equal the direct loop over every face pair."""
    case = make()
    dim = case.config.solver.DIM
    a = vsadapt.balance_flags(case.topo, dim)
    b = _balance_flags_every_pair(case.topo, dim)
    assert np.array_equal(a, b)
    ids, reps = vsadapt.grid_ids(case.topo, dim)
    assert len(reps) <= case.n_cells and ids.max() == len(reps) - 1


def test_balance_rounds_are_planned_before_any_device_call(monkeypatch):
    """vs_balance! plans every round on the host before the first injection on the device: a
    host error in a later round leaves the device untouched (no remap was issued)."""
    case = cases.riemann_2d(n=8, nv=8, vs_levels=3)
    calls = []

    class Dev:
        def vs_remap(self, t, src, ch):
            calls.append(int(np.asarray(ch).sum()))

    plans = vsadapt.plan_balance_rounds(case.topo, 2)
    assert len(plans) >= 2
    topo, rounds = vsadapt.balance_rounds(Dev(), case.topo, 2)
    assert rounds == len(plans) and len(calls) == rounds
    for k in ("vs_offset", "vs_level", "vs_mid"):
        assert np.array_equal(np.asarray(topo[k]), np.asarray(plans[-1][0][k])), k
    real = vsadapt.plan_balance
    count = {"n": 0}

    def failing(topo, dim):
        count["n"] += 1
        if count["n"] == 2:
            raise RuntimeError("planning failed")
        return real(topo, dim)

    monkeypatch.setattr(vsadapt, "plan_balance", failing)
    calls.clear()
    with pytest.raises(RuntimeError, match="planning failed"):
        vsadapt.balance_rounds(Dev(), case.topo, 2)
    assert calls == []
