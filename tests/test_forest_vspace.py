"""Host-side flattening: the forest (p4est stand-in) and the velocity grids.

Forest: face counts, neighbour-state symmetry, FULL 2:1 balance, face coverage of every cell
by flux pairs, p4est leaf order, periodic images. Velocity grids: the v=0 corner rule, Julia
range midpoints, RMT child order and key order, and -- the indexing contract the device path
relies on -- the key-interval overlap pairs equal the reference's two-pointer walk
(src/Flux/Slope.jl:29-64, src/Flux/Flux.jl:151-344) on random refined grids.
"""
from fractions import Fraction

import numpy as np
import pytest

from kitamr_hip import cases
from kitamr_hip.forest import FAT, Forest
from kitamr_hip.vspace import VelocityGrid, check_vs_setting


def test_uniform_mesh_face_counts():
    case = cases.riemann_2d(n=64, nv=4)
    t = case.topo
    domain = np.sum(t["pair_there"] < 0)
    inner = np.sum(t["pair_there"] >= 0)
    assert (inner, domain) == (8064, 256)
    st = t["nb_state"].reshape(-1, 4)
    assert set(np.unique(st)) == {0, 1}
    assert np.sum(st == 0) == 256


def _leaf_boxes(f):
    size = 1 << (f.Lmax - f.level)
    return f.corner, size


@pytest.mark.parametrize("dim", [2, 3])
def test_amr_balance_and_neighbour_symmetry(dim):
    if dim == 2:
        case = cases.riemann_2d(n=8, nv=4, ps_levels=3, halo=(2.0, 1.5, 1.0))
    else:
        case = cases.box_3d(n=4, nv=4, ps_levels=2)
    t = case.topo
    n = case.n_cells
    F = 1 << (dim - 1)
    st = t["nb_state"].reshape(n, 2 * dim)
    off = t["nb_offset"]
    nb = t["nb_cell"]
    lv = t["ps_level"]
    for c in range(n):
        for f in range(2 * dim):
            lst = nb[off[c * 2 * dim + f]:off[c * 2 * dim + f + 1]]
            s = st[c, f]
            opp = f ^ 1
            if s == 0:
                assert len(lst) == 0
                continue
            if s == 1:
                assert len(lst) == 1 and lv[lst[0]] == lv[c]
                o = lst[0]
                assert st[o, opp] == 1
                assert nb[off[o * 2 * dim + opp]] == c
            elif s == F:
                assert len(lst) == F and all(lv[x] == lv[c] + 1 for x in lst)
                for o in lst:
                    assert st[o, opp] == -1 and nb[off[o * 2 * dim + opp]] == c
            else:
                assert s == -1 and len(lst) == 1 and lv[lst[0]] == lv[c] - 1
                o = lst[0]
                assert st[o, opp] == F and c in nb[off[o * 2 * dim + opp]:off[o * 2 * dim + opp + 1]]
    # FULL balance: leaves sharing a corner differ by at most one level
    f = case.forest
    corner, size = _leaf_boxes(f)
    for c in range(0, n, max(1, n // 200)):
        lo, hi = corner[c] - 1, corner[c] + size[c]
        touch = np.all((corner <= hi[None, :]) & (corner + size[:, None] - 1 >= lo[None, :]), axis=1)
        assert np.all(np.abs(f.level[touch] - f.level[c]) <= 1)


@pytest.mark.parametrize("maker", [
    lambda: cases.riemann_2d(n=8, nv=4, ps_levels=2, halo=(2.0, 1.0)),
    lambda: cases.wall_channel(n=8, nv=4, ps_levels=1),
    lambda: cases.box_3d(n=4, nv=4, ps_levels=1),
])
def test_pairs_cover_every_face_once(maker):
    """Sum over flux pairs of |area| on each side of a cell equals that face's area, and every
    pair's midpoint lies on the face plane of its here cell."""
    case = maker()
    t = case.topo
    dim = case.config.solver.DIM
    n = case.n_cells
    ds = t["cell_ds"].reshape(-1, dim)
    mid = t["cell_mid"].reshape(-1, dim)
    src = np.concatenate([np.arange(n), t["image_src"]]).astype(int)
    cover = np.zeros((n, 2 * dim))
    for p in range(t["n_pairs"]):
        h, th, d = t["pair_here"][p], t["pair_there"][p], t["pair_dir"][p]
        rot, a = t["pair_rot"][p], t["pair_area"][p]
        fh = 2 * d + (1 if rot < 0 else 0)
        cover[h, fh] += abs(a)
        pm = t["pair_mid"].reshape(-1, dim)[p]
        assert pm[d] == mid[h, d] - 0.5 * rot * ds[h, d]
        if th >= 0:
            cover[src[th], fh ^ 1] += abs(a)
    for c in range(n):
        for f in range(2 * dim):
            area = np.prod(ds[c, FAT[dim][f // 2]])
            assert cover[c, f] == pytest.approx(area, rel=1e-14)


def test_leaves_tile_domain_in_p4est_order():
    case = cases.riemann_2d(n=8, nv=4, ps_levels=3, halo=(2.0, 1.5, 1.0))
    f = case.forest
    corner, size = _leaf_boxes(f)
    assert np.sum(size.astype(float) ** 2) == (8 << f.Lmax) ** 2
    # p4est order: by tree (x fastest over the brick), then Morton inside the tree
    tree = (corner >> f.Lmax)
    tid = tree[:, 0] + 8 * tree[:, 1]
    assert np.all(np.diff(tid) >= 0)
    # cell midpoints from quad_to_cell are exact dyadic values here
    t = case.topo
    mid = t["cell_mid"][:case.n_cells].reshape(-1, 2)
    h = 1.0 / (8 << f.Lmax)
    assert np.allclose(mid, (corner + size[:, None] / 2.0) * h, rtol=0, atol=0)
    assert np.array_equal(t["ps_level"], f.level)


def test_periodic_images():
    case = cases.wall_channel(n=8, nv=4)
    t = case.topo
    n = case.n_cells
    assert t["n_images"] == 8             # one per face across the periodic y boundary
    mid = t["cell_mid"].reshape(-1, 2)
    for k, s in enumerate(t["image_src"]):
        im = mid[n + k]
        assert im[0] == mid[s, 0]
        assert abs(abs(im[1] - mid[s, 1]) - 1.0) < 1e-15     # shifted by the domain length


# ---------------------------------------------------------------- velocity grids
def test_corner_rule():
    check_vs_setting([-5, 5, -5, 5], [28, 28], 2)
    with pytest.raises(ValueError, match="root-grid corner"):
        check_vs_setting([-5, 5, -5, 5], [27, 28], 2)
    with pytest.raises(ValueError):
        check_vs_setting([5, -5, -5, 5], [28, 28], 2)


def test_root_midpoints_follow_julia_range():
    g = VelocityGrid(2, [-5.0, 5.0, -5.0, 5.0], [28, 28])
    xs = g.midpoint[:28, 0]
    assert xs[0] == float(Fraction(-135, 28)) == -4.821428571428571
    assert xs[14] == float(Fraction(5, 28))
    assert np.array_equal(g.midpoint[28:56, 1], np.full(28, xs[1]))   # first dimension fastest
    assert g.weight[0] == (10.0 / 28) * (10.0 / 28)


def test_refinement_order_and_keys():
    rng = np.random.default_rng(1)
    g = VelocityGrid(2, [-4.0, 4.0, -4.0, 4.0], [4, 4], maxlevel=3)
    w0 = g.weight.sum()
    for _ in range(3):
        flags = (rng.random(g.n) < 0.3) & (g.level < 3)
        g.refine(flags)
    assert g.weight.sum() == pytest.approx(w0, rel=1e-15)
    k = g.keys()
    assert np.all(np.diff(k) > 0)
    # a refined cell's children follow RMT (x fastest): (-,-), (+,-), (-,+), (+,+)
    g2 = VelocityGrid(2, [-4.0, 4.0, -4.0, 4.0], [2, 2], maxlevel=1)
    g2.refine(np.array([True, False, False, False]))
    ch = g2.midpoint[:4]
    assert ch[0][0] < ch[1][0] and ch[0][1] == ch[1][1] and ch[2][1] > ch[0][1]


def walk_pairs(li_arr, lj_arr, dim):
    """Pairs (i, j, den) in the order of the two-pointer walk of diff_vs!
    (src/Flux/Slope.jl:29-64), written as in the reference."""
    out = []
    index, flag = 0, 0.0
    for i in range(len(li_arr)):
        li = li_arr[i]
        if li == lj_arr[index]:
            out.append((i, index, 1.0))
            index += 1
        elif li < lj_arr[index]:
            while flag != 1.0:
                den = 2.0 ** (dim * (lj_arr[index] - li))
                out.append((i, index, den))
                flag += 1 / den
                index += 1
            flag = 0.0
        else:
            out.append((i, index, 1.0))
            flag += 1 / 2.0 ** (dim * (li - lj_arr[index]))
            if flag == 1.0:
                index += 1
                flag = 0.0
    return out


def overlap_pairs(ga, gb, dim, lmax):
    """The device formulation: for each cell i of grid a, the cells of grid b whose key interval
    overlaps i's (predecessor search, then the finer cells inside i's interval)."""
    ka, kb = ga.keys(), gb.keys()
    out = []
    for i in range(ga.n):
        li = int(ga.level[i])
        j = int(np.searchsorted(kb, ka[i], side="right")) - 1
        lj = int(gb.level[j])
        if lj <= li:
            out.append((i, j, 1.0))
            continue
        end = ka[i] + (1 << (dim * (lmax - li)))
        while j < gb.n and kb[j] < end:
            out.append((i, j, 2.0 ** (dim * (int(gb.level[j]) - li))))
            j += 1
    return out


@pytest.mark.parametrize("dim,seed", [(2, 0), (2, 1), (2, 2), (3, 3), (3, 4)])
def test_walk_equals_key_overlap(dim, seed):
    rng = np.random.default_rng(seed)
    lmax = 3 if dim == 2 else 2
    roots = [2] * dim
    q = [-2.0, 2.0] * dim
    for trial in range(20):
        grids = []
        for _ in range(2):
            g = VelocityGrid(dim, q, roots, maxlevel=lmax)
            for _ in range(lmax):
                flags = (rng.random(g.n) < 0.35) & (g.level < lmax)
                if flags.any():
                    g.refine(flags)
            grids.append(g)
        a, b = grids
        walk = walk_pairs(a.level.astype(int), b.level.astype(int), dim)
        over = overlap_pairs(a, b, dim, lmax)
        assert walk == over, (trial, len(walk), len(over))
