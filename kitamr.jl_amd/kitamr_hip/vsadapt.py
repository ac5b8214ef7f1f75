"""Velocity-space adaptation with the distribution kept on the device.

vs_adaptive_mesh_refinement! (src/Velocity_space/AMR.jl:168-181) in four moves: the device
evaluates the criteria (``kah_vs_resolution`` / ``kah_vs_flags``), this module rebuilds the
grids from one flag byte per velocity cell (refine_grid_stream! / coarsen_grid_stream!,
src/Velocity_space/Rebuild.jl:44-163, on levels, midpoints and weights only, plus the source
of every new velocity cell), the device remaps f onto the new grids (``kah_vs_remap``) and
applies the conservative I-projection to the changed cells (``kah_vs_project``); the
optional vs_balance! rounds (src/Velocity_space/Balance.jl) plan on the grids and inject f on
the device the same way. f never leaves HBM.
"""
from __future__ import annotations

import numpy as np

from .ib import refresh_cuts
from .vspace import RMT


def cell_grids(topo, dim):
    """(level, midpoint [n, dim], weight) of every cell's velocity grid."""
    off = np.asarray(topo["vs_offset"])
    uni = int(topo.get("vs_uniform", 0))
    lev = np.asarray(topo["vs_level"])
    mid = np.asarray(topo["vs_mid"])
    wt = np.asarray(topo["vs_weight"])
    out = []
    for c in range(len(off) - 1):
        n = int(off[c + 1] - off[c])
        g = 0 if uni else int(off[c])
        out.append((lev[g:g + n].copy(), mid[g * dim:(g + n) * dim].reshape(dim, n).T.copy(),
                    wt[g:g + n].copy()))
    return out


def with_grids(topo, grids, dim):
    """The topology with per-cell velocity grids replaced (vs_uniform off)."""
    t = dict(topo)
    sizes = np.array([len(g[0]) for g in grids], dtype=np.int64)
    t["vs_offset"] = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    t["vs_level"] = np.concatenate([g[0] for g in grids]).astype(np.int8)
    t["vs_mid"] = np.concatenate([np.ascontiguousarray(g[1].T).reshape(-1) for g in grids])
    t["vs_weight"] = np.concatenate([g[2] for g in grids]).astype(np.float64)
    t["vs_uniform"] = 0
    # the cut velocity cells of the IB faces follow the donors' new grids
    return refresh_cuts(t)


def root_ds(topo, dim):
    """Root cell size per dimension, as vs_refine! forms it (quadrature extent / trees)."""
    return np.array([(topo["vs_hi"][d] - topo["vs_lo"][d]) / topo["vs_trees_num"][d]
                     for d in range(dim)])


def refine_stream(lev, mid, wt, flags, ds, dim):
    """refine_grid_stream! (Rebuild.jl:44-85): children in RMT order, the parent's f
    injected. Returns (changed, level, midpoint, weight, src) with src the old index of each
    new cell."""
    sel = np.asarray(flags, dtype=bool)
    n = len(lev)
    if not sel.any():
        return False, lev, mid, wt, np.arange(n, dtype=np.int32)
    nch = 1 << dim
    counts = np.where(sel, nch, 1)
    src = np.repeat(np.arange(n), counts)
    first = np.concatenate([[0], np.cumsum(counts)[:-1]])
    child = np.arange(len(src)) - np.repeat(first, counts)
    ref = sel[src]
    L = lev[src].astype(np.int64)
    nlev = (L + ref).astype(np.int8)
    nwt = np.where(ref, wt[src] / nch, wt[src])
    step = 0.5 * (ds[None, :] / (2.0 ** (L + 1))[:, None])
    nmid = np.where(ref[:, None], mid[src] + step * RMT[dim][child % nch], mid[src])
    return True, nlev, nmid, nwt, src.astype(np.int32)


def coarsen_stream(lev, mid, wt, ok, ds, maxlevel, dim):
    """coarsen_grid_stream! (Rebuild.jl:96-163): a 2^dim-aligned same-level sibling group is
    merged iff every member is willing, with the same alignment walk. Returns (changed,
    level, midpoint, weight, src) with src = old index of a kept cell or -(first+1) for a
    merged group."""
    ok = np.asarray(ok, dtype=bool)
    n = len(lev)
    if not ok.any():
        return False, lev, mid, wt, np.arange(n, dtype=np.int32)
    nch = 1 << dim
    flag = [0.0] * (maxlevel + 1)
    nl, nm, nw, src = [], [], [], []
    changed = False
    index = 0
    while index < n:
        fl = int(lev[index])
        if fl > 0:
            if (flag[fl] % 1 == 0.0 and index + nch - 1 <= n - 1 and
                    all(int(lev[index + g]) == fl for g in range(1, nch))):
                if all(ok[index + g] for g in range(nch)):
                    nl.append(fl - 1)
                    nw.append(wt[index] * nch)
                    nm.append([mid[index, d] - 0.5 * (ds[d] / 2.0 ** fl) * RMT[dim][0, d]
                               for d in range(dim)])
                    src.append(-(index + 1))
                    changed = True
                else:
                    for g in range(nch):
                        nl.append(int(lev[index + g]))
                        nw.append(wt[index + g])
                        nm.append(list(mid[index + g]))
                        src.append(index + g)
                index += nch
                for i in range(1, fl):
                    flag[i] += 1 / 2 ** (dim * (fl - i))
            else:
                nl.append(fl)
                nw.append(wt[index])
                nm.append(list(mid[index]))
                src.append(index)
                for i in range(1, fl + 1):
                    flag[i] += 1 / 2 ** (dim * (fl - i + 1))
                index += 1
        else:
            nl.append(fl)
            nw.append(wt[index])
            nm.append(list(mid[index]))
            src.append(index)
            index += 1
    return (changed, np.asarray(nl, dtype=np.int8), np.asarray(nm, dtype=np.float64).reshape(-1, dim),
            np.asarray(nw, dtype=np.float64), np.asarray(src, dtype=np.int32))


def plan_refine(topo, flags, dim):
    """New grids, sources and changed cells of the refine pass."""
    off = np.asarray(topo["vs_offset"])
    ds = root_ds(topo, dim)
    grids, srcs, changed = [], [], []
    for c, (lv, md, wt) in enumerate(cell_grids(topo, dim)):
        ch, nl, nm, nw, s = refine_stream(lv, md, wt, flags[off[c]:off[c + 1]] & 1, ds, dim)
        grids.append((nl, nm, nw))
        srcs.append(s)
        changed.append(ch)
    return with_grids(topo, grids, dim), np.concatenate(srcs), np.array(changed, dtype=np.uint8)


def plan_coarsen(topo, flags, dim):
    """New grids, sources and changed cells of the coarsen pass."""
    off = np.asarray(topo["vs_offset"])
    ds = root_ds(topo, dim)
    maxlevel = int(topo["vs_maxlevel"])
    grids, srcs, changed = [], [], []
    for c, (lv, md, wt) in enumerate(cell_grids(topo, dim)):
        ch, nl, nm, nw, s = coarsen_stream(lv, md, wt, (flags[off[c]:off[c + 1]] >> 1) & 1, ds,
                                           maxlevel, dim)
        grids.append((nl, nm, nw))
        srcs.append(s)
        changed.append(ch)
    return with_grids(topo, grids, dim), np.concatenate(srcs), np.array(changed, dtype=np.uint8)


def grid_keys(topo, lev, mid, dim):
    """cellkey (src/Velocity_space/Neighbor.jl:61-124: root_lin * 2^(DIM*Lmax) + morton of
    the lower corner on the finest lattice) of the velocity cells of one grid."""
    L = int(topo["vs_maxlevel"])
    lo = np.array([topo["vs_lo"][d] for d in range(dim)])
    h_fine = root_ds(topo, dim) / 2.0 ** L
    cell = h_fine[None, :] * (2.0 ** (L - lev.astype(np.int64)))[:, None]
    g = np.round((mid - 0.5 * cell - lo[None, :]) / h_fine[None, :]).astype(np.int64)
    stride = 1
    root_lin = np.zeros(len(lev), dtype=np.int64)
    morton = np.zeros(len(lev), dtype=np.int64)
    for d in range(dim):
        root_lin += (g[:, d] >> L) * stride
        stride *= int(topo["vs_trees_num"][d])
        q = g[:, d] & ((1 << L) - 1)
        for b in range(L):
            morton |= ((q >> b) & 1) << (b * dim + d)
    return (root_lin << (dim * L)) | morton


def grid_ids(topo, dim):
    """Id of each cell's velocity grid: cells with identical grids (levels and midpoints, byte
    for byte) share one id. Returns (ids, representative cell of each id)."""
    off = np.asarray(topo["vs_offset"])
    n = len(off) - 1
    if int(topo.get("vs_uniform", 0)):
        return np.zeros(n, dtype=np.int64), np.zeros(1, dtype=np.int64)
    lev = np.ascontiguousarray(np.asarray(topo["vs_level"]))
    mid = np.ascontiguousarray(np.asarray(topo["vs_mid"], dtype=np.float64))
    try:
        import xxhash
        digest = lambda b: xxhash.xxh3_128_digest(b)
    except ImportError:   # pragma: no cover - xxhash is part of the image
        import hashlib
        digest = lambda b: hashlib.blake2b(b, digest_size=16).digest()
    ids = np.empty(n, dtype=np.int64)
    seen, reps = {}, []
    for c in range(n):
        a, b = int(off[c]), int(off[c + 1])
        lv, md = lev[a:b], mid[a * dim:b * dim]
        key = digest(lv.tobytes()) + digest(md.tobytes())
        cand = seen.setdefault(key, [])
        for g in cand:   # equal digests are checked byte for byte
            r = reps[g]
            ra, rb = int(off[r]), int(off[r + 1])
            if rb - ra == b - a and np.array_equal(lev[ra:rb], lv) and \
                    np.array_equal(mid[ra * dim:rb * dim], md):
                ids[c] = g
                break
        else:
            ids[c] = len(reps)
            cand.append(len(reps))
            reps.append(c)
    return ids, np.asarray(reps, dtype=np.int64)


def balance_flags(topo, dim):
    """Refine flags of one balance round (vs_balance_local!, src/Velocity_space/Balance.jl:
    72-104): a velocity cell of a fluid cell is flagged when a velocity cell of one of its
    fluid face neighbours that lies inside it is two or more levels finer (vs_balance_pair!,
    :12-64, takes the maximum neighbour level over the cell's extent). Faces to a domain
    boundary or to a solid cell (SolidNeighbor) are skipped, as are cells inside a solid.
    Every flagged cell has to be refined in any balanced refinement of the current grids, so
    repeating rounds reaches the same grids as the reference's cell-by-cell passes. A face
    across a periodic boundary names the image record (n_cells + k) of its other cell; the
    balance reads the image's source cell, which p4est would return as the neighbour.
    The finest neighbour level inside each velocity cell depends only on the two grids, so it
    is computed once per distinct ordered pair of grids; neighbours on the cell's own grid
    are skipped (the cell inside each velocity cell is that cell itself, one level apart at
    most)."""
    off = np.asarray(topo["vs_offset"])
    benc = np.asarray(topo["bound_enc"])
    n = len(off) - 1
    img = np.asarray(topo.get("image_src", []), dtype=np.int64)
    L = int(topo["vs_maxlevel"])
    flags = np.zeros(int(off[-1]), dtype=np.uint8)
    if int(topo.get("vs_uniform", 0)):
        return flags
    gid, reps = grid_ids(topo, dim)
    lev_all = np.asarray(topo["vs_level"])
    mid_all = np.asarray(topo["vs_mid"], dtype=np.float64)

    def grid(g):
        c = int(reps[g])
        a, b = int(off[c]), int(off[c + 1])
        lv = lev_all[a:b]
        return lv.astype(np.int64), grid_keys(topo, lv, mid_all[a * dim:b * dim].reshape(dim, b - a).T, dim)

    here = np.asarray(topo["pair_here"], dtype=np.int64)
    there = np.asarray(topo["pair_there"], dtype=np.int64)
    keep = there >= 0
    here, there = here[keep], there[keep]
    if len(img):
        here = np.where(here >= n, img[np.clip(here - n, 0, len(img) - 1)], here)
        there = np.where(there >= n, img[np.clip(there - n, 0, len(img) - 1)], there)
    fluid = (benc[here] >= 0) & (benc[there] >= 0)
    here, there = here[fluid], there[fluid]
    cross = gid[here] != gid[there]
    # (cell, neighbour grid) of every cross-grid face side, both directions
    cells = np.concatenate([here[cross], there[cross]])
    nbg = np.concatenate([gid[there[cross]], gid[here[cross]]])
    cache, gcache = {}, {}

    def g_of(g):
        if g not in gcache:
            gcache[g] = grid(g)
        return gcache[g]

    need = {}
    for c, gb in sorted(set(zip(cells.tolist(), nbg.tolist()))):
        ga = int(gid[c])
        key = (ga, gb)
        if key not in cache:
            la, ka = g_of(ga)
            lb, kb = g_of(gb)
            own = np.searchsorted(ka, kb, side="right") - 1
            inside = own >= 0
            own_c = np.where(inside, own, 0)
            inside &= kb < ka[own_c] + (np.int64(1) << (dim * (L - la[own_c])))
            nd = np.full(len(la), -1, dtype=np.int64)
            np.maximum.at(nd, own_c[inside], lb[inside])
            cache[key] = nd
        if c in need:
            np.maximum(need[c], cache[key], out=need[c])
        else:
            need[c] = cache[key].copy()
    for c, nd in need.items():
        la = g_of(int(gid[c]))[0]
        flags[off[c]:off[c + 1]] = nd > la + 1
    return flags


def plan_balance(topo, dim):
    """New grids, sources and changed cells of one balance round (refine_grid_stream! on the
    flagged cells, the parent's f injected)."""
    return plan_refine(topo, balance_flags(topo, dim), dim)


def plan_balance_rounds(topo, dim):
    """The balance rounds of vs_balance! planned on the host alone (they depend on the grids,
    not on f): a list of (topology, sources, changed) per round, until a round flags nothing."""
    plans = []
    while True:
        t, src, ch = plan_balance(topo, dim)
        if not ch.any():
            return plans
        plans.append((t, src, ch))
        topo = t


def balance_rounds(dev, topo, dim):
    """vs_balance! on a context: every round planned on the host first, then each injected on
    the device (``kah_vs_remap``), so that a host error leaves the device untouched. Returns
    (topology, rounds)."""
    plans = plan_balance_rounds(topo, dim)
    for t, src, ch in plans:
        dev.vs_remap(t, src, ch)
    return (plans[-1][0] if plans else topo), len(plans)


def _inside_rows(topo):
    """Velocity rows of the cells inside a solid (InsideSolidData: no velocity space in the
    reference, so never adapted), or None."""
    inside = topo.get("ib_inside")
    if inside is None:
        return None
    deep = np.asarray(inside, dtype=bool).copy()
    deep[np.asarray(topo.get("ib_solid", []), dtype=np.int64)] = False
    if not deep.any():
        return None
    off = np.asarray(topo["vs_offset"])
    return np.concatenate([np.arange(off[c], off[c + 1]) for c in np.flatnonzero(deep)])


def vs_adapt(dev, topo, params, dim, balance=False):
    """One vs_adaptive_mesh_refinement!(ka; vs_balance=balance) on a context: returns the new
    topology and the per-cell changed flags of the refine and coarsen passes. vr is computed
    once, before refining, and used by both passes, as the reference does. Cells inside an
    immersed solid keep their grids (src/Velocity_space/AMR.jl:47, :93). With balance, the
    balance rounds run after the coarsen pass and before the conservative correction, which
    (as in the reference) projects only the cells the refine and coarsen passes changed."""
    vr = dev.vs_resolution()
    flags = dev.vs_flags(params, vr)
    rows = _inside_rows(topo)
    if rows is not None:
        flags[rows] = 0
    t1, src1, ch1 = plan_refine(topo, flags, dim)
    if ch1.any():
        dev.vs_remap(t1, src1, ch1)
    else:
        t1 = topo
    flags = dev.vs_flags(params, vr)
    rows = _inside_rows(t1)
    if rows is not None:
        flags[rows] = 0
    t2, src2, ch2 = plan_coarsen(t1, flags, dim)
    # the balance rounds depend on the grids alone: planned before the coarsen remap, so that
    # a host error there leaves the device as the refine pass left it
    plans = plan_balance_rounds(t2, dim) if balance else []
    # coarsen_grid_stream! rebuilds every grid, so every cell's slopes are zero afterwards
    dev.vs_remap(t2, src2, np.ones(len(ch2), dtype=np.uint8))
    for tb, srcb, chb in plans:
        dev.vs_remap(tb, srcb, chb)
    if plans:
        t2 = plans[-1][0]
    va = (ch1 | ch2).astype(bool)
    dev.vs_project(np.flatnonzero(va).astype(np.int32))
    return t2, va
