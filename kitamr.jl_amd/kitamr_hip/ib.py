"""Immersed-boundary geometry on the host: a circle in a 2D mesh or a sphere in a 3D mesh.

The geometry is computed once per mesh and handed to the library with the topology; the
per-step work (update_solid_cell!, update_solid_neighbor! and the IB face fluxes) runs on the
device. This module restates, for ``Circle`` and ``Sphere`` (src/Boundary/Circle.jl,
Types.jl:83-130):

* the cell roles: solid_flag / boundary_flag (Circle.jl:1-23): a cell whose midpoint is in
  the solid is a solid cell (bound_enc < 0) when one of its four neighbour points crosses the
  circle, else it lies inside the solid (InsideSolidData: no faces, no update); a fluid cell
  with a solid face neighbour is a donor (bound_enc > 0,
  initialize_solid_neighbor!, src/Boundary/Immersed_boundary.jl:258-275);
* the face list: a face between a solid cell and a fluid cell is owned by the fluid cell
  (initialize_full_face!, src/Solver/Initialize.jl:61-84), faces between two solid cells are
  dropped;
* calc_intersect (Circle.jl:33-52): the intersection point and normal of each donor/solid
  face;
* the image stencil of each face (image_df, Immersed_boundary.jl:367-377) and the fluid
  neighbours of each solid cell (update_solid_cell!, :142-146): the first neighbour of each
  face id, in face-id order, when it is a fluid cell;
* the cut velocity cells (initialize_cutted_velocity_cell, :224-275, with cut_rect and
  cut_cube, src/Velocity_space/Cut_cell.jl:13-200): the velocity cells that the line or plane
  u.n = 0 splits, with the sizes of their two parts. Here the gas part (u.n < 0) is measured
  by clipping: in 2D the square against the half plane; in 3D each face of the cube against
  the half space, with the volume from the divergence theorem, V = (1/3) sum over faces of
  (x.N) times the clipped face area (the cut face lies on a plane through the origin and adds
  nothing). The solid part (u.n > 0, the side the wall Maxwellian fills) is the rest.
"""
from __future__ import annotations

import math

import numpy as np

EPS = 1e-12           # src/Abstract/Types.jl:3
NMT = {2: np.array([[-1, 0], [1, 0], [0, -1], [0, 1]], dtype=np.float64),   # Types.jl:12
       3: np.array([[-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]],
                   dtype=np.float64)}


def solid_flag(mid, center, radius, solid=True):
    """Midpoints inside the solid: xor(|x - c| > r, solid) (Circle.jl:17-19)."""
    out = np.linalg.norm(mid - np.asarray(center)[None, :], axis=1) > radius
    return np.logical_xor(out, solid)


def boundary_flag(mid, ds, center, radius):
    """A neighbour point x + ds*NMT[i] lies on the other side of the circle or sphere
    (Circle.jl:1-16)."""
    c = np.asarray(center)
    nm = NMT[mid.shape[1]]
    flag = np.zeros(len(mid), dtype=np.int64)
    for i in range(len(nm)):
        p = mid + ds * nm[i][None, :]
        flag += np.where(np.linalg.norm(p - c[None, :], axis=1) > radius, 1, -1)
    return np.abs(flag) != len(nm)


def calc_intersect(f_mid, s_mid, center, radius, solid=True):
    """Intersection point and unit normal of the donor/solid face (Circle.jl:33-52). As in
    the reference, the point is measured from the circle's centre (it is the geometric
    point when the centre is the origin) and the normal points out of the solid."""
    c, r = center, radius
    sgn = r if solid else -r
    if abs(f_mid[0] - s_mid[0]) < EPS:
        t = math.acos((f_mid[0] - c[0]) / r)
        if f_mid[1] > c[1]:
            ap = [r * math.cos(t), r * math.sin(t)]
        else:
            ap = [r * math.cos(t), -r * math.sin(t)]
    else:
        t = math.asin((f_mid[1] - c[1]) / r)
        if f_mid[0] > c[0]:
            ap = [r * math.cos(t), r * math.sin(t)]
        else:
            ap = [-r * math.cos(t), r * math.sin(t)]
    ap = np.array(ap)
    return ap, ap / sgn


def calc_intersect_sphere(f_mid, s_mid, center, radius, solid=True):
    """Intersection point and normal of the donor/solid face of a Sphere (Circle.jl:54-65):
    the point moves along the face's direction onto the sphere; as in the reference the
    normal takes the donor's midpoint components across the face direction (the unit outward
    radius when the centre is the origin)."""
    c, r = np.asarray(center, dtype=np.float64), radius
    f = np.asarray(f_mid, dtype=np.float64)
    s = np.asarray(s_mid, dtype=np.float64)
    d = int(np.flatnonzero(np.abs(f - s) > EPS)[0])
    rm = f - c
    x, y = rm[(d + 1) % 3], rm[(d + 2) % 3]
    rz = math.sqrt(r * r - x * x - y * y)
    dz = rz - rm[d] if abs((rz - rm[d]) / (s[d] - f[d])) < 1 else -rz - rm[d]
    ap = f.copy()
    nrm = f.copy()
    ap[d] = dz + f[d]
    nrm[d] = dz + rm[d]
    return ap, nrm / (r if solid else -r)


def _clip_area(x0, x1, y0, y1, n, b=0.0):
    """Area of the rectangle [x0,x1]x[y0,y1] on the side n.y < b (Sutherland-Hodgman with
    one half plane, then the shoelace formula)."""
    poly = [(x0, y0), (x1, y0), (x1, y1), (x0, y1)]
    out = []
    for k in range(4):
        p, q = poly[k], poly[(k + 1) % 4]
        sp = p[0] * n[0] + p[1] * n[1] - b
        sq = q[0] * n[0] + q[1] * n[1] - b
        if sp < 0:
            out.append(p)
        if (sp < 0) != (sq < 0) and sp != sq:
            t = sp / (sp - sq)
            out.append((p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1])))
    if len(out) < 3:
        return 0.0
    a = 0.0
    for k in range(len(out)):
        p, q = out[k], out[(k + 1) % len(out)]
        a += p[0] * q[1] - q[0] * p[1]
    return 0.5 * abs(a)


def cut_cells_2d(normal, vs_mid, vs_level, du):
    """Cut velocity cells of a 2D grid for the line u.n = 0: (index, gas area, solid area),
    index ascending. No cut cells when n is within 1e-6 of an axis (the line then runs along
    cell edges, Immersed_boundary.jl:225)."""
    n = np.asarray(normal, dtype=np.float64)
    idx, gas, sol = [], [], []
    if np.any(np.abs(n) < 1e-6):
        return np.zeros(0, np.int32), np.zeros(0), np.zeros(0)
    for i in range(len(vs_level)):
        ddu = np.asarray(du) / 2.0 ** int(vs_level[i])
        m = vs_mid[i]
        if abs(m[0] * n[0] + m[1] * n[1]) > 0.5 * math.sqrt(ddu[0] * ddu[0] + ddu[1] * ddu[1]):
            continue
        x0, x1 = m[0] - 0.5 * ddu[0], m[0] + 0.5 * ddu[0]
        y0, y1 = m[1] - 0.5 * ddu[1], m[1] + 0.5 * ddu[1]
        full = (x1 - x0) * (y1 - y0)
        g = _clip_area(x0, x1, y0, y1, n)
        if g <= 0.0 or g >= full:
            continue
        idx.append(i)
        gas.append(g)
        sol.append(full - g)
    return np.asarray(idx, np.int32), np.asarray(gas), np.asarray(sol)


def _gas_volume(lo, hi, n):
    """Volume of the box [lo, hi] on the side u.n < 0: (1/3) sum over the six faces of
    (x.N) times the face area on that side (the cut face lies on u.n = 0 and adds nothing)."""
    v = 0.0
    for i in range(3):
        j, k = (i + 1) % 3, (i + 2) % 3
        for c, sgn in ((lo[i], -1.0), (hi[i], 1.0)):
            a = _clip_area(lo[j], hi[j], lo[k], hi[k], (n[j], n[k]), -n[i] * c)
            v += sgn * c * a
    return v / 3.0


def cut_cells_3d(normal, vs_mid, vs_level, du):
    """Cut velocity cells of a 3D grid for the plane u.n = 0: (index, gas volume, solid
    volume), index ascending. No cut cells when two components of n are within 1e-6 of zero
    (Immersed_boundary.jl:252-254)."""
    n = np.asarray(normal, dtype=np.float64)
    idx, gas, sol = [], [], []
    if np.count_nonzero(np.abs(n) < 1e-6) > 1:
        return np.zeros(0, np.int32), np.zeros(0), np.zeros(0)
    for i in range(len(vs_level)):
        ddu = np.asarray(du) / 2.0 ** int(vs_level[i])
        m = vs_mid[i]
        if abs(m[0] * n[0] + m[1] * n[1] + m[2] * n[2]) > \
                0.5 * math.sqrt(ddu[0] * ddu[0] + ddu[1] * ddu[1] + ddu[2] * ddu[2]):
            continue
        lo, hi = m - 0.5 * ddu, m + 0.5 * ddu
        full = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2])
        g = _gas_volume(lo, hi, n)
        if g <= 1e-14 * full or g >= (1.0 - 1e-14) * full:
            continue
        idx.append(i)
        gas.append(g)
        sol.append(full - g)
    return np.asarray(idx, np.int32), np.asarray(gas), np.asarray(sol)


def _cell_grid(topo, c, dim):
    """(level, midpoint [n, dim]) of cell c's velocity grid."""
    off = np.asarray(topo["vs_offset"])
    n = int(off[c + 1] - off[c])
    g = 0 if int(topo.get("vs_uniform", 0)) else int(off[c])
    lev = np.asarray(topo["vs_level"])[g:g + n]
    mid = np.asarray(topo["vs_mid"])[g * dim:(g + n) * dim].reshape(dim, n).T
    return lev, mid


def _root_du(topo, dim):
    lo = np.asarray(topo["vs_lo"], dtype=np.float64)[:dim]
    hi = np.asarray(topo["vs_hi"], dtype=np.float64)[:dim]
    return (hi - lo) / np.asarray(topo["vs_trees_num"][:dim], dtype=np.float64)


def cut_lists(topo, dim):
    """The cut velocity cells of every IB face on its donor's current grid (offsets, index,
    gas, solid), as initialize_solid_neighbor! rebuilds them after each adaptation
    (src/Solver/AMR.jl:62-63)."""
    du = _root_du(topo, dim)
    nrm = np.asarray(topo["ib_normal"]).reshape(-1, dim)
    here = np.asarray(topo["pair_here"])
    off, idx, gas, sol = [0], [], [], []
    for k, p in enumerate(np.asarray(topo["ib_face_pair"])):
        lev, mid = _cell_grid(topo, int(here[p]), dim)
        ci, cg, cs = (cut_cells_2d if dim == 2 else cut_cells_3d)(nrm[k], mid, lev, du)
        idx.extend(ci.tolist())
        gas.extend(cg.tolist())
        sol.extend(cs.tolist())
        off.append(len(idx))
    return (np.asarray(off, np.int64), np.asarray(idx, np.int32), np.asarray(gas, np.float64),
            np.asarray(sol, np.float64))


def refresh_cuts(topo):
    """The topology with the cut cells recomputed for the current donor grids (after a
    velocity-space adaptation changed them); topologies without IB faces pass through."""
    if not len(topo.get("ib_face_pair", [])):
        return topo
    t = dict(topo)
    t["ib_cut_off"], t["ib_cut_index"], t["ib_cut_gas"], t["ib_cut_solid"] = \
        cut_lists(t, int(t["dim"]))
    return t


def immerse_ball(topo, center=(0.0, 0.0), radius=1.0, solid=True, bc=(1.0, 0.0, 0.0, 1.0)):
    """The topology with a circular (2D) or spherical (3D) immersed boundary: bound_enc, the
    face list and the IB arrays of kah_topology. Needs equal-level cells on both sides of
    every solid/fluid face and no solid cell on the domain boundary or a periodic image."""
    dim = int(topo["dim"])
    if dim not in (2, 3):
        raise ValueError("immerse_ball: 2D or 3D meshes")
    n = int(topo["n_cells"])
    mid = np.asarray(topo["cell_mid"]).reshape(-1, dim)[:n]
    ds = np.asarray(topo["cell_ds"]).reshape(-1, dim)[:n]
    level = np.asarray(topo["ps_level"])
    img = np.asarray(topo.get("image_src", np.zeros(0, np.int64)))
    inside = solid_flag(mid, center, radius, solid)
    ghost = inside & boundary_flag(mid, ds, center, radius)
    benc = np.zeros(n, dtype=np.int32)
    benc[inside] = -1

    # faces: the fluid side owns a solid/fluid face; solid/solid faces are dropped
    keys = ("pair_here", "pair_there", "pair_dir", "pair_rot", "pair_area", "pair_bc")
    P = {k: np.array(topo[k], copy=True) for k in keys}
    pmid = np.asarray(topo["pair_mid"]).reshape(-1, dim).copy()
    keep = np.ones(len(P["pair_here"]), dtype=bool)
    for p in range(len(keep)):
        h, t = int(P["pair_here"][p]), int(P["pair_there"][p])
        if t >= n:
            if inside[h] or inside[int(img[t - n])]:
                raise ValueError("immerse_ball: a solid cell touches a periodic image")
            continue
        if t < 0:
            if inside[h]:
                raise ValueError("immerse_ball: a solid cell lies on the domain boundary")
            continue
        if inside[h] and inside[t]:
            keep[p] = False
            continue
        if inside[h] or inside[t]:
            s, f = (h, t) if inside[h] else (t, h)
            if not ghost[s]:
                raise ValueError(f"immerse_ball: fluid cell {f} touches cell {s} inside "
                                 "the solid")
            if level[s] != level[f]:
                raise ValueError("immerse_ball: solid/fluid faces need equal cell levels")
            if inside[h]:
                P["pair_here"][p], P["pair_there"][p] = t, h
                P["pair_rot"][p] = -P["pair_rot"][p]
                P["pair_area"][p] = -P["pair_area"][p]
    out = dict(topo)
    for k in keys:
        out[k] = P[k][keep]
    out["pair_mid"] = pmid[keep].reshape(-1)
    out["n_pairs"] = int(keep.sum())

    nbs = np.asarray(topo["nb_state"]).reshape(n, 2 * dim)
    nboff = np.asarray(topo["nb_offset"])
    nbc = np.asarray(topo["nb_cell"])

    def first_nb(c, f):
        if nbs[c, f] == 0:
            return -1
        return int(nbc[nboff[c * 2 * dim + f]])

    def fluid_nbs(c):
        lst = []
        for f in range(2 * dim):
            x = first_nb(c, f)
            if x >= 0 and not inside[x]:
                lst.append(x)
        return lst

    # donors and their IB faces (initialize_solid_neighbor!, face-id order)
    pair_of = {}
    for p in range(out["n_pairs"]):
        pair_of[(int(out["pair_here"][p]), int(out["pair_there"][p]))] = p
    faces = dict(pair=[], aux=[], normal=[], img_off=[0], img=[])
    for c in range(n):
        if inside[c]:
            continue
        for f in range(2 * dim):
            s = first_nb(c, f)
            if s < 0 or not inside[s]:
                continue
            benc[c] = 1
            if (c, s) not in pair_of:
                raise ValueError(f"immerse_ball: no face between donor {c} and solid {s}")
            faces["pair"].append(pair_of[(c, s)])
            if dim == 2:
                ap, nrm = calc_intersect(mid[c], mid[s], center, radius, solid)
            else:
                ap, nrm = calc_intersect_sphere(mid[c], mid[s], center, radius, solid)
            faces["aux"].append(ap)
            faces["normal"].append(nrm)
            st = fluid_nbs(c) + [c]
            faces["img"].extend(st)
            faces["img_off"].append(len(faces["img"]))
    solid_cells = np.flatnonzero(ghost)
    snb_off, snb = [0], []
    for s in solid_cells:
        lst = fluid_nbs(int(s))
        if not lst:
            raise ValueError(f"immerse_ball: solid cell {s} has no fluid face neighbour")
        snb.extend(lst)
        snb_off.append(len(snb))
    out["bound_enc"] = benc
    out["ib_solid"] = solid_cells.astype(np.int64)
    out["ib_solid_nb_off"] = np.asarray(snb_off, np.int64)
    out["ib_solid_nb"] = np.asarray(snb, np.int64)
    out["ib_face_pair"] = np.asarray(faces["pair"], np.int64)
    nf = len(faces["pair"])
    out["ib_aux_point"] = np.asarray(faces["aux"], np.float64).reshape(-1)
    out["ib_normal"] = np.asarray(faces["normal"], np.float64).reshape(-1)
    out["ib_bc"] = np.tile(np.asarray(bc, np.float64), nf)
    out["ib_image_off"] = np.asarray(faces["img_off"], np.int64)
    out["ib_image_cell"] = np.asarray(faces["img"], np.int64)
    out["ib_cut_off"], out["ib_cut_index"], out["ib_cut_gas"], out["ib_cut_solid"] = \
        cut_lists(out, dim)
    out["ib_inside"] = inside
    return out


immerse_circle = immerse_ball   # Circle (2D)
immerse_sphere = immerse_ball   # Sphere (3D)
