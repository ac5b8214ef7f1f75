"""Host-side physical-space forest: the part of KitAMR's set-up that p4est and the Julia
flattening provide before the device path runs.

It builds a brick of quadtrees/octrees (one tree per base cell), refines it, applies the
FULL 2:1 balance KitAMR requests from p4est (src/Physical_space/AMR.jl:1004-1058), orders
the leaves like p4est (trees in connectivity order, Morton order inside a tree), computes
cell midpoints/sizes the way ``quad_to_cell`` does (src/P4est/P4est_wrap.jl:142-191), finds
face neighbours with the reference's state encoding (src/Mesh/Neighbor.jl:3-135) and emits
the face list with the rules of ``initialize_faces!`` (src/Solver/Initialize.jl:5-158).

The output is the flattened topology consumed by ``kah_set_topology``.
"""
from __future__ import annotations

import numpy as np

P4EST_MAXLEVEL = {2: 30, 3: 19}
FAT = {2: [[1], [0]], 3: [[1, 2], [0, 2], [0, 1]]}  # face area table, src/Abstract/Types.jl:13


def _morton_interleave(coords, bits):
    """Morton code of integer coordinates (first dimension in the lowest bit)."""
    coords = np.asarray(coords, dtype=np.int64)
    dim = coords.shape[1]
    code = np.zeros(coords.shape[0], dtype=np.int64)
    for b in range(bits):
        for d in range(dim):
            code |= ((coords[:, d] >> b) & 1) << (b * dim + d)
    return code


class Forest:
    """A brick forest with uniform base trees and leaves stored as integer boxes.

    Leaves are kept as (level, lower corner in finest-level units over the whole brick).
    ``max_level`` fixes the finest resolution that refinement may reach.
    """

    def __init__(self, dim, trees_num, geometry, periodic=None, max_level=0):
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        self.dim = dim
        self.trees_num = tuple(int(x) for x in trees_num)
        self.geometry = [float(x) for x in geometry]
        self.periodic = tuple(bool(x) for x in (periodic or (False,) * dim))
        if len(self.trees_num) != dim or len(self.geometry) != 2 * dim:
            raise ValueError("trees_num/geometry do not match dim")
        for d in range(dim):
            if self.periodic[d] and self.trees_num[d] < 2:
                raise ValueError("periodic direction needs at least two trees")
        self.Lmax = int(max_level)
        self._set_trees()
        # start with every tree as one level-0 leaf
        tc = self.tree_coords
        self.level = np.zeros(len(tc), dtype=np.int64)
        self.corner = (tc << self.Lmax).astype(np.int64)

    # ------------------------------------------------------------------ trees
    def _set_trees(self):
        dim, nt = self.dim, self.trees_num
        grids = np.meshgrid(*[np.arange(n) for n in nt], indexing="ij")
        coords = np.stack([g.reshape(-1) for g in grids], axis=1)
        if any(self.periodic):
            # p4est brick connectivity (src/Mesh/Connectivity.jl:168-203): trees follow a
            # Morton curve over the brick; vertices are integer lattice points scaled
            # as  i*(hi-lo)/N + lo.
            bits = max(int(np.ceil(np.log2(max(n, 1)))) for n in nt) + 1
            order = np.argsort(_morton_interleave(coords, bits), kind="stable")
        else:
            # Cartesian_connectivity (src/Mesh/Connectivity.jl:104-149): first
            # dimension fastest, vertices lo + (i-1)*dx.
            key = np.zeros(len(coords), dtype=np.int64)
            stride = 1
            for d in range(dim):
                key += coords[:, d] * stride
                stride *= nt[d]
            order = np.argsort(key, kind="stable")
        self.tree_coords = coords[order].astype(np.int64)
        lin = np.zeros(len(coords), dtype=np.int64)
        stride = 1
        for d in range(dim):
            lin += self.tree_coords[:, d] * stride
            stride *= nt[d]
        self.tree_of_lin = np.empty(len(coords), dtype=np.int64)
        self.tree_of_lin[lin] = np.arange(len(coords))

    def _vertex(self, d, i):
        """Coordinate of lattice vertex i along d, as the connectivity stores it."""
        lo, hi = self.geometry[2 * d], self.geometry[2 * d + 1]
        n = self.trees_num[d]
        i = np.asarray(i, dtype=np.float64)
        if any(self.periodic):
            return i * (hi - lo) / n + lo
        dx = (hi - lo) / n
        return lo + i * dx

    # ------------------------------------------------------------------ geometry
    def leaf_geometry(self, level=None, corner=None):
        """Midpoint and size of leaves as p4est_qcoord_to_vertex gives them."""
        level = self.level if level is None else level
        corner = self.corner if corner is None else corner
        dim = self.dim
        ML = P4EST_MAXLEVEL[dim]
        root = float(1 << ML)
        tree = corner >> self.Lmax
        local = corner - (tree << self.Lmax)
        q = local << (ML - self.Lmax)
        half = (1 << (ML - level)) // 2
        # tree corner vertices: index bits (x lowest) select lower/upper vertex per dim
        vlo = [self._vertex(d, tree[:, d]) for d in range(dim)]
        vhi = [self._vertex(d, tree[:, d] + 1) for d in range(dim)]

        def qv(qc):
            w1 = [qc[:, d].astype(np.float64) / root for d in range(dim)]
            w0 = [1.0 - w1[d] for d in range(dim)]
            out = [np.zeros(len(qc)) for _ in range(dim)]
            for corner_id in range(1 << dim):
                bits = [(corner_id >> d) & 1 for d in range(dim)]
                # factor built as in p4est: z*y first, then *x
                if dim == 2:
                    fac = (w1[1] if bits[1] else w0[1]) * (w1[0] if bits[0] else w0[0])
                else:
                    fy = (w1[2] if bits[2] else w0[2]) * (w1[1] if bits[1] else w0[1])
                    fac = fy * (w1[0] if bits[0] else w0[0])
                for e in range(dim):
                    vert = vhi[e] if bits[e] else vlo[e]
                    out[e] = out[e] + fac * vert
            return np.stack(out, axis=1)

        quad = qv(q)
        mid = qv(q + half[:, None])
        ds = 2.0 * (mid - quad)
        return mid, ds

    # ------------------------------------------------------------------ refinement
    def refine(self, flag_fn, max_level=None):
        """Refine, level by level, every leaf for which flag_fn(mid, ds, level) is true."""
        max_level = self.Lmax if max_level is None else min(max_level, self.Lmax)
        for _ in range(max_level):
            mid, ds = self.leaf_geometry()
            flags = np.asarray(flag_fn(mid, ds, self.level), dtype=bool) & (self.level < max_level)
            if not flags.any():
                break
            self._split(flags)
        return self

    def _split(self, flags):
        dim = self.dim
        keep_l = self.level[~flags]
        keep_c = self.corner[~flags]
        pl = self.level[flags]
        pc = self.corner[flags]
        half = (np.int64(1) << (self.Lmax - pl - 1))
        kids_l, kids_c = [], []
        for child in range(1 << dim):
            off = np.stack([((child >> d) & 1) * half for d in range(dim)], axis=1)
            kids_l.append(pl + 1)
            kids_c.append(pc + off)
        self.level = np.concatenate([keep_l] + kids_l)
        self.corner = np.concatenate([keep_c] + kids_c)

    def _extent(self):
        return np.array([n << self.Lmax for n in self.trees_num], dtype=np.int64)

    def _lookup_tables(self):
        tabs = {}
        for L in range(self.Lmax + 1):
            sel = self.level == L
            keys = self._key(self.corner[sel], L)
            order = np.argsort(keys)
            tabs[L] = (keys[order], np.nonzero(sel)[0][order])
        return tabs

    def _key(self, corner, L):
        ext = self._extent()
        k = np.zeros(len(corner), dtype=np.int64)
        stride = 1
        for d in range(self.dim):
            k += (corner[:, d] >> (self.Lmax - L)) * stride
            stride *= int(ext[d] >> (self.Lmax - L))
        return k

    def _find(self, tabs, pts, L):
        """Index of the leaf at exactly level L whose box starts at pts (aligned), or -1."""
        keys, idx = tabs[L]
        k = self._key(pts, L)
        pos = np.searchsorted(keys, k)
        pos = np.minimum(pos, len(keys) - 1) if len(keys) else pos
        out = np.full(len(pts), -1, dtype=np.int64)
        if len(keys):
            hit = keys[pos] == k
            out[hit] = idx[pos[hit]]
        return out

    def _containing(self, tabs, pts):
        """Level and index of the leaf containing each finest-level point."""
        lev = np.full(len(pts), -1, dtype=np.int64)
        ind = np.full(len(pts), -1, dtype=np.int64)
        for L in range(self.Lmax + 1):
            todo = ind < 0
            if not todo.any():
                break
            aligned = (pts[todo] >> (self.Lmax - L)) << (self.Lmax - L)
            f = self._find(tabs, aligned, L)
            sub = np.nonzero(todo)[0]
            got = f >= 0
            ind[sub[got]] = f[got]
            lev[sub[got]] = L
        return lev, ind

    def _wrap(self, pts):
        """Apply periodicity; returns (points, inside-mask)."""
        ext = self._extent()
        pts = pts.copy()
        inside = np.ones(len(pts), dtype=bool)
        for d in range(self.dim):
            if self.periodic[d]:
                pts[:, d] %= ext[d]
            else:
                inside &= (pts[:, d] >= 0) & (pts[:, d] < ext[d])
        return pts, inside

    def balance(self):
        """FULL (face+edge+corner) 2:1 balance, as p4est_balance_ext(..., CONNECT_FULL)."""
        dim = self.dim
        offsets = np.array(np.meshgrid(*[[-1, 0, 1]] * dim, indexing="ij")).reshape(dim, -1).T
        offsets = offsets[np.any(offsets != 0, axis=1)]
        while True:
            tabs = self._lookup_tables()
            size = (np.int64(1) << (self.Lmax - self.level))
            need = np.zeros(len(self.level), dtype=bool)
            for off in offsets:
                # a point just outside the leaf in direction off
                pt = self.corner + np.where(off > 0, size[:, None], np.where(off < 0, -1, 0))
                pt, inside = self._wrap(pt)
                lev, ind = self._containing(tabs, pt[inside])
                src = np.nonzero(inside)[0]
                bad = lev < self.level[src] - 1
                need[ind[bad]] = True
            if not need.any():
                break
            self._split(need)
        return self

    # ------------------------------------------------------------------ order
    def finalize_order(self):
        """Sort leaves like p4est: by tree (connectivity order), Morton inside the tree."""
        tree_c = self.corner >> self.Lmax
        lin = np.zeros(len(tree_c), dtype=np.int64)
        stride = 1
        for d in range(self.dim):
            lin += tree_c[:, d] * stride
            stride *= self.trees_num[d]
        tree = self.tree_of_lin[lin]
        local = self.corner - (tree_c << self.Lmax)
        mort = _morton_interleave(local, max(self.Lmax, 1))
        order = np.lexsort((mort, tree))
        self.level = self.level[order]
        self.corner = self.corner[order]
        self.tree = tree[order]
        return self

    # ------------------------------------------------------------------ topology
    def build_topology(self, domain_bc):
        """Neighbour lists, flux pairs and periodic images of the balanced forest.

        domain_bc: list of 2*dim entries (faceid order xmin,xmax,ymin,...) giving the index of
        the boundary condition of each domain side (ignored on periodic sides).
        """
        self.finalize_order()
        dim = self.dim
        n = len(self.level)
        mid, ds = self.leaf_geometry()
        tabs = self._lookup_tables()
        size = (np.int64(1) << (self.Lmax - self.level))
        nfine = 1 << (dim - 1)

        states = np.zeros((n, 2 * dim), dtype=np.int32)
        nb_lists = [[None] * (2 * dim) for _ in range(n)]
        wrap_face = np.zeros((n, 2 * dim), dtype=bool)
        ext = self._extent()
        for f in range(2 * dim):
            d, hi = f // 2, f % 2
            step = np.where(hi, size, -size)
            pt = self.corner.copy()
            pt[:, d] += step
            raw_out = (pt[:, d] < 0) | (pt[:, d] >= ext[d])
            wrap_face[:, f] = raw_out & self.periodic[d]
            ptw, inside = self._wrap(pt)
            same = np.full(n, -1, dtype=np.int64)
            same[inside] = self._find_same(tabs, ptw[inside], self.level[inside])
            # coarser neighbour: leaf one level up containing the adjacent point
            coarse = np.full(n, -1, dtype=np.int64)
            probe = ptw.copy()
            if not hi:
                probe[:, d] = ptw[:, d] + size - 1   # any point of the adjacent box
            cand = inside & (same < 0) & (self.level > 0)
            if cand.any():
                L1 = self.level[cand] - 1
                shift = self.Lmax - L1
                aligned = (probe[cand] >> shift[:, None]) << shift[:, None]
                coarse[cand] = self._find_same(tabs, aligned, L1)
            for c in range(n):
                if not inside[c]:
                    states[c, f] = 0
                    nb_lists[c][f] = []
                elif same[c] >= 0:
                    states[c, f] = 1
                    nb_lists[c][f] = [int(same[c])]
                elif coarse[c] >= 0:
                    states[c, f] = -1
                    nb_lists[c][f] = [int(coarse[c])]
                else:
                    states[c, f] = nfine
                    nb_lists[c][f] = None  # filled below
            fine_cells = np.nonzero(inside & (same < 0) & (coarse < 0))[0]
            if len(fine_cells):
                h = size[fine_cells] // 2
                base = ptw[fine_cells].copy()
                if not hi:
                    base[:, d] += h  # the fine layer adjacent to our low face
                trans = [e for e in range(dim) if e != d]
                kids = []
                for sub in range(nfine):
                    p = base.copy()
                    for b, e in enumerate(trans):
                        p[:, e] += ((sub >> b) & 1) * h
                    kids.append(self._find_same(tabs, p, self.level[fine_cells] + 1))
                kids = np.stack(kids, axis=1)
                if (kids < 0).any():
                    raise RuntimeError("forest is not 2:1 balanced")
                for r, c in enumerate(fine_cells):
                    nb_lists[c][f] = [int(x) for x in kids[r]]

        # ---- flux pairs (initialize_faces!, src/Solver/Initialize.jl:5-158)
        geo = self.geometry
        ph, pt_, pd, prot, parea, pmid, pbc = [], [], [], [], [], [], []
        img_src, img_mid = [], []

        def add_image(src, m):
            img_src.append(src)
            img_mid.append(m)
            return n + len(img_src) - 1

        def face_area(c, d):
            a = 1.0
            first = True
            for e in FAT[dim][d]:
                a = ds[c, e] if first else a * ds[c, e]
                first = False
            return a

        for c in range(n):
            for f in range(2 * dim):
                d = f // 2
                rot = -1.0 if f % 2 else 1.0          # get_rot(faceid) = (-1)^(faceid-1)
                st = states[c, f]
                if st == 0:
                    m = mid[c].copy()
                    m[d] = m[d] - 0.5 * rot * ds[c, d]
                    ph.append(c); pt_.append(-1); pd.append(d); prot.append(rot)
                    parea.append(rot * face_area(c, d)); pmid.append(m); pbc.append(domain_bc[f])
                elif st == 1:
                    owner = (f % 2 == 1) != bool(wrap_face[c, f])
                    if not owner:
                        continue
                    t = nb_lists[c][f][0]
                    m = mid[c].copy()
                    m[d] = m[d] - 0.5 * rot * ds[c, d]
                    there = t
                    if m[d] == geo[2 * d] or m[d] == geo[2 * d + 1]:
                        tm = m.copy()
                        tm[d] = tm[d] - 0.5 * rot * ds[c, d]
                        there = add_image(t, tm)
                    ph.append(c); pt_.append(there); pd.append(d); prot.append(rot)
                    parea.append(face_area(c, d) * rot); pmid.append(m); pbc.append(-1)
                elif st == nfine:
                    nbs = nb_lists[c][f]
                    ms = []
                    for t in nbs:
                        m = mid[t].copy()
                        m[d] = mid[c, d] - 0.5 * rot * ds[c, d]
                        ms.append(m)
                    periodic = ms[0][d] == geo[2 * d] or ms[0][d] == geo[2 * d + 1]
                    for t, m in zip(nbs, ms):
                        there = t
                        if periodic:
                            tm = m.copy()
                            tm[d] = tm[d] - 0.5 * rot * ds[t, d]
                            there = add_image(t, tm)
                        ph.append(c); pt_.append(there); pd.append(d); prot.append(rot)
                        parea.append(face_area(c, d) / float(nfine) * rot)
                        pmid.append(m); pbc.append(-1)
                # st == -1: the coarse neighbour owns the hanging face

        nb_off = np.zeros(n * 2 * dim + 1, dtype=np.int64)
        flat = []
        for c in range(n):
            for f in range(2 * dim):
                lst = nb_lists[c][f]
                flat.extend(lst)
                nb_off[c * 2 * dim + f + 1] = nb_off[c * 2 * dim + f] + len(lst)

        img_src = np.asarray(img_src, dtype=np.int64)
        img_mid = np.asarray(img_mid, dtype=np.float64).reshape(-1, dim)
        tree_extent = (geo[1] - geo[0]) / self.trees_num[0]
        ps_level = np.round(np.log2(tree_extent / ds[:, 0])).astype(np.int32)   # cell_level
        topo = dict(
            dim=dim,
            n_cells=n,
            n_images=len(img_src),
            image_src=img_src,
            cell_mid=np.concatenate([mid, img_mid]) if len(img_src) else mid,
            cell_ds=np.concatenate([ds, ds[img_src]]) if len(img_src) else ds,
            bound_enc=np.zeros(n, dtype=np.int32),
            ps_level=ps_level,
            nb_state=states.reshape(-1).astype(np.int32),
            nb_offset=nb_off,
            nb_cell=np.asarray(flat, dtype=np.int64),
            n_pairs=len(ph),
            pair_here=np.asarray(ph, dtype=np.int64),
            pair_there=np.asarray(pt_, dtype=np.int64),
            pair_dir=np.asarray(pd, dtype=np.int32),
            pair_rot=np.asarray(prot, dtype=np.float64),
            pair_area=np.asarray(parea, dtype=np.float64),
            pair_mid=np.asarray(pmid, dtype=np.float64).reshape(-1, dim),
            pair_bc=np.asarray(pbc, dtype=np.int32),
        )
        return topo

    def _find_same(self, tabs, pts, levels):
        """_find for a per-point level array."""
        out = np.full(len(pts), -1, dtype=np.int64)
        levels = np.asarray(levels)
        for L in np.unique(levels):
            sel = levels == L
            if 0 <= L <= self.Lmax:
                out[sel] = self._find(tabs, pts[sel], int(L))
        return out
