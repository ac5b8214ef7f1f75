"""Velocity-space grids (VsData) on the host.

A grid is a regular root lattice of ``vs_trees_num`` cells on the ``quadrature`` box, each
root possibly refined in place into 2^DIM children in RMT order (x fastest), so a grid is
always sorted by ``cellkey = root_lin * 2^(DIM*Lmax) + morton(corner)``
(src/Velocity_space/Neighbor.jl:9-24). That order is the indexing contract that the
cross-grid walks of the reference rely on and that the device path turns into key-interval
overlaps.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np

# refine midpoint table RMT, src/Abstract/Types.jl:10 (child c, dimension d)
RMT = {
    2: np.array([[-1, -1], [1, -1], [-1, 1], [1, 1]], dtype=np.float64),
    3: np.array([[-1, -1, -1], [1, -1, -1], [-1, 1, -1], [1, 1, -1],
                 [-1, -1, 1], [1, -1, 1], [-1, 1, 1], [1, 1, 1]], dtype=np.float64),
}


def check_vs_setting(quadrature, vs_trees_num, dim):
    """check_vs_setting, src/Solver/Types.jl:335-353: v = 0 must be a root-grid corner."""
    if len(quadrature) != 2 * dim:
        raise ValueError(f"`quadrature` must have length 2*DIM = {2 * dim} (got {len(quadrature)})")
    if len(vs_trees_num) != dim:
        raise ValueError(f"`vs_trees_num` must have length DIM = {dim} (got {len(vs_trees_num)})")
    for d in range(dim):
        lo, hi, n = quadrature[2 * d], quadrature[2 * d + 1], vs_trees_num[d]
        if not hi > lo:
            raise ValueError(f"velocity-space dimension {d + 1}: `quadrature` range [{lo}, {hi}] must be increasing.")
        if n < 1:
            raise ValueError(f"velocity-space dimension {d + 1}: `vs_trees_num[{d + 1}]` = {n} must be >= 1.")
        k = -lo * n / (hi - lo)
        i = round(k)
        if not (0 <= i <= n and abs(k - i) <= 1e-8 * max(1, n)):
            raise ValueError(
                f"Invalid velocity-space setting in dimension {d + 1}: the origin v = 0 must lie on a "
                f"root-grid corner. With range [{lo}, {hi}] and `vs_trees_num[{d + 1}]` = {n}, 0 falls at "
                f"fractional cell index {k} (it must be an integer in 0:{n}).")


def _jl_range(lo, hi, n):
    """Midpoints of n root cells on [lo, hi] as Julia's range(lo+h, hi-h, n) yields them.

    Julia builds such a range from the exact rationals it recovers for the end points and
    rounds each element once; this evaluates the same rational and rounds it once.
    """
    flo, fhi = Fraction(lo), Fraction(hi)
    return np.array([float(flo + (2 * i + 1) * (fhi - flo) / (2 * n)) for i in range(n)])


class VelocityGrid:
    """One VsData grid: level, midpoint [n, DIM] and weight of each velocity cell."""

    def __init__(self, dim, quadrature, vs_trees_num, maxlevel=0):
        check_vs_setting(quadrature, vs_trees_num, dim)
        self.dim = dim
        self.quadrature = [float(q) for q in quadrature]
        self.vs_trees_num = [int(n) for n in vs_trees_num]
        self.maxlevel = int(maxlevel)
        # initialize_vs_data, src/Velocity_space/Initialize.jl:9-37
        self.ds = np.array([2 * ((self.quadrature[2 * d + 1] - self.quadrature[2 * d]) /
                                 (2 * self.vs_trees_num[d])) for d in range(dim)])
        ranges = [_jl_range(self.quadrature[2 * d], self.quadrature[2 * d + 1],
                            self.vs_trees_num[d]) for d in range(dim)]
        mesh = np.meshgrid(*ranges, indexing="ij")            # first dimension fastest
        self.midpoint = np.stack([m.reshape(-1, order="F") for m in mesh], axis=1)
        w = self.ds[0]
        for d in range(1, dim):
            w = w * self.ds[d]
        self.weight = np.full(len(self.midpoint), w)
        self.level = np.zeros(len(self.midpoint), dtype=np.int8)

    @property
    def n(self):
        return len(self.level)

    def copy(self):
        g = VelocityGrid.__new__(VelocityGrid)
        g.__dict__.update({k: (v.copy() if isinstance(v, np.ndarray) else v)
                           for k, v in self.__dict__.items()})
        g.quadrature = list(self.quadrature)
        g.vs_trees_num = list(self.vs_trees_num)
        return g

    def refine(self, flags):
        """refine_grid_stream!, src/Velocity_space/Rebuild.jl:44-85 (children in RMT order)."""
        flags = np.asarray(flags, dtype=bool)
        if (self.level[flags] >= self.maxlevel).any():
            raise ValueError("refinement beyond AMR_VS_MAXLEVEL")
        nc = 1 << self.dim
        counts = np.where(flags, nc, 1)
        src = np.repeat(np.arange(self.n), counts)
        first = np.repeat(np.cumsum(counts) - counts, counts)
        child = np.arange(len(src)) - first
        ref = flags[src]
        L = self.level[src].astype(np.int64)
        # child midpoint = parent + 0.5*(ds/2^(L+1))*RMT[child], rounded in that order
        step = 0.5 * (self.ds[None, :] / (2.0 ** (L + 1))[:, None])
        mid = self.midpoint[src]
        self.midpoint = np.where(ref[:, None], mid + step * RMT[self.dim][child % nc], mid)
        self.weight = np.where(ref, self.weight[src] / nc, self.weight[src])
        self.level = (L + ref).astype(np.int8)
        return self

    def keys(self):
        """cellkey of every cell (src/Velocity_space/Neighbor.jl:61-124)."""
        dim, L = self.dim, self.maxlevel
        lo = np.array([self.quadrature[2 * d] for d in range(dim)])
        h_fine = self.ds / 2.0 ** L
        cell = h_fine[None, :] * (2.0 ** (L - self.level.astype(np.int64)))[:, None]
        g = np.round((self.midpoint - 0.5 * cell - lo[None, :]) / h_fine[None, :]).astype(np.int64)
        stride = 1
        root_lin = np.zeros(self.n, dtype=np.int64)
        morton = np.zeros(self.n, dtype=np.int64)
        for d in range(dim):
            root_lin += (g[:, d] >> L) * stride
            stride *= self.vs_trees_num[d]
            q = g[:, d] & ((1 << L) - 1)
            for b in range(L):
                morton |= ((q >> b) & 1) << (b * dim + d)
        return (root_lin << (dim * L)) | morton

    def mid_colmajor(self):
        """midpoint in the Julia column-major [n, DIM] memory order (d-major)."""
        return np.ascontiguousarray(self.midpoint.T).reshape(-1)


def refine_toward(grid, prim, rounds, frac=0.3, rng=None):
    """Adapted grid for tests and workloads: refine, round by round, the cells that carry the
    largest share of the Maxwellian of prim (and a seeded random sprinkle), up to maxlevel."""
    rng = np.random.default_rng(rng)
    g = grid.copy()
    dim = g.dim
    for _ in range(rounds):
        c2 = np.sum((g.midpoint - np.asarray(prim[1:1 + dim])[None, :]) ** 2, axis=1)
        m = np.exp(-prim[dim + 1] * c2) * g.weight
        cut = np.quantile(m, 1.0 - frac)
        flags = (m >= cut) | (rng.random(g.n) < 0.05)
        flags &= g.level < g.maxlevel
        if not flags.any():
            break
        g.refine(flags)
    return g


# ---------------------------------------------------------------- initial adapted grids
MAXWELLIAN_INIT_FLOOR = 1e-3   # src/Velocity_space/Criteria.jl:123
VS_LOHNER_EPS_ABS = 1e-3       # src/Velocity_space/Criteria.jl:214


def _maxwellian_1d(c, du, lam):
    """_maxwellian_1d (src/Velocity_space/Criteria.jl:126-132): the integrals of exp(-lam s^2)
    and s^2 exp(-lam s^2) over [c - du/2, c + du/2], vectorised over c and du."""
    from scipy.special import erf
    sq = np.sqrt(lam)
    a = c - 0.5 * du
    b = c + 0.5 * du
    I0 = 0.5 * np.sqrt(np.pi / lam) * (erf(sq * b) - erf(sq * a))
    I2 = (a * np.exp(-lam * a * a) - b * np.exp(-lam * b * b)) / (2 * lam) + I0 / (2 * lam)
    return I0, I2


def maxwellian_refine_flags(grid, prim, K, mode="lohner", coeffi_lohner=0.6):
    """maxwellian_refine_flag (src/Velocity_space/Criteria.jl:165-198) of every cell of the
    grid, level < maxlevel: refine where the exact cell mass or energy of the Maxwellian of
    prim is above MAXWELLIAN_INIT_FLOOR of the total, or (lohner mode) where the midpoint rule
    misses the cell mass by more than coeffi_lohner (maxwellian_quad_error, :144-157)."""
    dim = grid.dim
    lam, rho = float(prim[-1]), float(prim[0])
    U = np.asarray(prim[1:1 + dim], dtype=np.float64)
    du = grid.ds[None, :] / (2.0 ** grid.level.astype(np.float64))[:, None]
    I0 = np.empty((grid.n, dim))
    I2 = np.empty((grid.n, dim))
    for d in range(dim):
        I0[:, d], I2[:, d] = _maxwellian_1d(grid.midpoint[:, d] - U[d], du[:, d], lam)
    pref = rho * (lam / np.pi) ** (dim / 2)
    prod0 = I0[:, 0].copy()
    for d in range(1, dim):
        prod0 *= I0[:, d]
    drho = pref * prod0
    dc2 = np.zeros(grid.n)
    for d in range(dim):
        p = np.ones(grid.n)
        for e in range(dim):
            if e != d:
                p *= I0[:, e]
        dc2 += I2[:, d] * p
    dc2 *= pref
    if K is not None:    # 2D2F: the b population carries K/(2 lam) of internal energy
        dE = 0.5 * (dc2 + (K / (2 * lam)) * drho)
        eint = rho * (dim + K) / (4 * lam)
    else:
        dE = 0.5 * dc2
        eint = rho * dim / (4 * lam)
    flags = np.maximum(drho / rho, dE / eint) > MAXWELLIAN_INIT_FLOOR
    if mode == "lohner":
        V = np.prod(du, axis=1)
        c2 = np.sum((grid.midpoint - U[None, :]) ** 2, axis=1)
        mmid = pref * np.exp(-lam * c2)
        err = np.abs(drho - mmid * V) / np.maximum(drho, VS_LOHNER_EPS_ABS * pref * V)
        flags |= err > coeffi_lohner
    return flags & (grid.level < grid.maxlevel)


def initial_vs_grid(base, prim, K, mode="lohner", coeffi_lohner=0.6):
    """initialize_vs_data's adapted grid (src/Velocity_space/Initialize.jl:31-34): maxlevel
    rounds of initial_vs_adaptive_mesh_refinement! (src/Velocity_space/AMR.jl:183-200) on the
    Maxwellian of prim. K is the gas's internal degrees of freedom for 2D2F, None for 3D1F."""
    g = base.copy()
    for _ in range(g.maxlevel):
        fl = maxwellian_refine_flags(g, prim, K, mode, coeffi_lohner)
        if not fl.any():
            break
        g.refine(fl)
    return g


def refine_to_target(grid, prim, n_target, rng, n_cap=None):
    """Seeded refinement of a grid toward the Maxwellian peak of prim until it holds at least
    n_target cells (or no cell can be refined, or one more split would pass n_cap): round by
    round the unrefined cells carrying the largest Maxwellian mass (with a seeded 25 % jitter)
    are split, as many as the missing count needs. A workload generator, not a reference
    criterion."""
    rng = np.random.default_rng(rng)
    g = grid.copy()
    dim = g.dim
    nch = (1 << dim) - 1
    U = np.asarray(prim[1:1 + dim], dtype=np.float64)
    lam = float(prim[-1])
    while g.n < n_target:
        cand = g.level < g.maxlevel
        if not cand.any():
            break
        c2 = np.sum((g.midpoint - U[None, :]) ** 2, axis=1)
        pri = np.exp(-lam * c2) * g.weight * (1.0 + 0.25 * rng.random(g.n))
        pri[~cand] = -1.0
        k = min(int(cand.sum()), -(-(n_target - g.n) // nch))
        if n_cap is not None:
            k = min(k, (int(n_cap) - g.n) // nch)
        if k <= 0:
            break
        sel = np.argpartition(-pri, k - 1)[:k]
        flags = np.zeros(g.n, dtype=bool)
        flags[sel] = True
        g.refine(flags)
    return g
