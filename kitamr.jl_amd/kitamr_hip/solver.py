"""Host-side mirror of KitAMR.jl's solver API for the per-step phase-space update.

The names follow the reference (src/Solver/Types.jl, src/Gas/Types.jl,
src/Boundary/Types.jl, src/Solver/Solver.jl): ``Solver``, ``Gas``, ``Configure``,
``Domain``, ``Uniform``, ``PCoordFn``, ``initialize``, and the per-step calls ``slope``,
``flux``, ``iterate`` (Julia's ``slope!``, ``flux!``, ``iterate!``), ``limit_dt``
(``limit_Δt!``) and ``solve`` (``solve!``). The step itself runs in the HIP library
through the C ABI; the host keeps the forest, the face list and the status counters.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import abi
from .forest import Forest
from .vspace import VelocityGrid, check_vs_setting

PI = math.pi


# ---------------------------------------------------------------- reference types
def ref_vhs_vis(Kn, omega, alpha=1.0, T_ref=1.0):
    """ref_vhs_vis(Kn, omega, alpha, T_ref), src/Gas/Model.jl:5-9."""
    mu0 = 5.0 * (alpha + 1.0) * (alpha + 2.0) * math.sqrt(PI) / (
        4.0 * alpha * (5.0 - 2.0 * omega) * (7.0 - 2.0 * omega)) * Kn
    return mu0 * T_ref ** (0.5 - omega)


@dataclass
class Gas:
    """Gas, src/Gas/Types.jl:18-51. The default mu_ref keeps the reference's call
    ``ref_vhs_vis(Kn, αᵣ, ωᵣ)`` whose arguments land in (omega, alpha)."""
    Kn: float = 0.05
    Pr: float = 2.0 / 3.0
    K: float = 1.0
    gamma: float = 5.0 / 3.0
    omega: float = 0.5
    alpha_ref: float = 1.0
    omega_ref: float = 0.81
    mu_ref: Optional[float] = None
    T_ref: float = 1.0

    def __post_init__(self):
        if self.mu_ref is None:
            self.mu_ref = ref_vhs_vis(self.Kn, self.alpha_ref, self.omega_ref)


@dataclass
class Solver:
    """Solver, src/Solver/Types.jl:59-149 (the fields the per-step path reads)."""
    DIM: int
    NDF: int
    AMR_PS_MAXLEVEL: int
    AMR_VS_MAXLEVEL: int
    flux: str = "CAIDVM"
    time_marching: str = "CAIDVM_Marching"
    CFL: float = 0.4
    PS_DYNAMIC_AMR: bool = False
    VS_DYNAMIC_AMR: bool = False
    TOLERANCE: float = 1e-6
    ST_CHECK_INTERVAL: int = 100
    REDUNDANT_STEPS_NUM: int = 100
    max_sim_time: float = math.inf

    def __post_init__(self):
        if self.flux not in abi.FLUX:
            raise ValueError(f"flux must be one of {list(abi.FLUX)}")
        if self.time_marching not in abi.MARCH:
            raise ValueError(f"time_marching must be one of {list(abi.MARCH)}")
        if (self.DIM, self.NDF) not in ((2, 2), (3, 1)):
            raise ValueError("supported (DIM, NDF): (2, 2) and (3, 1)")


@dataclass
class Domain:
    """Domain{T}(id, bc), src/Boundary/Types.jl:12-21; kind is the boundary-condition type."""
    kind: str
    id: int
    bc: Optional[Sequence[float]] = None


@dataclass
class Uniform:
    ic: Sequence[float]


@dataclass
class PCoordFn:
    fn: Callable  # fn(midpoint) -> prim


@dataclass
class PCoordField:
    """A PCoordFn evaluated on all cells at once: fn(midpoints [n, DIM]) -> prim [n, DIM+2]."""
    fn: Callable


@dataclass
class Configure:
    solver: Solver
    gas: Gas
    geometry: Sequence[float]
    trees_num: Sequence[int]
    quadrature: Sequence[float]
    vs_trees_num: Sequence[int]
    IC: object
    domain: List[Domain]
    precision: str = "fp64"


# ---------------------------------------------------------------- KitCore (host, numpy)
def get_conserved(prim, gamma):
    """get_conserved_2D / 3D, lib/KitCore/2D.jl:1-8, 3D.jl:1-11 (rows are cells)."""
    prim = np.atleast_2d(prim)
    dim = prim.shape[1] - 2
    w = np.empty_like(prim)
    w[:, 0] = prim[:, 0]
    for d in range(dim):
        w[:, 1 + d] = prim[:, 0] * prim[:, 1 + d]
    v2 = prim[:, 1] * prim[:, 1]
    for d in range(1, dim):
        v2 = v2 + prim[:, 1 + d] * prim[:, 1 + d]
    w[:, dim + 1] = 0.5 * prim[:, 0] / prim[:, dim + 1] / (gamma - 1.0) + 0.5 * prim[:, 0] * v2
    return w


def discrete_maxwell(mid, prim, K):
    """discrete_maxwell_2D2F / 3D1F, lib/KitCore/2D2F.jl:1-14, 3D1F.jl:1-13.
    mid: [n, DIM]; returns [NDF, n]."""
    dim = mid.shape[1]
    c2 = (mid[:, 0] - prim[1]) * (mid[:, 0] - prim[1])
    for d in range(1, dim):
        c2 = c2 + (mid[:, d] - prim[1 + d]) * (mid[:, d] - prim[1 + d])
    if dim == 2:
        h = prim[0] * (prim[3] / PI) * np.exp(-prim[3] * c2)
        return np.stack([h, h * K / (2.0 * prim[3])])
    return (prim[0] * (prim[4] / PI) ** 1.5 * np.exp(-prim[4] * c2))[None, :]


def discrete_maxwell_cells(grids, prim, K, ndf):
    """discrete_maxwell of every cell's prim on its own grid, concatenated in cell order
    ([NDF, n_c] blocks), evaluated in chunks of cells with the same arithmetic as
    discrete_maxwell."""
    out = []
    n = len(grids)
    c0 = 0
    while c0 < n:
        c1 = c0 + 1
        tot = grids[c0].n
        while c1 < n and tot < (1 << 24):
            tot += grids[c1].n
            c1 += 1
        sizes = np.array([grids[c].n for c in range(c0, c1)])
        mid = np.concatenate([grids[c].midpoint for c in range(c0, c1)])
        p = np.repeat(prim[c0:c1], sizes, axis=0)
        dim = mid.shape[1]
        c2 = (mid[:, 0] - p[:, 1]) * (mid[:, 0] - p[:, 1])
        for d in range(1, dim):
            c2 = c2 + (mid[:, d] - p[:, 1 + d]) * (mid[:, d] - p[:, 1 + d])
        if dim == 2:
            h = p[:, 0] * (p[:, 3] / PI) * np.exp(-p[:, 3] * c2)
            blocks = np.stack([h, h * K / (2.0 * p[:, 3])])
        else:
            blocks = (p[:, 0] * (p[:, 4] / PI) ** 1.5 * np.exp(-p[:, 4] * c2))[None, :]
        ends = np.cumsum(sizes)
        for k in range(len(sizes)):
            out.append(blocks[:, ends[k] - sizes[k]:ends[k]].reshape(-1))
        c0 = c1
    return np.concatenate(out) if out else np.zeros(0)


# ---------------------------------------------------------------- case construction
@dataclass
class Case:
    """Flattened topology + initial state of one flow case."""
    config: Configure
    topo: dict
    w: np.ndarray
    prim: np.ndarray
    df: np.ndarray
    dt_xi: float
    forest: Forest = None
    grids: list = field(default_factory=list)

    @property
    def n_cells(self):
        return int(self.topo["n_cells"])

    @property
    def phase_cells(self):
        off = self.topo["vs_offset"]
        benc = self.topo["bound_enc"]
        n = np.diff(off)
        return int(n[benc >= 0].sum())


def dt_xi(config):
    """Status(config).Δt_ξ, src/Solver/Types.jl:509-520."""
    s = config.solver
    dim = s.DIM
    ds = [(config.geometry[2 * i + 1] - config.geometry[2 * i]) / config.trees_num[i] /
          2 ** s.AMR_PS_MAXLEVEL for i in range(dim)]
    q = config.quadrature
    U = [max(q[2 * i + 1], abs(q[2 * i])) - (q[2 * i + 1] - q[2 * i]) / config.vs_trees_num[i] /
         2 ** s.AMR_VS_MAXLEVEL / 2 for i in range(dim)]
    return s.CFL * min(ds[i] / U[i] for i in range(dim))


def build_case(config, ps_refine=None, vs_grid_fn=None, ps_max_level=None, vs_grids_fn=None):
    """The flattened state initialize(config) would hand to the device.

    ps_refine(mid, ds, level) -> bool[] marks leaves to refine (static pre-refinement, then
    FULL 2:1 balance); vs_grid_fn(cell, prim, base_grid) -> VelocityGrid gives adapted
    per-cell velocity grids (None: every cell carries the base grid); vs_grids_fn(prim [n,
    DIM+2], mid [n, DIM], base_grid) -> list of VelocityGrid does the same for all cells at
    once (for large meshes).
    """
    s = config.solver
    dim = s.DIM
    check_vs_setting(config.quadrature, config.vs_trees_num, dim)
    periodic = []
    for d in range(dim):
        a, b = config.domain[2 * d].kind, config.domain[2 * d + 1].kind
        periodic.append(a == "Period" and b == "Period")
    lmax = s.AMR_PS_MAXLEVEL if ps_max_level is None else ps_max_level
    forest = Forest(dim, config.trees_num, config.geometry, periodic, max_level=lmax)
    if ps_refine is not None and lmax > 0:
        forest.refine(ps_refine)
        forest.balance()
    bc_kind, bc_prim, dom_bc = [], [], []
    for f in range(2 * dim):
        dm = config.domain[f]
        if dm.kind == "Period":
            dom_bc.append(-1)
            continue
        if dm.kind not in abi.BC:
            raise ValueError(f"unsupported domain boundary condition {dm.kind}")
        dom_bc.append(len(bc_kind))
        bc_kind.append(abi.BC[dm.kind])
        bc_prim.append(list(dm.bc) if dm.bc is not None else [0.0] * (dim + 2))
    topo = forest.build_topology(dom_bc)
    n = topo["n_cells"]
    mid = topo["cell_mid"][:n]
    # initial primitive state, initialize_ps! (src/Solver/Initialize.jl:343-380)
    if isinstance(config.IC, Uniform):
        prim = np.tile(np.asarray(config.IC.ic, dtype=np.float64), (n, 1))
    elif isinstance(config.IC, PCoordField):
        prim = np.ascontiguousarray(config.IC.fn(np.asarray(mid)), dtype=np.float64)
    else:
        prim = np.array([config.IC.fn(mid[c]) for c in range(n)], dtype=np.float64)
    w = get_conserved(prim, config.gas.gamma)
    base = VelocityGrid(dim, config.quadrature, config.vs_trees_num, s.AMR_VS_MAXLEVEL)
    ndf = s.NDF
    grids = []
    if vs_grid_fn is None and vs_grids_fn is None:
        n0 = base.n
        topo["vs_uniform"] = 1
        topo["vs_offset"] = np.arange(n + 1, dtype=np.int64) * n0
        topo["vs_level"] = base.level
        topo["vs_mid"] = base.mid_colmajor()
        topo["vs_weight"] = base.weight
        # f per cell: [NDF, n0] blocks in cell order
        df = np.empty((n, ndf, n0))
        uniq, inv = np.unique(prim, axis=0, return_inverse=True)
        for k, p in enumerate(uniq):
            df[inv.reshape(-1) == k] = discrete_maxwell(base.midpoint, p, config.gas.K)
        df = df.reshape(-1)
    elif vs_grids_fn is not None:
        grids = vs_grids_fn(prim, np.asarray(mid), base)
        sizes = np.array([g.n for g in grids], dtype=np.int64)
        topo["vs_uniform"] = 0
        topo["vs_offset"] = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        topo["vs_level"] = np.concatenate([g.level for g in grids]).astype(np.int8)
        topo["vs_mid"] = np.concatenate([g.mid_colmajor() for g in grids])
        topo["vs_weight"] = np.concatenate([g.weight for g in grids])
        df = discrete_maxwell_cells(grids, prim, config.gas.K, ndf)
    else:
        off = [0]
        lev, mids, wts, dfs = [], [], [], []
        for c in range(n):
            g = vs_grid_fn(c, prim[c], base)
            grids.append(g)
            off.append(off[-1] + g.n)
            lev.append(g.level)
            mids.append(g.mid_colmajor())
            wts.append(g.weight)
            dfs.append(discrete_maxwell(g.midpoint, prim[c], config.gas.K).reshape(-1))
        topo["vs_uniform"] = 0
        topo["vs_offset"] = np.asarray(off, dtype=np.int64)
        topo["vs_level"] = np.concatenate(lev).astype(np.int8)
        topo["vs_mid"] = np.concatenate(mids)
        topo["vs_weight"] = np.concatenate(wts)
        df = np.concatenate(dfs)
    topo["vs_maxlevel"] = s.AMR_VS_MAXLEVEL
    topo["vs_trees_num"] = list(config.vs_trees_num)
    topo["vs_lo"] = [config.quadrature[2 * d] for d in range(dim)]
    topo["vs_hi"] = [config.quadrature[2 * d + 1] for d in range(dim)]
    topo["bc_kind"] = np.asarray(bc_kind, dtype=np.int32)
    topo["bc_prim"] = np.asarray(bc_prim, dtype=np.float64).reshape(-1)
    return Case(config, topo, w.reshape(-1), prim.reshape(-1), df, dt_xi(config), forest, grids)


def kah_params(config, device=0):
    s, g = config.solver, config.gas
    return abi.make_params(s.DIM, s.NDF, flux=s.flux, marching=s.time_marching,
                           precision=config.precision, K=g.K, gamma=g.gamma, Pr=g.Pr,
                           mu_ref=g.mu_ref, omega=g.omega,
                           st_check_interval=s.ST_CHECK_INTERVAL, device=device)


# ---------------------------------------------------------------- the KA object and the loop
@dataclass
class Residual:
    """Residual, src/Solver/Types.jl (sumRes/sumAvg accumulators and the residual)."""
    sumRes: np.ndarray
    sumAvg: np.ndarray
    residual: np.ndarray
    redundant_step: int = 0


@dataclass
class Status:
    dt: float
    dt_xi: float
    sim_time: float = 0.0
    step: int = 0
    residual: Residual = None


class KA:
    """Host handle of one case on the device (the reference's KA for this path)."""

    def __init__(self, case, device=0, options=None):
        self.case = case
        self.config = case.config
        self.ctx = abi.HipContext(kah_params(case.config, device))
        for k, v in (options or {}).items():
            self.ctx.set_option(k, v)
        self.ctx.set_topology(case.topo)
        self.ctx.set_state(case.w, case.prim, case.df)
        nw = case.config.solver.DIM + 2
        self.status = Status(case.dt_xi, case.dt_xi,
                             residual=Residual(np.zeros(nw), np.zeros(nw), np.ones(nw)))

    def close(self):
        self.ctx.close()


def initialize(config, device=0, **kw):
    """initialize(config), src/Solver/Initialize.jl:554-573: forest, faces, state on device."""
    return KA(build_case(config, **kw), device)


def limit_dt(ka):
    """limit_Δt!, src/Theory/Iterate.jl:27-39 (no animation output on this path)."""
    st = ka.status
    dt = st.dt_xi
    dt_end = ka.config.solver.max_sim_time - st.sim_time
    if dt_end > 0.0:
        dt = min(dt, dt_end)
    st.dt = dt


def slope(ka):
    """slope!(p4est, ka), src/Flux/Slope.jl:1047-1070."""
    ka.ctx.slope(ka.status.dt)


def flux(ka):
    """flux!(p4est, ka), src/Flux/Flux.jl:458-487."""
    ka.ctx.flux(ka.status.dt)


def _after_iterate(ka, r, a):
    st = ka.status
    s = ka.config.solver
    if st.step % s.ST_CHECK_INTERVAL == 0:
        # residual_comm!, src/Solver/Finalize.jl:12-24 (N = all quadrants)
        N = ka.case.n_cells
        st.residual.residual = np.sqrt(r * N) / (a + 1e-12)
    if s.time_marching == "Euler":
        st_dt = st.dt
        st.dt = st.dt_xi
    else:
        st_dt = st.dt
    st.step += 1
    st.sim_time += st_dt


def iterate(ka):
    """iterate!(p4est, ka), src/Theory/Iterate.jl:5-16."""
    r, a = ka.ctx.iterate(ka.status.dt, ka.status.step)
    _after_iterate(ka, r, a)


def step(ka):
    """slope! + flux! + iterate! of one step through the fused device path."""
    r, a = ka.ctx.step(ka.status.dt, ka.status.step)
    _after_iterate(ka, r, a)


def check_for_convergence(ka):
    """check_for_convergence, src/Solver/Finalize.jl:79-82."""
    res = ka.status.residual
    if np.max(res.residual) < ka.config.solver.TOLERANCE:
        res.redundant_step += 1
    return res.redundant_step > ka.config.solver.REDUNDANT_STEPS_NUM


def reached_max_time(ka):
    return ka.status.sim_time >= ka.config.solver.max_sim_time * (1 - 1e-12)


def solve(ka, max_steps=None, fused=True):
    """solve!(p4est, ka), src/Solver/Solver.jl:44-87, without adaptation steps."""
    n = 0
    while True:
        limit_dt(ka)
        if fused:
            step(ka)
        else:
            slope(ka)
            flux(ka)
            iterate(ka)
        n += 1
        if check_for_convergence(ka) or reached_max_time(ka):
            break
        if max_steps is not None and n >= max_steps:
            break
    return ka
