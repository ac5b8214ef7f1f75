"""Pins of the oracle's CAIDVM limiter, Euler marching, inflow/outflow domain fluxes and
Löhner velocity-space criterion from what the mathematics fixes (no retyped formulas):

- CAIDVM on linear fields (src/Flux/CAIDVM.jl:99-141): for f = a(u) + b(u).x with slopes that
  are exact and f far above the slopes' reach, the limiter is 1 and every face value is the
  field at the departure point x_f - u dt, so each cell's micro flux is the divergence of a
  linear field, -V u.b, and its macroscopic flux the moments of that; with the
  InterpolatedOutflow ghost (CAIDVM.jl:71-94) the boundary cells obey it too (the mirrored
  ghost extrapolates a linear field exactly), with UniformOutflow (:53-67) they do not.
- the limiter itself (positivity_preserving_reconstruct, CAIDVM.jl:127-141): on a two-cell
  strip the face value of each upwind velocity is read back from the cells' micro fluxes;
  where f is below the reach of the slope over the cell (half the sum of |slope| times the
  cell size) the limited reconstruction touches zero at the cell's corner, so the effective
  limiter theta satisfies f - theta * reach = 0 (up to the eps/EPS offsets), lies in (0, 1),
  keeps the face value non-negative and sends it to f as f -> 0; elsewhere theta = 1.
- iterate!(Euler) (src/Theory/Iterate.jl:131-162): a uniform periodic state f = M + d, with d
  even in u and without mass or energy, keeps w, and d decays by exactly tau/(tau + dt) per
  step (the BGK-Shakhov target of a state with zero heat flux is M itself).
- UniformOutflow and SupersonicInflow (CAIDVM.jl:28-67): a uniform state stays uniform; an
  inflow at the interior state changes nothing; an inflow at another state changes the mass
  of each cell next to it by dt/V * area * sum over incoming velocities of w u_n (M_bc - f),
  and no other cell's.
- the Löhner indicator (src/Velocity_space/Criteria.jl:219-290, refine flag of
  src/Velocity_space/AMR.jl:26-60 with the contribution floor switched off): a field linear in
  u has zero second differences, so at threshold 0 only the cells next to the vacuum edge of
  the grid are flagged, also across coarse/fine velocity levels (non-uniform weights); a
  one-cell spike of height A has eta = 2A / (2A (1 + 1e-2) + 2 A 1e-3) at the spike and
  A / (A (1 + 1e-2) + 2 A 1e-3) at its face neighbours, which thresholds just above and below
  each value separate exactly.
- the contribution criteria (contribution_refine_flag, Criteria.jl:4-6, 17-24, 50-55): with
  h constant on a uniform grid, b = 0 and U = 0, a cell's share of the mass is 1/N and its share
  of the internal energy |u|^2 / sum |u|^2; the local criterion flags exactly the cells whose
  larger share exceeds ADAPT_COEFFI_VS_LOCAL, and the global one the cells whose mass or
  energy content exceeds ADAPT_COEFFI_VS_GLOBAL times Velocity_Resolution, so thresholds
  between the distinct shares of an 8x8 grid select all cells, the 32 outer ones or the 4
  corners.
"""
import math

import numpy as np
import pytest

import oracle as O
from kitamr_hip import abi, cases
from kitamr_hip.solver import (Configure, Domain, Gas, Solver, Uniform, build_case,
                               discrete_maxwell, kah_params)
from kitamr_hip.vspace import VelocityGrid

MACH_EPS = np.finfo(np.float64).eps


def _box(n=(6, 6), nv=8, vmax=4.0, bcs=("UniformOutflow",) * 4, bc_prim=None,
         marching="CAIDVM_Marching", geometry=(0.0, 1.0, 0.0, 1.0), Kn=0.05, prim0=None):
    solver = Solver(DIM=2, NDF=2, AMR_PS_MAXLEVEL=0, AMR_VS_MAXLEVEL=0, flux="CAIDVM",
                    time_marching=marching, CFL=0.5)
    gas = Gas(K=1.0, gamma=5.0 / 3.0, Pr=2.0 / 3.0, Kn=Kn, omega=0.81, omega_ref=0.81)
    dom = []
    for i, k in enumerate(bcs):
        if k in ("UniformOutflow", "InterpolatedOutflow", "Period"):
            dom.append(Domain(k, i + 1))
        else:
            dom.append(Domain(k, i + 1, list(bc_prim)))
    cfg = Configure(solver, gas, list(geometry), list(n), [-vmax, vmax, -vmax, vmax],
                    [nv, nv], Uniform(prim0 or [1.0, 0.0, 0.0, 1.0]), dom)
    return build_case(cfg)


def _grid(case):
    t = case.topo
    n0 = int(t["vs_offset"][1])
    mid = np.asarray(t["vs_mid"]).reshape(2, n0).T
    return mid, np.asarray(t["vs_weight"])[:n0]


def _moments(u, wt, h, b):
    return np.array([np.sum(wt * h), np.sum(wt * u[:, 0] * h), np.sum(wt * u[:, 1] * h),
                     0.5 * np.sum(wt * ((u[:, 0] ** 2 + u[:, 1] ** 2) * h + b))])


def _set(case, df):
    """Oracle with df [n_cells, 2, n0] and w/prim its moments."""
    n = case.n_cells
    u, wt = _grid(case)
    w = np.array([_moments(u, wt, df[c, 0], df[c, 1]) for c in range(n)])
    prim = np.array([O.get_prim(2, w[c], case.config.gas.gamma) for c in range(n)])
    o = O.Oracle(kah_params(case.config), case.topo)
    o.set_state(w.reshape(-1), prim.reshape(-1), df.reshape(-1))
    return o, w


# ------------------------------------------------------------------ CAIDVM on linear fields
def _linear_field(case, seed=3):
    rng = np.random.default_rng(seed)
    u, wt = _grid(case)
    n0 = len(wt)
    n = case.n_cells
    mid = np.asarray(case.topo["cell_mid"])[:2 * n].reshape(n, 2)
    a = rng.uniform(50.0, 60.0, (2, n0))            # far above the slopes' reach
    b = rng.uniform(-1.0, 1.0, (2, 2, n0))          # [dim, k, v]
    df = a[None] + np.einsum("cd,dkv->ckv", mid, b)
    return df, b, u, wt, mid


@pytest.mark.parametrize("bc", ["InterpolatedOutflow", "UniformOutflow"])
def test_caidvm_linear_field_flux_is_the_divergence(bc):
    case = _box(n=(6, 6), bcs=(bc,) * 4)
    df, b, u, wt, mid = _linear_field(case)
    o, _ = _set(case, df)
    o.slope()
    o.flux(0.01)
    psf, vf = o.get_flux()
    n = case.n_cells
    V = np.prod(np.asarray(case.topo["cell_ds"])[:2 * n].reshape(n, 2), axis=1)
    vf = vf.reshape(n, 2, -1)
    psf = psf.reshape(n, 4)
    # -V u.b per velocity cell and component
    expect = -V[:, None, None] * (u[:, 0][None, None] * b[0][None] +
                                  u[:, 1][None, None] * b[1][None])
    scale = np.max(np.abs(expect))
    inner = ((mid[:, 0] > 1 / 6) & (mid[:, 0] < 5 / 6) & (mid[:, 1] > 1 / 6) &
             (mid[:, 1] < 5 / 6))
    err = np.max(np.abs(vf - expect), axis=(1, 2)) / scale
    assert err[inner].max() < 1e-12, err[inner].max()
    if bc == "InterpolatedOutflow":
        assert err.max() < 1e-12, err.max()          # boundary cells too
    else:
        assert err[~inner].max() > 1e-3              # the copied ghost is not an extrapolation
    # macroscopic flux = the moments of the micro flux (both sides' half fluxes)
    for c in np.flatnonzero(inner if bc == "UniformOutflow" else np.ones(n, bool)):
        m = _moments(u, wt, vf[c, 0], vf[c, 1])
        assert np.max(np.abs(psf[c] - m)) <= 1e-11 * np.max(np.abs(m)) + 1e-14, c


def test_caidvm_limiter_touches_zero_at_the_corner():
    """Two cells on [0,2]x[0,1], UniformOutflow around: the left cell's micro flux is
    area*u*f_L through its outflow face (the ghost copies f, so both signs) minus
    area*u*f_face through the shared face, the y faces cancelling; so f_face per velocity."""
    case = _box(n=(2, 1), nv=8, geometry=(0.0, 2.0, 0.0, 1.0))
    u, wt = _grid(case)
    n0 = len(wt)
    rng = np.random.default_rng(11)
    # slopes along x and y; f of the upwind cell from far above the reach to tiny
    sx = rng.uniform(-2.0, 2.0, (2, 2, n0))          # [cell, k, v]
    sy = rng.uniform(-0.5, 0.5, (2, 2, n0))
    reach = 0.5 * (1.0 * np.abs(sx) + 1.0 * np.abs(sy))
    frac = np.exp(rng.uniform(np.log(1e-6), np.log(3.0), (2, 2, n0)))
    df = reach * frac
    o, _ = _set(case, df)
    sdf = np.zeros((2, 2, 2, n0))                    # [cell, dim, k, v]
    sdf[:, 0] = sx
    sdf[:, 1] = sy
    o.set_sdf(sdf.reshape(-1))
    dt = 0.02
    o.flux(dt)
    _, vf = o.get_flux()
    vf = vf.reshape(2, 2, n0)
    area = 1.0
    ux, uy = u[:, 0], u[:, 1]
    # left cell 0: outflow face at x=0 contributes +area*ux*f0 for every u (in and out),
    # shared face at x=1 contributes -area*ux*f_face; y faces: -area*uy*f0 + area*uy*f0 = 0
    f_face = (area * ux * df[0] - vf[0]) / (area * ux)
    for side, (cell, sel, xf_minus_xc) in enumerate([(0, ux > 0, 0.5), (1, ux < 0, -0.5)]):
        for k in range(2):
            f = df[cell, k, sel]
            dx = xf_minus_xc - ux[sel] * dt
            dy = -uy[sel] * dt
            s_dx = sx[cell, k, sel] * dx + sy[cell, k, sel] * dy
            theta = (f_face[k, sel] - f) / s_dx
            r = reach[cell, k, sel]
            big = f - MACH_EPS >= r + 1e-10 * r + 1e-12
            assert np.all(np.abs(theta[big] - 1.0) < 1e-9), "theta != 1 above the reach"
            small = f < r * (1 - 1e-9)
            assert small.sum() > 10
            t = theta[small]
            assert np.all((t > 0) & (t < 1))
            # the limited reconstruction is zero at the cell corner (up to eps, EPS)
            corner = f[small] - t * r[small]
            assert np.all(np.abs(corner) <= 1e-10 * np.maximum(f[small], r[small]) + 1e-13)
    # x-only slopes keep every face value non-negative; f -> 0 gives f_face -> f
    df2 = df.copy()
    df2[0, :, :4] = 1e-14
    o2, _ = _set(case, df2)
    sdf2 = sdf.copy()
    sdf2[:, 1] = 0.0
    o2.set_sdf(sdf2.reshape(-1))
    o2.flux(dt)
    _, vf2 = o2.get_flux()
    vf2 = vf2.reshape(2, 2, n0)
    ff = (area * ux * df2[0] - vf2[0]) / (area * ux)
    pos = ux > 0
    assert np.all(ff[:, pos] >= -1e-15)
    tiny = np.zeros(n0, bool)
    tiny[:4] = True
    sel = pos & tiny
    if sel.any():
        assert np.all(np.abs(ff[:, sel] - df2[0][:, sel]) <= 1e-12)


# ------------------------------------------------------------------ Euler
def test_euler_fixed_point_and_exact_decay():
    prim0 = [1.3, 0.0, 0.0, 1.0 / 1.2]
    case = _box(n=(4, 4), nv=12, vmax=6.0, bcs=("Period",) * 4, marching="Euler",
                prim0=prim0)
    u, wt = _grid(case)
    n0 = len(wt)
    K = case.config.gas.K
    M = discrete_maxwell(u, prim0, K)                   # [2, n0]
    # an even perturbation (u -> -u on the symmetric grid) with no mass or internal energy
    r2 = u[:, 0] ** 2 + u[:, 1] ** 2
    g = np.exp(-0.5 * r2) * np.cos(1.3 * u[:, 0]) * np.cos(0.7 * u[:, 1])
    dh = 1e-3 * g
    db = 1e-3 * g * 0.4
    basis = np.stack([np.ones(n0), r2])
    # remove mass and energy: dh += c0 + c1 r2 fitted with weights
    m = np.array([np.sum(wt * dh), np.sum(wt * (r2 * dh + db))])
    A = np.array([[np.sum(wt), np.sum(wt * r2)],
                  [np.sum(wt * r2), np.sum(wt * r2 * r2)]])
    c = np.linalg.solve(A, -m)
    dh = dh + c @ basis
    n = case.n_cells
    df = np.broadcast_to(np.stack([M[0] + dh, M[1] + db]), (n, 2, n0)).copy()
    o, w0 = _set(case, df)
    # the target Maxwellian is the one of the discrete moments (at rest: the grid is even)
    prim = O.get_prim(2, w0[0], case.config.gas.gamma)
    assert abs(prim[1]) < 1e-15 and abs(prim[2]) < 1e-15
    tau = O.get_tau(2, prim, case.config.gas.mu_ref, case.config.gas.omega)
    dt = case.dt_xi
    Mp = discrete_maxwell(u, prim, K)
    d_h, d_b = df[0, 0] - Mp[0], df[0, 1] - Mp[1]
    for s in range(3):
        o.step(dt, s)
        st = o.get_state()
        f = st["df"].reshape(n, 2, n0)
        fac = (tau / (tau + dt)) ** (s + 1)
        for cc in range(n):
            assert np.max(np.abs(f[cc, 0] - Mp[0] - fac * d_h)) <= 1e-13 * M[0].max()
            assert np.max(np.abs(f[cc, 1] - Mp[1] - fac * d_b)) <= 1e-13 * M[1].max()
        assert np.allclose(st["w"].reshape(n, 4), w0, rtol=1e-13, atol=1e-15)


# ------------------------------------------------------------------ inflow / outflow
def test_uniform_outflow_keeps_a_uniform_state():
    # a grid wide and fine enough that the discrete moments of the Maxwellian are its
    # continuous ones to rounding, so the Maxwellian is the relaxation's fixed point
    prim0 = [0.8, 0.4, -0.3, 1.0 / 0.9]
    case = _box(n=(5, 5), nv=32, vmax=8.0, prim0=prim0)
    o = O.Oracle(kah_params(case.config), case.topo)
    o.set_state(case.w, case.prim, case.df)
    for s in range(3):
        o.step(case.dt_xi, s)
    st = o.get_state()
    assert np.max(np.abs(st["df"] - case.df)) <= 1e-14 * np.max(case.df)
    assert np.allclose(st["w"], case.w, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("which", [0, 1, 2, 3])
def test_supersonic_inflow_at_the_interior_state_changes_nothing(which):
    prim0 = [1.1, 0.6, 0.2, 1.0 / 1.3]
    bcs = ["UniformOutflow"] * 4
    bcs[which] = "SuperSonicInflow"
    case = _box(n=(5, 5), nv=32, vmax=8.0, bcs=bcs, bc_prim=prim0, prim0=prim0)
    o = O.Oracle(kah_params(case.config), case.topo)
    o.set_state(case.w, case.prim, case.df)
    for s in range(3):
        o.step(case.dt_xi, s)
    st = o.get_state()
    assert np.max(np.abs(st["df"] - case.df)) <= 1e-13 * np.max(case.df)
    assert np.allclose(st["w"], case.w, rtol=1e-13, atol=1e-15)


def test_supersonic_inflow_mass_change_is_the_incoming_flux_difference():
    prim0 = [1.0, 0.3, 0.0, 1.0]
    pbc = [1.7, 1.2, 0.1, 1.0 / 1.4]
    case = _box(n=(5, 5), nv=32, vmax=8.0,
                bcs=("SuperSonicInflow", "UniformOutflow", "UniformOutflow", "UniformOutflow"),
                bc_prim=pbc, prim0=prim0)
    u, wt = _grid(case)
    K = case.config.gas.K
    f = discrete_maxwell(u, prim0, K)
    Mb = discrete_maxwell(u, pbc, K)
    n = case.n_cells
    mid = np.asarray(case.topo["cell_mid"])[:2 * n].reshape(n, 2)
    ds = np.asarray(case.topo["cell_ds"])[:2 * n].reshape(n, 2)
    o = O.Oracle(kah_params(case.config), case.topo)
    o.set_state(case.w, case.prim, case.df)
    dt = case.dt_xi
    o.step(dt, 0)
    rho = o.get_state()["w"].reshape(n, 4)[:, 0]
    inc = u[:, 0] > 0                                  # entering through xmin
    flux_in = np.sum(wt[inc] * u[inc, 0] * (Mb[0][inc] - f[0][inc]))
    for c in range(n):
        V = ds[c, 0] * ds[c, 1]
        area = ds[c, 1]
        want = prim0[0] + (dt / V * area * flux_in if mid[c, 0] < ds[c, 0] else 0.0)
        assert abs(rho[c] - want) <= 1e-13 * max(abs(want), 1.0), (c, rho[c], want)


# ------------------------------------------------------------------ Löhner criterion
def _vs_case(grid_fn, nv=8, vmax=4.0, lmax=2):
    solver = Solver(DIM=2, NDF=2, AMR_PS_MAXLEVEL=0, AMR_VS_MAXLEVEL=lmax, CFL=0.5)
    gas = Gas(K=1.0, Kn=0.05, omega=0.81, omega_ref=0.81)
    dom = [Domain("UniformOutflow", i + 1) for i in range(4)]
    cfg = Configure(solver, gas, [0.0, 1.0, 0.0, 1.0], [1, 1], [-vmax, vmax, -vmax, vmax],
                    [nv, nv], Uniform([1.0, 0.0, 0.0, 1.0]), dom)
    return build_case(cfg, vs_grid_fn=lambda c, prim, base: grid_fn(base))


def _flags(case, df, tau):
    g = case.grids[0]
    w = _moments(g.midpoint, g.weight, df[0], df[1])
    prim = O.get_prim(2, w, case.config.gas.gamma)
    o = O.Oracle(kah_params(case.config), case.topo)
    o.set_state(w, prim, df.reshape(-1))
    params = abi.vs_adapt_params("lohner", coeff_local=1e30, coeff_lohner=tau)
    fl = o.vs_flags(params, np.array([1.0, 1.0]))
    o.close()
    return fl & 1


def _edge_cells(g):
    """Cells with no velocity cell across one of their faces (the vacuum side)."""
    lo, hi = np.asarray(g.quadrature[0::2]), np.asarray(g.quadrature[1::2])
    size = g.ds[None, :] / (2.0 ** g.level.astype(float))[:, None]
    return np.any((np.abs(g.midpoint - 0.5 * size - lo) < 1e-12) |
                  (np.abs(g.midpoint + 0.5 * size - hi) < 1e-12), axis=1)


@pytest.mark.parametrize("layout", ["uniform", "band"])
def test_lohner_linear_field_flags_only_the_vacuum_edge(layout):
    def grid_fn(base):
        g = base.copy()
        if layout == "band":    # a band of columns one level finer: coarse/fine faces along u
            g.refine((g.midpoint[:, 0] > -1.5) & (g.midpoint[:, 0] < 2.5))
        return g
    case = _vs_case(grid_fn)
    g = case.grids[0]
    u = g.midpoint
    # exactly representable linear fields (dyadic coefficients and midpoints)
    h = 8.0 + 0.5 * u[:, 0] + (0.25 * u[:, 1] if layout == "uniform" else 0.0)
    b = 4.0 - 0.25 * u[:, 0]
    fl = _flags(case, np.stack([h, b]), 0.0)
    edge = _edge_cells(g)
    assert np.array_equal(fl.astype(bool), edge & (g.level < g.maxlevel))
    if layout == "band":
        assert (g.level == 1).sum() > 0 and (fl[g.level == 1] == 0).any()


def test_lohner_spike_values_separate_at_their_thresholds():
    case = _vs_case(lambda base: base.copy(), nv=8)
    g = case.grids[0]
    n = g.n
    A = 3.0
    i0 = 3 * 8 + 4                                    # an interior cell
    h = np.zeros(n)
    h[i0] = A
    b = np.zeros(n)
    eta_spike = 2 * A / (2 * A * (1 + 1e-2) + 2 * A * 1e-3)
    eta_nb = A / (A * (1 + 1e-2) + 2 * A * 1e-3)
    nbs = {i0 - 1, i0 + 1, i0 - 8, i0 + 8}
    for tau, want in [(0.5 * (eta_spike + eta_nb), {i0}),
                      (eta_nb * (1 - 1e-9), {i0} | nbs),
                      (eta_spike * (1 + 1e-9), set()),
                      (eta_spike * (1 - 1e-9), {i0}),
                      (eta_nb * (1 + 1e-9), {i0})]:
        fl = _flags(case, np.stack([h, b]), tau)
        assert set(np.flatnonzero(fl).tolist()) == want, (tau, np.flatnonzero(fl))


# ------------------------------------------------------------------ contribution criteria
def _flags_contribution(case, df, cl, cg, vr):
    g = case.grids[0]
    w = _moments(g.midpoint, g.weight, df[0], df[1])
    prim = O.get_prim(2, w, case.config.gas.gamma)
    o = O.Oracle(kah_params(case.config), case.topo)
    o.set_state(w, prim, df.reshape(-1))
    params = abi.vs_adapt_params("contribution", coeff_local=cl, coeff_global=cg)
    fl = o.vs_flags(params, np.asarray(vr, dtype=np.float64))
    o.close()
    return fl & 1


def test_contribution_criterion_flags_the_cells_holding_a_share():
    case = _vs_case(lambda base: base.copy(), nv=8, vmax=4.0)
    g = case.grids[0]
    n = g.n
    assert n == 64 and np.all(g.weight == 1.0)
    h, b = np.ones(n), np.zeros(n)
    u2 = np.sum(g.midpoint ** 2, axis=1)                  # U = 0 by symmetry
    share_e = u2 / u2.sum()                               # share of the internal energy
    corners = set(np.flatnonzero(u2 == u2.max()).tolist())
    outer = set(np.flatnonzero(share_e > 1.0 / n).tolist())
    assert len(corners) == 4 and len(outer) == 32
    top2 = np.unique(share_e)[-2:]
    df = np.stack([h, b])
    for cl, want in [(1.0 / n * (1 - 1e-6), set(range(n))),          # every cell's mass share
                     (1.0 / n * (1 + 1e-6), outer),                   # energy shares above 1/N
                     (0.5 * (top2[0] + top2[1]), corners),
                     (top2[1] * (1 + 1e-6), set())]:
        fl = _flags_contribution(case, df, cl, 1e30, [1.0, 1.0])
        assert set(np.flatnonzero(fl).tolist()) == want, (cl, np.flatnonzero(fl))
    # global: mass content h w = 1 everywhere, energy content |u|^2/2: with vr = (1, 1) the
    # threshold 10 sits between the corners' 12.25 and the next 9.25
    e = 0.5 * u2
    assert e.max() == 12.25 and np.unique(e)[-2] == 9.25
    fl = _flags_contribution(case, df, 1e30, 10.0, [1.0, 1.0])
    assert set(np.flatnonzero(fl).tolist()) == corners
    fl = _flags_contribution(case, df, 1e30, 0.999, [1.0, 1.0])      # mass content above
    assert fl.all()
