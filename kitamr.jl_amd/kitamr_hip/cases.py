"""Flow cases of the per-step workloads (all seeded / deterministic).

riemann_2d    2D four-quadrant Riemann problem on [0,1]^2 (UniformOutflow on all sides),
              optionally with static physical refinement toward the quadrant interfaces
              (a density-gradient criterion) and optionally with adapted velocity grids.
wall_channel  the reference's own smoke case (test/runtests.jl:1-42): Maxwellian walls in x,
              periodic in y.
box_3d        3D1F analogue on a small octree brick.
periodic_box_2d  the Riemann initial state (or a uniform state) with periodic boundaries on
              all sides, where mass, momentum and energy are conserved exactly.
cylinder_2d   flow past a circular immersed boundary (example/cylinder/cylinder.jl):
              supersonic inflow at xmin, outflow elsewhere, a Maxwellian wall on the circle.
sphere_3d     the 3D1F analogue with a spherical immersed boundary (example/sphere/sphere.jl).
"""
from __future__ import annotations

import math
import os

import numpy as np

from .solver import Configure, Domain, Gas, PCoordFn, Solver, Uniform, build_case
from .vspace import refine_toward

# quadrant states (rho, U, V, T) with prim = [rho, U, V, lambda = 1/T]
RIEMANN_STATES = {
    "ne": (1.0, 0.0, 0.0, 1.0),
    "nw": (0.5, 0.5, 0.0, 0.8),
    "sw": (0.125, 0.0, 0.0, 0.8),
    "se": (0.5, 0.0, 0.5, 0.8),
}


def riemann_prim(mid, x0=0.5, y0=0.5):
    x, y = mid[0], mid[1]
    if x >= x0 and y >= y0:
        r, u, v, T = RIEMANN_STATES["ne"]
    elif x < x0 and y >= y0:
        r, u, v, T = RIEMANN_STATES["nw"]
    elif x < x0 and y < y0:
        r, u, v, T = RIEMANN_STATES["sw"]
    else:
        r, u, v, T = RIEMANN_STATES["se"]
    return [r, u, v, 1.0 / T]


def interface_refine(halo_cells=(6.0, 4.0, 2.5), x0=0.5, y0=0.5, base_h=1.0 / 64):
    """Refine leaves whose centre lies within halo_cells[level]*base_h of an interface of the
    initial density field (the quadrant boundaries x = x0, y = y0): the static stand-in for the
    Löhner density-gradient sensor of src/Physical_space/Criteria.jl:14-22."""
    def flag(mid, ds, level):
        dist = np.minimum(np.abs(mid[:, 0] - x0), np.abs(mid[:, 1] - y0))
        lim = np.array([halo_cells[min(int(l), len(halo_cells) - 1)] for l in level]) * base_h
        return dist < lim
    return flag


def riemann_2d(n=64, nv=28, vmax=5.0, ps_levels=0, halo=(6.0, 4.0, 2.5), Kn=1e-2, omega=0.5,
               CFL=0.8, flux="CAIDVM", marching="CAIDVM_Marching", vs_levels=0,
               vs_adapt_rounds=0, seed=20261202, bc="UniformOutflow", precision="fp64",
               vs_variants=0):
    """vs_variants > 0: the adapted grid of cell c is the one of seed + (c mod vs_variants)
    for its initial state (built once per variant and state), so large meshes build fast
    while neighbouring cells still carry different grids."""
    solver = Solver(DIM=2, NDF=2, AMR_PS_MAXLEVEL=ps_levels, AMR_VS_MAXLEVEL=vs_levels,
                    flux=flux, time_marching=marching, CFL=CFL)
    gas = Gas(K=1.0, gamma=5.0 / 3.0, Pr=2.0 / 3.0, Kn=Kn, omega=omega)
    if bc == "UniformOutflow":
        dom = [Domain("UniformOutflow", i + 1) for i in range(4)]
    else:
        dom = [Domain(bc, i + 1, [1.0, 0.0, 0.0, 1.0]) for i in range(4)]
    cfg = Configure(solver, gas, [0.0, 1.0, 0.0, 1.0], [n, n], [-vmax, vmax, -vmax, vmax],
                    [nv, nv], PCoordFn(riemann_prim), dom, precision=precision)
    refine = interface_refine(halo, base_h=1.0 / n) if ps_levels > 0 else None
    vs_fn = None
    if vs_levels > 0:
        cache = {}

        def vs_fn(c, prim, base, _seed=seed):
            key = c if vs_variants <= 0 else (tuple(np.asarray(prim).tolist()), c % vs_variants)
            if key not in cache:
                r = c if vs_variants <= 0 else c % vs_variants
                cache[key] = refine_toward(base.copy(), prim,
                                           rounds=vs_adapt_rounds or vs_levels, rng=_seed + r)
            g = cache[key]
            if vs_variants <= 0:
                del cache[key]
            return g
    return build_case(cfg, ps_refine=refine, vs_grid_fn=vs_fn)


def periodic_box_2d(n=8, nv=8, uniform=False, ps_levels=0, vs_levels=0, precision="fp64",
                    seed=20261202):
    solver = Solver(DIM=2, NDF=2, AMR_PS_MAXLEVEL=ps_levels, AMR_VS_MAXLEVEL=vs_levels, CFL=0.8)
    gas = Gas(K=1.0, Kn=0.05, omega=0.5)
    dom = [Domain("Period", i + 1) for i in range(4)]
    ic = Uniform([1.0, 0.2, -0.1, 1.0]) if uniform else PCoordFn(riemann_prim)
    cfg = Configure(solver, gas, [0.0, 1.0, 0.0, 1.0], [n, n], [-5.0, 5.0, -5.0, 5.0],
                    [nv, nv], ic, dom, precision=precision)
    refine = None
    if ps_levels:
        def refine(mid, ds, level):
            return (np.abs(mid[:, 0] - 0.5) < 0.2) & (np.abs(mid[:, 1] - 0.5) < 0.3)
    vs_fn = None
    if vs_levels > 0:
        def vs_fn(c, prim, base, _seed=seed):
            return refine_toward(base.copy(), prim, rounds=vs_levels, rng=_seed + c)
    return build_case(cfg, ps_refine=refine, vs_grid_fn=vs_fn)


def wall_channel(n=16, nv=16, flux="CAIDVM", marching="CAIDVM_Marching", ps_levels=0):
    """test/runtests.jl:1-42."""
    solver = Solver(DIM=2, NDF=2, AMR_PS_MAXLEVEL=ps_levels, AMR_VS_MAXLEVEL=0, flux=flux,
                    time_marching=marching, max_sim_time=0.25)
    gas = Gas(K=0.0, Kn=0.075, omega=0.81, omega_ref=0.81)
    dom = [Domain("Maxwellian", 1, [1.0, 0.0, 0.0, 1.0]),
           Domain("Maxwellian", 2, [1.0, 1.0 * math.sqrt(5 / 6), 0.0, 1.0]),
           Domain("Period", 3), Domain("Period", 4)]
    cfg = Configure(solver, gas, [-0.5, 0.5, -0.5, 0.5], [n, n], [-5.0, 5.0, -5.0, 5.0],
                    [nv, nv], Uniform([1.0, 0.0, 0.0, 1.0]), dom)
    refine = None
    if ps_levels > 0:
        def refine(mid, ds, level):
            return (np.abs(mid[:, 0]) < 0.2) & (np.abs(mid[:, 1]) < 0.2)
    return build_case(cfg, ps_refine=refine)


def box_3d(n=4, nv=8, vmax=6.0, ps_levels=0, bcs=("UniformOutflow",) * 6, vs_levels=0,
           marching="CAIDVM_Marching", flux="CAIDVM", seed=7, precision="fp64", vs_variants=0):
    """3D1F box. vs_variants > 0: the adapted grid of cell c is built once per
    c mod vs_variants (from that cell's initial state), as in riemann_2d."""
    solver = Solver(DIM=3, NDF=1, AMR_PS_MAXLEVEL=ps_levels, AMR_VS_MAXLEVEL=vs_levels,
                    flux=flux, time_marching=marching, CFL=0.5)
    gas = Gas(K=0.0, Kn=0.05, omega=0.81)
    dom = []
    for i, k in enumerate(bcs):
        if k == "Period":
            dom.append(Domain("Period", i + 1))
        elif k == "UniformOutflow":
            dom.append(Domain(k, i + 1))
        else:
            dom.append(Domain(k, i + 1, [1.0, 0.3 if i == 0 else 0.0, 0.0, 0.0, 1.0]))

    def ic(mid):
        r = 1.0 + 0.5 * math.tanh(8 * (mid[0] - 0.5)) + 0.2 * math.sin(2 * math.pi * mid[1])
        return [r, 0.1 * mid[2], -0.05, 0.02 * mid[0], 1.0 / (1.0 + 0.2 * mid[1])]
    cfg = Configure(solver, gas, [0.0, 1.0] * 3, [n, n, n], [-vmax, vmax] * 3, [nv] * 3,
                    PCoordFn(ic), dom, precision=precision)
    refine = None
    if ps_levels > 0:
        def refine(mid, ds, level):
            return np.abs(mid[:, 0] - 0.5) < 0.2
    vs_fn = None
    if vs_levels > 0:
        cache = {}

        def vs_fn(c, prim, base, _seed=seed):
            if vs_variants <= 0:
                return refine_toward(base.copy(), prim, rounds=vs_levels, frac=0.15,
                                     rng=_seed + c)
            r = c % vs_variants
            if r not in cache:
                cache[r] = refine_toward(base.copy(), prim, rounds=vs_levels, frac=0.15,
                                         rng=_seed + r)
            return cache[r]
    return build_case(cfg, ps_refine=refine, vs_grid_fn=vs_fn)


def cylinder_2d(n=32, nv=16, vmax=10.0, L=4.0, radius=1.0, Ma=1.5, Kn=0.1, omega=0.81,
                ps_levels=0, band=0.6, flux="CAIDVM", marching="CAIDVM_Marching",
                precision="fp64", wall=(1.0, 0.0, 0.0, 1.0), CFL=0.8, vs_levels=0,
                vs_variants=6, seed=20261203):
    """Circle(Maxwellian, [0,0], radius, true, ., wall) in [-L, L]^2 with an inflow of Mach
    Ma (U = Ma*sqrt(5/6) at T = 1) from xmin; ps_levels > 0 refines the band |r - R| < band
    (and 2:1 balances), so that the cells on both sides of the circle share one level.
    vs_levels > 0: per-cell adapted velocity grids, seeded refinements toward the local
    Maxwellian peak, one of vs_variants per cell (c mod vs_variants), so that neighbouring
    cells carry different grids."""
    from .ib import immerse_circle
    solver = Solver(DIM=2, NDF=2, AMR_PS_MAXLEVEL=ps_levels, AMR_VS_MAXLEVEL=vs_levels,
                    flux=flux, time_marching=marching, CFL=CFL)
    gas = Gas(K=1.0, gamma=5.0 / 3.0, Pr=2.0 / 3.0, Kn=Kn, omega=omega, omega_ref=omega)
    U = Ma * math.sqrt(5.0 / 6.0)
    free = [1.0, U, 0.0, 1.0]
    dom = [Domain("SuperSonicInflow", 1, free)] + [Domain("UniformOutflow", i) for i in (2, 3, 4)]
    cfg = Configure(solver, gas, [-L, L, -L, L], [n, n], [-vmax, vmax, -vmax, vmax], [nv, nv],
                    Uniform(free), dom, precision=precision)
    refine = None
    if ps_levels > 0:
        def refine(mid, ds, level):
            return np.abs(np.hypot(mid[:, 0], mid[:, 1]) - radius) < band
    vs_fn = None
    if vs_levels > 0:
        cache = {}

        def vs_fn(c, prim, base, _seed=seed):
            r = c % vs_variants
            if r not in cache:
                cache[r] = refine_toward(base.copy(), prim, rounds=vs_levels, rng=_seed + r)
            return cache[r]
    case = build_case(cfg, ps_refine=refine, vs_grid_fn=vs_fn)
    case.topo = immerse_circle(case.topo, (0.0, 0.0), radius, True, wall)
    return case


def sphere_3d(n=12, nv=8, vmax=7.28, L=3.0, radius=1.0, Ma=1.5, Kn=0.05, omega=0.81,
              flux="CAIDVM", marching="CAIDVM_Marching", precision="fp64", CFL=0.5,
              vs_levels=0, vs_variants=6, seed=7):
    """Sphere(Maxwellian, [0,0,0], radius, true, ., [1,0,0,0,1]) in [-L, L]^3 with an inflow
    of Mach Ma (U = Ma*sqrt(5/6) at T = 1) from xmin, outflow elsewhere; vs_levels > 0:
    per-cell adapted velocity grids (one of vs_variants seeded refinements per cell, as in
    box_3d), else a shared grid."""
    from .ib import immerse_sphere
    solver = Solver(DIM=3, NDF=1, AMR_PS_MAXLEVEL=0, AMR_VS_MAXLEVEL=vs_levels, flux=flux,
                    time_marching=marching, CFL=CFL)
    gas = Gas(K=0.0, Kn=Kn, omega=omega, omega_ref=omega)
    U = Ma * math.sqrt(5.0 / 6.0)
    free = [1.0, U, 0.0, 0.0, 1.0]
    dom = [Domain("SuperSonicInflow", 1, free)] + [Domain("UniformOutflow", i)
                                                  for i in range(2, 7)]
    cfg = Configure(solver, gas, [-L, L] * 3, [n, n, n], [-vmax, vmax] * 3, [nv] * 3,
                    Uniform(free), dom, precision=precision)
    vs_fn = None
    if vs_levels > 0:
        cache = {}

        def vs_fn(c, prim, base, _seed=seed):
            r = c % vs_variants
            if r not in cache:
                cache[r] = refine_toward(base.copy(), prim, rounds=vs_levels, frac=0.15,
                                         rng=_seed + r)
            return cache[r]
    case = build_case(cfg, vs_grid_fn=vs_fn)
    case.topo = immerse_sphere(case.topo, (0.0, 0.0, 0.0), radius, True,
                               (1.0, 0.0, 0.0, 0.0, 1.0))
    return case


# ---------------------------------------------------------------- developed blunt-body flow
def bow_shock_state(mid, Ma, radius, gamma=5.0 / 3.0, Tw=1.0, delta=0.15, body="cylinder"):
    """A developed supersonic flow past a cylinder (2D) or sphere (3D) centred at the origin,
    flow along +x, as an analytic field: (prim [n, DIM+2], disturbance s [n] in [0, 1]).

    The bow shock follows Billig's correlations (standoff and vertex curvature of blunt
    bodies; cylinder: 0.386 exp(4.67/M^2) R and 1.386 exp(1.8/(M-1)^0.75) R, sphere:
    0.143 exp(3.24/M^2) R and 1.143 exp(0.54/(M-1)^1.2) R), asymptotic to the Mach angle. Behind
    it the Rankine-Hugoniot state of the local normal Mach number, smeared over a tanh of
    width delta; toward the front of the body the gas stagnates (U -> 0, T -> the stagnation
    temperature, isentropic density) and a thin layer relaxes it to the wall temperature;
    behind the body a slow, rarefied wake. s is the share of the stagnation heating, raised
    to 1 in the middle of the shock, where the distribution is furthest from one Maxwellian.
    Freestream: rho = 1, T = 1, U = Ma sqrt(gamma/2) (the reference's units, lambda = 1/T)."""
    mid = np.asarray(mid, dtype=np.float64)
    dim = mid.shape[1]
    x = mid[:, 0]
    y = np.sqrt(np.sum(mid[:, 1:] ** 2, axis=1))       # distance from the axis
    M = float(Ma)
    U1 = M * math.sqrt(gamma / 2.0)
    if body == "cylinder":
        standoff = radius * 0.386 * math.exp(4.67 / M ** 2)
        Rc = radius * 1.386 * math.exp(1.8 / (M - 1.0) ** 0.75)
    else:
        standoff = radius * 0.143 * math.exp(3.24 / M ** 2)
        Rc = radius * 1.143 * math.exp(0.54 / (M - 1.0) ** 1.2)
    beta = math.asin(1.0 / M)
    cot2, tan2 = 1.0 / math.tan(beta) ** 2, math.tan(beta) ** 2
    root = np.sqrt(1.0 + y * y * tan2 / Rc ** 2)
    xs = -(radius + standoff) + Rc * cot2 * (root - 1.0)
    slope = y / (Rc * root)                              # dx_s/dy
    sin_sig = 1.0 / np.sqrt(1.0 + slope * slope)
    Mn = np.maximum(M * sin_sig, 1.0)
    g = gamma
    r21 = (g + 1) * Mn ** 2 / ((g - 1) * Mn ** 2 + 2)
    T21 = (2 * g * Mn ** 2 - (g - 1)) * ((g - 1) * Mn ** 2 + 2) / ((g + 1) ** 2 * Mn ** 2)
    # velocity behind the shock: the normal component divided by the density ratio
    nx = 1.0 / np.sqrt(1.0 + slope * slope)
    ny = -slope * nx
    un = U1 * nx
    ux2 = U1 - un * nx + (un / r21) * nx
    ur2 = -un * ny + (un / r21) * ny                     # radial (away from the axis)
    T0 = 1.0 + 0.5 * (g - 1) * M ** 2
    rc = np.sqrt(np.sum(mid ** 2, axis=1))
    cos_t = np.where(rc > 0, x / np.maximum(rc, 1e-300), 0.0)
    front = np.maximum(-cos_t, 0.0)
    b = np.exp(-np.maximum(rc - radius, 0.0) / (0.6 * standoff + 0.05)) * front
    T2 = T21 + b * (T0 - T21)
    rho2 = r21 * (T2 / T21) ** (1.0 / (g - 1))
    ux2 = ux2 * (1.0 - b)
    ur2 = ur2 * (1.0 - b)
    # the wall's thin thermal layer
    kl = np.exp(-np.maximum(rc - radius, 0.0) / 0.05)
    T2 = T2 + kl * (Tw - T2)
    # wake
    wk = np.where(x > 0, np.exp(-(y / (radius * (1.0 + x / 4.0))) ** 2) * np.exp(-x / 6.0), 0.0)
    rho2 = rho2 * (1.0 - 0.7 * wk)
    ux2 = ux2 * (1.0 - 0.6 * wk)
    s_sh = 0.5 * (1.0 + np.tanh((x - xs) / delta))
    rho = 1.0 + s_sh * (rho2 - 1.0)
    ux = U1 + s_sh * (ux2 - U1)
    ur = s_sh * ur2
    T = 1.0 + s_sh * (T2 - 1.0)
    inside = rc < radius
    rho = np.where(inside, 1.0, rho)
    ux = np.where(inside, 0.0, ux)
    ur = np.where(inside, 0.0, ur)
    T = np.where(inside, Tw, T)
    prim = np.zeros((len(x), dim + 2))
    prim[:, 0] = rho
    prim[:, 1] = ux
    yy = np.maximum(y, 1e-300)
    for d in range(1, dim):
        prim[:, 1 + d] = np.where(y > 0, ur * mid[:, d] / yy, 0.0)
    prim[:, dim + 1] = 1.0 / T
    heat = np.clip((T - 1.0) / (T0 - 1.0), 0.0, 1.0)
    s = np.maximum(heat, 4.0 * s_sh * (1.0 - s_sh))
    s = np.where(inside, 0.0, s)
    return prim, s


def _target_grids_chunk(args):
    prims, targets, seeds, base, K, n_cap = args
    from .vspace import initial_vs_grid, refine_to_target
    out = []
    cache = {}
    for p, t, sd in zip(prims, targets, seeds):
        key = tuple(p.tolist())
        if key not in cache:
            cache[key] = initial_vs_grid(base, p, K)
        g = cache[key]
        if t > g.n:
            g = refine_to_target(g, p, int(t), int(sd), n_cap)
        out.append(g)
    return out


def target_grids(prim, s, base, K, n_min, n_max, seed, workers=0):
    """Per-cell adapted grids of a developed flow: the reference's initial grid of each cell's
    Maxwellian (initial_vs_adaptive_mesh_refinement!, src/Velocity_space/AMR.jl:183-200),
    refined further toward its peak (a seeded fallback) until it holds n_min (n_max/n_min)^s
    cells, where s in [0, 1] is the local disturbance. Cells with the same state and no
    further target share one grid object (the reference gives them equal grids)."""
    n = len(prim)
    targets = np.rint(n_min * (n_max / n_min) ** np.asarray(s)).astype(np.int64)
    seeds = seed + np.arange(n)
    # cells whose state equals another's and whose target the initial grid already meets
    # share that grid; the rest are built in parallel chunks
    workers = workers or min(16, os.cpu_count() or 1)
    chunks = []
    step = max(1, -(-n // (workers * 4)))
    for c0 in range(0, n, step):
        c1 = min(n, c0 + step)
        chunks.append((prim[c0:c1], targets[c0:c1], seeds[c0:c1], base, K, n_max))
    if workers > 1 and n > 4096:
        import multiprocessing as mp
        with mp.get_context("fork").Pool(workers) as pool:
            parts = pool.map(_target_grids_chunk, chunks)
    else:
        parts = [_target_grids_chunk(a) for a in chunks]
    grids = [g for part in parts for g in part]
    # one object per distinct grid content among cells of equal state
    return grids


def cylinder_c3(n=448, nv=20, vmax=15.0, L=8.0, radius=0.5, Ma=5.0, Kn=0.1, omega=0.81,
                vs_levels=2, n_min=400, n_max=4000, precision="fp64", flux="CAIDVM",
                marching="CAIDVM_Marching", CFL=0.4, seed=20261202, workers=0, walls=False,
                immersed=True):
    """The velocity-adaptive cylinder (example/cylinder/cylinder.jl): Circle(Maxwellian,
    [0,0], radius, true, ., [1,0,0,1]) in [-L, L]^2 on n x n leaves, supersonic inflow at xmin,
    outflow elsewhere, a developed bow-shock state (bow_shock_state) and per-cell adapted
    velocity grids on a 20x20 base on [-vmax, vmax]^2 with vs_levels levels: each cell's
    initial reference grid, refined toward n_min..n_max cells with the local disturbance.
    walls=True closes the box with Maxwellian walls (at the freestream density and
    temperature, at rest) instead of the inflow and outflow boundaries; immersed=False leaves
    the circle out (same leaves, state and velocity grids, every cell fluid). CFL 0.4 is the
    reference's default (src/Solver/Types.jl:122), which example/cylinder/cylinder.jl keeps."""
    from .ib import immerse_circle
    from .solver import PCoordField
    solver = Solver(DIM=2, NDF=2, AMR_PS_MAXLEVEL=0, AMR_VS_MAXLEVEL=vs_levels, flux=flux,
                    time_marching=marching, CFL=CFL)
    gas = Gas(K=1.0, gamma=5.0 / 3.0, Pr=2.0 / 3.0, Kn=Kn, omega=omega, omega_ref=omega)
    U = Ma * math.sqrt(5.0 / 6.0)
    free = [1.0, U, 0.0, 1.0]
    dom = [Domain("SuperSonicInflow", 1, free)] + [Domain("UniformOutflow", i) for i in (2, 3, 4)]
    if walls:
        dom = [Domain("Maxwellian", i, [1.0, 0.0, 0.0, 1.0]) for i in (1, 2, 3, 4)]
    field = PCoordField(lambda m: bow_shock_state(m, Ma, radius)[0])
    cfg = Configure(solver, gas, [-L, L, -L, L], [n, n], [-vmax, vmax, -vmax, vmax], [nv, nv],
                    field, dom, precision=precision)

    def grids_fn(prim, mid, base):
        _, s = bow_shock_state(mid, Ma, radius)
        return target_grids(prim, s, base, gas.K, n_min, n_max, seed, workers)
    case = build_case(cfg, vs_grids_fn=grids_fn)
    if immersed:
        case.topo = immerse_circle(case.topo, (0.0, 0.0), radius, True, (1.0, 0.0, 0.0, 1.0))
    return case


def sphere_c4(n=48, nv=16, vmax=10.0, L=6.0, radius=0.5, Ma=3.0, Kn=0.05, omega=0.81,
              vs_levels=1, n_min=4096, n_max=30000, precision="fp32f", flux="CAIDVM",
              marching="CAIDVM_Marching", CFL=0.4, seed=20261204, workers=0, walls=False,
              immersed=True):
    """The sphere (example/sphere/sphere.jl): Sphere(Maxwellian, [0,0,0], radius, true, .,
    [1,0,0,0,1]) in [-L, L]^3 on n^3 leaves, Mach-Ma inflow at xmin, outflow elsewhere, a
    developed bow-shock state and per-cell adapted grids on a 16^3 base on [-vmax, vmax]^3
    (3D1F): each cell's initial reference grid refined toward n_min..n_max cells with the
    local disturbance. walls=True closes the box with Maxwellian walls (freestream density and
    temperature, at rest); immersed=False leaves the sphere out (same leaves, state, grids)."""
    from .ib import immerse_sphere
    from .solver import PCoordField
    solver = Solver(DIM=3, NDF=1, AMR_PS_MAXLEVEL=0, AMR_VS_MAXLEVEL=vs_levels, flux=flux,
                    time_marching=marching, CFL=CFL)
    gas = Gas(K=0.0, Kn=Kn, omega=omega, omega_ref=omega)
    U = Ma * math.sqrt(5.0 / 6.0)
    free = [1.0, U, 0.0, 0.0, 1.0]
    dom = [Domain("SuperSonicInflow", 1, free)] + [Domain("UniformOutflow", i)
                                                  for i in range(2, 7)]
    if walls:
        dom = [Domain("Maxwellian", i, [1.0, 0.0, 0.0, 0.0, 1.0]) for i in range(1, 7)]
    field = PCoordField(lambda m: bow_shock_state(m, Ma, radius, body="sphere")[0])
    cfg = Configure(solver, gas, [-L, L] * 3, [n, n, n], [-vmax, vmax] * 3, [nv] * 3, field,
                    dom, precision=precision)

    def grids_fn(prim, mid, base):
        _, s = bow_shock_state(mid, Ma, radius, body="sphere")
        return target_grids(prim, s, base, None, n_min, n_max, seed, workers)
    case = build_case(cfg, vs_grids_fn=grids_fn)
    if immersed:
        case.topo = immerse_sphere(case.topo, (0.0, 0.0, 0.0), radius, True,
                                   (1.0, 0.0, 0.0, 0.0, 1.0))
    return case
