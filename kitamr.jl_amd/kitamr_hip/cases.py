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
