"""The stated workloads on the device.

- C1 (the CPU-runnable configuration): the 2D four-quadrant Riemann problem on exactly 64x64
  leaves with a 28x28 velocity grid on [-5,5]^2, Kn = 1e-2, omega = 0.5, FP64, CAIDVM +
  CAIDVM_Marching, 100 fused steps against the oracle at 1e-12 (relative, per array / per
  cell row), with the residual sums of the check steps.
  At the stated CFL 0.8 the reference's step is unstable: the time step is limited per
  direction (Dt_xi = CFL min(ds/U), src/Solver/Types.jl:518), so at the corner velocities
  (|u| = |v| = 4.82) the unsplit update has CFL_x + CFL_y = 1.6 > 1, and the negative
  round-off values of f there grow about 2x per step from step ~50 (the limiter takes |f|,
  src/Flux/CAIDVM.jl:127-141, so it does not stop them) until the state is not finite at step
  ~63. The oracle and the device agree to 1e-12 up to step 60 and both turn non-finite at the
  same step. At CFL 0.4 (example/Riemann_problem/rp_2D.jl:29 and the reference's default) all
  100 steps stay finite and agree at 1e-12.
- C3 (the velocity-adaptive cylinder, 200,704 leaves, grids of 568-4000 cells): one step of
  the FP64 form against the oracle at full size (w, prim, df of the live cells at 1e-12); the
  FP32 form's moments within 1e-5 (per cell, relative) of the FP64 form's after 3 steps; and on
  the same leaves, state and velocity grids closed by Maxwellian walls, the total mass kept to
  1e-12 in both forms over 3 steps (every face's mass flux enters its two cells with opposite
  signs and the wall densities zero the walls' mass fluxes, so the total changes by round-off
  only). The circle is left out of that check: the reference's immersed boundary is a
  ghost-cell method whose faces between donors and solid cells exchange mass with the solid
  cells (the wall density zeroes the mass flux at the intersection point, not at the cell
  face), so the fluid's mass is not an invariant with it (the oracle loses 2.5e-4 of it in one
  step on a 32x32 closed box).
- C4 (the sphere: 48^3 = 110,592 leaves in [-6,6]^3, grids of 4k-30k cells on a 16^3 base,
  3D1F): at full size, the FP32 form's moments within 1e-5 (per cell) of the FP64 form's after
  3 steps, every SolidNeighbor carrying zero net mass through the wall (1e-12 of the sum of
  |terms| in FP64, 1e-6 in the FP32 form, whose stored f is float), and on the same leaves,
  state and grids closed by Maxwellian walls without the sphere, the total mass kept to 1e-12
  in both forms; and one FP64 step of the same sphere on 16^3 leaves
  on [-2,2]^3 (the sphere resolved as at full size, 26M phase-space cells) against the
  oracle at 1e-12 (live cells, SolidNeighbor df).
"""
import numpy as np
import pytest

from _util import pair, rel_err, rel_err_rows, say
from kitamr_hip import abi, cases
from kitamr_hip.solver import kah_params

TOL = 1e-12


def _first_nonfinite(dev, orc, dt, s0, s1):
    """Steps both sides from s0 to s1; the first step after which each state is not finite."""
    first = {"dev": None, "orc": None}
    for s in range(s0, s1):
        dev.step(dt, s, residual=False)
        orc.step(dt, s, residual=False)
        for name, side in (("dev", dev), ("orc", orc)):
            if first[name] is None and not np.isfinite(side.get_state()["w"]).all():
                first[name] = s + 1
        if first["dev"] is not None and first["orc"] is not None:
            break
    return first


@pytest.mark.gpu
@pytest.mark.parametrize("CFL", [0.8, 0.4])
def test_c1_riemann_64x64_28x28_100_steps(CFL):
    case = cases.riemann_2d(n=64, nv=28, CFL=CFL)
    assert case.n_cells == 4096 and case.phase_cells == 4096 * 784
    say(f"[C1 CFL={CFL}] cells={case.n_cells} P={case.phase_cells} dt={case.dt_xi:.6e}")
    assert abs(case.dt_xi - CFL * (1 / 64) / (5 - 10 / 56)) < 1e-15
    dev, orc = pair(case)
    dt = case.dt_xi
    last = 60 if CFL > 0.5 else 100
    for s in range(last):
        check = (s + 1) % 20 == 0
        rg = dev.step(dt, s, residual=check)
        ro = orc.step(dt, s, residual=check)
        if check:
            assert rel_err(rg[0], ro[0]) <= 1e-10 and rel_err(rg[1], ro[1]) <= TOL, s
            g, o = dev.get_state(), orc.get_state()
            e = {k: rel_err(g[k], o[k]) if k == "df" else rel_err_rows(g[k], o[k], 4)
                 for k in ("df", "w", "prim")}
            say(f"[C1 CFL={CFL}] step {s + 1}: {e}")
            assert max(e.values()) <= TOL, (s, e)
            assert np.isfinite(g["w"]).all()
    if last < 100:
        first = _first_nonfinite(dev, orc, dt, last, 100)
        say(f"[C1 CFL={CFL}] first non-finite step: {first}")
        assert first["orc"] is not None and first["dev"] == first["orc"], first
    dev.close()
    orc.close()


_C3 = {}


def _c3(precision="fp64", walls=False):
    """The C3 case (built once per boundary kind; the storage form only changes the
    parameters). walls=True: the same leaves, state and grids closed by Maxwellian walls,
    without the circle."""
    import dataclasses
    if walls not in _C3:
        _C3.clear()    # one full-size case held at a time
        _C3[walls] = cases.cylinder_c3(n=448, walls=walls, immersed=not walls)
    c = _C3[walls]
    return dataclasses.replace(c, config=dataclasses.replace(c.config, precision=precision))


def _dev(case):
    d = abi.HipContext(kah_params(case.config))
    d.set_topology(case.topo)
    d.set_state(case.w, case.prim, case.df)
    return d


@pytest.mark.gpu
def test_c3_full_size_step_fp64_equals_oracle():
    case = _c3()
    sizes = np.diff(case.topo["vs_offset"])
    say(f"[C3] cells={case.n_cells} P={case.phase_cells} grids {sizes.min()}-{sizes.max()}")
    assert sizes.min() >= 400 and sizes.max() <= 4000 and case.n_cells == 448 * 448
    dev, orc = pair(case)
    dev.step(case.dt_xi, 0, residual=False)
    orc.step(case.dt_xi, 0, residual=False)
    g, o = dev.get_state(qf=False), orc.get_state()
    fluid = np.asarray(case.topo["bound_enc"]) >= 0
    for k in ("w", "prim"):
        e = rel_err_rows(np.asarray(g[k]).reshape(-1, 4)[fluid],
                         np.asarray(o[k]).reshape(-1, 4)[fluid], 4)
        say(f"[C3] one step {k}: {e:.3g}")
        assert e <= TOL, k
    live = fluid.copy()
    live[np.asarray(case.topo["ib_solid"])] = True     # rebuilt by update_solid_cell!
    off = np.asarray(case.topo["vs_offset"]) * 2
    rows = np.repeat(live, np.diff(off))
    e = rel_err(np.asarray(g["df"])[:off[-1]][rows], np.asarray(o["df"])[:off[-1]][rows])
    say(f"[C3] one step df ({int(live.sum())} live cells): {e:.3g}")
    assert e <= TOL
    dev.close()
    orc.close()


@pytest.mark.gpu
def test_c3_fp32_form_moments_track_fp64():
    out = {}
    for prec in ("fp64", "fp32f"):
        case = _c3(prec)
        say(f"[C3] {prec}: cells={case.n_cells} P={case.phase_cells}")
        d = _dev(case)
        for s in range(3):
            d.step(case.dt_xi, s, residual=False)
        out[prec] = d.get_state(df=False)
        d.close()
    fluid = np.asarray(case.topo["bound_enc"]) >= 0
    for k in ("w", "prim"):
        a = out["fp32f"][k].reshape(-1, 4)[fluid]
        b = out["fp64"][k].reshape(-1, 4)[fluid]
        e = rel_err_rows(a, b, 4)
        say(f"[C3] fp32 vs fp64 {k}: {e:.3g}")
        assert e <= 1e-5, k
    q32 = out["fp32f"]["qf"].reshape(-1, 2)[fluid]
    q64 = out["fp64"]["qf"].reshape(-1, 2)[fluid]
    e = float(np.max(np.abs(q32 - q64)) / np.max(np.abs(q64)))
    say(f"[C3] fp32 vs fp64 q: {e:.3g}")
    assert e <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32f"])
def test_c3_closed_box_keeps_mass(precision):
    case = _c3(precision, walls=True)
    say(f"[C3 closed] {precision}: cells={case.n_cells} P={case.phase_cells}")
    n = case.n_cells
    assert (np.asarray(case.topo["bound_enc"]) >= 0).all()
    V = np.prod(np.asarray(case.topo["cell_ds"]).reshape(-1, 2)[:n], axis=1)
    fluid = np.asarray(case.topo["bound_enc"]) >= 0
    d = _dev(case)
    m0 = float((np.asarray(case.w).reshape(n, 4)[fluid, 0] * V[fluid]).sum())
    for s in range(3):
        d.step(case.dt_xi, s, residual=False)
        w = d.get_state(df=False, prim=False, qf=False)["w"].reshape(n, 4)
        m = float((w[fluid, 0] * V[fluid]).sum())
        say(f"[C3 closed] {precision} step {s + 1}: relative mass change {(m - m0) / m0:.3g}")
        assert abs(m - m0) <= 1e-12 * m0, (s, (m - m0) / m0)
    d.close()


_C4 = {}


def _c4(precision="fp32f", walls=False):
    """The C4 case (built once per boundary kind; the storage form only changes the
    parameters). walls=True: the same leaves, state and grids closed by Maxwellian walls,
    without the sphere."""
    import dataclasses
    if walls not in _C4:
        _C4.clear()
        _C4[walls] = cases.sphere_c4(n=48, walls=walls, immersed=not walls)
    c = _C4[walls]
    return dataclasses.replace(c, config=dataclasses.replace(c.config, precision=precision))


def _wall_mass_flux_ok(case, dev, tol, dim):
    """Every SolidNeighbor's net mass flux through the wall, sum of w u.n f over its donor's
    grid, against the sum of its |terms|."""
    aux = dev.get_ib()
    t = case.topo
    off = np.asarray(t["vs_offset"])
    mid = np.asarray(t["vs_mid"])
    wts = np.asarray(t["vs_weight"])
    nr = np.asarray(t["ib_normal"]).reshape(-1, dim)
    worst = 0.0
    for k, p in enumerate(t["ib_face_pair"]):
        h = int(t["pair_here"][p])
        a, b = int(off[h]), int(off[h + 1])
        vm = mid[dim * a:dim * b].reshape(dim, b - a).T
        fl = wts[a:b] * (vm @ nr[k]) * np.asarray(aux[k][0])
        worst = max(worst, abs(fl.sum()) / np.abs(fl).sum())
    say(f"[C4] worst relative wall mass flux {worst:.3g}")
    return worst <= tol


@pytest.mark.gpu
def test_c4_full_size_fp32_form_tracks_fp64_and_wall():
    out = {}
    for prec, tol in (("fp32f", 1e-6), ("fp64", 1e-12)):
        case = _c4(prec)
        sizes = np.diff(np.asarray(case.topo["vs_offset"]))
        fluid = np.asarray(case.topo["bound_enc"]) >= 0
        say(f"[C4] {prec}: cells={case.n_cells} P={case.phase_cells} "
            f"grids {sizes[fluid].min()}-{sizes[fluid].max()}")
        assert case.n_cells == 48 ** 3
        assert sizes[fluid].min() >= 4000 and sizes[fluid].max() <= 30000
        d = _dev(case)
        for s in range(3):
            d.step(case.dt_xi, s, residual=False)
        assert _wall_mass_flux_ok(case, d, tol, 3), prec
        out[prec] = d.get_state(df=False)
        d.close()
    fluid = np.asarray(case.topo["bound_enc"]) >= 0
    for k in ("w", "prim"):
        a = out["fp32f"][k].reshape(-1, 5)[fluid]
        b = out["fp64"][k].reshape(-1, 5)[fluid]
        assert np.isfinite(a).all() and (a[:, 0] > 0).all()
        e = rel_err_rows(a, b, 5)
        say(f"[C4] fp32 vs fp64 {k}: {e:.3g}")
        assert e <= 1e-5, k
    q32 = out["fp32f"]["qf"].reshape(-1, 3)[fluid]
    q64 = out["fp64"]["qf"].reshape(-1, 3)[fluid]
    e = float(np.max(np.abs(q32 - q64)) / np.max(np.abs(q64)))
    say(f"[C4] fp32 vs fp64 q: {e:.3g}")
    assert e <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32f"])
def test_c4_closed_box_keeps_mass(precision):
    case = _c4(precision, walls=True)
    say(f"[C4 closed] {precision}: cells={case.n_cells} P={case.phase_cells}")
    n = case.n_cells
    assert (np.asarray(case.topo["bound_enc"]) >= 0).all()
    V = np.prod(np.asarray(case.topo["cell_ds"]).reshape(-1, 3)[:n], axis=1)
    d = _dev(case)
    m0 = float((np.asarray(case.w).reshape(n, 5)[:, 0] * V).sum())
    for s in range(3):
        d.step(case.dt_xi, s, residual=False)
        w = d.get_state(df=False, prim=False, qf=False)["w"].reshape(n, 5)
        m = float((w[:, 0] * V).sum())
        say(f"[C4 closed] {precision} step {s + 1}: relative mass change {(m - m0) / m0:.3g}")
        assert abs(m - m0) <= 1e-12 * m0, (s, (m - m0) / m0)
    d.close()


@pytest.mark.gpu
def test_c4_reduced_sphere_step_equals_oracle():
    _C4.clear()
    # 16^3 leaves on [-2,2]^3: the sphere is resolved as at full size (ds = 0.25)
    case = cases.sphere_c4(n=16, L=2.0, precision="fp64")
    assert len(case.topo["ib_face_pair"]) > 0
    say(f"[C4 reduced] cells={case.n_cells} P={case.phase_cells}")
    dev, orc = pair(case)
    dev.step(case.dt_xi, 0, residual=False)
    orc.step(case.dt_xi, 0, residual=False)
    g, o = dev.get_state(qf=False), orc.get_state()
    fluid = np.asarray(case.topo["bound_enc"]) >= 0
    live = fluid.copy()
    live[np.asarray(case.topo["ib_solid"])] = True
    for k in ("w", "prim"):
        e = rel_err_rows(np.asarray(g[k]).reshape(-1, 5)[live],
                         np.asarray(o[k]).reshape(-1, 5)[live], 5)
        say(f"[C4 reduced] {k}: {e:.3g}")
        assert e <= TOL, k
    off = np.asarray(case.topo["vs_offset"])
    rows = np.repeat(live, np.diff(off))
    e = rel_err(np.asarray(g["df"])[:off[-1]][rows], np.asarray(o["df"])[:off[-1]][rows])
    say(f"[C4 reduced] df: {e:.3g}")
    assert e <= TOL
    e = rel_err(dev.get_ib(flat=True), orc.get_ib(flat=True))
    say(f"[C4 reduced] SolidNeighbor df: {e:.3g}")
    assert e <= TOL
    dev.close()
    orc.close()
