"""Pins the oracle's CIP_Marching (iterate!(::Type{CIP_Marching}), src/Theory/I-projection.jl:
161-192) and its conserved I-projection (:55-158) without the reference:

* the projection of one Maxwellian onto the moments of another recovers the second one
  exactly (their ratio exp(a + b.u + c|u|^2/2) lies in the span of psi = (1, u, |u|^2/2)),
  2D2F (with b's internal energy removed from the energy target) and 3D1F;
* a distribution that already has the target moments is returned unchanged, bit for bit
  (|G| < tol at lambda = 0);
* after projecting a perturbed distribution its moments equal the targets to the Newton
  tolerance; negative entries are shaved to positive values first;
* a CIP step leaves a uniform Maxwellian state uniform and conserves mass, momentum and
  energy on a periodic box; the step equals its parts recomputed from the point kernels.
"""
import numpy as np
import pytest

import oracle as O
from kitamr_hip import cases
from kitamr_hip.solver import Configure, Domain, Gas, Solver, Uniform, build_case, kah_params
from kitamr_hip.vspace import VelocityGrid


def grid2(n=28, vmax=5.0):
    return VelocityGrid(2, [-vmax, vmax, -vmax, vmax], [n, n])


def grid3(n=16, vmax=6.0):
    return VelocityGrid(3, [-vmax, vmax] * 3, [n, n, n])


@pytest.mark.parametrize("p1,p2", [
    ((1.0, 0.1, -0.2, 1.0), (1.05, 0.15, -0.18, 0.95)),
    ((0.5, 0.0, 0.0, 1.25), (0.45, -0.05, 0.02, 1.3)),
])
def test_projection_2d_recovers_target_maxwellian(p1, p2):
    g = grid2()
    mid = g.mid_colmajor()
    n = g.n
    m1 = O.discrete_maxwell(2, 2, mid, p1, 1.0)
    m2 = O.discrete_maxwell(2, 2, mid, p2, 1.0)
    W = O.micro_to_macro(2, 2, mid, m2, g.weight)
    df = np.concatenate([m1[:n], m2[n:]])        # h of the first, b of the target
    out = O.i_projection(2, 2, mid, g.weight, df, W)
    assert np.array_equal(out[n:], m2[n:])        # b untouched
    err = np.max(np.abs(out[:n] - m2[:n])) / np.max(np.abs(m2[:n]))
    assert err < 1e-9, err
    # a wrong sign of a psi term or a dropped e_int would miss the target by >> 1e-9
    Wo = O.micro_to_macro(2, 2, mid, out, g.weight)
    assert np.allclose(Wo, W, rtol=1e-10, atol=1e-12)


def test_projection_3d_recovers_target_maxwellian():
    g = grid3()
    mid = g.mid_colmajor()
    p1 = (1.0, 0.1, 0.0, -0.1, 1.0)
    p2 = (0.97, 0.12, 0.03, -0.08, 1.04)
    m1 = O.discrete_maxwell(3, 1, mid, p1, 0.0)
    m2 = O.discrete_maxwell(3, 1, mid, p2, 0.0)
    W = O.micro_to_macro(3, 1, mid, m2, g.weight)
    out = O.i_projection(3, 1, mid, g.weight, m1, W)
    err = np.max(np.abs(out - m2)) / np.max(np.abs(m2))
    assert err < 1e-9, err


def test_projection_keeps_a_consistent_distribution_bitwise():
    g = grid2()
    mid = g.mid_colmajor()
    m = O.discrete_maxwell(2, 2, mid, (0.8, 0.2, 0.1, 1.1), 1.0)
    W = O.micro_to_macro(2, 2, mid, m, g.weight)
    out = O.i_projection(2, 2, mid, g.weight, m, W)
    assert np.array_equal(out, m)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_projection_matches_moments_of_perturbed_distribution(seed):
    rng = np.random.default_rng(seed)
    g = grid2(24)
    mid = g.mid_colmajor()
    n = g.n
    m = O.discrete_maxwell(2, 2, mid, (1.0, 0.05, -0.1, 1.0), 1.0)
    W = O.micro_to_macro(2, 2, mid, m, g.weight)
    df = m * (1.0 + 0.05 * rng.standard_normal(2 * n))
    out = O.i_projection(2, 2, mid, g.weight, df, W)
    Wo = O.micro_to_macro(2, 2, mid, out, g.weight)
    assert np.allclose(Wo, W, rtol=1e-10, atol=1e-12), Wo - W
    # the projected h is the perturbed h times exp(lambda.psi): the log-ratio is quadratic
    u, v = mid[:n], mid[n:]
    lr = np.log(out[:n] / df[:n])
    A = np.stack([np.ones(n), u, v, 0.5 * (u * u + v * v)], axis=1)
    coef, res, *_ = np.linalg.lstsq(A, lr, rcond=None)
    assert np.max(np.abs(A @ coef - lr)) < 1e-9


def test_projection_shaves_negative_entries():
    g = grid2(20)
    mid = g.mid_colmajor()
    n = g.n
    m = O.discrete_maxwell(2, 2, mid, (1.0, 0.0, 0.0, 1.0), 1.0)
    W = O.micro_to_macro(2, 2, mid, m, g.weight)
    df = m.copy()
    tail = np.argsort(df[:n])[:12]               # the 12 smallest h values go negative
    df[tail] = -0.5 * df[tail] - 1e-9
    out = O.i_projection(2, 2, mid, g.weight, df, W)
    assert np.all(out[:n] > 0)
    # shave formula (:74-84): f <- (f - 1.1*min)/(fp_min - 1.1*min)*fp_min, applied before
    # the exp(lambda.psi) factor, so the ratio out/shaved is exp of a quadratic
    f_min = 1.1 * df[:n].min()
    fp_min = df[:n][df[:n] > 0].min()
    shaved = np.where(df[:n] < 0, (df[:n] - f_min) / (fp_min - f_min) * fp_min, df[:n])
    u, v = mid[:n], mid[n:]
    lr = np.log(out[:n] / shaved)
    A = np.stack([np.ones(n), u, v, 0.5 * (u * u + v * v)], axis=1)
    coef, *_ = np.linalg.lstsq(A, lr, rcond=None)
    assert np.max(np.abs(A @ coef - lr)) < 1e-9
    Wo = O.micro_to_macro(2, 2, mid, out, g.weight)
    assert np.allclose(Wo, W, rtol=1e-10, atol=1e-12)


def _cip_case(uniform, n=8, nv=16):
    solver = Solver(DIM=2, NDF=2, AMR_PS_MAXLEVEL=0, AMR_VS_MAXLEVEL=0,
                    time_marching="CIP_Marching", CFL=0.8)
    gas = Gas(K=1.0, Kn=0.05, omega=0.5)
    dom = [Domain("Period", i + 1) for i in range(4)]
    ic = Uniform([1.0, 0.2, -0.1, 1.0]) if uniform else cases.PCoordFn(cases.riemann_prim)
    cfg = Configure(solver, gas, [0.0, 1.0, 0.0, 1.0], [n, n], [-5.0, 5.0, -5.0, 5.0],
                    [nv, nv], ic, dom)
    return build_case(cfg)


def test_cip_step_keeps_uniform_state():
    case = _cip_case(uniform=True)
    o = O.Oracle(kah_params(case.config), case.topo)
    o.set_state(case.w, case.prim, case.df)
    for s in range(3):
        o.step(case.dt_xi, s)
    st = o.get_state()
    d = np.asarray(st["df"]).reshape(case.n_cells, -1)
    assert np.max(np.abs(d - d[0])) <= 1e-13 * np.max(np.abs(d))
    w = np.asarray(st["w"]).reshape(case.n_cells, 4)
    assert np.max(np.abs(w - w[0])) <= 1e-13
    o.close()


def test_cip_step_conserves_on_periodic_box():
    case = _cip_case(uniform=False)
    n = case.n_cells
    V = np.prod(case.topo["cell_ds"][:n].reshape(n, 2), axis=1)
    o = O.Oracle(kah_params(case.config), case.topo)
    o.set_state(case.w, case.prim, case.df)
    tot0 = (np.asarray(case.w).reshape(n, 4) * V[:, None]).sum(0)
    for s in range(5):
        o.step(case.dt_xi, s)
    st = o.get_state()
    tot = (st["w"].reshape(n, 4) * V[:, None]).sum(0)
    assert np.allclose(tot, tot0, rtol=1e-13, atol=1e-14), tot - tot0
    assert np.all(np.isfinite(st["df"]))
    o.close()


def test_cip_iterate_equals_its_parts():
    """One iterate of a cell recomputed from the point kernels: w += flux*dt/V, prim_c,
    f* = f + dt/V*vflux, q(f*, prim_c), F = M(prim_c) + shakhov, I-projection of f* onto
    w, relaxation with tau(prim_c)."""
    case = _cip_case(uniform=False, n=4, nv=12)
    o = O.Oracle(kah_params(case.config), case.topo)
    o.set_state(case.w, case.prim, case.df)
    o.slope()
    o.flux(case.dt_xi)
    flux, vflux = o.get_flux()
    s0 = o.get_state()
    o.iterate(case.dt_xi, 1)
    s1 = o.get_state()
    g = case.config.gas
    mid = np.asarray(case.topo["vs_mid"])
    wt = np.asarray(case.topo["vs_weight"])
    nv = len(wt)
    dt = case.dt_xi
    c = 5
    V = np.prod(np.asarray(case.topo["cell_ds"]).reshape(-1, 2)[c])
    w = s0["w"][4 * c:4 * c + 4] + flux[4 * c:4 * c + 4] * dt / V
    pc = O.get_prim(2, w, g.gamma)
    f = s0["df"][2 * nv * c:2 * nv * (c + 1)] + dt / V * vflux[2 * nv * c:2 * nv * (c + 1)]
    q = O.heat_flux(2, 2, mid, f, wt, pc)
    F = O.discrete_maxwell(2, 2, mid, pc, g.K)
    F = F + O.shakhov_part(2, 2, mid, F, pc, q, g.Pr, g.K)
    fp = O.i_projection(2, 2, mid, wt, f, w)
    tau = O.get_tau(2, pc, g.mu_ref, g.omega)
    expect = fp * (tau / (tau + dt)) + dt / (tau + dt) * F
    got = s1["df"][2 * nv * c:2 * nv * (c + 1)]
    assert np.max(np.abs(got - expect)) <= 1e-14 * np.max(np.abs(expect))
    assert np.allclose(s1["qf"][2 * c:2 * c + 2], q, rtol=1e-14, atol=1e-16)
    assert np.allclose(s1["prim"][4 * c:4 * c + 4], pc, rtol=1e-15, atol=0)
    o.close()
