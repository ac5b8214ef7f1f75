"""Closed-form pins of the oracle's KitCore restatement (lib/KitCore/*.jl).

The reference ships no golden vectors, so the point functions are pinned by what the
mathematics fixes: discrete moments of a sampled Maxwellian against the continuous moments
(error bounded by the truncated tail), exact inverses, symmetry zeros, and closed-form
constants. A dropped term (the b-moment in the energy, the K in B, a wrong sign in a peculiar
velocity) moves one of these by orders of magnitude more than the stated tolerance.
"""
import math

import numpy as np
import pytest

import oracle as O
from kitamr_hip.vspace import VelocityGrid


def grid2(n=28, vmax=5.0):
    return VelocityGrid(2, [-vmax, vmax, -vmax, vmax], [n, n])


def grid3(n=24, vmax=6.0):
    return VelocityGrid(3, [-vmax, vmax] * 3, [n, n, n])


QUADRANTS = [(1.0, 0.0, 0.0, 1.0), (0.5, 0.5, 0.0, 1.25), (0.125, 0.0, 0.0, 1.25),
             (0.5, 0.0, 0.5, 1.25)]


@pytest.mark.parametrize("prim", QUADRANTS)
@pytest.mark.parametrize("K", [0.0, 1.0, 2.0])
def test_maxwellian_moments_2d(prim, K):
    """micro_to_macro(discrete_maxwell(prim)) == get_conserved(prim) up to the tail/aliasing
    bound of the midpoint rule on [-5,5]^2 (C1's velocity grid)."""
    g = grid2()
    gamma = (K + 4.0) / (K + 2.0)          # 2D: gamma = (K+4)/(K+2) with K internal dof
    M = O.discrete_maxwell(2, 2, g.mid_colmajor(), prim, K)
    w = O.micro_to_macro(2, 2, g.mid_colmajor(), M, g.weight)
    w_exact = O.get_conserved(2, prim, gamma)
    lam, U, V = prim[3], prim[1], prim[2]
    # relative tail of the Gaussian outside |u|<=5 in each dimension; the energy row carries
    # c^2 ~ 2*a^2 more weight there (a = distance of the cut from the centre)
    a = math.sqrt(lam) * (5 - max(abs(U), abs(V)))
    tail = math.erfc(a) * (2 + 2 * a * a)
    assert np.allclose(w, w_exact, rtol=4 * tail + 1e-13, atol=1e-13), (w - w_exact, tail)
    # and much tighter than a dropped-term error: e.g. the b-moment is 0.25*K*rho/lambda
    if K > 0:
        assert abs(w[3] - w_exact[3]) < 1e-3 * (0.25 * K * prim[0] / lam)


def test_maxwellian_b_is_h_times_K_over_2lambda():
    g = grid2(8)
    prim = [0.7, 0.3, -0.2, 1.1]
    M = O.discrete_maxwell(2, 2, g.mid_colmajor(), prim, 1.0).reshape(2, -1)
    assert np.array_equal(M[1], M[0] * 1.0 / (2.0 * prim[3]))
    # h itself: rho*(lambda/pi)*exp(-lambda c^2) at one cell, by hand
    u, v = g.midpoint[5]
    h = prim[0] * (prim[3] / math.pi) * math.exp(-prim[3] * ((u - prim[1]) ** 2 + (v - prim[2]) ** 2))
    assert M[0][5] == pytest.approx(h, rel=1e-15)


@pytest.mark.parametrize("prim", [(1.0, 0.0, 0.0, 0.0, 1.0), (0.4, 0.3, -0.2, 0.1, 0.8)])
def test_maxwellian_moments_3d(prim):
    g = grid3()
    M = O.discrete_maxwell(3, 1, g.mid_colmajor(), prim, 0.0)
    w = O.micro_to_macro(3, 1, g.mid_colmajor(), M, g.weight)
    w_exact = O.get_conserved(3, prim, 5.0 / 3.0)
    a = math.sqrt(prim[4]) * (6 - max(abs(x) for x in prim[1:4]))
    tail = math.erfc(a) * (3 + 2 * a * a)
    alias = 2 * math.exp(-math.pi ** 2 / (prim[4] * g.ds[0] ** 2))   # midpoint-rule aliasing
    assert np.allclose(w, w_exact, rtol=6 * tail + 6 * alias + 1e-13, atol=1e-13), (w - w_exact)


@pytest.mark.parametrize("dim", [2, 3])
def test_prim_conserved_inverse(dim):
    rng = np.random.default_rng(3)
    for _ in range(50):
        prim = np.concatenate([[rng.uniform(0.1, 3)], rng.uniform(-2, 2, dim), [rng.uniform(0.2, 4)]])
        w = O.get_conserved(dim, prim, 5.0 / 3.0)
        back = O.get_prim(dim, w, 5.0 / 3.0)
        assert np.allclose(back, prim, rtol=1e-13, atol=1e-14)


def test_heat_flux_of_centred_maxwellian_is_zero():
    """On a grid symmetric about U=V=0 the heat flux of a centred Maxwellian vanishes, so the
    Shakhov part vanishes too (g+ = M)."""
    g = grid2(28)
    prim = [1.0, 0.0, 0.0, 1.0]
    M = O.discrete_maxwell(2, 2, g.mid_colmajor(), prim, 1.0)
    q = O.heat_flux(2, 2, g.mid_colmajor(), M, g.weight, prim)
    assert np.max(np.abs(q)) < 1e-15
    S = O.shakhov_part(2, 2, g.mid_colmajor(), M, prim, q, 2.0 / 3.0, 1.0)
    assert np.max(np.abs(S)) < 1e-16


def test_heat_flux_closed_form_shifted_distribution():
    """f = M*(1 + eps*(u-U)(c^2 - 2/lambda ...)) style check: the heat flux is linear in f and
    odd in (u-U); a distribution mirrored in u gives the opposite q_x and the same q_y."""
    g = grid2(28)
    prim = [1.0, 0.0, 0.0, 1.0]
    rng = np.random.default_rng(5)
    f = rng.uniform(0, 1, 2 * g.n)
    # mirror u -> -u: grid index (i, j) -> (n-1-i, j), first dimension fastest
    n = 28
    idx = np.arange(g.n)
    i, j = idx % n, idx // n
    mirror = (n - 1 - i) + n * j
    fm = f.reshape(2, -1)[:, mirror].reshape(-1)
    q = O.heat_flux(2, 2, g.mid_colmajor(), f, g.weight, prim)
    qm = O.heat_flux(2, 2, g.mid_colmajor(), fm, g.weight, prim)
    assert qm[0] == pytest.approx(-q[0], rel=1e-12)
    assert qm[1] == pytest.approx(q[1], rel=1e-12)


def test_shakhov_closed_form_point():
    g = grid2(4)
    prim = [0.9, 0.2, -0.1, 1.3]
    q = [0.05, -0.02]
    Pr, K = 2.0 / 3.0, 1.0
    M = O.discrete_maxwell(2, 2, g.mid_colmajor(), prim, K).reshape(2, -1)
    S = O.shakhov_part(2, 2, g.mid_colmajor(), M.reshape(-1), prim, q, Pr, K).reshape(2, -1)
    for c in range(g.n):
        u, v = g.midpoint[c]
        cu, cv = u - prim[1], v - prim[2]
        c2 = cu * cu + cv * cv
        pre = 0.8 * (1 - Pr) * prim[3] ** 2 / prim[0] * (cu * q[0] + cv * q[1])
        assert S[0][c] == pytest.approx(pre * (2 * prim[3] * c2 + K - 5) * M[0][c], rel=1e-14, abs=1e-300)
        assert S[1][c] == pytest.approx(pre * (2 * prim[3] * c2 + K - 3) * M[1][c], rel=1e-14, abs=1e-300)


def test_shakhov_conserves_mass_momentum_energy():
    """The Shakhov correction carries no mass, momentum or energy (continuous identity; on
    the discrete grid up to the tail)."""
    g = grid2(48, 6.0)
    prim = [1.0, 0.3, -0.2, 1.0]
    M = O.discrete_maxwell(2, 2, g.mid_colmajor(), prim, 1.0)
    S = O.shakhov_part(2, 2, g.mid_colmajor(), M, prim, [0.1, 0.05], 2.0 / 3.0, 1.0)
    w = O.micro_to_macro(2, 2, g.mid_colmajor(), S, g.weight)
    assert np.max(np.abs(w)) < 1e-8


def test_tau_and_vhs_viscosity():
    prim = [0.8, 0.1, 0.0, 1.25]
    mu, omega = 0.03, 0.81
    assert O.get_tau(2, prim, mu, omega) == pytest.approx(mu * 2.0 * 1.25 ** (1 - omega) / 0.8, rel=1e-15)
    # ref_vhs_vis(Kn, omega, alpha): 5(a+1)(a+2)sqrt(pi) / (4a(5-2w)(7-2w)) Kn
    Kn = 0.01
    a, w = 1.0, 0.81
    exact = 5 * (a + 1) * (a + 2) * math.sqrt(math.pi) / (4 * a * (5 - 2 * w) * (7 - 2 * w)) * Kn
    assert O.ref_vhs_vis(Kn, w, a) == pytest.approx(exact, rel=1e-15)
    # Gas(...) passes (Kn, alpha_ref, omega_ref) into (Kn, omega, alpha): the default mu_ref is
    # ~0.927*Kn instead of the intended ~0.731*Kn (src/Gas/Types.jl:46 vs src/Gas/Model.jl:5)
    swapped = O.ref_vhs_vis(Kn, 1.0, 0.81)
    assert swapped / Kn == pytest.approx(
        5 * 1.81 * 2.81 * math.sqrt(math.pi) / (4 * 0.81 * 3.0 * 5.0), rel=1e-14)   # ~0.927
    assert exact / Kn == pytest.approx(0.7310, abs=5e-4)
    from kitamr_hip.solver import Gas
    assert Gas(Kn=Kn).mu_ref == pytest.approx(swapped, rel=1e-15)
