"""CPU tests of the oracle's Hamiltonian/momentum constraints
(GR/constraints.py:161-173).  No compiled reference exists for them
(bssn/CodeGen/physcon.cpp does not build), so they are pinned by what the
mathematics fixes: flat space, a closed-form momentum constraint, and the
sixth-order convergence of the Hamiltonian constraint to zero on
Brill-Lindquist data (time-symmetric vacuum data with a harmonic conformal
factor, so the exact constraint vanishes)."""
import numpy as np

import pyoracle as po
from sympygr_amd import data


def _block(n, lo, h):
    return (n, n, n), (lo,) * 3, (lo + (n - 1) * h,) * 3


def test_flat_space_constraints_vanish():
    sz, pmin, pmax = _block(17, -1.0, 0.125)
    u = np.zeros((24, 17 ** 3))
    for f in (0, 1, 12, 15, 17):
        u[f] = 1.0
    c = po.constraints_block(u, sz, pmin, pmax, 0)
    assert np.abs(c).max() == 0.0


def test_momentum_constraint_closed_form():
    """Flat metric, chi = 1, constant At and Gt, K linear:
    M_i = -Gt^j A_ij - 2/3 d_i K exactly (every other term vanishes)."""
    n, h = 17, 0.125
    sz, pmin, pmax = _block(n, -1.0, h)
    x, y, z = data.block_coords(data._Np(), sz, pmin, pmax)
    u = np.zeros((24, n ** 3))
    for f in (0, 1, 12, 15, 17):
        u[f] = 1.0
    A = np.array([[0.3, -0.2, 0.1], [-0.2, 0.5, 0.4], [0.1, 0.4, -0.8]])
    slots = {(0, 0): 18, (0, 1): 19, (0, 2): 20, (1, 1): 21, (1, 2): 22, (2, 2): 23}
    for (i, j), f in slots.items():
        u[f] = A[i, j]
    G = np.array([0.7, -0.3, 0.2])
    for i in range(3):
        u[3 + i] = G[i]
    dK = np.array([0.11, -0.07, 0.05])
    u[2] = (0.4 + dK[0] * x + dK[1] * y + dK[2] * z).reshape(-1)
    c = po.constraints_block(u, sz, pmin, pmax, 0).reshape(4, n, n, n)[:, 3:-3, 3:-3, 3:-3]
    for i in range(3):
        expect = -(G @ A[i]) - 2.0 / 3.0 * dK[i]
        np.testing.assert_allclose(c[1 + i], expect, rtol=1e-12, atol=1e-14)
    # H = -A_ij A^ij + 2/3 K^2 on the flat metric
    K = u[2].reshape(n, n, n)[3:-3, 3:-3, 3:-3]
    np.testing.assert_allclose(c[0], -np.sum(A * A) + 2.0 / 3.0 * K * K, rtol=1e-12)


def _bl_hamiltonian_at(point, h):
    n = 15
    lo = [point[d] - 7 * h for d in range(3)]
    sz = (n, n, n)
    pmin = tuple(lo)
    pmax = tuple(lo[d] + (n - 1) * h for d in range(3))
    x, y, z = data.block_coords(data._Np(), sz, pmin, pmax)
    vals = [None] * 24
    data.brill_lindquist_fill(data._Np(), x, y, z, vals)
    u = np.stack([v.reshape(-1) for v in vals])
    c = po.constraints_block(u, sz, pmin, pmax, 0)
    return c[0].reshape(n, n, n)[7, 7, 7]


def test_hamiltonian_constraint_converges_at_sixth_order():
    p = (0.7, 0.9, 0.6)  # away from the punctures at (+-2, 0, 0)
    h1 = _bl_hamiltonian_at(p, 0.1)
    h2 = _bl_hamiltonian_at(p, 0.05)
    ratio = abs(h1) / abs(h2)
    assert abs(h1) < 1e-4
    assert 40.0 < ratio < 100.0, (h1, h2, ratio)


# ---------------------------------------------------------------------------
# Non-flat conformal metric, Brill-Lindquist conformal factor, polynomial
# trace-free A~ and K.  The expected constraints are those of the physical
# data gamma_ij = gt_ij / chi, A_ij = At_ij / chi and K, evaluated by the
# textbook route (physical Christoffels and Ricci scalar of gamma, physical
# covariant divergence of A^j_i), with exact derivatives from sympy, so they
# share nothing with the oracle's conformal decomposition or its stencils:
#   H = R(gamma) - A_ij A^ij + 2/3 K^2,   M_i = D_j A^j_i - 2/3 d_i K
# (A_ij = At_ij / chi, A^j_i = gt^jk At_ki).  The conformal forms of
# GR/constraints.py:164-173 equal these when det gt = 1 (gt = L L^T with L
# unit lower triangular), gt^ij At_ij = 0 and Gt^i = gt^jk C2^i_jk, which the
# data satisfy exactly.  The oracle must converge to them at sixth order.


def _analytic_data():
    import sympy as sp
    x, y, z = sp.symbols("x y z")
    X = (x, y, z)
    L = sp.Matrix([[1, 0, 0],
                   [sp.Rational(1, 10) * x * y + sp.Rational(1, 20) * z ** 2 + sp.Rational(1, 10), 1, 0],
                   [-sp.Rational(2, 25) * x + sp.Rational(3, 100) * y * z, sp.Rational(3, 50) * y ** 2 - sp.Rational(1, 25) * x * z + sp.Rational(1, 20), 1]])
    gt = L * L.T
    igt = (L.inv()).T * L.inv()
    r1 = sp.sqrt((x - 2) ** 2 + y ** 2 + z ** 2)
    r2 = sp.sqrt((x + 2) ** 2 + y ** 2 + z ** 2)
    psi = 1 + sp.Rational(1, 2) / (2 * r1) + sp.Rational(1, 2) / (2 * r2)
    chi = psi ** -4
    Bm = sp.Matrix([[sp.Rational(3, 10) + sp.Rational(1, 10) * x * y, -sp.Rational(1, 5) * z + sp.Rational(1, 20) * x ** 2, sp.Rational(1, 10) + sp.Rational(1, 10) * y],
                    [0, sp.Rational(1, 2) - sp.Rational(1, 5) * x * z, sp.Rational(2, 5) * x + sp.Rational(1, 20) * y * z],
                    [0, 0, -sp.Rational(4, 5) + sp.Rational(3, 20) * y ** 2]])
    Bm = Bm + Bm.T - sp.diag(*[Bm[i, i] for i in range(3)])
    tr = sum(igt[k, l] * Bm[k, l] for k in range(3) for l in range(3))
    At = Bm - gt * tr / 3
    K = sp.Rational(3, 10) + sp.Rational(1, 10) * x - sp.Rational(1, 5) * y * z + sp.Rational(1, 20) * x ** 2
    # conformal Christoffels C2^i_jk and their contraction Gt^i
    C2 = [[[sum(igt[i, l] * (sp.diff(gt[l, j], X[k]) + sp.diff(gt[l, k], X[j]) - sp.diff(gt[j, k], X[l]))
                for l in range(3)) / 2 for k in range(3)] for j in range(3)] for i in range(3)]
    Gt = [sum(igt[j, k] * C2[i][j][k] for j in range(3) for k in range(3)) for i in range(3)]
    return X, gt, igt, chi, At, K, Gt


def _fields(X, gt, chi, At, K, Gt, xg, yg, zg):
    import sympy as sp
    from sympygr_amd.codegen.bssn_eqs import SYM
    f = [None] * 24
    ones = np.ones_like(xg)
    f[0] = ones
    f[1] = sp.lambdify(X, chi, "numpy")(xg, yg, zg) * ones
    f[2] = sp.lambdify(X, K, "numpy")(xg, yg, zg) * ones
    for i in range(3):
        f[3 + i] = sp.lambdify(X, Gt[i], "numpy")(xg, yg, zg) * ones
        f[6 + i] = 0.0 * ones
        f[9 + i] = 0.0 * ones
    for (i, j), s in SYM.items():
        f[12 + s] = sp.lambdify(X, gt[i, j], "numpy")(xg, yg, zg) * ones
        f[18 + s] = sp.lambdify(X, At[i, j], "numpy")(xg, yg, zg) * ones
    return np.stack([v.reshape(-1) for v in f])


def _expected_physical(X, gt, igt, chi, At, K, p):
    """H and M_i of the physical data at point p, textbook formulas."""
    import sympy as sp
    sub = dict(zip(X, p))
    g = gt / chi
    ev = lambda e: float(e.subs(sub).evalf(30))  # noqa: E731
    G = np.array([[ev(g[i, j]) for j in range(3)] for i in range(3)])
    dG = np.array([[[ev(sp.diff(g[i, j], X[k])) for k in range(3)] for j in range(3)] for i in range(3)])
    ddG = np.array([[[[ev(sp.diff(g[i, j], X[k], X[l])) for l in range(3)] for k in range(3)] for j in range(3)]
                    for i in range(3)])
    Gi = np.linalg.inv(G)
    # first-kind Gamma_kij = (d_i g_kj + d_j g_ki - d_k g_ij) / 2 and its derivative
    # (dG[a][b][c] = d_c g_ab, ddG[a][b][c][d] = d_c d_d g_ab)
    C1 = 0.5 * (np.einsum("kji->kij", dG) + dG - np.einsum("ijk->kij", dG))
    dC1 = 0.5 * (np.einsum("kjil->kijl", ddG) + ddG - np.einsum("ijkl->kijl", ddG))
    dGi = -np.einsum("ka,abl,bm->kml", Gi, dG, Gi)
    C2 = np.einsum("km,mij->kij", Gi, C1)
    dC2 = np.einsum("kml,mij->kijl", dGi, C1) + np.einsum("km,mijl->kijl", Gi, dC1)
    # R_ij = d_k C2^k_ij - d_j C2^k_ik + C2^k_kl C2^l_ij - C2^k_jl C2^l_ik
    Ric = (np.einsum("kijk->ij", dC2) - np.einsum("kikj->ij", dC2) + np.einsum("kkl,lij->ij", C2, C2)
           - np.einsum("kjl,lik->ij", C2, C2))
    R = np.einsum("ij,ij->", Gi, Ric)
    Amix = (igt * At)  # A^j_i = gt^jk At_ki (the 1/chi factors cancel)
    Aphys = np.array([[ev(At[i, j] / chi) for j in range(3)] for i in range(3)])
    AA = np.einsum("ij,ik,jl,kl->", Aphys, Gi, Gi, Aphys)
    Kp = ev(K)
    H = R - AA + 2.0 / 3.0 * Kp * Kp
    Amx = np.array([[ev(Amix[j, i]) for i in range(3)] for j in range(3)])          # [j][i] = A^j_i
    div = np.array([sum(ev(sp.diff(Amix[j, i], X[j])) for j in range(3)) for i in range(3)])
    DA = div + np.einsum("jjl,li->i", C2, Amx) - np.einsum("lji,jl->i", C2, Amx)
    dK = np.array([ev(sp.diff(K, X[i])) for i in range(3)])
    return np.concatenate([[H], DA - 2.0 / 3.0 * dK])


def _oracle_constraints_at(data_, p, h, n=15):
    X, gt, igt, chi, At, K, Gt = data_
    c = n // 2
    lo = [p[d] - c * h for d in range(3)]
    sz = (n, n, n)
    pmin = tuple(lo)
    pmax = tuple(lo[d] + (n - 1) * h for d in range(3))
    xg, yg, zg = data.block_coords(data._Np(), sz, pmin, pmax)
    u = _fields(X, gt, chi, At, K, Gt, xg, yg, zg)
    out = po.constraints_block(u, sz, pmin, pmax, 0)
    return out.reshape(4, n, n, n)[:, c, c, c]


def test_constraints_converge_to_textbook_values_on_non_flat_data():
    d = _analytic_data()
    X, gt, igt, chi, At, K, Gt = d
    p = (0.7, 0.9, 0.6)
    expect = _expected_physical(X, gt, igt, chi, At, K, p)
    e1 = np.abs(_oracle_constraints_at(d, p, 0.1) - expect)
    e2 = np.abs(_oracle_constraints_at(d, p, 0.05) - expect)
    print("expected", expect, "err h", e1, "err h/2", e2)
    assert e2.max() < 1e-6
    for q in range(4):
        assert 40.0 < e1[q] / e2[q] < 100.0, (q, e1, e2)


def test_every_momentum_term_is_live_in_the_convergence_test():
    """Each term of M_i in GR/constraints.py:164-173 is non-zero at the test
    point and at least 1e4 times the h/2 discretisation error, so a dropped
    term or a flipped sign (d_k A~_ij, the Christoffel term, Gt^j A~_ij,
    3/2 A~ d chi / chi, 2/3 d K) cannot pass the convergence test."""
    import sympy as sp
    d = _analytic_data()
    X, gt, igt, chi, At, K, Gt = d
    p = (0.7, 0.9, 0.6)
    sub = dict(zip(X, p))
    ev = lambda e: float(sp.sympify(e).subs(sub).evalf(30))  # noqa: E731
    C2 = [[[sum(igt[i, l] * (sp.diff(gt[l, j], X[k]) + sp.diff(gt[l, k], X[j]) - sp.diff(gt[j, k], X[l]))
                for l in range(3)) / 2 for k in range(3)] for j in range(3)] for i in range(3)]
    err = np.abs(_oracle_constraints_at(d, p, 0.05) - _expected_physical(X, gt, igt, chi, At, K, p))[1:].max()
    for i in range(3):
        terms = {
            "dA": sum(ev(igt[j, k] * sp.diff(At[i, j], X[k])) for j in range(3) for k in range(3)),
            "C2A": -sum(ev(igt[j, k] * C2[m][k][i] * At[j, m]) for j in range(3) for k in range(3) for m in range(3)),
            "GtA": -sum(ev(Gt[j] * At[i, j]) for j in range(3)),
            "dchi": -1.5 * sum(ev(igt[j, k] * At[k, i] * sp.diff(chi, X[j]) / chi) for j in range(3) for k in range(3)),
            "dK": -2.0 / 3.0 * ev(sp.diff(K, X[i])),
        }
        for name, v in terms.items():
            assert abs(v) > 1e4 * err, (i, name, v, err)
