"""The Weyl scalar Psi4 of GR/constraints.py:46-155 as restated in
sympygr_amd/codegen/bssn_eqs.py (psi4_stages), checked on exact data where
the answer is fixed by the physics, with exact derivatives from sympy (no
stencils):

* a curved slice t = T(x) of Minkowski space: the Weyl tensor vanishes, so
  Re Psi4 = 0 while its Ricci, extrinsic-curvature and curvature-derivative
  parts are each non-zero (a wrong sign or factor between them fails); the
  reference's Im Psi4 differs from zero by exactly one term it writes with a
  flipped sign (checked as such);
* Schwarzschild in isotropic coordinates at t = 0: the tetrad is the
  spacetime's principal frame and Psi4 = 0, while the Ricci tensor is not;
* linearised plane waves along z at a point near the +z axis: the wave
  moving outward (+z) gives |Psi4| = |d^2 h/dt^2| (to first order), the
  ingoing one (-z) gives 0 to first order; a flipped sign of the
  extrinsic-curvature-derivative part would swap the two.

The Hamiltonian and momentum constraints vanish on the Minkowski slice too.
The oracle's oracle_constraints_psi4 is then pinned to the statement on
polynomial fields, which its stencils differentiate exactly.
"""
import numpy as np
import pytest
import sympy as sp

from sympygr_amd.codegen.bssn_eqs import E3, SYM, Inputs, constraint_stages, sym_slot

X = sp.symbols("x y z")


def _bssn(gamma, Kij, det, igamma):
    """BSSN variables of physical data gamma_ij, K_ij (sympy matrices) with
    det gamma and the inverse metric given in closed form."""
    chi = det ** sp.Rational(-1, 3)
    gt = chi * gamma
    igt = igamma / chi
    K = sum(igamma[i, j] * Kij[i, j] for i in E3 for j in E3)
    At = chi * (Kij - gamma * K / 3)
    C2 = [[[sum(igt[i, l] * (sp.diff(gt[l, j], X[k]) + sp.diff(gt[l, k], X[j]) - sp.diff(gt[j, k], X[l]))
                for l in E3) / 2 for k in E3] for j in E3] for i in E3]
    Gt = [sum(igt[j, k] * C2[i][j][k] for j in E3 for k in E3) for i in E3]
    return chi, gt, igt, K, At, Gt


def _inputs(chi, gt, igt, K, At, Gt, p):
    """Values of every input symbol of the constraint statement at point p."""
    sub = dict(zip(X, p))
    ev = lambda e: sp.Float(sp.sympify(e).subs(sub).evalf(40), 40)  # noqa: E731
    I = Inputs()
    v = {sp.Symbol("px"): sp.Float(p[0], 40), sp.Symbol("py"): sp.Float(p[1], 40), sp.Symbol("pz"): sp.Float(p[2], 40)}
    v[I.u["chi"]] = ev(chi)
    v[I.u["K"]] = ev(K)
    for i in E3:
        v[I.u[f"Gt{i}"]] = ev(Gt[i])
        v[I.d1("chi", i)] = ev(sp.diff(chi, X[i]))
        v[I.d1("K", i)] = ev(sp.diff(K, X[i]))
        for j in E3:
            v[I.d1(f"Gt{i}", j)] = ev(sp.diff(Gt[i], X[j]))
            if j >= i:
                v[I.d2("chi", i, j)] = ev(sp.diff(chi, X[i], X[j]))
    for (i, j), s in SYM.items():
        v[I.u[f"gt{s}"]] = ev(gt[i, j])
        v[I.u[f"At{s}"]] = ev(At[i, j])
        v[I.lap(f"gt{s}")] = ev(sum(igt[l, m] * sp.diff(gt[i, j], X[l], X[m]) for l in E3 for m in E3))
        for k in E3:
            v[I.d1(f"gt{s}", k)] = ev(sp.diff(gt[i, j], X[k]))
            v[I.d1(f"At{s}", k)] = ev(sp.diff(At[i, j], X[k]))
    return v


def _evaluate(v, parts=False):
    S = constraint_stages(psi4=True, parts=parts)
    vals = dict(v)
    for _, pairs in S.stages:
        for s, e in pairs:
            vals[s] = sp.Float(e.xreplace(vals).evalf(40), 40)
    return {n: float(e.xreplace(vals).evalf(40)) for n, e in S.outputs}


def _minkowski_slice():
    """The slice t = T(x, y, z) of Minkowski space (T a cubic, |grad T| < 1
    near the test points): gamma_ij = delta_ij - T_i T_j and K_ij = -T_ij /
    sqrt(1 - |grad T|^2) (the sign of K does not change Psi4's vanishing)."""
    x, y, z = X
    T = (x * y + sp.Rational(1, 2) * y * z - sp.Rational(3, 10) * x * z + sp.Rational(1, 5) * x ** 2
         + sp.Rational(1, 10) * x * y * z) / 5
    dT = [sp.diff(T, a) for a in X]
    gamma = sp.Matrix(3, 3, lambda i, j: (1 if i == j else 0) - dT[i] * dT[j])
    det = 1 - sum(d * d for d in dT)
    igamma = sp.Matrix(3, 3, lambda i, j: (1 if i == j else 0) + dT[i] * dT[j] / det)
    Kij = sp.Matrix(3, 3, lambda i, j: -sp.diff(T, X[i], X[j]) / sp.sqrt(det))
    return gamma, Kij, det, igamma


@pytest.mark.parametrize("p", [(0.7, -0.4, 0.9), (-0.5, 0.8, 0.3)])
def test_minkowski_slice_gives_zero_psi4_and_constraints(p):
    """H, M_i and Re Psi4 vanish (to rounding, against parts of size 1e-3 to
    1e-2, the magnetic part included).  Im Psi4 equals exactly
    -(Re m.A Im m.d chi + Im m.A Re m.d chi) / chi^2: GR/constraints.py:150
    writes the d chi terms of its last imaginary part with the opposite sign
    of the real part's (:149), and with the real part's signs Im Psi4
    vanishes too.  This build keeps the reference's statement."""
    o = _evaluate(_inputs(*_bssn(*_minkowski_slice()), p), parts=True)
    print({k: f"{x:.3e}" for k, x in o.items()})
    for k in ("psi4_p1r", "psi4_p2r", "psi4_p3r", "psi4_p4r", "psi4_p1i", "psi4_p2i", "psi4_p4i"):
        assert abs(o[k]) > 1e-4, k
    for k in ("psi4_real", "ham", "mom0", "mom1", "mom2"):
        assert abs(o[k]) < 1e-15, (k, o[k])
    term = o["psi4_ic"] ** 2 * (o["psi4_mrA"] * o["psi4_mid"] + o["psi4_miA"] * o["psi4_mrd"])
    assert abs(term) > 1e-5
    assert abs(o["psi4_img"] + term) < 1e-15 * max(1.0, abs(term)) * 10, (o["psi4_img"], term)


def test_schwarzschild_gives_zero_psi4():
    """Time-symmetric Schwarzschild data, psi = 1 + M/(2r) with M = 1:
    chi = psi^-4, gt = delta, A~ = K = Gt = 0.  H = 0 and Psi4 = 0, while the
    Ricci part of Psi4 is built from a Ricci tensor of order one."""
    r = sp.sqrt(X[0] ** 2 + X[1] ** 2 + X[2] ** 2)
    chi = (1 + sp.Rational(1, 2) / r) ** -4
    z3 = sp.zeros(3, 3)
    o = _evaluate(_inputs(chi, sp.eye(3), sp.eye(3), sp.Integer(0), z3, [0, 0, 0], (0.9, -0.6, 1.3)), parts=True)
    print({k: f"{x:.3e}" for k, x in o.items()})
    assert abs(o["ham"]) < 1e-13
    assert abs(o["psi4_real"]) < 1e-13 and abs(o["psi4_img"]) < 1e-13


def _plane_wave(direction, A=sp.Rational(1, 10000)):
    """gamma = delta + h(t - s z) with h_xx = -h_yy = A sin(z0 - s z ...) at t = 0,
    K_ij = -1/2 d_t gamma_ij; s = +1 outgoing (+z), -1 ingoing."""
    s = direction
    k = sp.Rational(3, 2)
    ph = k * (-s * X[2])          # phase k (t - s z) at t = 0
    h = A * sp.sin(ph)
    dht = A * k * sp.cos(ph)      # d_t h at t = 0
    gamma = sp.Matrix([[1 + h, 0, 0], [0, 1 - h, 0], [0, 0, 1]])
    Kij = sp.Matrix([[-dht / 2, 0, 0], [0, dht / 2, 0], [0, 0, 0]])
    igamma = sp.Matrix([[1 / (1 + h), 0, 0], [0, 1 / (1 - h), 0], [0, 0, 1]])
    return gamma, Kij, 1 - h * h, igamma, A * k * k


def test_plane_waves_outgoing_seen_ingoing_not():
    p = (0.03, 0.04, 2.0)
    outs = {}
    for s in (1, -1):
        gamma, Kij, det, igamma, amp = _plane_wave(s)
        o = _evaluate(_inputs(*_bssn(gamma, Kij, det, igamma), p))
        outs[s] = complex(o["psi4_real"], o["psi4_img"])
        print(s, outs[s], float(amp))
    ddh = float(sp.Rational(1, 10000) * sp.Rational(9, 4) * abs(sp.sin(sp.Rational(3, 2) * -2.0)))
    # outgoing: |Psi4| = |d^2 h / dt^2| (up to O(A^2)); ingoing: zero to first order
    assert abs(abs(outs[1]) - ddh) < 1e-3 * ddh, (abs(outs[1]), ddh)
    assert abs(outs[-1]) < 1e-3 * ddh, (abs(outs[-1]), ddh)


def test_committed_psi4_algebra_is_up_to_date():
    import os
    from sympygr_amd.codegen import __main__ as gen
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    _, text = gen.generate_constraints_psi4(write=False)
    for rel in ("sympygr_amd/csrc/gen/bssn_constraints_psi4.inc", "oracle/gen/bssn_constraints_psi4.inc"):
        with open(os.path.join(root, rel)) as f:
            assert f.read() == text, f"{rel} is stale: run python -m sympygr_amd.codegen"


def _poly_fields():
    """Polynomial BSSN fields of degree <= 3 (chi > 0 and gt positive definite
    on the test block): the oracle's sixth-order stencils differentiate them
    exactly, so its Psi4 must equal the statement with exact derivatives."""
    x, y, z = X
    chi = 0.8 + 0.1 * x - 0.05 * y * z + 0.03 * x * x * z
    gt = sp.Matrix([[1.1 + 0.05 * x * y, 0.02 * z - 0.01 * x * x, 0.03 * y],
                    [0.02 * z - 0.01 * x * x, 0.95 - 0.04 * z * z, 0.01 * x * y * z],
                    [0.03 * y, 0.01 * x * y * z, 1.02 + 0.06 * x]])
    At = sp.Matrix([[0.02 * y, 0.01 * x - 0.03 * z, 0.015],
                    [0.01 * x - 0.03 * z, -0.01 * x * z, 0.02 * y * y],
                    [0.015, 0.02 * y * y, 0.005 * x * y - 0.01]])
    K = 0.1 - 0.2 * x * y + 0.05 * z ** 3
    Gt = [0.02 * x * z, -0.03 + 0.01 * y, 0.04 * x * y * z]
    return chi, gt, K, At, Gt


def test_oracle_psi4_equals_statement_on_polynomial_fields():
    """oracle_constraints_psi4 at two interior nodes of a 13^3 block vs the
    statement at those nodes' coordinates pmin + (i, j, k) h; a wrong
    coordinate, derivative slot or output row fails by orders of magnitude."""
    import pyoracle as po
    chi, gt, K, At, Gt = _poly_fields()
    n, h = 13, 0.05
    pmin = (0.35, -0.7, 0.6)
    pmax = tuple(p + (n - 1) * h for p in pmin)
    kk, jj, ii = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    grid = [pmin[0] + ii * h, pmin[1] + jj * h, pmin[2] + kk * h]
    u = np.zeros((24, n ** 3))
    num = lambda e: np.broadcast_to(sp.lambdify(X, e, "numpy")(*grid), (n, n, n)).reshape(-1)  # noqa: E731
    u[0] = 1.0
    u[1] = num(chi)
    u[2] = num(K)
    for i in E3:
        u[3 + i] = num(Gt[i])
    for (i, j), s in SYM.items():
        u[12 + s] = num(gt[i, j])
        u[18 + s] = num(At[i, j])
    got = po.constraints_block(u, (n, n, n), pmin, pmax, 0, psi4=True).reshape(6, n, n, n)
    plain = po.constraints_block(u, (n, n, n), pmin, pmax, 0).reshape(4, n, n, n)
    np.testing.assert_allclose(got[:4, 3:-3, 3:-3, 3:-3], plain[:, 3:-3, 3:-3, 3:-3], rtol=1e-12, atol=1e-13)
    assert np.all(got[:, :3] == 0) and np.all(got[:, :, :, -3:] == 0)
    for (i, j, k) in [(3, 4, 9), (8, 6, 3)]:
        p = (pmin[0] + i * h, pmin[1] + j * h, pmin[2] + k * h)
        igt = gt.subs(dict(zip(X, p))).evalf(40).inv()
        o = _evaluate(_inputs(chi, gt, igt, K, At, Gt, p))
        want = np.array([o[n_] for n_ in ("ham", "mom0", "mom1", "mom2", "psi4_real", "psi4_img")])
        have = got[:, k, j, i]
        print((i, j, k), want, have)
        assert np.all(np.abs(want[4:]) > 1e-4), want
        np.testing.assert_allclose(have, want, rtol=1e-10, atol=1e-13)  # seen: <= 5e-12
