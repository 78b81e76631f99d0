"""The oracle's RK4 stepper (pyoracle.rk4_tableau_step) pinned by what the
mathematics fixes, independently of the RHS it steps (CPU only):

* linear ODE u' = lam u: one step multiplies u by the degree-4 Taylor
  polynomial of exp(z), z = lam dt (the stability polynomial of the classic
  tableau: each stage weight and node enters one of its coefficients);
* linear system u' = A u with A a rotation generator: one step applies
  I + hA + (hA)^2/2 + (hA)^3/6 + (hA)^4/24;
* non-linear u' = u^2 (exact solution u0 / (1 - u0 t)): the one-step error
  falls 32-fold when dt halves (fifth-order local error);
* mask/fill: columns outside the mask keep their values after `fill`.

The GPU RK4 tests compare the library against this stepper, so a wrong
stage weight or node in both would fail here.  The last test plants such a
mistake in the oracle's tableau and checks that the closed form catches it.
"""
import numpy as np
import pytest

import pyoracle as po


def _taylor4(z):
    return 1.0 + z + z ** 2 / 2.0 + z ** 3 / 6.0 + z ** 4 / 24.0


def _linear_amplification_error():
    lam = np.array([-0.7, 0.3, -2.5, 1.1, -0.05])
    dt = 0.3
    u0 = np.array([[1.0, -2.0, 0.5, 3.0, 7.0]])
    got = po.rk4_tableau_step(lambda v: lam * v, u0, dt)
    expect = _taylor4(lam * dt) * u0
    return np.abs(got - expect).max() / np.abs(expect).max()


def test_linear_ode_amplification_is_the_taylor_polynomial():
    assert _linear_amplification_error() <= 1e-15


def test_rotation_system_is_the_matrix_taylor_polynomial():
    w, dt = 1.3, 0.2
    A = np.array([[0.0, -w], [w, 0.0]])
    u0 = np.array([[1.0, 0.25, -0.5], [0.0, 2.0, 0.75]])  # 3 independent columns
    got = po.rk4_tableau_step(lambda v: A @ v, u0, dt)
    hA = dt * A
    M = np.eye(2) + hA + hA @ hA / 2.0 + hA @ hA @ hA / 6.0 + hA @ hA @ hA @ hA / 24.0
    np.testing.assert_allclose(got, M @ u0, rtol=0, atol=1e-15)
    # and it is a fourth-order approximation of the rotation
    c, s = np.cos(w * dt), np.sin(w * dt)
    assert np.abs(got - np.array([[c, -s], [s, c]]) @ u0).max() < (w * dt) ** 5


def test_nonlinear_local_error_is_fifth_order():
    u0 = np.array([[0.8]])
    errs = []
    for dt in (0.05, 0.025, 0.0125):
        got = po.rk4_tableau_step(lambda v: v * v, u0, dt)[0, 0]
        errs.append(abs(got - 0.8 / (1.0 - 0.8 * dt)))
    for a, b in zip(errs, errs[1:]):
        assert 30.0 < a / b < 38.0, errs


def test_mask_and_fill():
    u0 = np.arange(12, dtype=np.float64).reshape(2, 6) + 1.0
    mask = np.array([True, False, True, True, False, True])
    seen = []

    def fill(v):
        v = v.copy()
        v[:, ~mask] = -1.0
        seen.append(v.copy())
        return v
    got = po.rk4_tableau_step(lambda v: 0.5 * v, u0, 0.1, mask=mask, fill=fill)
    assert len(seen) == 4 and all((s[:, ~mask] == -1.0).all() for s in seen)
    np.testing.assert_array_equal(got[:, ~mask], -1.0)  # the filled values of u, untouched
    np.testing.assert_allclose(got[:, mask], _taylor4(0.05) * u0[:, mask], rtol=1e-15)


@pytest.mark.parametrize("which,val", [("w", (6.0, 3.0, 6.0, 6.0)), ("w", (6.0, 6.0, 3.0, 6.0)),
                                       ("c", (2.0, 1.0, 1.0)), ("c", (2.0, 2.0, 2.0))])
def test_planted_tableau_mistake_is_caught(monkeypatch, which, val):
    monkeypatch.setattr(po, "RK4_WEIGHT_DIV" if which == "w" else "RK4_NODE_DIV", val)
    assert _linear_amplification_error() > 1e-6
