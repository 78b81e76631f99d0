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
