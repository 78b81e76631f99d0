"""TVB dissipation in the oracle (bssn/src/derivs.cpp:1348-2054 with the
speeds of bssn/src/rhs.cpp:984-1028), on the CPU: the characteristic speeds
and every single sweep bit-equal the compiled reference's (hashes in
tests/golden/tvb.json), the values the restatement chooses where the
reference leaves the array's previous contents equal the reference's interior
form (on shifted input), the interior forms annihilate the polynomials their
difference order fixes, and the fixtures tell TVB3, TVB5 and Kreiss-Oliger
apart."""
import hashlib
import json
import os

import numpy as np
import pytest

import pyoracle as po
from helpers import GOLDEN, case_gauge, case_input, case_params, fixture_mask, golden, interior, manifest, normwise


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def hashes():
    with open(os.path.join(GOLDEN, "tvb.json")) as f:
        return json.load(f)


def fake():
    m = manifest()["fake_input"]
    return np.load(os.path.join(GOLDEN, m["file"])), m["n"]


def written(which, n, bflag):
    """Interior nodes (k, j, i) a reference TVB sweep writes: disstvb3 skips the
    first and last interior node before an unflagged face."""
    m = np.ones((n - 6,) * 3, bool)
    if which < 3:
        for side in range(2):
            if not bflag & (1 << (2 * which + side)):
                sl = [slice(None)] * 3
                sl[2 - which] = 0 if side == 0 else n - 7
                m[tuple(sl)] = False
    return m


def test_speeds_bit_equal_reference():
    u, sz, _, _, _ = case_input("tvb3_staged_b63")
    n = sz[0]
    lam = po.spacetime_speeds(u, sz)
    got = sha(np.stack([lam[d].reshape(n, n, n)[3:-3, 3:-3, 3:-3] for d in range(3)]))
    assert got == hashes()["speeds_gauge"]


@pytest.mark.parametrize("which", range(6))
def test_sweeps_bit_equal_reference(which):
    uf, n = fake()
    lam = po.spacetime_speeds(uf, (n, n, n))
    h = hashes()
    for bflag in (0, 63, 21, 42):
        out = po.deriv(12 + which, uf[13], 0.1, (n, n, n), bflag, beta=lam[which % 3])
        a = out.reshape(n, n, n)[3:-3, 3:-3, 3:-3]
        assert sha(np.where(written(which, n, bflag), a, 0.0)) == h[f"sweep_{which}_{bflag}"], (which, bflag)


@pytest.mark.parametrize("axis", range(3))
def test_tvb3_next_to_unflagged_face_is_the_interior_form(axis):
    """The reference leaves disstvb3's first and last interior node before an
    unflagged face unwritten; the restatement takes the interior form there,
    which the reference itself evaluates one node further in on the input
    shifted by one node."""
    uf, n = fake()
    lam = po.spacetime_speeds(uf, (n, n, n))
    out = po.deriv(12 + axis, uf[13], 0.1, (n, n, n), 0, beta=lam[axis]).reshape(n, n, n)
    h = hashes()
    for side, plane in ((0, 3), (1, n - 4)):
        sl = [slice(3, -3)] * 3
        sl[2 - axis] = plane
        assert sha(out[tuple(sl)]) == h[f"tvb3_shift_{axis}_{side}"], (axis, side)


def _poly_block(n, deg, axis):
    x = np.arange(n, dtype=np.float64) * 0.1 - 0.7
    shape = [1, 1, 1]
    shape[2 - axis] = n
    return np.broadcast_to((x ** deg).reshape(shape), (n, n, n)).ravel().copy()


@pytest.mark.parametrize("axis", range(3))
def test_interior_forms_on_polynomials(axis):
    """With constant speed c, TVB3's interior form is c (-1/(12h)) times the
    fourth difference and TVB5's is c s times the sixth (s = 2/(75h) in x and
    y, -1/(64h) in z, as the reference writes them): zero on cubics and
    quintics, 24 h^4 and 720 h^6 times the leading coefficient on x^4, x^6."""
    n, h, c = 17, 0.1, 0.7
    lam = np.full(n ** 3, c)
    inner = (slice(5, -5),) * 3
    for deg, expect3 in ((3, 0.0), (4, c * (-1.0 / 12.0 / h) * 24 * h ** 4)):
        out = po.deriv(12 + axis, _poly_block(n, deg, axis), h, (n, n, n), 0, beta=lam).reshape(n, n, n)
        assert np.allclose(out[inner], expect3, rtol=1e-9, atol=1e-9), (deg, axis)
    s5 = -1.0 / 64.0 / h if axis == 2 else 2.0 / 75.0 / h
    for deg, expect5 in ((5, 0.0), (6, c * s5 * 720 * h ** 6)):
        out = po.deriv(15 + axis, _poly_block(n, deg, axis), h, (n, n, n), 0, beta=lam).reshape(n, n, n)
        assert np.allclose(out[inner], expect5, rtol=1e-8, atol=1e-9), (deg, axis)


def test_fixtures_tell_the_dissipations_apart():
    """The TVB fixtures differ from the Kreiss-Oliger right-hand side and from
    each other, so matching them pins the dissipation, not just the algebra."""
    u, sz, pmin, pmax, c = case_input("tvb3_staged_b63")
    p, up = case_params(c)
    kw = case_gauge(c)
    kw.pop("dissipation")
    ko = interior(po.rhs_block(u, sz, pmin, pmax, 63, po.make_params(upwind=up, ko_sigma=p["ko_sigma"], **kw)), sz)
    t3, t5 = golden("tvb3_staged_b63"), golden("tvb5_staged_b63")
    assert normwise(ko, t3).max() > 1e-3
    assert normwise(ko, t5).max() > 1e-3
    assert normwise(t3, t5).max() > 1e-3


def test_ccz4_with_tvb_is_refused():
    u, sz, pmin, pmax, _ = case_input("tvb3_staged_b63")
    u25 = np.vstack([u, np.zeros((1, u.shape[1]))])
    with pytest.raises(RuntimeError):
        po.rhs_block(u25, sz, pmin, pmax, 63, po.make_params(ccz4=True, dissipation="tvb3"))


def _plan_rc(prm, sz=19):
    import sympygr_amd as sgr
    blk = sgr.SgrBlock()
    blk.offset = 0
    for k in range(3):
        blk.sz[k] = sz
        blk.pmin[k] = 0.0
        blk.pmax[k] = 1.8
    blk.bflag = 0
    import ctypes
    plan = ctypes.c_void_p()
    return sgr.lib().sgr_plan_create(ctypes.byref(plan), ctypes.byref(blk), 1, ctypes.byref(prm), 0)


def test_library_refuses_unsupported_dissipation():
    """Checked before any device call: an unknown dissipation value
    (SGR_E_INVALID), TVB with CCZ4 (SGR_E_UNSUPPORTED), TVB in the
    allocation-free entry (SGR_E_UNSUPPORTED)."""
    import ctypes
    import sympygr_amd as sgr
    prm = sgr.params(dissipation="tvb3")
    prm.dissipation = 3
    assert _plan_rc(prm) == -1
    assert "dissipation" in sgr.last_error()
    assert _plan_rc(sgr.params(dissipation="tvb5", formulation="ccz4")) == -2
    assert "BSSN" in sgr.last_error()
    n = 19
    arrs = [np.zeros(n ** 3) for _ in range(48)]
    PD = ctypes.POINTER(ctypes.c_double)
    ins = (PD * 24)(*[a.ctypes.data_as(PD) for a in arrs[:24]])
    outs = (PD * 24)(*[a.ctypes.data_as(PD) for a in arrs[24:]])
    rc = sgr.lib().sgr_rhs_block_noalloc(outs, ins, ctypes.c_uint64(0), (ctypes.c_double * 3)(0, 0, 0),
                                         (ctypes.c_double * 3)(1.8, 1.8, 1.8), (ctypes.c_uint32 * 3)(n, n, n), 0,
                                         ctypes.byref(sgr.params(dissipation="tvb3")), None)
    assert rc == -2
    assert "TVB" in sgr.last_error()
