"""CPU tests of the product library: it loads, exports every symbol the
public headers declare, and rejects bad input with error codes (no GPU is
needed for any of these paths)."""
import ctypes
import os
import re

import pytest

import sympygr_amd as sgr
from sympygr_amd.data import BlockList

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("sympygr_amd.h", "bssn_dropin.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(sgr_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_loads_and_exports_header_symbols():
    L = sgr.lib()
    names = declared_functions()
    assert {"sgr_plan_create", "sgr_rhs", "sgr_bssnrhs_host", "sgr_last_error"} <= names
    for n in names:
        assert hasattr(L, n), n
    assert sgr.lib().sgr_version().startswith(b"sympygr_amd")


DROPIN_MANGLED = "_Z7bssnrhsPPdPPKdRKjS2_S2_PS4_S5_"


def test_dropin_symbol_exported():
    """void bssnrhs(double**, const double**, const unsigned&, const double*,
    const double*, const unsigned*, const unsigned&) with C++ linkage, plus the
    bssn:: parameter globals it reads."""
    L = ctypes.CDLL(sgr.LIB_PATH)
    assert hasattr(L, DROPIN_MANGLED)
    assert ctypes.c_double.in_dll(L, "_ZN4bssn13KO_DISS_SIGMAE").value == 0.01
    assert list((ctypes.c_uint * 4).in_dll(L, "_ZN4bssn11BSSN_LAMBDAE")) == [1, 1, 1, 1]


def test_staged_dropin_and_gauge_globals_exported():
    """bssnrhs_sep (bssn/include/rhs.h:39-42) with the same signature, and the
    gauge-variant globals with the reference defaults (parameters.cpp:50-51,123)."""
    L = ctypes.CDLL(sgr.LIB_PATH)
    assert hasattr(L, "_Z11bssnrhs_sepPPdPPKdRKjS2_S2_PS4_S5_")
    assert hasattr(L, "sgr_dropin_configure_gauge")
    assert ctypes.c_double.in_dll(L, "_ZN4bssn11BSSN_ETA_R0E").value == 1.31
    assert list((ctypes.c_double * 2).in_dll(L, "_ZN4bssn14BSSN_ETA_POWERE")) == [2.0, 2.0]
    assert list((ctypes.c_uint * 3).in_dll(L, "_ZN4bssn7BSSN_XIE")) == [0, 0, 0]


def test_params_defaults_match_reference():
    p = sgr.params(adv="centred", precision="fp64", ko_sigma=0.01)
    assert list(p.lambda_) == [1, 1, 1, 1]
    assert list(p.lambda_f) == [1.0, 0.0]
    assert (p.eta_const, p.eta_r0, p.eta_damping_exp) == (2.0, 50.0, 1.0)
    q = sgr.SgrParams()
    sgr.lib().sgr_params_default(ctypes.byref(q), 0, 0)
    assert q.ko_sigma == 0.01  # KO_DISS_SIGMA default, bssn/src/parameters.cpp:107


def _plan_error(blocks, prm=None):
    with pytest.raises(sgr.SgrError) as e:
        sgr.Plan(blocks, prm or sgr.params())
    return e.value


def _one(sz=(22, 22, 22), pmin=(0, 0, 0), pmax=(1, 1, 1), bflag=0):
    bl = BlockList()
    bl.add(sz, pmin, pmax, bflag)
    return bl


@pytest.mark.parametrize("sz", [(12, 22, 22), (22, 6, 22), (22, 22, 0)])
def test_rejects_too_small_blocks(sz):
    e = _plan_error(_one(sz=sz))
    assert e.code == -1 and "< 13" in str(e)


def test_accepts_any_interior_size_from_7():
    """Interior sizes that do not divide into tiles (9 = 1 mod 8, 17) pass
    validation; without a GPU the plan then fails on the device query."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    for n in (13, 15, 23, 26):
        assert _plan_error(_one(sz=(n, 22, n))).code == -3


def test_rejects_bad_bflag_and_box():
    assert _plan_error(_one(bflag=64)).code == -1
    assert _plan_error(_one(pmin=(1, 0, 0), pmax=(1, 1, 1))).code == -1
    assert _plan_error(_one(pmin=(float("nan"), 0, 0))).code == -1


def test_rejects_bad_params():
    p = sgr.params()
    p.adv_mode = 7
    assert _plan_error(_one(), p).code == -1
    p = sgr.params()
    p.precision = 3
    assert _plan_error(_one(), p).code == -1
    p = sgr.params()
    p.gauge = 2
    assert _plan_error(_one(), p).code == -1
    p = sgr.params()
    p.eta_mode = -1
    assert _plan_error(_one(), p).code == -1


def test_gauge_defaults_and_fused_kernel_restriction(monkeypatch):
    p = sgr.params()
    assert (p.gauge, p.eta_mode, list(p.xi), p.eta_func_r0, list(p.eta_power)) == (0, 0, [0, 0, 0], 1.31, [2.0, 2.0])
    monkeypatch.setenv("SGR_KERNEL", "fused")
    e = _plan_error(_one(), sgr.params(gauge="rochester"))
    assert e.code == -2 and "standard gauge" in str(e)


def test_null_arguments_are_errors_not_crashes():
    L = sgr.lib()
    assert L.sgr_plan_create(None, None, 0, None, 0) == -1
    h = ctypes.c_void_p()
    assert L.sgr_plan_create(ctypes.byref(h), None, 3, ctypes.byref(sgr.params()), 0) == -1
    assert L.sgr_rhs(None, None, None, None) == -1
    assert L.sgr_plan_get_info(None, None) == -1
    assert "null" in sgr.last_error()


def test_no_gpu_is_reported_as_hip_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    e = _plan_error(_one())
    assert e.code == -3
