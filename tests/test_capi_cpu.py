"""CPU tests of the product library: it loads, exports every symbol the
public headers declare, and rejects bad input with error codes (no GPU is
needed for any of these paths)."""
import ctypes
import os
import re

import numpy as np
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


def test_gauge_defaults():
    p = sgr.params()
    assert (p.gauge, p.eta_mode, list(p.xi), p.eta_func_r0, list(p.eta_power)) == (0, 0, [0, 0, 0], 1.31, [2.0, 2.0])


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


def test_tensor_arguments_are_checked_before_any_kernel_sees_them():
    """Every plan entry point takes its field pointers through one checker:
    wrong shape, dtype, device or layout is a ValueError on the host, never a
    pointer handed to a kernel."""
    import torch
    from sympygr_amd import _field_ptrs, _rows
    with pytest.raises(ValueError, match=r"\(24, total\)"):
        _field_ptrs(torch.zeros((23, 10), dtype=torch.float64), 10, 0)
    with pytest.raises(ValueError, match="float64"):
        _field_ptrs(torch.zeros((24, 10), dtype=torch.float32), 10, 0)
    with pytest.raises(ValueError, match="cuda:0"):
        _field_ptrs(torch.zeros((24, 10), dtype=torch.float64), 10, 0)
    with pytest.raises(ValueError, match=r"\(4, total\)"):
        _rows(torch.zeros((24, 10), dtype=torch.float64), 4, 10, 0)
    with pytest.raises(ValueError, match="torch tensor"):
        _field_ptrs(np.zeros((24, 10)), 10, 0)


def test_device_memory_limit_refuses_a_plan_before_allocating(monkeypatch):
    """SGR_DEVICE_MEM_LIMIT_GIB caps what one plan may take: a plan that
    needs more is refused with SGR_E_NOMEM before any device call (so this
    runs without a GPU), naming its need and the limit."""
    bl = BlockList()
    bl.add((38, 38, 38), (-1.0,) * 3, (1.0,) * 3, 0)
    monkeypatch.setenv("SGR_DEVICE_MEM_LIMIT_GIB", "0.0001")
    e = _plan_error(bl)
    assert e.code == -4 and "SGR_DEVICE_MEM_LIMIT_GIB" in str(e) and "GiB" in str(e)


def test_application_definitions_of_bssn_globals_override_the_library(tmp_path):
    """A program that defines the bssn:: parameter globals (as Dendro's
    parameters.cpp does) and links libsympygr_amd.so: the drop-in reads the
    program's values, not the library's weak defaults."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "dropin_link_check")
    libdir = os.path.dirname(sgr.LIB_PATH)
    r = subprocess.run(["g++", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        "-I", os.path.join(root, "include"), os.path.join(root, "tests", "dropin_link_check.cpp"),
                        "-L", libdir, "-lsympygr_amd", f"-Wl,-rpath,{libdir}", "-L/opt/rocm/lib",
                        "-Wl,-rpath,/opt/rocm/lib", "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    v = out.stdout.split()
    assert [int(x) for x in v[:4]] == [1, 0, 1, 0]
    assert [float(x) for x in v[4:13]] == [0.75, 0.25, 1.25, 30.0, 2.0, 0.37, 1.5, 1.0, 3.0]
    assert [int(x) for x in v[13:]] == [2, 1, 3]


def test_dropin_failure_leaves_nan_and_a_status_instead_of_throwing():
    """bssnrhs cannot report errors (void, reference signature), and may run
    inside an OpenMP loop: a failed call (here: no GPU, or an invalid block)
    writes NaN into the block's interior RHS, leaves its ghosts untouched and
    keeps the first failure for sgr_dropin_status, which clears it."""
    import torch
    L = ctypes.CDLL(sgr.LIB_PATH)
    fn = getattr(L, "_Z7bssnrhsPPdPPKdRKjS2_S2_PS4_S5_")
    PD = ctypes.POINTER(ctypes.c_double)
    fn.argtypes = [ctypes.POINTER(PD), ctypes.POINTER(PD), ctypes.POINTER(ctypes.c_uint), PD, PD,
                   ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_uint)]
    fn.restype = None
    L.sgr_dropin_status.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    L.sgr_dropin_status(None, 0)  # clear
    n = 13
    u = np.ones((24, n ** 3))
    out = np.full_like(u, 3.0)
    pi = (PD * 24)(*[u[f].ctypes.data_as(PD) for f in range(24)])
    po_ = (PD * 24)(*[out[f].ctypes.data_as(PD) for f in range(24)])
    # bflag 64 is invalid everywhere, GPU or not
    fn(po_, pi, ctypes.byref(ctypes.c_uint(0)), (ctypes.c_double * 3)(0, 0, 0), (ctypes.c_double * 3)(1, 1, 1),
       (ctypes.c_uint * 3)(n, n, n), ctypes.byref(ctypes.c_uint(64)))
    v = out.reshape(24, n, n, n)
    assert np.isnan(v[:, 3:-3, 3:-3, 3:-3]).all()
    ghost = np.ones((n, n, n), dtype=bool)
    ghost[3:-3, 3:-3, 3:-3] = False
    assert (v[:, ghost] == 3.0).all()
    buf = ctypes.create_string_buffer(512)
    assert L.sgr_dropin_status(buf, 512) == -1 and b"bflag" in buf.value
    assert L.sgr_dropin_status(buf, 512) == 0 and buf.value == b""
    if not torch.cuda.is_available():
        out[:] = 3.0
        fn(po_, pi, ctypes.byref(ctypes.c_uint(0)), (ctypes.c_double * 3)(0, 0, 0), (ctypes.c_double * 3)(1, 1, 1),
           (ctypes.c_uint * 3)(n, n, n), ctypes.byref(ctypes.c_uint(0)))
        assert np.isnan(v[:, 3:-3, 3:-3, 3:-3]).all()
        assert L.sgr_dropin_status(buf, 512) == -3
