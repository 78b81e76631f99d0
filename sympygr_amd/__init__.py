"""sympygr_amd: the BSSN right-hand side of paralab/SymPyGR's Dendro path on
AMD Instinct MI355X.

The compute lives in libsympygr_amd.so (hand-written HIP for gfx950 behind the
C ABI of include/sympygr_amd.h).  This module is a thin ctypes binding used by
the tests and the benchmark; torch only supplies device memory and streams.

    import sympygr_amd as sgr
    plan = sgr.Plan(blocks, sgr.params(adv="upwind", precision="fp64"))
    plan.rhs(out, u)          # u, out: torch.float64 cuda tensors (24, total)

``bssnrhs(out, in, offset, pmin, pmax, sz, bflag, params)`` mirrors the
reference entry point (bssn/include/rhs.h:33-36) on host arrays.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libsympygr_amd.so")
NUM_VARS = 24

VAR_NAMES = ["alpha", "chi", "K", "Gt0", "Gt1", "Gt2", "beta0", "beta1", "beta2", "B0", "B1", "B2",
             "gt0", "gt1", "gt2", "gt3", "gt4", "gt5", "At0", "At1", "At2", "At3", "At4", "At5"]

ADV = {"centred": 0, "upwind": 1}
PREC = {"fp64": 0, "mixed": 1}
ERRORS = {0: "SGR_OK", -1: "SGR_E_INVALID", -2: "SGR_E_UNSUPPORTED", -3: "SGR_E_HIP", -4: "SGR_E_NOMEM",
          -5: "SGR_E_NONFINITE"}

P_D = ctypes.POINTER(ctypes.c_double)


class SgrParams(ctypes.Structure):
    _fields_ = [
        ("lambda_", ctypes.c_uint32 * 4),
        ("lambda_f", ctypes.c_double * 2),
        ("eta_const", ctypes.c_double),
        ("eta_r0", ctypes.c_double),
        ("eta_damping_exp", ctypes.c_double),
        ("ko_sigma", ctypes.c_double),
        ("adv_mode", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("gauge", ctypes.c_int32),
        ("eta_mode", ctypes.c_int32),
        ("xi", ctypes.c_uint32 * 3),
        ("eta_func_r0", ctypes.c_double),
        ("eta_power", ctypes.c_double * 2),
        ("formulation", ctypes.c_int32),
        ("kappa", ctypes.c_double * 3),
        ("dissipation", ctypes.c_int32),
    ]


class SgrBlock(ctypes.Structure):
    _fields_ = [
        ("offset", ctypes.c_uint64),
        ("sz", ctypes.c_uint32 * 3),
        ("bflag", ctypes.c_uint32),
        ("pmin", ctypes.c_double * 3),
        ("pmax", ctypes.c_double * 3),
    ]


class SgrPlanInfo(ctypes.Structure):
    _fields_ = [
        ("blocks", ctypes.c_uint64),
        ("tiles", ctypes.c_uint64),
        ("interior_points", ctypes.c_uint64),
        ("workspace_bytes", ctypes.c_uint64),
        ("derivative_slots", ctypes.c_uint32),
        ("chunks", ctypes.c_uint32),
        ("pipelined", ctypes.c_uint32),
        ("tile_queue", ctypes.c_uint32),
    ]


class SgrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None

# exported functions: name -> (restype, argtypes)
_SIGS = {
    "sgr_params_default": (None, [ctypes.POINTER(SgrParams), ctypes.c_int32, ctypes.c_int32]),
    "sgr_plan_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(SgrBlock), ctypes.c_uint32,
                                       ctypes.POINTER(SgrParams), ctypes.c_int]),
    "sgr_plan_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sgr_plan_get_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SgrPlanInfo)]),
    "sgr_rhs": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(P_D), ctypes.POINTER(P_D), ctypes.c_void_p]),
    "sgr_rhs_block": (ctypes.c_int, [ctypes.POINTER(P_D), ctypes.POINTER(P_D), ctypes.c_uint64, P_D, P_D,
                                     ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32, ctypes.POINTER(SgrParams),
                                     ctypes.c_int]),
    "sgr_rhs_block_noalloc": (ctypes.c_int, [ctypes.POINTER(P_D), ctypes.POINTER(P_D), ctypes.c_uint64, P_D, P_D,
                                             ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32,
                                             ctypes.POINTER(SgrParams), ctypes.c_void_p]),
    "sgr_bssnrhs_host": (ctypes.c_int, [ctypes.POINTER(P_D), ctypes.POINTER(P_D), ctypes.c_uint64, P_D, P_D,
                                        ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32, ctypes.POINTER(SgrParams),
                                        ctypes.c_int]),
    "sgr_rhs_events": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(P_D), ctypes.POINTER(P_D), ctypes.c_void_p,
                                      ctypes.POINTER(ctypes.c_void_p)]),
    "sgr_constraints": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(P_D), ctypes.POINTER(P_D), ctypes.c_void_p]),
    "sgr_constraints_psi4": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(P_D), ctypes.POINTER(P_D),
                                            ctypes.c_void_p]),
    "sgr_ghost_fill": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(P_D), ctypes.c_void_p]),
    "sgr_hbm_copy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]),
    "sgr_ghost_table_stats": (ctypes.c_int, [ctypes.POINTER(SgrBlock), ctypes.c_uint32,
                                             ctypes.POINTER(ctypes.c_uint64)]),
    "sgr_rk4_stage": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.POINTER(P_D),
                                     ctypes.POINTER(P_D), ctypes.POINTER(P_D), ctypes.POINTER(P_D), ctypes.c_void_p]),
    "sgr_rk4_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(P_D), ctypes.POINTER(P_D),
                                    ctypes.POINTER(P_D), ctypes.POINTER(P_D), ctypes.c_void_p]),
    "sgr_rk4_graph_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(P_D),
                                            ctypes.POINTER(P_D), ctypes.POINTER(P_D), ctypes.POINTER(P_D),
                                            ctypes.POINTER(ctypes.c_void_p)]),
    "sgr_graph_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "sgr_graph_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sgr_check_finite": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(P_D), ctypes.c_void_p]),
    "sgr_last_error": (ctypes.c_size_t, [ctypes.c_char_p, ctypes.c_size_t]),
    "sgr_version": (ctypes.c_char_p, []),
}


def lib():
    """Load libsympygr_amd.so (fails loudly when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m sympygr_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().sgr_last_error(buf, 1024)
    return buf.value.decode()


def check(rc):
    if rc != 0:
        raise SgrError(rc, last_error())
    return rc


GAUGE = {"standard": 0, "rochester": 1}
ETA = {"const": 0, "func": 1}
FORM = {"bssn": 0, "ccz4": 1}
DISS = {"ko": 0, "tvb3": 1, "tvb5": 2}
NUM_VARS_CCZ4 = 25
VAR_NAMES_CCZ4 = VAR_NAMES + ["Theta"]


def params(adv="upwind", precision="fp64", ko_sigma=0.1, eta_const=2.0, eta_r0=50.0, eta_damping_exp=1.0,
           lambda_=(1, 1, 1, 1), lambda_f=(1.0, 0.0), gauge="standard", eta="const", xi=(0, 0, 0),
           eta_func_r0=1.31, eta_power=(2.0, 2.0), formulation="bssn", kappa=(0.02, 0.0, 0.5),
           dissipation="ko") -> SgrParams:
    """sgr_params: the reference's parameter defaults with the given
    choices.  formulation "ccz4" evolves the CCZ4 system (25 fields, Theta
    last) with damping kappa = (kappa1, kappa2, kappa3).  dissipation "ko"
    (Kreiss-Oliger), "tvb3" or "tvb5" (DISSIPATION_TYPE 0, 1, 2)."""
    p = SgrParams()
    lib().sgr_params_default(ctypes.byref(p), ADV[adv], PREC[precision])
    p.formulation = FORM[formulation]
    p.dissipation = DISS[dissipation]
    for i in range(3):
        p.kappa[i] = kappa[i]
    for i in range(4):
        p.lambda_[i] = lambda_[i]
    p.lambda_f[0], p.lambda_f[1] = lambda_f
    p.eta_const, p.eta_r0, p.eta_damping_exp, p.ko_sigma = eta_const, eta_r0, eta_damping_exp, ko_sigma
    p.gauge, p.eta_mode = GAUGE[gauge], ETA[eta]
    for i in range(3):
        p.xi[i] = xi[i]
    p.eta_func_r0 = eta_func_r0
    p.eta_power[0], p.eta_power[1] = eta_power
    return p


def block_array(bl):
    """ctypes array of SgrBlock from a sympygr_amd.data.BlockList."""
    arr = (SgrBlock * max(len(bl), 1))()
    for b in range(len(bl)):
        arr[b].offset = bl.offsets[b]
        for d in range(3):
            arr[b].sz[d] = bl.sz[b][d]
            arr[b].pmin[d] = bl.pmin[b][d]
            arr[b].pmax[d] = bl.pmax[b][d]
        arr[b].bflag = bl.bflag[b]
    return arr


def hbm_copy(dst, src, stream=None):
    """Copy device tensor src into dst (same byte size) with the library's
    plain copy kernel."""
    import torch
    nbytes = src.numel() * src.element_size()
    if dst.numel() * dst.element_size() != nbytes:
        raise ValueError("size mismatch")
    s = stream if stream is not None else torch.cuda.current_stream(src.device).cuda_stream
    check(lib().sgr_hbm_copy(ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src.data_ptr()), nbytes,
                             ctypes.c_void_p(s)))


def ghost_table_stats(bl):
    """Host-only check of the ghost refresh table for a BlockList:
    {"patches", "nodes", "interpolated"}; raises SgrError (UNSUPPORTED) when
    the refresh would be refused."""
    st = (ctypes.c_uint64 * 3)()
    check(lib().sgr_ghost_table_stats(block_array(bl), len(bl), st))
    return {"patches": st[0], "nodes": st[1], "interpolated": st[2]}


def _rows(t, rows, need, device):
    """Device pointers of the `rows` rows of a (rows, >= need) tensor after
    checking everything the kernels assume of it: float64, contiguous, on
    cuda:`device`, each row at least `need` values long (the kernels read and
    write 8-byte values over the whole block table)."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError("expected a torch tensor")
    if t.dim() != 2 or t.shape[0] != rows:
        raise ValueError(f"expected a ({rows}, total) tensor, got shape {tuple(t.shape)}")
    if t.dtype != torch.float64:
        raise ValueError(f"expected float64, got {t.dtype}")
    if not t.is_cuda or t.device.index != device:
        raise ValueError(f"tensor must live on cuda:{device}, not {t.device}")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    if t.shape[1] < need:
        raise ValueError(f"rows hold {t.shape[1]} values, the block table needs {need}")
    base = t.data_ptr()
    stride = t.stride(0) * t.element_size()
    return (P_D * rows)(*[ctypes.cast(base + f * stride, P_D) for f in range(rows)])


def _field_ptrs(t, need, device):
    """24 device pointers of the rows of a checked (24, >= need) tensor."""
    return _rows(t, NUM_VARS, need, device)


class Plan:
    """A validated block table on one device plus its workspace."""

    def __init__(self, blocks, prm: SgrParams, device: int = 0):
        self._h = ctypes.c_void_p()
        self.blocks = blocks
        self.params = prm
        self._barr = block_array(blocks)
        check(lib().sgr_plan_create(ctypes.byref(self._h), self._barr, len(blocks), ctypes.byref(prm), device))
        self.device = device
        self.nfields = NUM_VARS_CCZ4 if prm.formulation == FORM["ccz4"] else NUM_VARS

    def info(self) -> SgrPlanInfo:
        i = SgrPlanInfo()
        check(lib().sgr_plan_get_info(self._h, ctypes.byref(i)))
        return i

    def _f(self, t):
        """Checked field pointers of a (nfields, total) tensor for this plan
        (24 fields, or 25 for CCZ4)."""
        return _rows(t, self.nfields, self.blocks.total, self.device)

    def rhs(self, out, u, stream=None):
        """Asynchronous RHS on `stream` (default: torch's current stream)."""
        po, pi = self._f(out), self._f(u)
        check(lib().sgr_rhs(self._h, po, pi, ctypes.c_void_p(self._stream(u, stream))))

    def rhs_events(self, out, u, events, stream=None):
        """sgr_rhs plus hipEvent records around each of the two kernels;
        `events` is a sequence of 4 raw hipEvent_t handles (ints)."""
        po, pi = self._f(out), self._f(u)
        ev = (ctypes.c_void_p * 4)(*events)
        check(lib().sgr_rhs_events(self._h, po, pi, ctypes.c_void_p(self._stream(u, stream)), ev))

    @staticmethod
    def _stream(t, stream):
        import torch
        return stream if stream is not None else torch.cuda.current_stream(t.device).cuda_stream

    def constraints(self, u, out4, stream=None):
        """Hamiltonian (out4[0]) and momentum (out4[1:4]) constraints at the
        interior points; out4 is a (4, total) float64 tensor."""
        po = _rows(out4, 4, self.blocks.total, self.device)
        check(lib().sgr_constraints(self._h, self._f(u), po, ctypes.c_void_p(self._stream(u, stream))))

    def constraints_psi4(self, u, out6, stream=None):
        """Constraints (out6[0:4], as constraints()) and the real (out6[4])
        and imaginary (out6[5]) parts of Psi4 at the interior points, at the
        coordinates pmin + i h of each block; out6 is a (6, total) float64
        tensor."""
        po = _rows(out6, 6, self.blocks.total, self.device)
        check(lib().sgr_constraints_psi4(self._h, self._f(u), po, ctypes.c_void_p(self._stream(u, stream))))

    def ghost_fill(self, u, stream=None):
        """Ghost refresh of u in place (copies where grids align, 6-point
        Lagrange interpolation from the owning block elsewhere)."""
        check(lib().sgr_ghost_fill(self._h, self._f(u), ctypes.c_void_p(self._stream(u, stream))))

    def rk4_stage(self, stage, dt, u_n, u_stage, acc, u_next, stream=None):
        check(lib().sgr_rk4_stage(self._h, stage, dt, self._f(u_n), self._f(u_stage), self._f(acc),
                                  self._f(u_next), ctypes.c_void_p(self._stream(u_n, stream))))

    def rk4_step(self, dt, u, s_a, s_b, acc, stream=None):
        """One RK4 step of u in place (ghost refresh + 4 fused stages)."""
        check(lib().sgr_rk4_step(self._h, dt, self._f(u), self._f(s_a), self._f(s_b), self._f(acc),
                                 ctypes.c_void_p(self._stream(u, stream))))

    def rk4_graph(self, dt, u, s_a, s_b, acc):
        return Graph(self, dt, u, s_a, s_b, acc)

    def check_finite(self, u, stream=None):
        """Return code of sgr_check_finite: 0, or SGR_E_NONFINITE (-5) when a
        padded input value of some block is NaN or infinite."""
        return lib().sgr_check_finite(self._h, self._f(u), ctypes.c_void_p(self._stream(u, stream)))

    def close(self):
        if self._h:
            lib().sgr_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bssnrhs(out, u, offset, pmin, pmax, sz, bflag, prm: SgrParams, device=0):
    """Mirror of the reference bssnrhs (bssn/include/rhs.h:33-36) on host numpy
    arrays u, out of shape (24, N): reads the padded block at `offset`, writes
    the interior RHS into `out`, leaves ghost entries untouched."""
    import numpy as np
    for a in (u, out):
        if a.dtype != np.float64 or not a.flags.c_contiguous or a.shape[0] != NUM_VARS:
            raise ValueError("expected C-contiguous float64 arrays of shape (24, N)")
    pi = (P_D * NUM_VARS)(*[u[f].ctypes.data_as(P_D) for f in range(NUM_VARS)])
    po = (P_D * NUM_VARS)(*[out[f].ctypes.data_as(P_D) for f in range(NUM_VARS)])
    check(lib().sgr_bssnrhs_host(po, pi, offset, (ctypes.c_double * 3)(*pmin), (ctypes.c_double * 3)(*pmax),
                                 (ctypes.c_uint32 * 3)(*sz), bflag, ctypes.byref(prm), device))
    return out


def rhs_block(out, u, offset, pmin, pmax, sz, bflag, prm: SgrParams, device=0):
    """sgr_rhs_block: the reference bssnrhs (bssn/include/rhs.h:33-36) on
    device tensors u, out of shape (24, N), synchronous; reads the padded
    block at `offset`, writes its interior RHS into `out`."""
    need = offset + sz[0] * sz[1] * sz[2]
    check(lib().sgr_rhs_block(_field_ptrs(out, need, device), _field_ptrs(u, need, device), offset,
                              (ctypes.c_double * 3)(*pmin),
                              (ctypes.c_double * 3)(*pmax), (ctypes.c_uint32 * 3)(*sz), bflag,
                              ctypes.byref(prm), device))
    return out


def rhs_block_noalloc(out, u, offset, pmin, pmax, sz, bflag, prm: SgrParams, stream=None):
    """sgr_rhs_block_noalloc: one block on the current device, asynchronous on
    `stream` (default: torch's current stream), no device allocation (its
    scratch is the library's fixed per-device array; overlapping calls on
    one GPU race)."""
    import torch
    nf = NUM_VARS_CCZ4 if prm.formulation == FORM["ccz4"] else NUM_VARS
    need = offset + sz[0] * sz[1] * sz[2]
    dev = torch.cuda.current_device()
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    check(lib().sgr_rhs_block_noalloc(_rows(out, nf, need, dev), _rows(u, nf, need, dev), offset,
                                      (ctypes.c_double * 3)(*pmin), (ctypes.c_double * 3)(*pmax),
                                      (ctypes.c_uint32 * 3)(*sz), bflag, ctypes.byref(prm), ctypes.c_void_p(s)))
    return out


class Graph:
    """A captured RK4 step (HIP graph) of one plan and fixed buffers."""

    def __init__(self, plan, dt, u, s_a, s_b, acc):
        self._g = ctypes.c_void_p()
        self._keep = (plan, u, s_a, s_b, acc)
        check(lib().sgr_rk4_graph_create(plan._h, dt, plan._f(u), plan._f(s_a), plan._f(s_b), plan._f(acc),
                                         ctypes.byref(self._g)))

    def launch(self, stream=None):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(lib().sgr_graph_launch(self._g, ctypes.c_void_p(s)))

    def close(self):
        """Destroy the graph and release the plan and buffers it kept alive."""
        if self._g:
            lib().sgr_graph_destroy(self._g)
            self._g = ctypes.c_void_p()
        self._keep = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
