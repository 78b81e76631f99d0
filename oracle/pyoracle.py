"""Test infrastructure: ctypes access to the CPU oracle (liboracle.so) and,
where it has been built, the compiled reference (_ref/libbssnref_{c,u}.so).

Only tests, the smoke check and the benchmark's CPU-baseline leg import this.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
P_D = ctypes.POINTER(ctypes.c_double)


class OracleParams(ctypes.Structure):
    _fields_ = [
        ("lambda_", ctypes.c_uint32 * 4),
        ("lambda_f", ctypes.c_double * 2),
        ("eta_const", ctypes.c_double),
        ("eta_r0", ctypes.c_double),
        ("eta_damping_exp", ctypes.c_double),
        ("ko_sigma", ctypes.c_double),
        ("upwind", ctypes.c_int32),
        ("rochester", ctypes.c_int32),
        ("eta_func", ctypes.c_int32),
        ("xi", ctypes.c_uint32 * 3),
        ("eta_func_r0", ctypes.c_double),
        ("eta_power", ctypes.c_double * 2),
        ("ccz4", ctypes.c_int32),
        ("kappa", ctypes.c_double * 3),
        ("dissipation", ctypes.c_int32),
    ]

DISSIPATION = {"ko": 0, "tvb3": 1, "tvb5": 2}


def make_params(upwind=True, ko_sigma=0.1, eta_const=2.0, eta_r0=50.0, eta_damping_exp=1.0,
                lambda_=(1, 1, 1, 1), lambda_f=(1.0, 0.0), gauge="standard", eta="const", xi=(0, 0, 0),
                eta_func_r0=1.31, eta_power=(2.0, 2.0), ccz4=False, kappa=(0.02, 0.0, 0.5), dissipation="ko"):
    p = OracleParams()
    p.dissipation = DISSIPATION[dissipation]
    p.ccz4 = 1 if ccz4 else 0
    for i in range(3):
        p.kappa[i] = kappa[i]
    p.rochester = 1 if gauge == "rochester" else 0
    p.eta_func = 1 if eta == "func" else 0
    for i in range(3):
        p.xi[i] = xi[i]
    p.eta_func_r0 = eta_func_r0
    p.eta_power[0], p.eta_power[1] = eta_power
    for i in range(4):
        p.lambda_[i] = lambda_[i]
    p.lambda_f[0], p.lambda_f[1] = lambda_f
    p.eta_const, p.eta_r0, p.eta_damping_exp = eta_const, eta_r0, eta_damping_exp
    p.ko_sigma = ko_sigma
    p.upwind = 1 if upwind else 0
    return p


_lib = None


def lib():
    global _lib
    if _lib is None:
        # SGR_ORACLE_LIB: another build of the same sources (the sanitizer test)
        path = os.environ.get("SGR_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle oracle`")
        L = ctypes.CDLL(path)
        L.oracle_bssnrhs.argtypes = [ctypes.POINTER(P_D), ctypes.POINTER(P_D), ctypes.c_uint64, P_D, P_D,
                                     ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32, ctypes.POINTER(OracleParams)]
        L.oracle_bssnrhs_blocks.argtypes = [ctypes.POINTER(P_D), ctypes.POINTER(P_D), ctypes.c_uint32,
                                            ctypes.POINTER(ctypes.c_uint64), P_D, P_D,
                                            ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                            ctypes.POINTER(OracleParams), ctypes.c_int]
        L.oracle_constraints.argtypes = [ctypes.POINTER(P_D), ctypes.POINTER(P_D), ctypes.c_uint64, P_D, P_D,
                                         ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32]
        L.oracle_constraints_psi4.argtypes = L.oracle_constraints.argtypes
        L.oracle_ghost_fill.argtypes = [ctypes.POINTER(P_D), ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64),
                                        ctypes.POINTER(ctypes.c_uint32), P_D, P_D, ctypes.POINTER(ctypes.c_uint32)]
        L.oracle_deriv.argtypes = [ctypes.c_int, P_D, P_D, ctypes.c_double, ctypes.POINTER(ctypes.c_uint32), P_D,
                                   ctypes.c_uint32, ctypes.c_int]
        L.oracle_spacetime_speeds.argtypes = [ctypes.POINTER(P_D), ctypes.POINTER(P_D), ctypes.c_uint64,
                                              ctypes.POINTER(ctypes.c_uint32)]
        _lib = L
    return _lib


def _ptrs(arr):
    return (P_D * arr.shape[0])(*[arr[f].ctypes.data_as(P_D) for f in range(arr.shape[0])])


def rhs_block(u, sz, pmin, pmax, bflag, params, out=None):
    """Oracle RHS of one block. u: (24, npts) float64 C-contiguous."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    if out is None:
        out = np.zeros_like(u)
    rc = lib().oracle_bssnrhs(_ptrs(out), _ptrs(u), 0, (ctypes.c_double * 3)(*pmin), (ctypes.c_double * 3)(*pmax),
                              (ctypes.c_uint32 * 3)(*sz), bflag, ctypes.byref(params))
    if rc != 0:
        raise RuntimeError(f"oracle_bssnrhs failed: {rc}")
    return out


def rhs_blocks(u, bl, params, nthreads=1, out=None):
    """Oracle RHS over a BlockList (u: (24, bl.total))."""
    nb = len(bl)
    if out is None:
        out = np.zeros_like(u)
    offs = (ctypes.c_uint64 * nb)(*bl.offsets)
    pmin = (ctypes.c_double * (3 * nb))(*[v for p in bl.pmin for v in p])
    pmax = (ctypes.c_double * (3 * nb))(*[v for p in bl.pmax for v in p])
    sz = (ctypes.c_uint32 * (3 * nb))(*[v for s in bl.sz for v in s])
    bf = (ctypes.c_uint32 * nb)(*bl.bflag)
    rc = lib().oracle_bssnrhs_blocks(_ptrs(out), _ptrs(u), nb, offs, pmin, pmax, sz, bf, ctypes.byref(params), nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle_bssnrhs_blocks failed: {rc}")
    return out


def constraints_block(u, sz, pmin, pmax, bflag, psi4=False):
    """(4, npts): Hamiltonian and momentum constraints at interior points;
    psi4=True: (6, npts), with Re/Im Psi4 in rows 4 and 5."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros((6 if psi4 else 4, u.shape[1]))
    fn = lib().oracle_constraints_psi4 if psi4 else lib().oracle_constraints
    rc = fn(_ptrs(out), _ptrs(u), 0, (ctypes.c_double * 3)(*pmin), (ctypes.c_double * 3)(*pmax),
            (ctypes.c_uint32 * 3)(*sz), bflag)
    if rc != 0:
        raise RuntimeError(f"oracle_constraints failed: {rc}")
    return out


def deriv(which, u, h, sz, bflag, beta=None, upwind=True):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros_like(u)
    b = None if beta is None else np.ascontiguousarray(beta, dtype=np.float64).ctypes.data_as(P_D)
    rc = lib().oracle_deriv(which, out.ctypes.data_as(P_D), u.ctypes.data_as(P_D), h, (ctypes.c_uint32 * 3)(*sz),
                            b, bflag, 1 if upwind else 0)
    if rc != 0:
        raise RuntimeError("oracle_deriv failed")
    return out


def spacetime_speeds(u, sz):
    """(3, npts): the TVB characteristic speeds at every node of one block."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    lam = np.zeros((3, u.shape[1]))
    if lib().oracle_spacetime_speeds(_ptrs(lam), _ptrs(u), 0, (ctypes.c_uint32 * 3)(*sz)) != 0:
        raise RuntimeError("oracle_spacetime_speeds failed")
    return lam


# ---------------------------------------------------------------- reference
def ref_path(variant):
    return os.path.join(HERE, "_ref", f"libbssnref_{variant}.so")


def ref_available(variant="u"):
    return os.path.exists(ref_path(variant))


_ref = {}


def ref_lib(variant):
    if variant not in _ref:
        L = ctypes.CDLL(ref_path(variant))
        L.ref_bssnrhs.argtypes = [ctypes.POINTER(P_D), ctypes.POINTER(P_D), ctypes.c_uint, P_D, P_D,
                                  ctypes.POINTER(ctypes.c_uint), ctypes.c_uint, ctypes.POINTER(ctypes.c_uint),
                                  P_D, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.ref_deriv.argtypes = [ctypes.c_int, P_D, P_D, ctypes.c_double, ctypes.POINTER(ctypes.c_uint), P_D,
                                ctypes.c_uint]
        L.ref_fake_initial_data.argtypes = [ctypes.c_double] * 3 + [P_D]
        L.ref_bssnrhs_sep.argtypes = L.ref_bssnrhs.argtypes
        L.ref_set_gauge_params.argtypes = [ctypes.c_double, P_D, ctypes.POINTER(ctypes.c_uint)]
        L.ref_set_dissipation.argtypes = [ctypes.c_uint]
        L.ref_tvb.argtypes = [ctypes.c_int, P_D, P_D, P_D, ctypes.c_double, ctypes.POINTER(ctypes.c_uint),
                              ctypes.c_uint]
        L.ref_spacetime_speeds.argtypes = [ctypes.POINTER(P_D), ctypes.POINTER(P_D), ctypes.POINTER(ctypes.c_uint)]
        _ref[variant] = L
    return _ref[variant]


def ref_rhs_block(variant, u, sz, pmin, pmax, bflag, params, staged=False):
    """The compiled reference on one block (bssnrhs, or bssnrhs_sep when
    staged).  The gauge switches are the build variant's; the gauge
    parameters (BSSN_ETA_R0, BSSN_ETA_POWER, BSSN_XI) come from params."""
    L = ref_lib(variant)
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros_like(u)
    lam = (ctypes.c_uint * 4)(*params.lambda_)
    lf = (ctypes.c_double * 2)(*params.lambda_f)
    L.ref_set_gauge_params(params.eta_func_r0, (ctypes.c_double * 2)(*params.eta_power),
                           (ctypes.c_uint * 3)(*params.xi))
    fn = L.ref_bssnrhs_sep if staged else L.ref_bssnrhs
    # DISSIPATION_TYPE is read by bssnrhs_sep only (rhs.cpp:729-757)
    L.ref_set_dissipation(params.dissipation if staged else 0)
    fn(_ptrs(out), _ptrs(u), 0, (ctypes.c_double * 3)(*pmin), (ctypes.c_double * 3)(*pmax),
       (ctypes.c_uint * 3)(*sz), bflag, lam, lf, params.eta_const, params.eta_r0,
       params.eta_damping_exp, params.ko_sigma)
    L.ref_set_dissipation(0)
    return out


def ref_rhs_blocks(variant, u, bl, params, nthreads=1, out=None):
    """The compiled reference over a BlockList on `nthreads` host threads
    (ref_bssnrhs_blocks: an OpenMP loop over blocks, one bssnrhs call each,
    as the reference's own example driver runs it)."""
    L = ref_lib(variant)
    if not hasattr(L.ref_bssnrhs_blocks, "_typed"):
        L.ref_bssnrhs_blocks.argtypes = [ctypes.POINTER(P_D), ctypes.POINTER(P_D), ctypes.c_uint,
                                         ctypes.POINTER(ctypes.c_ulonglong), P_D, P_D, ctypes.POINTER(ctypes.c_uint),
                                         ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_uint), P_D,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int]
        L.ref_bssnrhs_blocks._typed = True
    nb = len(bl)
    if out is None:
        out = np.zeros_like(u)
    L.ref_set_gauge_params(params.eta_func_r0, (ctypes.c_double * 2)(*params.eta_power),
                           (ctypes.c_uint * 3)(*params.xi))
    L.ref_bssnrhs_blocks(_ptrs(out), _ptrs(u), nb, (ctypes.c_ulonglong * nb)(*bl.offsets),
                         (ctypes.c_double * (3 * nb))(*[v for p in bl.pmin for v in p]),
                         (ctypes.c_double * (3 * nb))(*[v for p in bl.pmax for v in p]),
                         (ctypes.c_uint * (3 * nb))(*[v for s_ in bl.sz for v in s_]),
                         (ctypes.c_uint * nb)(*bl.bflag), (ctypes.c_uint * 4)(*params.lambda_),
                         (ctypes.c_double * 2)(*params.lambda_f), params.eta_const, params.eta_r0,
                         params.eta_damping_exp, params.ko_sigma, nthreads)
    return out


def ref_deriv(variant, which, u, h, sz, bflag, beta=None):
    L = ref_lib(variant)
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros_like(u)
    b = None if beta is None else np.ascontiguousarray(beta, dtype=np.float64).ctypes.data_as(P_D)
    L.ref_deriv(which, out.ctypes.data_as(P_D), u.ctypes.data_as(P_D), h, (ctypes.c_uint * 3)(*sz), b, bflag)
    return out


def ref_tvb(variant, which, u, lam, h, sz, bflag):
    """One reference TVB sweep: which 0..2 disstvb3_{x,y,z}, 3..5
    disstvb5_{x,y,z}, with the speeds `lam` along the axis.  Nodes the sweep
    does not write stay NaN."""
    L = ref_lib(variant)
    u = np.ascontiguousarray(u, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    out = np.full_like(u, np.nan)
    L.ref_tvb(which, out.ctypes.data_as(P_D), u.ctypes.data_as(P_D), lam.ctypes.data_as(P_D), h,
              (ctypes.c_uint * 3)(*sz), bflag)
    return out


def ref_spacetime_speeds(variant, u, sz):
    """(3, npts): the reference's max_spacetime_speeds; nodes it does not
    write (ghost nodes) stay NaN."""
    L = ref_lib(variant)
    u = np.ascontiguousarray(u, dtype=np.float64)
    lam = np.full((3, u.shape[1]), np.nan)
    L.ref_spacetime_speeds(_ptrs(lam), _ptrs(u), (ctypes.c_uint * 3)(*sz))
    return lam


def ref_fake_initial_data(variant, sz, pmin, pmax):
    L = ref_lib(variant)
    n = sz[0] * sz[1] * sz[2]
    u = np.zeros((24, n))
    h = [(pmax[d] - pmin[d]) / (sz[d] - 1) for d in range(3)]
    v = (ctypes.c_double * 24)()
    for k in range(sz[2]):
        for j in range(sz[1]):
            for i in range(sz[0]):
                L.ref_fake_initial_data(pmin[0] + i * h[0], pmin[1] + j * h[1], pmin[2] + k * h[2], v)
                u[:, i + sz[0] * (j + sz[1] * k)] = np.frombuffer(v, dtype=np.float64)
    return u


# ----------------------------------------------------------------- RK4 + ghosts
def ghost_fill_uniform(u, bl, nb):
    """Same-level ghost refresh of an nb^3 uniform tiling (blocks in z-y-x
    order, equal sizes): every ghost point inside the domain takes the value of
    the interior point of the block that owns that lattice position.  Written
    independently of the library's neighbour table (global lattice indices)."""
    import itertools
    n = bl.sz[0][0]
    m = n - 6
    out = u.copy()
    idx = np.arange(n)
    for b in range(len(bl)):
        bz, rem = divmod(b, nb * nb)
        by, bx = divmod(rem, nb)
        blk = np.array([bx, by, bz])
        o = bl.offsets[b]
        view = out[:, o:o + n ** 3].reshape(24, n, n, n)  # [f][k][j][i]
        for i, j, k in itertools.product(range(n), repeat=3):
            if 3 <= i < n - 3 and 3 <= j < n - 3 and 3 <= k < n - 3:
                continue
            g = blk * m + np.array([i, j, k]) - 3          # global lattice index
            ob = g // m
            if np.any(ob < 0) or np.any(ob >= nb):
                continue
            li = g - ob * m + 3
            src = ob[0] + nb * (ob[1] + nb * ob[2])
            so = bl.offsets[src]
            view[:, k, j, i] = u[:, so + li[0] + n * (li[1] + n * li[2])]
    return out


def ghost_fill(u, bl):
    """Ghost refresh of any block set (oracle_ghost_fill); returns a copy."""
    nb = len(bl)
    out = np.ascontiguousarray(u, dtype=np.float64).copy()
    offs = (ctypes.c_uint64 * nb)(*bl.offsets)
    pmin = (ctypes.c_double * (3 * nb))(*[v for p in bl.pmin for v in p])
    pmax = (ctypes.c_double * (3 * nb))(*[v for p in bl.pmax for v in p])
    sz = (ctypes.c_uint32 * (3 * nb))(*[v for s in bl.sz for v in s])
    bf = (ctypes.c_uint32 * nb)(*bl.bflag)
    rc = lib().oracle_ghost_fill(_ptrs(out), nb, offs, sz, pmin, pmax, bf)
    if rc != 0:
        raise RuntimeError(f"oracle_ghost_fill failed: {rc}")
    return out


# Classic RK4 tableau (Kutta 1901; e.g. Hairer, Norsett & Wanner, Solving
# ODEs I, table 1.2): stage nodes c = (1/2, 1/2, 1) and weights
# b = (1/6, 1/3, 1/3, 1/6), written as divisors of dt so that the products
# round as the library's stage update (dt/2, dt/6, ...) does.
RK4_NODE_DIV = (2.0, 2.0, 1.0)
RK4_WEIGHT_DIV = (6.0, 3.0, 3.0, 6.0)


def rk4_tableau_step(rhs, u, dt, mask=None, fill=None):
    """One classic RK4 step of du/dt = rhs(u) for an array u of shape
    (nfields, npoints).  Only the columns selected by `mask` (all when None)
    are updated; `fill(v)` (identity when None) refreshes the other columns of
    every stage argument, as a ghost refresh does, and is applied to u first.
    The accumulation order is the library's fused stage update:
    acc = u + dt/6 k1; acc += dt/3 k2; acc += dt/3 k3; u_new = acc + dt/6 k4,
    and the stage arguments are u + (dt/2) k1, u + (dt/2) k2, u + dt k3."""
    mk = slice(None) if mask is None else mask
    fill = fill if fill is not None else (lambda v: v)
    un = fill(np.array(u, dtype=np.float64, copy=True))
    k = rhs(un)
    acc = un.copy()
    acc[:, mk] = un[:, mk] + (dt / RK4_WEIGHT_DIV[0]) * k[:, mk]
    for s in range(1, 4):
        st = un.copy()
        st[:, mk] = un[:, mk] + (dt / RK4_NODE_DIV[s - 1]) * k[:, mk]
        k = rhs(fill(st))
        if s < 3:
            acc[:, mk] = acc[:, mk] + (dt / RK4_WEIGHT_DIV[s]) * k[:, mk]
    new = un.copy()
    new[:, mk] = acc[:, mk] + (dt / RK4_WEIGHT_DIV[3]) * k[:, mk]
    return new


def rk4_step(u, bl, nb, dt, params, nthreads=1):
    """Classic RK4 step (rk4_tableau_step) of the oracle RHS over a block set,
    with a ghost refresh before every stage; only interior points change.
    nb: the side of an nb^3 uniform tiling (lattice-index ghost copy), or
    None for any block set (oracle_ghost_fill)."""
    mk = np.zeros(u.shape[1], dtype=bool)
    for b in range(len(bl)):
        n = bl.sz[b]
        v = np.zeros((n[2], n[1], n[0]), dtype=bool)
        v[3:-3, 3:-3, 3:-3] = True
        mk[bl.offsets[b]:bl.offsets[b] + v.size] = v.reshape(-1)

    def fill(v):
        return ghost_fill_uniform(v, bl, nb) if nb is not None else ghost_fill(v, bl)
    return rk4_tableau_step(lambda v: rhs_blocks(v, bl, params, nthreads), u, dt, mask=mk, fill=fill)
