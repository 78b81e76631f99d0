#!/usr/bin/env python3
"""BSSN RHS throughput on one MI355X (Mpoints/s = interior points per second
for one RHS evaluation over the whole block set).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-mixed]

Workload: 8x8x8 blocks of 32^3 interior points + 3 ghost layers (38^3 padded)
tiling [-8,8]^3 with Brill-Lindquist two-puncture data (masses 0.5 at
(+-2,0,0), ghosts filled analytically), outer-face blocks carrying bflag,
6th-order stencils, shift-upwinded advection, KO sigma 0.1.  A step is one
RHS evaluation of all 512 blocks with the inputs resident in HBM.

`value` is the plain FP64 rate.  The same line carries the mixed-precision
rate (FP32 pointwise algebra) with its measured error against FP64, the
roofline of the RHS pass from HIP events recorded around each kernel inside
the timed loop, and the CPU oracle timed on a bounded sample of the same
blocks on this host.  Beside it (skipped with --no-octree): the batched
uneven octree (3032 leaves of interior 16/24/32/48 over [-16,16]^3, one RHS
in FP64 and mixed with the mixed error, the captured RK4 step with its ghost
refresh across leaves, the FP64-vs-mixed constraint drift over --drift-steps
RK4 steps) and one 192^3-interior block.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BSSN RHS Mpoints/s: plain FP64 vs batched mixed-prec at stated err; % roofline"

# Per interior point of an n^3 block (compulsory, fused single pass): read the
# 24 FP64 fields over the padded block once, write the 24 FP64 RHS values.
def alg_bytes_per_point(n: int) -> float:
    return 192.0 * ((n + 6) / n) ** 3 + 192.0


# Algorithmic flops per interior point, upwind mode: stencils
# 2*(45*6 + 72*5 + 33*7 + 33*6 + 72*7) = 3126, KO add 24*5 = 120, pointwise
# algebra 3917 (operation count of the reference's generated expression).
# In mixed precision the algebra's 3917 run in FP32, the rest in FP64.
FLOPS_STENCIL = 3126
FLOPS_KO_ADD = 120
FLOPS_ALGEBRA = 3917
FLOPS_UPWIND = FLOPS_STENCIL + FLOPS_KO_ADD + FLOPS_ALGEBRA  # 7163
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
FP32_PEAK_TFS = 157.3


def bound_ps_per_point(n: int, precision: str):
    """Roofline time per interior point of an n^3 block (picoseconds) for the
    HBM roof (algorithmic bytes) and the vector-ALU roof (FP64, plus FP32 for
    the mixed algebra), and which one binds."""
    t_hbm = alg_bytes_per_point(n) / HBM_PEAK_GBS * 1e3
    f64 = FLOPS_UPWIND if precision == "fp64" else FLOPS_STENCIL + FLOPS_KO_ADD
    f32 = 0 if precision == "fp64" else FLOPS_ALGEBRA
    t_alu = f64 / FP64_PEAK_TFS + f32 / FP32_PEAK_TFS
    return {"hbm": t_hbm, "valu": t_alu, "binding": "hbm" if t_hbm >= t_alu else "valu"}


def mix_roofline(size_points, seconds, precision):
    """Roofline fraction of one RHS over a mix of block sizes
    {n: interior points}: the bound is the sum over sizes of points times the
    binding roof's time per point (the harmonic bound over the mix);
    frac = bound time / measured time."""
    t_bound = t_hbm = t_alu = 0.0
    for n, pts in size_points.items():
        b = bound_ps_per_point(n, precision)
        t_bound += pts * max(b["hbm"], b["valu"]) * 1e-12
        t_hbm += pts * b["hbm"] * 1e-12
        t_alu += pts * b["valu"] * 1e-12
    return {"bound_s": t_bound, "frac": t_bound / seconds, "hbm_frac": t_hbm / seconds, "valu_frac": t_alu / seconds,
            "binding": "hbm" if t_hbm >= t_alu else "valu_fp64" + ("+fp32" if precision == "mixed" else ""),
            "bound_mpts": sum(size_points.values()) / t_bound / 1e6}


def measured_traffic(points):
    """HBM bytes of one RHS pass on this workload as last measured with
    rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (per dispatch, summed over the
    pass's kernels, each counter scaled by its reading for a 1 GiB k_hbm_copy);
    None when no measurement for this workload is present."""
    path = os.path.join(ROOT, "profiles", "pmc_bssn_rhs_512x32cubed_fp64.json")
    try:
        with open(path) as f:
            J = json.load(f)
    except (OSError, ValueError):
        return None, None
    if J.get("interior_points") != points:
        return None, None
    per = {k.split("<")[0]: round(v["read_bytes"] + v["write_bytes"]) for k, v in J["kernels"].items()}
    return J["rhs_pass_bytes"], {"per_kernel_bytes": per, "measured_at": J.get("git"),
                                 "calibration": J["calibration"]}


def measured_valu(points):
    """FP64 and FP32 vector-ALU work of one RHS pass as counted by the
    hardware (rocprofv3 --pmc SQ_INSTS_VALU_FLOPS_FP64 ..._FP32 and their
    transcendental parts, per dispatch, over the pass's kernels) with the
    kernels' durations from the same run: the utilisation of the FP64 and
    FP32 vector peaks.  None when no measurement for this workload exists."""
    path = os.path.join(ROOT, "profiles", "pmc_valu_bssn_rhs_512x32cubed.json")
    try:
        with open(path) as f:
            J = json.load(f)
    except (OSError, ValueError):
        return None
    if J.get("interior_points") != points:
        return None
    out = {"measured_at": J.get("git"), "source": os.path.relpath(path, ROOT)}
    for prec, K in J["precisions"].items():
        f64 = sum(k["fp64_flops"] for k in K.values())
        f32 = sum(k["fp32_flops"] for k in K.values())
        t = sum(k["duration_ms"] for k in K.values()) * 1e-3
        out[prec] = {"fp64_TFs": f64 / t / 1e12, "fp64_frac": f64 / t / 1e12 / FP64_PEAK_TFS,
                     "fp32_TFs": f32 / t / 1e12, "fp32_frac": f32 / t / 1e12 / FP32_PEAK_TFS,
                     "fp64_flops_per_point": f64 / points, "fp32_flops_per_point": f32 / points,
                     "per_kernel": {n: {"fp64_frac": k["fp64_flops"] / (k["duration_ms"] * 1e-3) / 1e12 / FP64_PEAK_TFS,
                                        "fp32_frac": k["fp32_flops"] / (k["duration_ms"] * 1e-3) / 1e12 / FP32_PEAK_TFS}
                                    for n, k in K.items()}}
    return out


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


class HipEvents:
    """hipEvent_t handles through libamdhip64 (events are recorded by the
    library around each kernel launch on the launch stream)."""

    def __init__(self):
        import torch  # noqa: F401  (loads the HIP runtime)
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                self.h = ctypes.CDLL(name)
                break
            except OSError:
                continue
        self.h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.h.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def create(self, n):
        evs = []
        for _ in range(n):
            e = ctypes.c_void_p()
            if self.h.hipEventCreate(ctypes.byref(e)) != 0:
                raise RuntimeError("hipEventCreate failed")
            evs.append(e.value)
        return evs

    def elapsed(self, a, b):
        ms = ctypes.c_float()
        if self.h.hipEventElapsedTime(ctypes.byref(ms), ctypes.c_void_p(a), ctypes.c_void_p(b)) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return ms.value

    def destroy(self, evs):
        for e in evs:
            self.h.hipEventDestroy(ctypes.c_void_p(e))


def timed_steps(plan, out, u, steps, warmup, hev):
    import torch
    for _ in range(warmup):
        plan.rhs(out, u)
    torch.cuda.synchronize()
    evs = [hev.create(4) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        plan.rhs_events(out, u, evs[s])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kd = sum(hev.elapsed(e[0], e[1]) for e in evs) / steps
    kp = sum(hev.elapsed(e[2], e[3]) for e in evs) / steps
    for e in evs:
        hev.destroy(e)
    return wall, kd, kp


def normwise_err(a, b):
    """per-field max|a-b| / max|b|"""
    import torch
    d = (a - b).abs().amax(dim=1)
    r = b.abs().amax(dim=1).clamp_min(1e-300)
    return (d / r)


def block_volume_weights(bl, device):
    """h^3 at every node of every block (flattened like the field arrays), for
    volume-weighted norms over blocks of different spacings."""
    import torch
    w = torch.empty(bl.total, dtype=torch.float64, device=device)
    for b in range(len(bl)):
        s = bl.sz[b]
        h = [(bl.pmax[b][d] - bl.pmin[b][d]) / (s[d] - 1) for d in range(3)]
        w[bl.offsets[b]:bl.offsets[b] + s[0] * s[1] * s[2]] = h[0] * h[1] * h[2]
    return w


def constraint_drift(bl, u0, steps, dt, every=10):
    """Evolve `u0` (not modified) for `steps` RK4 steps (captured RK4-step
    graph: 4 ghost refreshes + 4 fused stages) in FP64 and in mixed precision;
    volume-weighted L2 norms of the Hamiltonian and momentum constraints over
    the interior nodes every `every` steps, always evaluated by an FP64 plan."""
    import torch
    import sympygr_amd as sgr
    res = {"steps": steps, "dt": dt, "cfl": 0.25}
    w = block_volume_weights(bl, u0.device)
    vol = 0.0
    for b in range(len(bl)):
        s = bl.sz[b]
        h = [(bl.pmax[b][d] - bl.pmin[b][d]) / (s[d] - 1) for d in range(3)]
        vol += h[0] * h[1] * h[2] * (s[0] - 6) * (s[1] - 6) * (s[2] - 6)
    for prec in ("fp64", "mixed"):
        u = u0.clone()
        sa, sb, acc = u0.clone(), u0.clone(), u0.clone()
        c = torch.zeros((4, bl.total), dtype=torch.float64, device=u0.device)
        plan = sgr.Plan(bl, sgr.params(adv="upwind", precision=prec, ko_sigma=0.1))
        old = os.environ.get("SGR_WS_GIB")
        os.environ["SGR_WS_GIB"] = "8"  # the constraint plan runs in chunks beside the RK4 plan
        try:
            cplan = sgr.Plan(bl, sgr.params(adv="upwind", precision="fp64"))
        finally:
            if old is None:
                del os.environ["SGR_WS_GIB"]
            else:
                os.environ["SGR_WS_GIB"] = old
        graph = plan.rk4_graph(dt, u, sa, sb, acc)
        hist = []

        def norms():
            cplan.ghost_fill(u)
            cplan.constraints(u, c)
            l2 = ((c * c * w).sum(dim=1) / vol).sqrt()
            return {"ham": l2[0].item(), "mom": (l2[1:] ** 2).sum().sqrt().item()}
        hist.append({"step": 0, **norms()})
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s in range(1, steps + 1):
            graph.launch()
            if s % every == 0 or s == steps:
                torch.cuda.synchronize()
                hist.append({"step": s, **norms()})
                log(f"drift {prec}: step {s} H {hist[-1]['ham']:.4e} M {hist[-1]['mom']:.4e}")
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        finite = bool(torch.isfinite(u).all().item())
        res[prec] = {"history": hist, "finite": finite, "wall_s": wall}
        graph.close()
        del graph, plan, cplan, u, sa, sb, acc, c
        torch.cuda.empty_cache()
    hf, hm = res["fp64"]["history"][-1], res["mixed"]["history"][-1]
    res["ham_l2_rel_diff_mixed_vs_fp64"] = abs(hm["ham"] - hf["ham"]) / max(hf["ham"], 1e-300)
    res["mom_l2_rel_diff_mixed_vs_fp64"] = abs(hm["mom"] - hf["mom"]) / max(hf["mom"], 1e-300)
    return res


def octree_section(args, hev):
    """Batched uneven blocks: one RHS over all 3032 octree leaves (FP64 and
    mixed, with the mixed error), the captured RK4 step on them (ghost
    refresh across leaves of other sizes and levels), and the constraint
    drift of FP64 against mixed."""
    import torch
    import sympygr_amd as sgr
    from sympygr_amd import data
    sizes = tuple(args.octree_sizes)
    obl = data.octree_blocks(sizes=sizes)
    t0 = time.perf_counter()
    ou = data.fill_blocks(obl, backend="torch", device="cuda")
    torch.cuda.synchronize()
    opts = obl.interior_points()
    size_pts = {}
    for b in range(len(obl)):
        m = obl.sz[b][0] - 6
        size_pts[m] = size_pts.get(m, 0) + m ** 3
    log(f"octree: {len(obl)} uneven blocks, {opts} interior points, filled in {time.perf_counter() - t0:.1f} s")
    res = {"workload": f"{len(obl)} octree leaves over [-16,16]^3 refined toward the punctures, interior "
                       f"{'/'.join(str(x) for x in sizes)} (+3 ghost), Brill-Lindquist, upwind, KO sigma 0.1",
           "blocks": len(obl), "interior_points": opts, "padded_points": obl.total}
    nst = max(2, args.steps // 4)
    outs = {}
    for prec in ("fp64", "mixed"):
        op = sgr.Plan(obl, sgr.params(adv="upwind", precision=prec, ko_sigma=0.1))
        outs[prec] = torch.zeros_like(ou)
        w, _, _ = timed_steps(op, outs[prec], ou, nst, 1, hev)
        info = op.info()
        res[prec] = {"value": opts * nst / w / 1e6, "chunks": info.chunks, "pipelined": bool(info.pipelined),
                     "workspace_gib": info.workspace_bytes / 2 ** 30, "ms_per_step": 1e3 * w / nst,
                     "roofline": mix_roofline(size_pts, w / nst, prec)}
        log(f"octree {prec}: {res[prec]['value']:.1f} Mpts/s in {info.chunks} workspace chunks"
            f"{' (pipelined)' if info.pipelined else ''}")
        del op
    err = normwise_err(outs["mixed"], outs["fp64"])
    ref = outs["fp64"]
    rms = ((outs["mixed"] - ref).pow(2).mean(dim=1).sqrt() / ref.pow(2).mean(dim=1).sqrt().clamp_min(1e-300))
    res["mixed"]["max_normwise_rel_err_vs_fp64"] = err.max().item()
    res["mixed"]["rms_over_fields_of_normwise_rel_err"] = err.pow(2).mean().sqrt().item()
    res["mixed"]["per_field_normwise_rel_err"] = [float(f"{x:.3e}") for x in err.tolist()]
    res["mixed"]["per_field_rms_rel_err"] = [float(f"{x:.3e}") for x in rms.tolist()]
    res["mixed_speedup_vs_fp64"] = res["mixed"]["value"] / res["fp64"]["value"]
    log(f"octree mixed error: max {err.max().item():.2e}")
    del outs, ref, err, rms
    torch.cuda.empty_cache()

    hmin = min((obl.pmax[b][0] - obl.pmin[b][0]) / (obl.sz[b][0] - 1) for b in range(len(obl)))
    dt = 0.25 * hmin
    st = sgr.ghost_table_stats(obl)
    res["ghost_table"] = st
    res["rk4_step"] = {"dt": dt, "what": "captured graph: 4 ghost refreshes + 4 RHS with the RK4 update fused "
                                         "into the output store"}
    for prec in ("fp64", "mixed"):
        plan = sgr.Plan(obl, sgr.params(adv="upwind", precision=prec, ko_sigma=0.1))
        x = ou.clone()
        bufs = [ou.clone() for _ in range(3)]
        g = plan.rk4_graph(dt, x, *bufs)
        g.launch()
        torch.cuda.synchronize()
        k = 3
        t0 = time.perf_counter()
        for _ in range(k):
            g.launch()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / k
        t0 = time.perf_counter()
        for _ in range(k):
            plan.ghost_fill(x)
        torch.cuda.synchronize()
        gf = (time.perf_counter() - t0) / k
        res["rk4_step"][prec] = {"ms_per_step": 1e3 * wall, "rhs_mpts_equiv": 4 * opts / wall / 1e6,
                                 "ghost_fill_ms": 1e3 * gf}
        log(f"octree rk4 {prec}: {1e3 * wall:.1f} ms per step, ghost refresh {1e3 * gf:.1f} ms")
        g.close()
        del g, plan, x, bufs
        torch.cuda.empty_cache()

    drift = None
    if not args.no_drift:
        drift = constraint_drift(obl, ou, args.drift_steps, dt)
    del ou
    torch.cuda.empty_cache()
    return res, drift


def ccz4_section(hev, steps=10):
    """CCZ4 (25 fields, Theta last; sympygr_amd/codegen/ccz4_eqns.py) on one
    block of 192^3 interior points (198^3 padded): the perturbed flat
    background at amplitude 1e-2 plus a Gaussian lapse pulse (A = 0.1,
    width 1), kappa1 = 0.02, kappa2 = 0, kappa3 = 0.5; RHS in FP64 and mixed
    with the mixed error, and the constraint kernel.  Algorithmic bytes and
    flops are the BSSN figures times 25/24."""
    import torch
    import sympygr_amd as sgr
    from sympygr_amd import data
    nL = 198
    bl = data.BlockList()
    bl.add((nL, nL, nL), (-4.0,) * 3, (4.0,) * 3, 63)
    u = data.fill_blocks(bl, backend="torch", device="cuda", kind="ccz4", seed=7, amp=1e-2, lo=-4.0, L=8.0)
    pts = bl.interior_points()
    res = {"workload": "one block of 192^3 interior (198^3 padded), CCZ4, perturbed flat data (1e-2) plus a lapse "
                       "pulse 0.1 exp(-r^2), all faces flagged, upwind, KO sigma 0.1, kappa = (0.02, 0, 0.5)",
           "interior_points": pts}
    outs = {}
    for prec in ("fp64", "mixed"):
        plan = sgr.Plan(bl, sgr.params(adv="upwind", precision=prec, ko_sigma=0.1, formulation="ccz4",
                                       kappa=(0.02, 0.0, 0.5)))
        outs[prec] = torch.zeros_like(u)
        w, kd, kp = timed_steps(plan, outs[prec], u, steps, 2, hev)
        t = (kd + kp) * 1e-3
        rf = mix_roofline({nL - 6: pts}, t, prec)
        rf = {k: (v * 25 / 24 if k in ("frac", "hbm_frac", "valu_frac") else v) for k, v in rf.items()}
        rf["bound_mpts"] = rf["bound_mpts"] * 24 / 25
        res[prec] = {"value": pts * steps / w / 1e6, "ms_per_step": 1e3 * w / steps,
                     "kernel_ms": {"k_stencil": kd, "k_pointwise": kp}, "roofline": rf}
        log(f"ccz4 {prec}: {res[prec]['value']:.1f} Mpts/s (k_stencil {kd:.2f} ms, pointwise {kp:.2f} ms)")
        if prec == "fp64":
            c = torch.zeros((4, bl.total), dtype=torch.float64, device="cuda")
            plan.constraints(u, c)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(3):
                plan.constraints(u, c)
            torch.cuda.synchronize()
            tc = (time.perf_counter() - t0) / 3
            res["constraints"] = {"ms": 1e3 * tc, "mpts": pts / tc / 1e6}
            c = torch.zeros((6, bl.total), dtype=torch.float64, device="cuda")
            plan.constraints_psi4(u, c)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(3):
                plan.constraints_psi4(u, c)
            torch.cuda.synchronize()
            tc = (time.perf_counter() - t0) / 3
            res["constraints_psi4"] = {"ms": 1e3 * tc, "mpts": pts / tc / 1e6}
            del c
        del plan
    err = normwise_err(outs["mixed"], outs["fp64"])
    res["mixed"]["max_normwise_rel_err_vs_fp64"] = err.max().item()
    res["mixed"]["per_field_normwise_rel_err"] = [float(f"{x:.3e}") for x in err.tolist()]
    res["mixed_speedup_vs_fp64"] = res["mixed"]["value"] / res["fp64"]["value"]
    del outs, u
    torch.cuda.empty_cache()
    return res


def interior_blocks_leg(bl, u, steps, hev):
    """The interior-only leg of the 512-block workload: the 216 blocks with
    no flagged face (bflag 0), run as one plan over the same field arrays
    (their offsets into them), FP64 and mixed."""
    import torch
    import sympygr_amd as sgr
    from sympygr_amd import data
    sub = data.BlockList()
    for b in range(len(bl)):
        if bl.bflag[b] == 0:
            sub.offsets.append(bl.offsets[b])
            sub.sz.append(bl.sz[b])
            sub.pmin.append(bl.pmin[b])
            sub.pmax.append(bl.pmax[b])
            sub.bflag.append(0)
    sub.total = bl.total
    pts = sub.interior_points()
    res = {"blocks": len(sub), "interior_points": pts}
    out = torch.zeros_like(u)
    for prec in ("fp64", "mixed"):
        plan = sgr.Plan(sub, sgr.params(adv="upwind", precision=prec, ko_sigma=0.1))
        w, kd, kp = timed_steps(plan, out, u, steps, 2, hev)
        info = plan.info()
        # a pipelined plan's events bracket the whole pass (the two kernels overlap)
        kms = ({"rhs_pass": kd + kp} if info.pipelined else {"k_stencil": kd, "k_pointwise": kp})
        res[prec] = {"value": pts * steps / w / 1e6, "kernel_ms": kms, "chunks": info.chunks,
                     "pipelined": bool(info.pipelined),
                     "roofline": mix_roofline({sub.sz[0][0] - 6: pts}, (kd + kp) * 1e-3, prec)}
        log(f"interior blocks {prec}: {res[prec]['value']:.1f} Mpts/s"
            f"{f' ({info.chunks} pipelined chunks)' if info.pipelined else ''}")
        del plan
    del out
    return res


def tvb_leg(bl, u, steps, hev):
    """The 512-block workload with the TVB dissipation the reference's
    bssnrhs_sep selects with DISSIPATION_TYPE 1 and 2 (rhs.cpp:729-757)
    instead of Kreiss-Oliger, FP64: the speeds kernel and k_tvb run beside
    the two RHS passes (whole-call rate)."""
    import torch
    import sympygr_amd as sgr
    pts = bl.interior_points()
    res = {"workload": "the 512 x 32^3 blocks, FP64, upwind, sigma 0.1, TVB3 / TVB5 dissipation"}
    out = torch.zeros_like(u)
    for diss in ("tvb3", "tvb5"):
        plan = sgr.Plan(bl, sgr.params(adv="upwind", precision="fp64", ko_sigma=0.1, dissipation=diss))
        w, _, _ = timed_steps(plan, out, u, steps, 2, hev)
        res[diss] = {"value": pts * steps / w / 1e6, "ms_per_step": 1e3 * w / steps}
        log(f"{diss} dissipation fp64: {res[diss]['value']:.1f} Mpts/s, {res[diss]['ms_per_step']:.3f} ms per RHS")
        del plan
    del out
    return res


def dendro_sizes_section(hev, sizes=(9, 13, 17, 25, 33), steps=10):
    """Block sizes of Dendro's own meshes (2^k p + 1 interior points): one
    RHS in FP64 over a uniform tiling of about 12 M interior points per
    size, with the tuned tile shape of the size (the library's table) and
    with 8^3 tiles forced (SGR_TILE_SHAPE)."""
    import torch
    import sympygr_amd as sgr
    from sympygr_amd import data
    res = {}
    for n in sizes:
        nb = max(2, round((12e6 / n ** 3) ** (1.0 / 3.0)))
        bl = data.uniform_tiling(nb=nb, n=n, stagger=0.5)
        u = data.fill_blocks(bl, backend="torch", device="cuda")
        out = torch.zeros_like(u)
        pts = bl.interior_points()
        r = {"blocks": len(bl), "interior_points": pts}
        for tag, force in (("tuned", None), ("cube8", "8,8,8")):
            if force:
                os.environ["SGR_TILE_SHAPE"] = force
            try:
                plan = sgr.Plan(bl, sgr.params(adv="upwind", precision="fp64", ko_sigma=0.1))
            finally:
                os.environ.pop("SGR_TILE_SHAPE", None)
            w, kd, kp = timed_steps(plan, out, u, steps, 2, hev)
            r[tag] = {"value": pts * steps / w / 1e6, "tiles": plan.info().tiles,
                      "kernel_ms": {"k_stencil": kd, "k_pointwise": kp}}
            del plan
        r["gain"] = r["tuned"]["value"] / r["cube8"]["value"]
        log(f"dendro size {n}^3: tuned tiles {r['tuned']['value']:.1f} Mpts/s, 8^3 tiles {r['cube8']['value']:.1f}")
        res[str(n)] = r
        del u, out
        torch.cuda.empty_cache()
    return {"workload": "uniform tilings of n^3-interior blocks (about 12 M interior points each), Brill-Lindquist, "
                        "FP64, upwind, KO sigma 0.1", "sizes": res}


def host_dropin_rate(bl, u, nsample=16, reps=3):
    """The C++ drop-in path on HOST arrays (sgr_bssnrhs_host, as bssnrhs(...)
    is called by a Dendro driver): one call per block over the first
    `nsample` blocks, each staging the padded block to the GPU and copying
    the RHS back.  Interior points per second, PCIe included."""
    import numpy as np
    import sympygr_amd as sgr
    n = bl.offsets[nsample - 1] + bl.sz[nsample - 1][0] * bl.sz[nsample - 1][1] * bl.sz[nsample - 1][2]
    hu = np.ascontiguousarray(u[:, :n].cpu().numpy())
    out = np.zeros_like(hu)
    prm = sgr.params(adv="upwind", precision="fp64", ko_sigma=0.1)

    def sweep():
        for b in range(nsample):
            sgr.bssnrhs(out, hu, bl.offsets[b], bl.pmin[b], bl.pmax[b], bl.sz[b], bl.bflag[b], prm)
    sweep()  # the first call builds the staging context
    t0 = time.perf_counter()
    for _ in range(reps):
        sweep()
    t = time.perf_counter() - t0
    pts = sum((bl.sz[b][0] - 6) * (bl.sz[b][1] - 6) * (bl.sz[b][2] - 6) for b in range(nsample)) * reps
    return {"value": pts / t / 1e6, "unit": "Mpoints/s", "ms_per_call": 1e3 * t / (reps * nsample),
            "what": f"bssnrhs on host arrays (sgr_bssnrhs_host), one call per block, {nsample} blocks x {reps}, "
                    "each staging 24 x 38^3 doubles in and out over PCIe"}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def _sample_blocks(bl, u, nint, nface):
    """A BlockList of the first `nint` interior-only blocks and the first
    `nface` boundary blocks (z-y-x order) of bl, with their host data."""
    import numpy as np
    from sympygr_amd.data import BlockList
    inner = [b for b in range(len(bl)) if bl.bflag[b] == 0][:nint]
    faces = [b for b in range(len(bl)) if bl.bflag[b] != 0][:nface]
    sub = BlockList()
    cols = []
    for b in inner + faces:
        sub.add(bl.sz[b], bl.pmin[b], bl.pmax[b], bl.bflag[b])
        n = bl.sz[b][0] * bl.sz[b][1] * bl.sz[b][2]
        cols.append(u[:, bl.offsets[b]:bl.offsets[b] + n].cpu().numpy())
    return sub, np.ascontiguousarray(np.concatenate(cols, axis=1)), len(inner), len(faces)


def _host_threads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def reference_baseline(bl, u, target_s=10.0):
    """The compiled reference bssnrhs (oracle/_ref, upwind build: the
    reference's own sources, g++ -O2) on this host, as its example driver
    runs it (bssn/cuda_gr/src/gpuBSSNExample.cpp:267-300): an OpenMP loop over
    blocks, one bssnrhs call per block, on every host core this process may
    use; sample: 16 interior-only and 16 boundary blocks of the workload.
    Beside it the same code on one thread (4 + 4 blocks).  None when the
    reference has not been built."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    if not po.ref_available("u"):
        return None
    P = po.make_params(upwind=True, ko_sigma=0.1)
    res = {}
    for tag, threads, nint, nface in (("all", _host_threads(), 16, 16), ("one", 1, 4, 4)):
        sub, hu, ni, nf = _sample_blocks(bl, u, nint, nface)
        out = np.zeros_like(hu)
        po.ref_rhs_blocks("u", hu, sub, P, nthreads=threads, out=out)  # warm-up
        reps, t = 0, 0.0
        t0 = time.perf_counter()
        while t < target_s / 2 or reps < 2:
            po.ref_rhs_blocks("u", hu, sub, P, nthreads=threads, out=out)
            reps += 1
            t = time.perf_counter() - t0
        res[tag] = {"value": sub.interior_points() * reps / t / 1e6, "threads": threads, "blocks": (ni, nf),
                    "passes": reps, "wall_s": t}
    a, o = res["all"], res["one"]
    return {"value": a["value"], "unit": "Mpoints/s", "cores": a["threads"], "kind": "reference",
            "sample": f"{a['blocks'][0]} interior-only + {a['blocks'][1]} boundary blocks of the 512 (38^3), "
                      f"{a['passes']} passes in {a['wall_s']:.1f} s; the reference bssnrhs built from its sources "
                      f"(oracle/Makefile, upwind variant, g++ -O2 -fopenmp), OpenMP over blocks on {a['threads']} "
                      f"threads as bssn/cuda_gr/src/gpuBSSNExample.cpp:267-300 runs it; nproc {os.cpu_count()}, "
                      f"CPU: {_cpu_model()}",
            "one_thread": {"value": o["value"], "unit": "Mpoints/s",
                           "sample": f"{o['blocks'][0]} interior + {o['blocks'][1]} boundary blocks, "
                                     f"{o['passes']} passes in {o['wall_s']:.1f} s"}}


def cpu_baseline(bl, u, target_s=12.0):
    """The CPU oracle (this project's C++ restatement of the reference
    bssnrhs, same structure: full-array sweeps, pointwise loop, BCs, KO) on a
    bounded sample of the same blocks, OpenMP over blocks on this host."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    threads = _host_threads()
    sub, hu, ni, nf = _sample_blocks(bl, u, 16, 16)
    out = np.zeros_like(hu)
    P = po.make_params(upwind=True, ko_sigma=0.1)
    po.rhs_blocks(hu, sub, P, nthreads=threads, out=out)  # warm-up
    reps, t = 0, 0.0
    t0 = time.perf_counter()
    while t < target_s / 4 or reps < 2:
        po.rhs_blocks(hu, sub, P, nthreads=threads, out=out)
        reps += 1
        t = time.perf_counter() - t0
    pts = sub.interior_points() * reps
    return {"value": pts / t / 1e6, "unit": "Mpoints/s", "cores": threads, "kind": "port",
            "sample": f"{ni} interior-only + {nf} boundary blocks of the 512, {reps} RHS passes, {t:.1f} s wall; "
                      f"OpenMP over blocks on {threads} threads; nproc {os.cpu_count()}, CPU: {_cpu_model()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-host", action="store_true",
                    help="skip the host-array drop-in leg (PCIe-inclusive per-block calls)")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-drift", action="store_true",
                    help="skip the constraint-drift comparison of FP64 and mixed on the octree")
    ap.add_argument("--drift-steps", type=int, default=100)
    ap.add_argument("--octree-sizes", type=int, nargs="+", default=[16, 24, 32, 48],
                    help="leaf interior sizes of the octree section, drawn per leaf")
    ap.add_argument("--no-octree", action="store_true",
                    help="skip the batched uneven-octree measurements (RHS, RK4 step, drift) and the large block")
    ap.add_argument("--no-ccz4", action="store_true", help="skip the CCZ4 192^3-block leg")
    ap.add_argument("--no-tvb", action="store_true", help="skip the TVB-dissipation leg")
    ap.add_argument("--no-dendro-sizes", action="store_true",
                    help="skip the leg over Dendro's block sizes (9/13/17/25/33, tuned tiles against 8^3)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 1 GiB device-to-device copy (k_hbm_copy, reference point for HBM counters)")
    ap.add_argument("--counters", action="store_true",
                    help="hardware-counter runs: inputs filled on the host and copied over, no work on the "
                         "GPU but the RHS steps (and --copy-calibration)")
    ap.add_argument("--counters-prec", choices=("fp64", "mixed"), default="fp64",
                    help="precision of the RHS steps of a --counters run")
    args = ap.parse_args()
    if args.counters:
        args.no_mixed = args.no_drift = args.no_octree = args.no_cpu = args.no_ccz4 = args.no_dendro_sizes = True
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the RHS path is single-GPU here); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import sympygr_amd as sgr
    from sympygr_amd import data

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    sgr.lib()
    n = 32
    bl = data.uniform_tiling(nb=8, n=n, stagger=0.5)
    log(f"filling {len(bl)} blocks ({bl.total} padded points per field)")
    if args.counters:
        u = torch.from_numpy(data.fill_blocks(bl, backend="numpy")).cuda()
        out = torch.empty_like(u)
    else:
        u = data.fill_blocks(bl, backend="torch", device="cuda")
        out = torch.zeros_like(u)
    pts = bl.interior_points()
    hev = HipEvents()

    prec0 = args.counters_prec if args.counters else "fp64"
    plan = sgr.Plan(bl, sgr.params(adv="upwind", precision=prec0, ko_sigma=0.1))
    log(f"{prec0} plan: workspace {plan.info().workspace_bytes / 2**30:.1f} GiB; warmup {args.warmup}, "
        f"steps {args.steps}")
    wall, kd, kp = timed_steps(plan, out, u, args.steps, args.warmup, hev)
    value = pts * args.steps / wall / 1e6
    log(f"{prec0}: {value:.1f} Mpts/s, {1e3 * wall / args.steps:.3f} ms/step, k_stencil {kd:.3f} ms, "
        f"k_pointwise {kp:.3f} ms")
    ref_out = out if args.counters else out.clone()
    del plan

    mixed = None
    if not args.no_mixed:
        mplan = sgr.Plan(bl, sgr.params(adv="upwind", precision="mixed", ko_sigma=0.1))
        mout = torch.zeros_like(u)
        mwall, mkd, mkp = timed_steps(mplan, mout, u, args.steps, args.warmup, hev)
        mval = pts * args.steps / mwall / 1e6
        err = normwise_err(mout, ref_out)
        rms = ((mout - ref_out).pow(2).mean(dim=1).sqrt() / ref_out.pow(2).mean(dim=1).sqrt().clamp_min(1e-300))
        log(f"mixed: {mval:.1f} Mpts/s, k_stencil {mkd:.3f} ms, k_pointwise {mkp:.3f} ms, "
            f"max err {err.max().item():.2e}")
        mixed = {"value": mval, "ms_per_step": 1e3 * mwall / args.steps, "speedup_vs_fp64": mval / value,
                 "kernel_ms": {"k_stencil": mkd, "k_pointwise": mkp},
                 "roofline": mix_roofline({n: pts}, (mkd + mkp) * 1e-3, "mixed"),
                 "max_normwise_rel_err_vs_fp64": err.max().item(),
                 "per_field_normwise_rel_err": [float(f"{x:.3e}") for x in err.tolist()],
                 "per_field_rms_rel_err": [float(f"{x:.3e}") for x in rms.tolist()]}
        del mplan, mout

    interior = None
    if not args.counters:
        interior = interior_blocks_leg(bl, u, args.steps, hev)
    tvb = None
    if not (args.counters or args.no_tvb):
        tvb = tvb_leg(bl, u, max(2, args.steps // 2), hev)

    rhs_ms = kd + kp
    abytes = alg_bytes_per_point(n) * pts
    achieved_gbs = abytes / (rhs_ms * 1e-3) / 1e9
    achieved_tfs = FLOPS_UPWIND * pts / (rhs_ms * 1e-3) / 1e12
    dominant = "k_pointwise" if kp >= kd else "k_stencil"
    # per kernel: k_stencil does the stencils and the KO terms (FP64) and
    # reads the fields over the padded blocks; the pointwise kernel runs the
    # algebra
    st_flops = FLOPS_STENCIL + FLOPS_KO_ADD
    st_tfs = st_flops * pts / (kd * 1e-3) / 1e12
    st_gbs = 192.0 * ((n + 6) / n) ** 3 * pts / (kd * 1e-3) / 1e9
    pw_tfs = FLOPS_ALGEBRA * pts / (kp * 1e-3) / 1e12
    traffic, traffic_detail = measured_traffic(pts)
    copy_gbs = None
    if not args.counters:
        # for scale: the HBM rate of the library's plain copy kernel here
        csrc = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        cdst = torch.empty_like(csrc)
        sgr.hbm_copy(cdst, csrc)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            sgr.hbm_copy(cdst, csrc)
        e1.record()
        torch.cuda.synchronize()
        copy_gbs = 10 * 2 * (1 << 30) / (e0.elapsed_time(e1) * 1e-3) / 1e9
        del csrc, cdst
        log(f"copy kernel: {copy_gbs:.0f} GB/s (1 GiB read + 1 GiB written)")
    bnd = bound_ps_per_point(n, "fp64")
    roofline = {
        # at n = 32 the FP64 vector roof binds (91 ps/pt against 64 for HBM)
        "bound": "fp64_vector" if bnd["binding"] == "valu" else "hbm",
        "achieved": achieved_tfs, "peak": FP64_PEAK_TFS, "unit": "TFLOP/s", "frac": achieved_tfs / FP64_PEAK_TFS,
        "traffic": traffic, "traffic_detail": traffic_detail,
        "kernel": "RHS pass = k_stencil + k_pointwise (HIP events around each launch on "
                  "the launch stream)",
        "flops_per_point": FLOPS_UPWIND, "bound_ps_per_point": bnd,
        "hbm": {"achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved_gbs / HBM_PEAK_GBS,
                "algorithmic_bytes_per_point": alg_bytes_per_point(n),
                "traffic_rate_GBps": traffic / (rhs_ms * 1e-3) / 1e9 if traffic else None,
                "copy_kernel_GBps": copy_gbs},
        "dominant_kernel": dominant,
        "kernels": {
            "k_stencil": {"ms": kd, "fp64_TFs": st_tfs, "fp64_frac": st_tfs / FP64_PEAK_TFS,
                          "field_read_GBps": st_gbs, "hbm_frac": st_gbs / HBM_PEAK_GBS,
                          "flops_per_point": st_flops},
            "k_pointwise": {"ms": kp, "fp64_TFs": pw_tfs, "fp64_frac": pw_tfs / FP64_PEAK_TFS,
                            "flops_per_point": FLOPS_ALGEBRA},
        },
        "measured_valu": measured_valu(pts),
    }
    octree = drift = None
    if not args.no_octree:
        del out, ref_out
        torch.cuda.empty_cache()
        octree, drift = octree_section(args, hev)

    # a single large block (192^3 interior, 198^3 padded): the large-block,
    # single-launch shape (13824 tiles of one block)
    large = None
    if not args.no_octree:
        lbl = data.BlockList()
        nL = 198
        lbl.add((nL, nL, nL), (-4.0,) * 3, (4.0,) * 3, 63)
        lu = data.fill_blocks(lbl, backend="torch", device="cuda", kind="perturbed", seed=7, amp=1e-2,
                              lo=-4.0, L=8.0)
        lout = torch.zeros_like(lu)
        lpts = lbl.interior_points()
        large = {"workload": "one block of 192^3 interior (198^3 padded), perturbed flat data, all faces flagged",
                 "interior_points": lpts}
        for prec in ("fp64", "mixed"):
            lp = sgr.Plan(lbl, sgr.params(adv="upwind", precision=prec, ko_sigma=0.1))
            w, _, _ = timed_steps(lp, lout, lu, 10, 2, hev)
            large[prec] = lpts * 10 / w / 1e6
            del lp
        log(f"large block: fp64 {large['fp64']:.1f} Mpts/s, mixed {large['mixed']:.1f} Mpts/s")
        del lu, lout

    ccz4 = None
    if not args.no_ccz4:
        ccz4 = ccz4_section(hev)
    dendro = None
    if not args.no_dendro_sizes:
        dendro = dendro_sizes_section(hev)

    if args.copy_calibration:
        import numpy as np
        src = torch.from_numpy(np.ones(1 << 27)).cuda()
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        sgr.hbm_copy(dst, src)  # k_hbm_copy: reads and writes exactly 1 GiB
        torch.cuda.synchronize()
        del src, dst
    host = None
    if not (args.no_host or args.counters):
        host = host_dropin_rate(bl, u)
        log(f"host drop-in: {host['value']:.1f} Mpts/s, {host['ms_per_call']:.2f} ms per block call (PCIe included)")

    cpu = None
    if not args.no_cpu:
        # the compiled reference when oracle/_ref holds it, with this
        # project's multi-threaded restatement beside it; else the restatement
        log("cpu baseline (compiled reference and oracle on bounded samples)")
        port = cpu_baseline(bl, u)
        log(f"cpu oracle: {port['value']:.3f} Mpts/s on {port['cores']} threads")
        ref = reference_baseline(bl, u)
        if ref is not None:
            log(f"cpu reference: {ref['value']:.4f} Mpts/s on {ref['cores']} threads, "
                f"{ref['one_thread']['value']:.4f} on 1")
            cpu = dict(ref, port=port)
        else:
            cpu = port

    line = {
        "metric": METRIC, "value": value, "unit": "Mpoints/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: Brill-Lindquist two-puncture data, analytic ghosts",
        "config": {"workload": "512 blocks (8^3 tiling of [-8,8]^3) of 32^3 interior + 3 ghost (38^3), "
                               "FP64, 6th-order stencils, upwinded advection, KO sigma 0.1",
                   "blocks": len(bl), "interior_points": pts, "adv": "upwind", "precision": prec0},
        "roofline": roofline, "cpu_baseline": cpu, "mixed": mixed, "batched_octree": octree,
        "constraint_drift": drift, "large_block": large, "ccz4_large_block": ccz4, "interior_blocks": interior,
        "tvb_dissipation": tvb,
        "dendro_block_sizes": dendro,
        "host_dropin": host,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
