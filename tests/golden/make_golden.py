"""Write the golden fixtures from the compiled reference (oracle/_ref).

Run in a container that has the reference sources and a built oracle/_ref:

    make -C oracle ref && python tests/golden/make_golden.py

Outputs (all numbers, nothing copied from the reference):
  rhs_<case>.npy       interior RHS, shape (24, n-6, n-6, n-6), float64
  stencils.json        SHA-256 of every single-sweep output (bit-exact checks)
  tvb.json             SHA-256 of the TVB sweeps and characteristic speeds
  manifest.json        input recipes, parameters, flags, input hashes
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle as po  # noqa: E402
from sympygr_amd import data  # noqa: E402

C0 = dict(seed=1, n=22, amp=1e-3, pmin=[-1.05] * 3, pmax=[1.05] * 3)
FAKE = dict(n=13, pmin=[0.1, 0.2, 0.3], pmax=[1.3, 1.5, 1.4])
GAUGE = dict(seed=5, n=16, amp=1e-2, pmin=[-0.8, -0.75, -0.7], pmax=[0.7, 0.75, 0.8])
# gauge-variant parameters (BSSN_XI, BSSN_ETA_R0, BSSN_ETA_POWER)
GAUGE_PARAMS = dict(xi=[1, 1, 1], eta_func_r0=1.31, eta_power=[2.0, 2.0])
PARAMS = dict(ko_sigma=0.1, eta_const=2.0, eta_r0=50.0, eta_damping_exp=1.0, lambda_=[1, 1, 1, 1], lambda_f=[1.0, 0.0])
REF_FLAGS = "g++ -std=c++11 -O2 -ffp-contract=off -fopenmp (oracle/Makefile)"


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def interior(a, n):
    return np.ascontiguousarray(a.reshape(24, n, n, n)[:, 3:-3, 3:-3, 3:-3])


def c0_input():
    return data.perturbed_flat_block(seed=C0["seed"], n=C0["n"], amp=C0["amp"], pmin=tuple(C0["pmin"]),
                                     pmax=tuple(C0["pmax"]))


def fake_input():
    n = FAKE["n"]
    return po.ref_fake_initial_data("u", (n, n, n), FAKE["pmin"], FAKE["pmax"])


def gauge_input():
    return data.puncture_perturbed_block(seed=GAUGE["seed"], n=GAUGE["n"], amp=GAUGE["amp"],
                                         pmin=tuple(GAUGE["pmin"]), pmax=tuple(GAUGE["pmax"]))


def tvb_written(which, n, bflag):
    """Nodes of a TVB sweep's interior box that the reference writes, (k, j, i):
    disstvb3 leaves the first and last interior node before an unflagged face
    as they were (derivs.cpp:1374, 1478, 1585)."""
    m = np.ones((n - 6,) * 3, bool)
    if which < 3:
        ax = which
        for side in range(2):
            if not bflag & (1 << (2 * ax + side)):
                sl = [slice(None)] * 3
                sl[2 - ax] = 0 if side == 0 else n - 7
                m[tuple(sl)] = False
    return m


def tvb_sweep_hash(out, which, n, bflag):
    a = out.reshape(n, n, n)[3:-3, 3:-3, 3:-3]
    return sha(np.where(tvb_written(which, n, bflag), a, 0.0))


def write_tvb_hashes(ug, uf):
    """tvb.json: speeds of the gauge input at its interior nodes; the six TVB
    sweeps of the fake input's gt1 (field 13) with the oracle's speeds of the
    fake input along the sweep axis, h = 0.1, four face-flag sets; and, for the
    disstvb3 nodes next to unflagged faces, the reference's interior form one
    node further in on the input shifted by one node along the axis (the value
    the restatement uses there)."""
    t = {}
    n = GAUGE["n"]
    lam = po.ref_spacetime_speeds("u", ug, (n, n, n))
    t["speeds_gauge"] = sha(np.stack([lam[d].reshape(n, n, n)[3:-3, 3:-3, 3:-3] for d in range(3)]))
    n = FAKE["n"]
    lf = po.spacetime_speeds(uf, (n, n, n))
    for which in range(6):
        for bflag in (0, 63, 21, 42):
            out = po.ref_tvb("u", which, uf[13], lf[which % 3], 0.1, (n, n, n), bflag)
            t[f"sweep_{which}_{bflag}"] = tvb_sweep_hash(out, which, n, bflag)
    for ax in range(3):
        for side, shift, plane in ((0, 1, 4), (1, -1, n - 5)):
            us = np.roll(uf[13].reshape(n, n, n), shift, axis=2 - ax).ravel()
            ls = np.roll(lf[ax].reshape(n, n, n), shift, axis=2 - ax).ravel()
            out = po.ref_tvb("u", ax, us, ls, 0.1, (n, n, n), 0).reshape(n, n, n)
            sl = [slice(3, -3)] * 3
            sl[2 - ax] = plane
            t[f"tvb3_shift_{ax}_{side}"] = sha(out[tuple(sl)])
    with open(os.path.join(HERE, "tvb.json"), "w") as f:
        json.dump(t, f, indent=1, sort_keys=True)


def main():
    manifest = {"reference_build": REF_FLAGS, "params": PARAMS, "cases": {}}
    u0 = c0_input()
    manifest["c0_input"] = dict(C0, recipe="sympygr_amd.data.perturbed_flat_block", sha256=sha(u0))
    uf = fake_input()
    np.save(os.path.join(HERE, "fake_input_13.npy"), uf)
    manifest["fake_input"] = dict(FAKE, recipe="reference bssn::fake_initial_data (bssn/src/grUtils.cpp:15-57)",
                                  file="fake_input_13.npy", sha256=sha(uf))
    cases = [
        ("c0_upwind_b0", u0, C0, "u", 0), ("c0_upwind_b63", u0, C0, "u", 63),
        ("c0_centred_b63", u0, C0, "c", 63),
        ("fake_upwind_b0", uf, FAKE, "u", 0), ("fake_upwind_b63", uf, FAKE, "u", 63),
        ("fake_centred_b0", uf, FAKE, "c", 0), ("fake_centred_b21", uf, FAKE, "c", 21),
        ("fake_upwind_b42", uf, FAKE, "u", 42),
    ]
    for name, u, geo, var, bflag in cases:
        n = geo["n"]
        P = po.make_params(upwind=(var == "u"), **PARAMS)
        out = po.ref_rhs_block(var, u, (n, n, n), geo["pmin"], geo["pmax"], bflag, P)
        out = interior(out, n)
        np.save(os.path.join(HERE, f"rhs_{name}.npy"), out)
        manifest["cases"][name] = dict(input="c0" if geo is C0 else "fake", variant=var, bflag=bflag, n=n,
                                       file=f"rhs_{name}.npy", sha256=sha(out))
    # gauge variants of the reference build and its staged bssnrhs_sep, on a
    # 16^3 block with every field non-trivial
    ug = gauge_input()
    manifest["gauge_input"] = dict(GAUGE, recipe="sympygr_amd.data.puncture_perturbed_block", sha256=sha(ug))
    manifest["gauge_params"] = GAUGE_PARAMS
    gcases = [("gauge_standard_b63", "u", "standard", "const", False),
              ("gauge_etafunc_b63", "u_func", "standard", "func", False),
              ("gauge_rochester_b63", "u_roch", "rochester", "const", False),
              ("gauge_rochester_etafunc_b63", "u_roch_func", "rochester", "func", False),
              ("gauge_standard_staged_b63", "u", "standard", "const", True)]
    n = GAUGE["n"]
    for name, var, gauge, eta, staged in gcases:
        P = po.make_params(upwind=True, gauge=gauge, eta=eta, **PARAMS, **GAUGE_PARAMS)
        out = interior(po.ref_rhs_block(var, ug, (n, n, n), GAUGE["pmin"], GAUGE["pmax"], 63, P, staged=staged), n)
        np.save(os.path.join(HERE, f"rhs_{name}.npy"), out)
        manifest["cases"][name] = dict(input="gauge", variant=var, gauge=gauge, eta=eta,
                                       entry="bssnrhs_sep" if staged else "bssnrhs", bflag=63, n=n,
                                       file=f"rhs_{name}.npy", sha256=sha(out))
    # TVB dissipation of bssnrhs_sep (DISSIPATION_TYPE 1 and 2) on the gauge
    # input.  With every face flagged the reference's result is fully defined;
    # with faces 1, 3, 5 flagged (bflag 42) the nodes within 2 (TVB3) or 3
    # (TVB5) of an unflagged face read speeds the reference never wrote, and a
    # test compares the other nodes only.
    for diss in ("tvb3", "tvb5"):
        for bflag in (63, 42):
            name = f"{diss}_staged_b{bflag}"
            P = po.make_params(upwind=True, dissipation=diss, **PARAMS)
            out = interior(po.ref_rhs_block("u", ug, (n, n, n), GAUGE["pmin"], GAUGE["pmax"], bflag, P, staged=True), n)
            np.save(os.path.join(HERE, f"rhs_{name}.npy"), out)
            manifest["cases"][name] = dict(input="gauge", variant="u", dissipation=diss, entry="bssnrhs_sep",
                                           bflag=bflag, n=n, file=f"rhs_{name}.npy", sha256=sha(out))
    write_tvb_hashes(ug, uf)
    # single-sweep hashes on the fake input's gt1 (field 13) with beta from the
    # same input, 13^3 block, h = 0.1
    st = {}
    n = FAKE["n"]
    for var in ("u", "c"):
        for bflag in (0, 63, 21, 42):
            for which in range(12):
                beta = uf[6 + which % 3] if which // 3 == 2 else None
                out = po.ref_deriv(var, which, uf[13], 0.1, (n, n, n), bflag, beta=beta)
                st[f"{var}_{bflag}_{which}"] = sha(interior(np.stack([out] * 24), n)[0])
    with open(os.path.join(HERE, "stencils.json"), "w") as f:
        json.dump(st, f, indent=1, sort_keys=True)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print("wrote", len(cases) + len(gcases) + 4, "rhs fixtures,", len(st), "stencil hashes and tvb.json")


if __name__ == "__main__":
    main()
