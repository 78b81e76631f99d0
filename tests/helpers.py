"""Shared helpers for the tests (fixtures, norms, the C0 block)."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


def golden(name):
    return np.load(os.path.join(GOLDEN, f"rhs_{name}.npy"))


def interior(a, sz):
    nx, ny, nz = sz
    return a.reshape(24, nz, ny, nx)[:, 3:-3, 3:-3, 3:-3]


def normwise(a, b):
    """Per-field max|a-b| / max|b| (fields whose reference is zero use an
    absolute floor of 1e-300 in the denominator)."""
    a = a.reshape(24, -1)
    b = b.reshape(24, -1)
    return np.array([np.abs(a[f] - b[f]).max() / max(np.abs(b[f]).max(), 1e-300) for f in range(24)])


def case_input(name):
    """Input array and geometry of a golden case."""
    from sympygr_amd import data
    m = manifest()
    c = m["cases"][name]
    if c["input"] == "c0":
        g = m["c0_input"]
        u = data.perturbed_flat_block(seed=g["seed"], n=g["n"], amp=g["amp"], pmin=tuple(g["pmin"]),
                                      pmax=tuple(g["pmax"]))
    else:
        g = m["fake_input"]
        u = np.load(os.path.join(GOLDEN, g["file"]))
    n = g["n"]
    return u, (n, n, n), g["pmin"], g["pmax"], c


def case_params(c, **kw):
    m = manifest()
    p = dict(m["params"])
    p.update(kw)
    return p, c["variant"] == "u"
