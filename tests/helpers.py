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
    elif c["input"] == "gauge":
        g = m["gauge_input"]
        u = data.puncture_perturbed_block(seed=g["seed"], n=g["n"], amp=g["amp"], pmin=tuple(g["pmin"]),
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
    return p, c["variant"].startswith("u")


def case_gauge(c):
    """Gauge keywords of a golden case (standard gauge, constant eta unless
    the case was made with another variant of the reference build)."""
    kw = dict(gauge=c.get("gauge", "standard"), eta=c.get("eta", "const"))
    if "gauge" in c:
        g = manifest()["gauge_params"]
        kw.update(xi=tuple(g["xi"]), eta_func_r0=g["eta_func_r0"], eta_power=tuple(g["eta_power"]))
    if "dissipation" in c:
        kw["dissipation"] = c["dissipation"]
    return kw


def fixture_mask(c):
    """Interior nodes (k, j, i) where a golden case is defined.  The
    reference's TVB dissipation reads characteristic speeds it never wrote at
    the 2 (TVB3) or 3 (TVB5) nodes next to an unflagged face, and leaves the
    TVB3 value of the node next to one as it was; those nodes are left out."""
    m = np.ones((c["n"] - 6,) * 3, dtype=bool)
    depth = {"tvb3": 2, "tvb5": 3}.get(c.get("dissipation"), 0)
    for ax in range(3):
        for side in range(2):
            if depth and not c["bflag"] & (1 << (2 * ax + side)):
                sl = [slice(None)] * 3
                sl[2 - ax] = slice(0, depth) if side == 0 else slice(-depth, None)
                m[tuple(sl)] = False
    return m


def small_octree(sizes=(16, 24)):
    """22 leaves on levels 1..3 of [-2, 2]^3 refined around one point: leaves
    of interior 16 or 24, neighbours of other sizes and levels (4:1 spacing
    jumps included) - every kind of ghost patch in a set the oracle refreshes
    in well under a second."""
    from sympygr_amd import data
    return data.octree_blocks(domain=2.0, centres=((0.3, 0.2, 0.1),), c=0.01, lmin=1, lmax=3, sizes=sizes)


def ghost_mask(bl, inside_only=True):
    """Boolean mask over the field arrays: ghost nodes (inside the domain, i.e.
    not beyond a flagged face, when inside_only)."""
    mk = np.zeros(bl.total, dtype=bool)
    for b in range(len(bl)):
        s = bl.sz[b]
        v = np.ones((s[2], s[1], s[0]), dtype=bool)
        v[3:-3, 3:-3, 3:-3] = False
        if inside_only:
            f = bl.bflag[b]
            if f & 1:
                v[:, :, :3] = False
            if f & 2:
                v[:, :, -3:] = False
            if f & 4:
                v[:, :3, :] = False
            if f & 8:
                v[:, -3:, :] = False
            if f & 16:
                v[:3, :, :] = False
            if f & 32:
                v[-3:, :, :] = False
        mk[bl.offsets[b]:bl.offsets[b] + v.size] = v.reshape(-1)
    return mk


def coords(bl):
    """x, y, z of every node of every block, flattened like the field arrays."""
    from sympygr_amd import data
    X = np.empty((3, bl.total))
    for b in range(len(bl)):
        o, s = bl.offsets[b], bl.sz[b]
        n = s[0] * s[1] * s[2]
        x, y, z = data.block_coords(data._Np(), s, bl.pmin[b], bl.pmax[b])
        X[0, o:o + n], X[1, o:o + n], X[2, o:o + n] = x.reshape(-1), y.reshape(-1), z.reshape(-1)
    return X


def poly5(X, f):
    """A field-dependent polynomial of degree 5 in each coordinate."""
    x, y, z = X / 2.0
    a = 0.1 * (f + 1)
    return (1.0 + a * x - 0.3 * x ** 2 + 0.07 * x ** 5) * (0.5 - y ** 3 + a * y ** 4 - 0.2 * y ** 5) * \
        (2.0 + z - a * z ** 2 + 0.4 * z ** 5) + 0.25 * x ** 5 * z ** 4
