"""Synthetic inputs for the BSSN RHS workloads (host side, numpy or torch).

Every generator is written once against a tiny array-module shim so the same
formulas produce the committed fixtures (numpy, float64) and the benchmark's
device-resident inputs (torch on cuda:0), and a test checks that both agree.

Workloads
---------
* ``perturbed_flat_block`` -- one ghost-padded block: flat background
  (gt=delta, chi=alpha=1, rest 0) plus ``amp`` times a sum of 8 seeded random
  Fourier modes per field (integer wavenumbers 1-4 across the block, uniform
  phases), then gt renormalised to det 1 and At made trace-free.
* ``brill_lindquist`` -- two punctures of mass 0.5 at (+-2,0,0):
  psi = 1 + sum m/(2 r), chi = psi^-4, alpha = psi^-2, gt = delta, rest 0,
  r floored at 1e-6.
* ``uniform_tiling`` -- 8x8x8 blocks of 32^3 interior points tiling
  [-8,8]^3 (dx = 0.0625), bflag on the outer faces.
* ``octree_blocks`` -- an uneven block set refined toward the punctures.
* ``ccz4_fill`` (``fill_blocks(kind="ccz4")``) -- the CCZ4 test data: the
  perturbed flat background (amplitude 1e-2), Theta perturbed the same way
  (field 24), and a Gaussian lapse pulse alpha += A exp(-r^2/w^2), A = 0.1,
  w = 1.

Random numbers come from SplitMix64 keyed by (seed, field, mode), so any
implementation of the recipe reproduces the inputs bit for bit.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

NFIELDS = 24
GHOST = 3
MASK64 = (1 << 64) - 1

# field indices (bssn/include/grDef.h:31)
ALPHA, CHI, K = 0, 1, 2
GT = [12, 13, 14, 15, 16, 17]
AT = [18, 19, 20, 21, 22, 23]


def splitmix64(state: int):
    """Yield the SplitMix64 sequence starting after ``state``."""
    while True:
        state = (state + 0x9E3779B97F4A7C15) & MASK64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        yield z ^ (z >> 31)


def fourier_modes(seed: int, f: int, nmodes: int = 8):
    """(kx, ky, kz, phase) for each mode of field f."""
    g = splitmix64(((seed & 0xFFFFFFFF) << 32) ^ (f << 16) ^ 0x5EED)
    modes = []
    for _ in range(nmodes):
        kx = 1 + next(g) % 4
        ky = 1 + next(g) % 4
        kz = 1 + next(g) % 4
        ph = (next(g) >> 11) * (1.0 / (1 << 53)) * 2.0 * math.pi
        modes.append((kx, ky, kz, ph))
    return modes


class _Np:
    """numpy array-module shim"""
    def __init__(self):
        self.sin, self.sqrt, self.maximum, self.exp = np.sin, np.sqrt, np.maximum, np.exp

    @staticmethod
    def arange(n, dtype=None, like=None):
        return np.arange(n, dtype=np.float64)

    @staticmethod
    def zeros(shape, like=None):
        return np.zeros(shape, dtype=np.float64)

    @staticmethod
    def full_like(x, v):
        return np.full_like(x, v)

    @staticmethod
    def cbrt(x):
        return np.cbrt(x)


class _Torch:
    def __init__(self, device):
        import torch
        self.t = torch
        self.device = device
        self.sin, self.sqrt, self.exp = torch.sin, torch.sqrt, torch.exp

    def maximum(self, a, b):
        if not hasattr(b, "shape"):
            b = self.t.full_like(a, b)
        return self.t.maximum(a, b)

    def arange(self, n, dtype=None, like=None):
        return self.t.arange(n, dtype=self.t.float64, device=self.device)

    def zeros(self, shape, like=None):
        return self.t.zeros(shape, dtype=self.t.float64, device=self.device)

    def full_like(self, x, v):
        return self.t.full_like(x, v)

    def cbrt(self, x):
        return self.t.sign(x) * self.t.abs(x) ** (1.0 / 3.0)


def _xp(backend, device):
    if backend == "numpy":
        return _Np()
    return _Torch(device)


def block_coords(xp, sz, pmin, pmax):
    """x, y, z at every point of a padded block, x fastest (IDX = i + nx(j + ny k));
    h = (pmax - pmin)/(n - 1) and x = pmin + i h as in bssn/src/rhs.cpp:144-146,207-211."""
    nx, ny, nz = sz
    hx = (pmax[0] - pmin[0]) / (nx - 1)
    hy = (pmax[1] - pmin[1]) / (ny - 1)
    hz = (pmax[2] - pmin[2]) / (nz - 1)
    x1 = pmin[0] + xp.arange(nx) * hx
    y1 = pmin[1] + xp.arange(ny) * hy
    z1 = pmin[2] + xp.arange(nz) * hz
    x = x1[None, None, :] + 0.0 * y1[None, :, None] + 0.0 * z1[:, None, None]
    y = 0.0 * x1[None, None, :] + y1[None, :, None] + 0.0 * z1[:, None, None]
    z = 0.0 * x1[None, None, :] + 0.0 * y1[None, :, None] + z1[:, None, None]
    return x, y, z


def _normalise_metric_and_trace(u):
    """gt <- gt det(gt)^(-1/3); At <- At - gt (igt^{kl} At_kl)/3 (in place, numpy)."""
    g = [u[i] for i in GT]
    det = g[0] * (g[3] * g[5] - g[4] ** 2) - g[1] * (g[1] * g[5] - g[2] * g[4]) + g[2] * (g[1] * g[4] - g[2] * g[3])
    s = det ** (-1.0 / 3.0)
    for i in GT:
        u[i] = u[i] * s
    g = [u[i] for i in GT]
    det = g[0] * (g[3] * g[5] - g[4] ** 2) - g[1] * (g[1] * g[5] - g[2] * g[4]) + g[2] * (g[1] * g[4] - g[2] * g[3])
    inv = [(g[3] * g[5] - g[4] ** 2) / det, (g[2] * g[4] - g[1] * g[5]) / det, (g[1] * g[4] - g[2] * g[3]) / det,
           (g[0] * g[5] - g[2] ** 2) / det, (g[1] * g[2] - g[0] * g[4]) / det, (g[0] * g[3] - g[1] ** 2) / det]
    w = [1, 2, 2, 1, 2, 1]  # off-diagonal slots count twice in igt^{kl} A_kl
    tr = sum(w[s] * inv[s] * u[AT[s]] for s in range(6))
    for s in range(6):
        u[AT[s]] = u[AT[s]] - g[s] * tr / 3.0


def perturbed_flat_block(seed=1, n=22, amp=1e-3, pmin=(-1.05,) * 3, pmax=(1.05,) * 3):
    """One padded block (24, n^3) float64 (numpy)."""
    xp = _Np()
    sz = (n, n, n)
    x, y, z = block_coords(xp, sz, pmin, pmax)
    L = [pmax[d] - pmin[d] for d in range(3)]
    u = np.zeros((NFIELDS,) + x.shape)
    for f in range(NFIELDS):
        acc = np.zeros_like(x)
        for (kx, ky, kz, ph) in fourier_modes(seed, f):
            acc += np.sin(2.0 * math.pi * (kx * (x - pmin[0]) / L[0] + ky * (y - pmin[1]) / L[1]
                                            + kz * (z - pmin[2]) / L[2]) + ph)
        u[f] = amp * acc
    u[ALPHA] += 1.0
    u[CHI] += 1.0
    for s in (0, 3, 5):
        u[GT[s]] += 1.0
    _normalise_metric_and_trace(u)
    return u.reshape(NFIELDS, -1)


def brill_lindquist_fill(xp, x, y, z, out, masses=(0.5, 0.5), centres=((2.0, 0.0, 0.0), (-2.0, 0.0, 0.0)), rmin=1e-6):
    """Write the Brill-Lindquist fields into out[24] (views of shape like x)."""
    psi = 1.0
    for m, c in zip(masses, centres):
        r = xp.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
        psi = psi + m / (2.0 * xp.maximum(r, rmin))
    p2 = psi * psi
    for f in range(NFIELDS):
        out[f] = 0.0 * x
    out[CHI] = 1.0 / (p2 * p2)
    out[ALPHA] = 1.0 / p2
    for s in (0, 3, 5):
        out[GT[s]] = 0.0 * x + 1.0


def puncture_perturbed_block(seed=5, n=16, amp=1e-2, pmin=(-0.8, -0.75, -0.7), pmax=None, h=0.1):
    """perturbed_flat_block with chi and alpha multiplied by their
    Brill-Lindquist values (punctures at (+-2,0,0), outside the block): every
    field non-trivial, chi well below 1 (about 0.4), so the eta function
    (1 - chi^p0)^-p1 stays moderate."""
    if pmax is None:
        pmax = tuple(p + (n - 1) * h for p in pmin)
    u = perturbed_flat_block(seed=seed, n=n, amp=amp, pmin=pmin, pmax=pmax)
    x, y, z = block_coords(_Np(), (n, n, n), pmin, pmax)
    bl = [None] * NFIELDS
    brill_lindquist_fill(_Np(), x, y, z, bl)
    u[CHI] *= bl[CHI].reshape(-1)
    u[ALPHA] *= bl[ALPHA].reshape(-1)
    return u


@dataclass
class BlockList:
    """POD block table, one row per block (the batch entry point's input)."""
    offsets: list = field(default_factory=list)   # start of block in every field array
    sz: list = field(default_factory=list)        # (nx, ny, nz) padded
    pmin: list = field(default_factory=list)      # coordinates of the first padded point
    pmax: list = field(default_factory=list)      # coordinates of the last padded point
    bflag: list = field(default_factory=list)
    total: int = 0                                # points per field array

    def add(self, sz, pmin, pmax, bflag):
        self.offsets.append(self.total)
        self.sz.append(tuple(int(s) for s in sz))
        self.pmin.append(tuple(float(v) for v in pmin))
        self.pmax.append(tuple(float(v) for v in pmax))
        self.bflag.append(int(bflag))
        self.total += int(sz[0]) * int(sz[1]) * int(sz[2])

    def __len__(self):
        return len(self.offsets)

    def interior_points(self):
        return sum((s[0] - 6) * (s[1] - 6) * (s[2] - 6) for s in self.sz)


def uniform_tiling(nb=8, n=32, lo=-8.0, hi=8.0, stagger=0.5, faces=True):
    """nb^3 blocks of n^3 interior points tiling [lo,hi]^3.

    Interior node i of block b sits at lo + (n b + i + stagger) dx with
    dx = (hi-lo)/(nb n); stagger=0.5 keeps nodes off the punctures at
    (+-2,0,0).  Blocks on the outer boundary get the bflag bits of their outer
    faces (LEFT=0 RIGHT=1 DOWN=2 UP=3 BACK=4 FRONT=5, bssn/include/def.h:13-18)."""
    bl = BlockList()
    dx = (hi - lo) / (nb * n)
    m = n + 2 * GHOST
    for bz in range(nb):
        for by in range(nb):
            for bx in range(nb):
                b = (bx, by, bz)
                pmin = [lo + (n * b[d] - GHOST + stagger) * dx for d in range(3)]
                pmax = [pmin[d] + (m - 1) * dx for d in range(3)]
                flag = 0
                if faces:
                    for d in range(3):
                        if b[d] == 0:
                            flag |= 1 << (2 * d)
                        if b[d] == nb - 1:
                            flag |= 1 << (2 * d + 1)
                bl.add((m, m, m), pmin, pmax, flag)
    return bl


def octree_blocks(domain=16.0, sizes=(16, 24, 32, 48), centres=((2.0, 0.0, 0.0), (-2.0, 0.0, 0.0)), c=1.5,
                  lmin=2, lmax=7, stagger=0.5):
    """Deterministic uneven block set over [-domain, domain]^3 refined toward
    the punctures: an octant of side s is split while it is coarser than
    level lmin, or (up to level lmax) while its distance to the nearest
    puncture is below c*s.  With the defaults: 3032 leaves, finest side 0.25.
    Each leaf's interior size comes from `sizes`, chosen by SplitMix64 of its
    id; dx = side / size; nodes staggered by `stagger` cells.  bflag marks the
    faces on the domain boundary."""
    leaves = []

    def dist(lo, side):
        best = 1e30
        for p in centres:
            d2 = 0.0
            for k in range(3):
                q = min(max(p[k], lo[k]), lo[k] + side)
                d2 += (q - p[k]) ** 2
            best = min(best, math.sqrt(d2))
        return best

    def rec(lvl, lo, side):
        if lvl < lmin or (lvl < lmax and dist(lo, side) < c * side):
            h = side / 2.0
            for cz in range(2):
                for cy in range(2):
                    for cx in range(2):
                        rec(lvl + 1, (lo[0] + cx * h, lo[1] + cy * h, lo[2] + cz * h), h)
        else:
            leaves.append((lvl, lo, side))

    rec(0, (-domain,) * 3, 2.0 * domain)
    bl = BlockList()
    for bid, (lvl, lo, side) in enumerate(leaves):
        g = splitmix64(0xB10C ^ (bid << 8))
        n = sizes[next(g) % len(sizes)]
        dx = side / n
        m = n + 2 * GHOST
        pmin = [lo[d] + (stagger - GHOST) * dx for d in range(3)]
        pmax = [pmin[d] + (m - 1) * dx for d in range(3)]
        flag = 0
        for d in range(3):
            if lo[d] <= -domain + 1e-12:
                flag |= 1 << (2 * d)
            if lo[d] + side >= domain - 1e-12:
                flag |= 1 << (2 * d + 1)
        bl.add((m, m, m), pmin, pmax, flag)
    return bl


def perturbed_flat_fill(xp, x, y, z, out, seed=1, amp=1e-3, lo=-1.0, L=2.0):
    """Flat background plus amp * 8 seeded Fourier modes per field evaluated
    at global coordinates over a cube [lo, lo+L]^3 (consistent across
    blocks), metric renormalised to det 1 and At made trace-free."""
    for f in range(NFIELDS):
        acc = 0.0 * x
        for (kx, ky, kz, ph) in fourier_modes(seed, f):
            acc = acc + xp.sin(2.0 * math.pi * (kx * (x - lo) + ky * (y - lo) + kz * (z - lo)) / L + ph)
        out[f] = amp * acc
    out[ALPHA] = out[ALPHA] + 1.0
    out[CHI] = out[CHI] + 1.0
    for s in (0, 3, 5):
        out[GT[s]] = out[GT[s]] + 1.0
    g = [out[i] for i in GT]
    det = g[0] * (g[3] * g[5] - g[4] ** 2) - g[1] * (g[1] * g[5] - g[2] * g[4]) + g[2] * (g[1] * g[4] - g[2] * g[3])
    sc = 1.0 / xp.cbrt(det)
    for i in GT:
        out[i] = out[i] * sc
    g = [out[i] for i in GT]
    det = g[0] * (g[3] * g[5] - g[4] ** 2) - g[1] * (g[1] * g[5] - g[2] * g[4]) + g[2] * (g[1] * g[4] - g[2] * g[3])
    inv = [(g[3] * g[5] - g[4] ** 2) / det, (g[2] * g[4] - g[1] * g[5]) / det, (g[1] * g[4] - g[2] * g[3]) / det,
           (g[0] * g[5] - g[2] ** 2) / det, (g[1] * g[2] - g[0] * g[4]) / det, (g[0] * g[3] - g[1] ** 2) / det]
    w = [1, 2, 2, 1, 2, 1]
    tr = sum(w[s] * inv[s] * out[AT[s]] for s in range(6))
    for s in range(6):
        out[AT[s]] = out[AT[s]] - g[s] * tr / 3.0


def ccz4_fill(xp, x, y, z, out, seed=1, amp=1e-2, lo=-1.0, L=2.0, pulse=0.1, width=1.0):
    """CCZ4 data (25 fields, Theta last): perturbed_flat_fill for the 24 BSSN
    fields (the Gt slots hold Gamma-hat), Theta = amp * 8 seeded Fourier
    modes (key field 24), and the lapse pulse alpha += pulse exp(-r^2/width^2)."""
    perturbed_flat_fill(xp, x, y, z, out, seed=seed, amp=amp, lo=lo, L=L)
    acc = 0.0 * x
    for (kx, ky, kz, ph) in fourier_modes(seed, 24):
        acc = acc + xp.sin(2.0 * math.pi * (kx * (x - lo) + ky * (y - lo) + kz * (z - lo)) / L + ph)
    out[24] = amp * acc
    out[ALPHA] = out[ALPHA] + pulse * xp.exp(-(x * x + y * y + z * z) / (width * width))


def fill_blocks(bl: BlockList, backend="numpy", device="cpu", kind="bl", **kw):
    """Allocate a (24, total) float64 array and fill every padded block with
    the analytic data (ghosts included): kind "bl" (Brill-Lindquist) or
    "perturbed" (perturbed_flat_fill, keywords seed, amp, lo, L); or a
    (25, total) array for kind "ccz4" (ccz4_fill, same keywords plus pulse,
    width)."""
    xp = _xp(backend, device)
    nfields = 25 if kind == "ccz4" else NFIELDS
    u = xp.zeros((nfields, bl.total))
    for b in range(len(bl)):
        sz = bl.sz[b]
        x, y, z = block_coords(xp, sz, bl.pmin[b], bl.pmax[b])
        n = sz[0] * sz[1] * sz[2]
        vals = [None] * nfields
        if kind == "bl":
            brill_lindquist_fill(xp, x, y, z, vals)
        elif kind == "perturbed":
            perturbed_flat_fill(xp, x, y, z, vals, **kw)
        elif kind == "ccz4":
            ccz4_fill(xp, x, y, z, vals, **kw)
        else:
            raise ValueError(kind)
        o = bl.offsets[b]
        for f in range(nfields):
            u[f, o:o + n] = vals[f].reshape(-1)
    return u
