"""Synthetic workloads of the solve path (host numpy builders).

Each builder returns the geometry in the form ConcatPoissonSolver2D takes
(domains, conn), the cell-centre right-hand sides and the ghost data, sampled
the way Variable2D::set_value / set_buffer_value sample (cell centres at
offset + (i + 1/2) h, ghost centres at i = -1 / nx; variable2d.cpp:345-511).
"""
import math

import numpy as np

D, N, P, A = 0, 1, 2, 3
XNEG, XPOS, YNEG, YPOS = 0, 1, 2, 3


def centres(nx, ny, hx, hy, ox, oy):
    x = ox + (np.arange(nx) + 0.5) * hx
    y = oy + (np.arange(ny) + 0.5) * hy
    return np.meshgrid(x, y, indexing="ij")


def single(n=256, seed=0):
    """One n x n Dirichlet domain on [0,1]^2, f = -5 pi^2 sin(pi x) sin(2 pi y)
    + 1e-3 U(-1, 1), zero ghost data."""
    h = 1.0 / n
    X, Y = centres(n, n, h, h, 0.0, 0.0)
    rng = np.random.default_rng(seed)
    f = -5.0 * math.pi ** 2 * np.sin(math.pi * X) * np.sin(2 * math.pi * Y) + 1e-3 * rng.uniform(-1, 1, (n, n))
    return [(n, n, h, h, (D, D, D, D))], [], [f], {}


# boundary table of the cross: centre fully adjacent, arms mixing Dirichlet
# and Neumann so that all four y transforms and Neumann x ends occur
CROSS_BC = {
    "C": (A, A, A, A),
    "L": (D, A, N, N),
    "R": (A, N, D, D),
    "B": (D, D, N, A),
    "T": (N, D, A, N),
}


def cross_geometry(n=1024, arm=None):
    """Centre n x n on [0,1]^2 and four arms of depth `arm` (default n/2),
    h = 1/n; domain order C, L, R, B, T."""
    arm = n // 2 if arm is None else arm
    h = 1.0 / n
    doms = [
        (n, n, h, h, CROSS_BC["C"]),
        (arm, n, h, h, CROSS_BC["L"]),
        (arm, n, h, h, CROSS_BC["R"]),
        (n, arm, h, h, CROSS_BC["B"]),
        (n, arm, h, h, CROSS_BC["T"]),
    ]
    conn = [(0, XNEG, 1), (0, XPOS, 2), (0, YNEG, 3), (0, YPOS, 4)]
    la = arm * h
    offsets = [(0.0, 0.0), (-la, 0.0), (1.0, 0.0), (0.0, -la), (0.0, 1.0)]
    return doms, conn, offsets


def cross(n=1024, seed=1, arm=None, nrhs=1):
    """Cross geometry, f = sin(2 pi x) cos(pi y) + 1e-3 N(0,1) per right-hand
    side (seed + r), zero ghost data.  Fields have shape (nrhs, nx, ny)."""
    doms, conn, offs = cross_geometry(n, arm)
    fields = []
    for (nx, ny, hx, hy, _), (ox, oy) in zip(doms, offs):
        X, Y = centres(nx, ny, hx, hy, ox, oy)
        base = np.sin(2 * math.pi * X) * np.cos(math.pi * Y)
        fields.append(np.stack([base + 1e-3 * np.random.default_rng(seed + r).standard_normal((nx, ny))
                                for r in range(nrhs)]))
    return doms, conn, fields, {}


def lshape(n=2048, seed=4):
    """L-shaped geometry of three n x n Dirichlet domains, h = 1/n: domain 0
    on [0,1]^2, domain 1 to its right, domain 2 above it; right-hand side the
    multiscale sum of 32 seeded Fourier modes on the whole L."""
    h = 1.0 / n
    doms = [(n, n, h, h, (D, A, D, A)), (n, n, h, h, (A, D, D, D)), (n, n, h, h, (D, D, A, D))]
    conn = [(0, XPOS, 1), (0, YPOS, 2)]
    offs = [(0.0, 0.0), (1.0, 0.0), (0.0, 1.0)]
    fields = [multiscale_rhs(n, n, h, h, ox, oy, seed=seed) for ox, oy in offs]
    return doms, conn, fields, {}


def batched_sizes(count=512, seed=2, lo=64, hi=1024):
    """Side lengths log-uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    sz = np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), (count, 2)))).astype(int)
    return [(int(a), int(b)) for a, b in sz]


def multiscale_rhs(nx, ny, hx, hy, ox, oy, seed=4, modes=32):
    """Sum of `modes` seeded Fourier modes with amplitude N(0,1)/|k|^2."""
    rng = np.random.default_rng(seed)
    X, Y = centres(nx, ny, hx, hy, ox, oy)
    f = np.zeros((nx, ny))
    for _ in range(modes):
        kx, ky = rng.integers(1, 65, 2)
        a = rng.standard_normal() / float(kx * kx + ky * ky)
        px, py = rng.uniform(0, 2 * math.pi, 2)
        f += a * np.sin(kx * math.pi * X + px) * np.sin(ky * math.pi * Y + py)
    return f


def smooth_target(doms, offsets, seed=6, nrhs=1):
    """A seeded smooth field per domain (the adjoint's u*): per right-hand
    side sin(a pi x + p) cos(b pi y + q) with a, b in [1, 3] and phases in
    [0, 2 pi).  Arrays of shape (nrhs, nx, ny)."""
    rng = np.random.default_rng(seed)
    prm = [(rng.uniform(1, 3), rng.uniform(1, 3), rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi))
           for _ in range(nrhs)]
    out = []
    for (nx, ny, hx, hy, _), (ox, oy) in zip(doms, offsets):
        X, Y = centres(nx, ny, hx, hy, ox, oy)
        out.append(np.stack([np.sin(a * math.pi * X + p) * np.cos(b * math.pi * Y + q) for a, b, p, q in prm]))
    return out
