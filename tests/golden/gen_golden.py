"""Writes the golden fixtures under tests/golden/ from the assembled sparse
system (tests/discrete.py, scipy spsolve).  No oracle code and no reference code
is involved: the inputs are the ones the reference's validation programs
specify, the outputs come from a direct solve of the discrete problem.

  poisson_3x4.npz   case/validation/2d/poisson_solver_reference.cpp:3-95
                    (nx=3, ny=4, hx=1, hy=2 as ints, f(i,j) = i*ny+j+1), the 15
                    D/N combinations with a Dirichlet face.  The all-Neumann
                    case has an incompatible RHS there and no solution.
  concat_six.npz    case/validation/2d/concat_solver_reference.cpp:42-108
                    (6 domains, sizes 8x7/7/8/8/7/8, all Dirichlet,
                    f(i,j) = i*ny+j+1, h = 1).
  cross_32.npz      a 32-cell cross with the mixed boundary table and ghost data,
                    seeded random RHS.

Run: python tests/golden/gen_golden.py
"""
import itertools
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import discrete  # noqa: E402

D, N, A = 0, 1, 3


def fill(nx, ny):
    return (np.arange(nx * ny, dtype=np.float64) + 1.0).reshape(nx, ny)


def poisson_3x4():
    nx, ny, hx, hy = 3, 4, 1.0, 2.0
    combos, outs = [], []
    for bc in itertools.product([D, N], repeat=4):
        if all(b == N for b in bc):
            continue
        u = discrete.solve([(nx, ny, hx, hy, bc)], [], [fill(nx, ny) / (hx * hx)])[0]
        combos.append(bc)
        outs.append(u)
    np.savez(os.path.join(HERE, "poisson_3x4.npz"), bc=np.array(combos, dtype=np.int32),
             f=fill(nx, ny), u=np.array(outs), nx=nx, ny=ny, hx=hx, hy=hy)


def six_domains():
    nx1, ny1, nx2, nx3, ny4, ny5, nx6 = 8, 7, 7, 8, 8, 7, 8
    sizes = [(nx1, ny1), (nx2, ny1), (nx3, ny1), (nx1, ny4), (nx1, ny5), (nx6, ny1)]
    doms = [(a, b, 1.0, 1.0, (D, D, D, D)) for a, b in sizes]
    conn = [(0, 0, 1), (0, 1, 2), (0, 2, 3), (0, 3, 4), (2, 1, 5)]
    fs = [fill(a, b) for a, b in sizes]
    us = discrete.solve(doms, conn, fs)
    out = {"sizes": np.array(sizes, dtype=np.int32), "conn": np.array(conn, dtype=np.int32)}
    for k, (f, u) in enumerate(zip(fs, us)):
        out["f%d" % k] = f
        out["u%d" % k] = u
    np.savez(os.path.join(HERE, "concat_six.npz"), **out)


def cross_32():
    n = 32
    h = 1.0 / n
    bcs = [(A, A, A, A), (D, A, N, N), (A, N, D, D), (D, D, N, A), (N, D, A, N)]
    sizes = [(n, n), (n // 2, n), (n // 2, n), (n, n // 2), (n, n // 2)]
    doms = [(a, b, h, h, bc) for (a, b), bc in zip(sizes, bcs)]
    conn = [(0, 0, 1), (0, 1, 2), (0, 2, 3), (0, 3, 4)]
    rng = np.random.default_rng(32)
    fs = [rng.standard_normal(s) for s in sizes]
    bufs = {(1, 0): rng.standard_normal(n), (2, 2): rng.standard_normal(n // 2),
            (4, 1): rng.standard_normal(n // 2), (3, 2): rng.standard_normal(n),
            (1, 2): rng.standard_normal(n // 2), (4, 0): rng.standard_normal(n // 2)}
    us = discrete.solve(doms, conn, fs, bufs)
    out = {"sizes": np.array(sizes, dtype=np.int32), "bcs": np.array(bcs, dtype=np.int32),
           "conn": np.array(conn, dtype=np.int32), "h": h}
    for k, (f, u) in enumerate(zip(fs, us)):
        out["f%d" % k] = f
        out["u%d" % k] = u
    for (d, face), g in bufs.items():
        out["g%d_%d" % (d, face)] = g
    np.savez(os.path.join(HERE, "cross_32.npz"), **out)


if __name__ == "__main__":
    poisson_3x4()
    six_domains()
    cross_32()
    print("written")
