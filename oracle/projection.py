"""The pressure projection of an incompressible flow step on a staggered grid,
restated in numpy: the checker of fp_proj_* (include/focalors_poisson.h).

TEST INFRASTRUCTURE ONLY.  The caller this serves (the Focalors Navier–Stokes
solver, README_en.md:26) is not in the reference tree, so nothing here
restates reference code: the discrete operators are the textbook MAC ones,
and tests/test_oracle.py pins D G against the assembled composite 5-point
system of tests/discrete.py.

Per domain d (nx x ny cells): u[d] the x-face velocities ((nx+1) x ny, row i
is the face between cells i-1 and i), v[d] the y-face velocities
(nx x (ny+1)), p[d] the cell pressures.  Faces: 0 XNeg, 1 XPos, 2 YNeg,
3 YPos; boundary codes D 0, N 1, P 2, A 3.
"""
import numpy as np

import direct

D, N, P, A = 0, 1, 2, 3


def divergence(domains, u, v, scale=1.0):
    """scale * ((u[i+1] - u[i]) / hx + (v[:, j+1] - v[:, j]) / hy)."""
    out = []
    for (nx, ny, hx, hy, _), a, b in zip(domains, u, v):
        a = np.asarray(a).reshape(nx + 1, ny)
        b = np.asarray(b).reshape(nx, ny + 1)
        out.append(scale * ((a[1:] - a[:-1]) / hx + (b[:, 1:] - b[:, :-1]) / hy))
    return out


def _beyond(domains, conn, p, pghost, d, face):
    """The pressure beyond a face: None (Neumann), else a line (Dirichlet
    ghost values, the neighbour's adjacent cells, the opposite periodic line,
    zeros for an Adjacented face without a neighbour)."""
    bcs = direct.effective_bc(domains, conn)
    nx, ny = domains[d][0], domains[d][1]
    t = bcs[d][face]
    pd = np.asarray(p[d]).reshape(nx, ny)
    if t == D:
        g = (pghost or {}).get((d, face))
        return np.zeros(ny if face < 2 else nx) if g is None else np.asarray(g, dtype=np.float64)
    if t == P:
        return {0: pd[-1, :], 1: pd[0, :], 2: pd[:, -1], 3: pd[:, 0]}[face]
    if t == A:
        for a, dr, b in conn:
            if (a == d and dr == face) or (b == d and (dr ^ 1) == face):
                e = b if a == d else a
                pe = np.asarray(p[e]).reshape(domains[e][0], domains[e][1])
                return {0: pe[-1, :], 1: pe[0, :], 2: pe[:, -1], 3: pe[:, 0]}[face]
        # an Adjacented face without a neighbour: ghost 0, as the operator has it
        return np.zeros(ny if face < 2 else nx)
    return None


def correct(domains, conn, u, v, p, pghost=None, scale=1.0):
    """u -= scale Gx p, v -= scale Gy p; boundary faces with nothing beyond
    them (Neumann) unchanged.  Returns new arrays."""
    uo, vo = [], []
    for d, (nx, ny, hx, hy, _) in enumerate(domains):
        pd = np.asarray(p[d]).reshape(nx, ny)
        a = np.array(u[d], dtype=np.float64).reshape(nx + 1, ny)
        b = np.array(v[d], dtype=np.float64).reshape(nx, ny + 1)
        a[1:-1] -= scale * (pd[1:] - pd[:-1]) / hx
        b[:, 1:-1] -= scale * (pd[:, 1:] - pd[:, :-1]) / hy
        for face in range(4):
            out = _beyond(domains, conn, p, pghost, d, face)
            if out is None:
                continue
            if face == 0:
                a[0] -= scale * (pd[0] - out) / hx
            elif face == 1:
                a[-1] -= scale * (out - pd[-1]) / hx
            elif face == 2:
                b[:, 0] -= scale * (pd[:, 0] - out) / hy
            else:
                b[:, -1] -= scale * (out - pd[:, -1]) / hy
        uo.append(a)
        vo.append(b)
    return uo, vo


def project(domains, conn, u, v, dt, pghost=None, solve=None):
    """p solving D G p = D u / dt (by `solve(domains, conn, rhs, ghost)`,
    default the converged direct composite solve, with the Dirichlet ghost
    values as boundary data) and the corrected velocity."""
    rhs = divergence(domains, u, v, 1.0 / dt)
    p = (solve or direct.direct_solution)(domains, conn, rhs, pghost)
    uo, vo = correct(domains, conn, u, v, p, pghost, dt)
    return p, uo, vo
