"""The exact discrete problem the solve path defines, assembled as one sparse
matrix and solved directly (scipy).  Independent of the oracle: it never uses
transforms, Thomas sweeps, Schur blocks or GMRES, so it pins all of them.

Scaled 5-point equation per subdomain (hx^2 scaling, r = hx^2/hy^2):
  u(i-1,j) - 2u(i,j) + u(i+1,j) + r (u(i,j-1) - 2u(i,j) + u(i,j+1)) = hx^2 F(i,j)
Ghosts (reference: chasing_method_base.cpp:33-46, poisson_solver2d_base.cpp:28-62,
concat_solver2d.cpp:110-223):
  Dirichlet / Adjacented face: ghost 0 in the operator; Dirichlet data g moves to
    the RHS as -g/h^2 (h = hx on x faces, hy on y faces).
  Neumann face: ghost = adjacent cell; data g adds +g/h on XNeg/YNeg, -g/h on
    XPos/YPos.
  Adjacented face: the neighbour's adjacent cell enters with coefficient exactly 1.
  Periodic pair of faces: the opposite line's cell is the neighbour.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

D, N, P, A = 0, 1, 2, 3
XNEG, XPOS, YNEG, YPOS = 0, 1, 2, 3


def effective_bc(domains, conn):
    bcs = [list(d[4]) for d in domains]
    for a, dr, b in conn:
        bcs[a][dr] = A
        bcs[b][dr ^ 1] = A
    return bcs


def assemble(domains, conn, fields, buffers=None):
    """domains: (nx, ny, hx, hy, bc4); conn: (a, dir, b); fields: raw RHS f.
    Returns (matrix, rhs, offsets) for the hx^2-scaled global system."""
    bcs = effective_bc(domains, conn)
    offs = np.cumsum([0] + [d[0] * d[1] for d in domains])
    n = offs[-1]
    rows, cols, vals = [], [], []
    rhs = np.zeros(n)
    buffers = buffers or {}
    for d, (nx, ny, hx, hy, _) in enumerate(domains):
        r = hx * hx / (hy * hy)
        bc = bcs[d]
        f = np.asarray(fields[d], dtype=np.float64).reshape(nx, ny).copy()
        for face in range(4):
            g = buffers.get((d, face))
            if g is None:
                continue
            g = np.asarray(g, dtype=np.float64)
            if bc[face] == D:
                a = -1.0 / (hx * hx) if face < 2 else -1.0 / (hy * hy)
            elif bc[face] == N:
                h = hx if face < 2 else hy
                a = 1.0 / h if face in (XNEG, YNEG) else -1.0 / h
            else:
                continue
            if face == XNEG:
                f[0, :] += a * g
            elif face == XPOS:
                f[nx - 1, :] += a * g
            elif face == YNEG:
                f[:, 0] += a * g
            else:
                f[:, ny - 1] += a * g
        rhs[offs[d]:offs[d + 1]] = (hx * hx * f).ravel()
        idx = offs[d] + np.arange(nx * ny).reshape(nx, ny)
        diag = np.full((nx, ny), -2.0 - 2.0 * r)
        if bc[XNEG] == N:
            diag[0, :] += 1.0
        if bc[XPOS] == N:
            diag[nx - 1, :] += 1.0
        if bc[YNEG] == N:
            diag[:, 0] += r
        if bc[YPOS] == N:
            diag[:, ny - 1] += r
        rows.append(idx.ravel()); cols.append(idx.ravel()); vals.append(diag.ravel())
        rows.append(idx[1:, :].ravel()); cols.append(idx[:-1, :].ravel()); vals.append(np.ones((nx - 1) * ny))
        rows.append(idx[:-1, :].ravel()); cols.append(idx[1:, :].ravel()); vals.append(np.ones((nx - 1) * ny))
        rows.append(idx[:, 1:].ravel()); cols.append(idx[:, :-1].ravel()); vals.append(np.full(nx * (ny - 1), r))
        rows.append(idx[:, :-1].ravel()); cols.append(idx[:, 1:].ravel()); vals.append(np.full(nx * (ny - 1), r))
        # periodic axes: the first and last lines are neighbours
        if bc[XNEG] == P and bc[XPOS] == P:
            rows.append(idx[0, :]); cols.append(idx[nx - 1, :]); vals.append(np.ones(ny))
            rows.append(idx[nx - 1, :]); cols.append(idx[0, :]); vals.append(np.ones(ny))
        if bc[YNEG] == P and bc[YPOS] == P:
            rows.append(idx[:, 0]); cols.append(idx[:, ny - 1]); vals.append(np.full(nx, r))
            rows.append(idx[:, ny - 1]); cols.append(idx[:, 0]); vals.append(np.full(nx, r))
    for a, dr, b in conn:
        nxa, nya = domains[a][0], domains[a][1]
        nxb, nyb = domains[b][0], domains[b][1]
        ia = offs[a] + np.arange(nxa * nya).reshape(nxa, nya)
        ib = offs[b] + np.arange(nxb * nyb).reshape(nxb, nyb)
        if dr == XPOS:
            la, lb = ia[nxa - 1, :], ib[0, :]
        elif dr == XNEG:
            la, lb = ia[0, :], ib[nxb - 1, :]
        elif dr == YPOS:
            la, lb = ia[:, nya - 1], ib[:, 0]
        else:
            la, lb = ia[:, 0], ib[:, nyb - 1]
        assert len(la) == len(lb)
        rows += [la, lb]; cols += [lb, la]; vals += [np.ones(len(la)), np.ones(len(la))]
    M = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    return M, rhs, offs


def solve(domains, conn, fields, buffers=None):
    M, rhs, offs = assemble(domains, conn, fields, buffers)
    u = spla.spsolve(M.tocsc(), rhs)
    return [u[offs[d]:offs[d + 1]].reshape(domains[d][0], domains[d][1]) for d in range(len(domains))]
