"""The exact discrete problem of one 3-D subdomain (PoissonSolver3D,
pe/poisson/poisson_solver3d.cpp) assembled as one sparse matrix and solved
directly (scipy); independent of the transforms and sweeps it pins.

Scaled 7-point equation (hx^2 scaling, ry = hx^2/hy^2, rz = hx^2/hz^2):
  u(i-1) - 2u + u(i+1) + ry (u(j-1) - 2u + u(j+1)) + rz (u(k-1) - 2u + u(k+1)) = hx^2 F
Ghosts per face: Dirichlet / Adjacented 0, Neumann = the adjacent cell,
Periodic = the opposite face's cell.  Faces: x-, x+, y-, y+, z-, z+.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

D, N, P, A = 0, 1, 2, 3


def axis_operator(n, neg, pos):
    M = sp.lil_matrix((n, n))
    for i in range(n):
        M[i, i] = -2.0
        if i > 0:
            M[i, i - 1] = 1.0
        if i < n - 1:
            M[i, i + 1] = 1.0
    if neg == N:
        M[0, 0] += 1.0
    if pos == N:
        M[n - 1, n - 1] += 1.0
    if neg == P and pos == P:
        M[0, n - 1] += 1.0
        M[n - 1, 0] += 1.0
    return M.tocsr()


def assemble(nx, ny, nz, hx, hy, hz, bc):
    Ix, Iy, Iz = (sp.identity(n, format="csr") for n in (nx, ny, nz))
    Ax = axis_operator(nx, bc[0], bc[1])
    Ay = axis_operator(ny, bc[2], bc[3])
    Az = axis_operator(nz, bc[4], bc[5])
    ry, rz = hx * hx / (hy * hy), hx * hx / (hz * hz)
    return (sp.kron(sp.kron(Ax, Iy), Iz) + ry * sp.kron(sp.kron(Ix, Ay), Iz) + rz * sp.kron(sp.kron(Ix, Iy), Az)).tocsc()


def solve(f, hx, hy, hz, bc):
    """The solution of the already hx^2-scaled right-hand side f (nx, ny, nz)."""
    nx, ny, nz = f.shape
    return spla.spsolve(assemble(nx, ny, nz, hx, hy, hz, bc), f.ravel()).reshape(f.shape)


def refined(f, hx, hy, hz, bc, solve_fn, steps=2):
    """solve_fn's solution refined with residuals summed in long double (the
    discrete solution to about FP64 rounding)."""
    M = assemble(*f.shape, hx, hy, hz, bc).tocoo()
    vl = M.data.astype(np.longdouble)
    u = solve_fn(f)
    for _ in range(steps):
        y = np.zeros(f.size, dtype=np.longdouble)
        np.add.at(y, M.row, vl * u.ravel().astype(np.longdouble)[M.col])
        r = (f.ravel().astype(np.longdouble) - y).astype(np.float64).reshape(f.shape)
        u = u + solve_fn(r)
    return u


def face_index(loc, nx, ny, nz):
    """Flat indices of face `loc` (x-, x+, y-, y+, z-, z+) of a box, in the
    order of the ghost-data planes: x faces (j, k), y faces (i, k), z faces
    (i, j), last index fastest."""
    idx = np.arange(nx * ny * nz).reshape(nx, ny, nz)
    sl = {0: idx[0, :, :], 1: idx[nx - 1, :, :], 2: idx[:, 0, :], 3: idx[:, ny - 1, :], 4: idx[:, :, 0],
          5: idx[:, :, nz - 1]}[loc]
    return sl.ravel()


def effective_bc(boxes, conn):
    bcs = [list(b[6]) for b in boxes]
    for a, dr, b in conn:
        bcs[a][dr] = A
        bcs[b][dr ^ 1] = A
    return bcs


def assemble_composite(boxes, conn, fields, ghost=None):
    """boxes: (nx, ny, nz, hx, hy, hz, bc6); conn: (a, dir, b); fields: raw f.
    Returns (matrix, rhs, offsets) of the hx^2-scaled global 7-point system
    with the faces named in conn coupled with coefficient 1
    (concat_solver3d.cpp:97-157, 159-247)."""
    bcs = effective_bc(boxes, conn)
    offs = np.cumsum([0] + [b[0] * b[1] * b[2] for b in boxes])
    blocks, rhs = [], np.zeros(offs[-1])
    ghost = ghost or {}
    for d, (nx, ny, nz, hx, hy, hz, _) in enumerate(boxes):
        blocks.append(assemble(nx, ny, nz, hx, hy, hz, bcs[d]))
        f = np.asarray(fields[d], dtype=np.float64).ravel().copy()
        h = (hx, hy, hz)
        for face in range(6):
            g = ghost.get((d, face))
            if g is None or bcs[d][face] not in (D, N):
                continue
            hh = h[face // 2]
            a = -1.0 / hh ** 2 if bcs[d][face] == D else (-1.0 / hh if face & 1 else 1.0 / hh)
            f[face_index(face, nx, ny, nz)] += a * np.asarray(g).ravel()
        rhs[offs[d]:offs[d + 1]] = hx * hx * f
    M = sp.block_diag(blocks, format="lil")
    for a, dr, b in conn:
        na, nb = boxes[a][:3], boxes[b][:3]
        ia = offs[a] + face_index(dr, *na)
        ib = offs[b] + face_index(dr ^ 1, *nb)
        assert len(ia) == len(ib)
        for p, q in zip(ia, ib):
            M[p, q] += 1.0
            M[q, p] += 1.0
    return M.tocsc(), rhs, offs


def solve_composite(boxes, conn, fields, ghost=None):
    M, rhs, offs = assemble_composite(boxes, conn, fields, ghost)
    u = spla.spsolve(M, rhs)
    return [u[offs[d]:offs[d + 1]].reshape(boxes[d][:3]) for d in range(len(boxes))]
