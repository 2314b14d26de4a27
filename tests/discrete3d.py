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
