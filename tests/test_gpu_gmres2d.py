"""SchurMat2D and GMRESSolver2D over caller-supplied domain solvers on the
device (pe/concat/schur_mat2d.h:7-69, pe/concat/gmres_solver2d.h:12-57),
against the oracle's restatement of the same iteration (or_gmres_mock,
gmres_solver2d.cpp:123-252) and against dense algebra of the solvers'
operators."""
import numpy as np
import pytest

import oracle as O
from gpu_util import banner, device, rel, to_dev

pytestmark = pytest.mark.gpu
D, N, P, A = 0, 1, 2, 3
XNEG, XPOS, YNEG, YPOS = 0, 1, 2, 3


@pytest.fixture(scope="module")
def ctx():
    device()
    import focalors_poisson_amd as fpa
    return fpa.Context(0)


def mock_solver(fields):
    """DomainSolver2DTest of case/validation/2d/gmres_reference.cpp:
    f(i, j) *= i * ny + j + 1, in place on the device."""
    import torch
    nrhs, nx, ny = fields.shape
    w = torch.arange(1, nx * ny + 1, dtype=torch.float64, device=fields.device).view(1, nx, ny)
    fields.mul_(w)


def line_cells(loc, nx, ny):
    idx = np.arange(nx * ny).reshape(nx, ny)
    return {XNEG: idx[0, :], XPOS: idx[nx - 1, :], YNEG: idx[:, 0], YPOS: idx[:, ny - 1]}[loc]


@pytest.mark.parametrize("loc", [XNEG, XPOS, YNEG, YPOS])
def test_schur_block_of_mock_branch(ctx, loc):
    """shur_mat_reference.cpp: the Schur block of a mock branch is diagonal
    with the branch's scale on its line next to the root, and operator*
    touches only the root's line."""
    import torch
    import focalors_poisson_amd as fpa
    nx, ny = (4, 3) if loc < 2 else (3, 4)
    bnx, bny = (6, 3) if loc < 2 else (3, 6)
    banner("schur mock loc=%d" % loc)
    S = fpa.SchurMat2D(ctx, loc, bnx, bny, mock_solver)
    bscale = (np.arange(bnx * bny) + 1.0).reshape(bnx, bny)
    bline = {XNEG: bscale[bnx - 1, :], XPOS: bscale[0, :], YNEG: bscale[:, bny - 1], YPOS: bscale[:, 0]}[loc]
    assert np.array_equal(S.values(), np.diag(bline))
    f = to_dev((np.arange(nx * ny) + 1.0).reshape(nx, ny))
    R = S.apply(f).cpu().numpy().ravel()
    want = np.zeros(nx * ny)
    cells = line_cells(loc, nx, ny)
    want[cells] = bline * (cells + 1.0)
    assert np.array_equal(R, want)


@pytest.mark.parametrize("loc", [XNEG, XPOS, YNEG, YPOS])
def test_gmres_mock_matches_oracle_iteration(ctx, loc):
    """gmres_reference.cpp: root 3x5, a neighbour 4 deep at `loc`, the mock
    solver for both, m = 2, tol = 1e-3, maxIter = 2: the device iteration
    gives the oracle's solution and restart residuals."""
    import torch
    import focalors_poisson_amd as fpa
    nx, ny = 3, 5
    bnx, bny = (4, ny) if loc < 2 else (nx, 4)
    banner("gmres mock loc=%d" % loc)
    g = fpa.GMRESSolver2D(ctx, nx, ny, m=2, tol=1e-3, max_iter=2)
    g.set_solver(mock_solver)
    g.schur_mat_construct({loc: (bnx, bny, mock_solver)})
    b = to_dev((np.arange(nx * ny) + 1.0).reshape(nx, ny))
    info = g.solve(b, allow_not_converged=True)
    torch.cuda.synchronize()
    n = nx * ny
    scale = np.arange(n) + 1.0
    Sd = np.zeros((n, n))
    bscale = (np.arange(bnx * bny) + 1.0).reshape(bnx, bny)
    bline = {XNEG: bscale[bnx - 1, :], XPOS: bscale[0, :], YNEG: bscale[:, bny - 1], YPOS: bscale[:, 0]}[loc]
    for a, v in zip(line_cells(loc, nx, ny), bline):
        Sd[a, a] = v
    x, resvec, rc = O.gmres_mock(scale, Sd, scale.copy(), m=2, tol=1e-3, maxiter=2)
    assert rel(b.cpu().numpy().ravel(), x) < 1e-13
    assert len(info["resvec"]) == len(resvec)
    assert np.allclose(info["resvec"], resvec, rtol=1e-12, atol=0)
    assert (info["status"] == 0) == (rc == 0)


def test_gmres_over_poisson_solvers(ctx):
    """The reference's use: A^-1 the root's PoissonSolver2D (Adjacented faces
    toward two branches), the Schur blocks from the branches'
    PoissonSolver2D, on a 12 x 10 root with branches on its x- and y+ faces;
    the converged x solves (I - A^-1 S) x = A^-1 b, with A^-1 and S formed
    densely from the oracle's subdomain solves; an initial guess changes the
    start, not the answer."""
    import torch
    import focalors_poisson_amd as fpa
    h = 1.0 / 16
    nx, ny = 12, 10
    root_bc = (A, D, N, A)
    bl = (7, ny, (D, A, N, D))   # on the root's x- face
    bt = (nx, 6, (N, D, A, D))   # on the root's y+ face
    root = fpa.PoissonSolver2D(ctx, nx, ny, h, h, root_bc)
    sl = fpa.PoissonSolver2D(ctx, bl[0], bl[1], h, h, bl[2])
    st = fpa.PoissonSolver2D(ctx, bt[0], bt[1], h, h, bt[2])
    banner("gmres over poisson solvers")
    g = fpa.GMRESSolver2D(ctx, nx, ny, m=30, tol=1e-13, max_iter=20)
    g.set_solver(root)
    g.schur_mat_construct({XNEG: (bl[0], bl[1], sl), YPOS: (bt[0], bt[1], st)})
    rng = np.random.default_rng(8)
    b0 = rng.standard_normal((nx, ny)) * h * h
    b = to_dev(b0)
    info = g.solve(b)
    torch.cuda.synchronize()
    x = b.cpu().numpy().ravel()
    n = nx * ny

    def dense_inverse(rows, cols, bc):
        M = np.zeros((rows * cols, rows * cols))
        for q in range(rows * cols):
            e = np.zeros(rows * cols)
            e[q] = 1.0
            M[:, q] = O.sub_solve(e.reshape(rows, cols), h, h, bc).ravel()
        return M
    Ainv = dense_inverse(nx, ny, root_bc)
    S = np.zeros((n, n))
    for loc, (bnx, bny, bc) in ((XNEG, bl), (YPOS, bt)):
        Binv = dense_inverse(bnx, bny, bc)
        bcells = line_cells(loc ^ 1, bnx, bny)
        rcells = line_cells(loc, nx, ny)
        blk = Binv[np.ix_(bcells, bcells)]  # value(j, i) = t_a(line j) of the unit at line i
        S[np.ix_(rcells, rcells)] += blk
    lhs = x - Ainv @ (S @ x)
    assert rel(lhs, Ainv @ b0.ravel()) < 1e-11, info
    # the device blocks are the dense ones
    Sx = fpa.SchurMat2D(ctx, XNEG, bl[0], bl[1], sl)
    Bx = dense_inverse(bl[0], bl[1], bl[2])
    cells = line_cells(XPOS, bl[0], bl[1])
    assert rel(Sx.values(), Bx[np.ix_(cells, cells)]) < 1e-13
    # from an initial guess
    x0 = to_dev(rng.standard_normal((nx, ny)))
    g.set_initial_guess(x0)
    b2 = to_dev(b0)
    g.solve(b2)
    torch.cuda.synchronize()
    assert rel(b2.cpu().numpy().ravel(), x) < 1e-11


def test_failing_callback_is_an_error(ctx):
    import focalors_poisson_amd as fpa

    def bad(fields):
        raise RuntimeError("solver failed")
    g = fpa.GMRESSolver2D(ctx, 3, 5, m=2, tol=1e-3, max_iter=2)
    g.set_solver(bad)
    with pytest.raises(fpa.FPError):
        g.solve(to_dev(np.ones((3, 5))))
    with pytest.raises(fpa.FPError):
        fpa.SchurMat2D(ctx, XNEG, 4, 5, bad)
