"""The 3-D composite solve on the device (ConcatPoissonSolver3D::solve,
pe/concat/concat_solver3d.cpp:97-247): two boxes, an L of three, a
depth-three chain of unequal boxes and a centre with four arms, mixed
Dirichlet / Neumann outer faces with ghost data, several right-hand sides;
against the assembled 7-point composite system (sparse direct) and the CPU
oracle's restatement.  Interface GMRES to relative 1e-13; bar relative L2
<= 1e-12."""
import numpy as np
import pytest

import oracle as O
from gpu_util import banner, device, rel, to_dev
from test_oracle import composite3d_cases

pytestmark = pytest.mark.gpu
D, N, P, A = 0, 1, 2, 3


@pytest.fixture(scope="module")
def ctx():
    device()
    import focalors_poisson_amd as fpa
    return fpa.Context(0)


def ghost_data(boxes, faces, rng, nrhs=1):
    out = {}
    for d, face in faces:
        nx, ny, nz = boxes[d][:3]
        n = (ny * nz, ny * nz, nx * nz, nx * nz, nx * ny, nx * ny)[face]
        out[(d, face)] = rng.standard_normal((nrhs, n))
    return out


def gpu_cat3d(ctx, boxes, conn, fs, ghost, nrhs=1):
    import torch
    import focalors_poisson_amd as fpa
    cat = fpa.ConcatPoissonSolver3D(ctx, boxes, conn, m=60, tol=1e-13, tol_is_relative=True, max_iter=100)
    ts = [to_dev(f) for f in fs]
    info = cat.solve(ts, {k: to_dev(np.ascontiguousarray(v).ravel()) for k, v in ghost.items()})
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in ts], info, cat


@pytest.mark.parametrize("case", [0, 1, 2, 3])
def test_cat3d_small_vs_sparse_and_oracle(ctx, case):
    import discrete3d
    boxes, conn, gfaces = composite3d_cases()[case]
    rng = np.random.default_rng(case + 60)
    fs = [rng.standard_normal(b[:3]) for b in boxes]
    ghost = ghost_data(boxes, gfaces, rng)
    g1 = {k: v[0] for k, v in ghost.items()}
    banner("3-D composite case %d" % case)
    us, info, cat = gpu_cat3d(ctx, boxes, conn, fs, ghost)
    assert rel(us, discrete3d.solve_composite(boxes, conn, fs, g1)) < 1e-12, info
    ocat = O.Composite3D(boxes, conn, m=40, tol=1e-14, tol_is_relative=True, maxiter=200)
    assert cat.root == ocat.root
    assert rel(us, ocat.solve(fs, g1)) < 1e-12, info


def test_cat3d_multi_rhs(ctx):
    import torch
    import discrete3d
    boxes, conn, gfaces = composite3d_cases()[3]
    rng = np.random.default_rng(70)
    nrhs = 3
    fs = [rng.standard_normal((nrhs,) + b[:3]) for b in boxes]
    ghost = ghost_data(boxes, gfaces, rng, nrhs)
    banner("3-D composite three right-hand sides")
    us, info, _ = gpu_cat3d(ctx, boxes, conn, fs, ghost)
    for r in range(nrhs):
        ref = discrete3d.solve_composite(boxes, conn, [f[r] for f in fs], {k: v[r] for k, v in ghost.items()})
        assert rel([u[r] for u in us], ref) < 1e-12, r


def test_cat3d_two_boxes_32(ctx):
    """Two 32^3 boxes side by side (1024 face unknowns), manufactured smooth
    right-hand side: against the sparse system."""
    import discrete3d
    h = 1.0 / 32
    boxes = [(32, 32, 32, h, h, h, (D, A, D, N, D, D)), (32, 32, 32, h, h, h, (A, N, D, D, N, D))]
    conn = [(0, 1, 1)]
    x = (np.arange(32) + 0.5) * h
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    fs = [np.sin(np.pi * X) * np.cos(2 * np.pi * Y) * Z, np.cos(np.pi * X) * Y * Z]
    banner("3-D composite two 32^3 boxes")
    us, info, _ = gpu_cat3d(ctx, boxes, conn, fs, {})
    err = rel(us, discrete3d.solve_composite(boxes, conn, fs))
    banner("3-D composite two 32^3 boxes rel-L2 %.3e, GMRES %d" % (err, info["iterations"]))
    assert err < 1e-12, info


def test_cat3d_singular_unsupported(ctx):
    import focalors_poisson_amd as fpa
    h = 0.25
    boxes = [(4, 4, 4, h, h, h, (N, A, N, N, N, N)), (4, 4, 4, h, h, h, (A, N, N, N, N, N))]
    with pytest.raises(fpa.FPError) as e:
        fpa.ConcatPoissonSolver3D(ctx, boxes, [(0, 1, 1)])
    assert e.value.status == 6
