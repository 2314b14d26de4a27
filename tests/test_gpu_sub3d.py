"""The 3-D subdomain solver on the device (PoissonSolver3D::solve,
pe/poisson/poisson_solver3d.cpp:87-123): every boundary family of the
transforms in y and z and of the x sweep, singular gauges, several
right-hand sides, against the CPU oracle (or_sub_solve3d, pinned by the
assembled 7-point system) and, at 128^3, against the discrete solution
refined with long-double residuals.  Bar: relative L2 <= 1e-12."""
import numpy as np
import pytest

import oracle as O
from gpu_util import banner, device, rel, to_dev

pytestmark = pytest.mark.gpu
D, N, P, A = 0, 1, 2, 3
BC3 = [(D, D, D, D, D, D), (N, D, A, N, D, N), (D, A, N, N, N, D), (A, N, D, N, N, N), (P, P, D, N, N, D),
       (D, N, P, P, D, D), (N, D, D, A, P, P), (P, P, P, P, D, N)]


@pytest.fixture(scope="module")
def ctx():
    device()
    import focalors_poisson_amd as fpa
    return fpa.Context(0)


def gpu3d(ctx, f, hx, hy, hz, bc):
    import torch
    import focalors_poisson_amd as fpa
    s = fpa.PoissonSolver3D(ctx, *f.shape[-3:], hx, hy, hz, bc)
    t = to_dev(f)
    s.solve(t)
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("bc", BC3)
@pytest.mark.parametrize("shape", [(5, 4, 6), (16, 18, 10), (33, 20, 18), (64, 48, 40)])
def test_sub3d_vs_oracle(ctx, bc, shape):
    nx, ny, nz = shape
    rng = np.random.default_rng(nx * 7 + ny * 3 + nz + sum(bc))
    hx, hy, hz = 1.0 / nx, 1.2 / ny, 0.9 / nz
    f = rng.standard_normal(shape) * hx * hx
    banner("3-D nx=%d ny=%d nz=%d bc=%s" % (nx, ny, nz, bc))
    assert rel(gpu3d(ctx, f, hx, hy, hz, bc), O.sub_solve3d(f, hx, hy, hz, bc)) < 1e-12


@pytest.mark.parametrize("bc", [(N, N, N, N, N, N), (P, P, N, N, P, P), (N, N, P, P, N, N)])
def test_sub3d_singular(ctx, bc):
    rng = np.random.default_rng(5)
    f = rng.standard_normal((24, 16, 20))
    f -= f.mean()
    f *= 1.0 / 24 ** 2
    h = 1.0 / 24
    banner("3-D singular bc=%s" % (bc,))
    assert rel(gpu3d(ctx, f, h, h, h, bc), O.sub_solve3d(f, h, h, h, bc)) < 1e-12


def test_sub3d_multi_rhs(ctx):
    import torch
    import focalors_poisson_amd as fpa
    rng = np.random.default_rng(6)
    shape, bc = (20, 12, 14), (D, N, A, N, N, D)
    h = 1.0 / 20
    f = rng.standard_normal((3,) + shape) * h * h
    s = fpa.PoissonSolver3D(ctx, *shape, h, h, h, bc)
    t = to_dev(f)
    banner("3-D three right-hand sides")
    s.solve(t)
    torch.cuda.synchronize()
    got = t.cpu().numpy()
    for r in range(3):
        assert rel(got[r], O.sub_solve3d(f[r], h, h, h, bc)) < 1e-12, r


@pytest.mark.parametrize("bc", [(D, D, D, D, D, D), (D, N, N, D, N, N)])
def test_sub3d_128_vs_discrete(ctx, bc):
    """128^3 (2.1 M cells) against the discrete solution: the oracle's solve
    refined twice with long-double residuals of the 7-point system."""
    import discrete3d
    n = 128
    h = 1.0 / n
    rng = np.random.default_rng(7)
    f = rng.standard_normal((n, n, n)) * h * h
    banner("3-D 128^3 bc=%s: oracle" % (bc,))
    ref = discrete3d.refined(f, h, h, h, bc, lambda g: O.sub_solve3d(g, h, h, h, bc))
    err = rel(gpu3d(ctx, f, h, h, h, bc), ref)
    banner("3-D 128^3 bc=%s rel-L2 vs the discrete solution %.3e" % (bc, err))
    assert err < 1e-12
