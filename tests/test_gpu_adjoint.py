"""Adjoint of the composite solve (fp_cat_vjp) against central finite
differences along seeded directions and against the oracle's gradient
(L = 1/2 ||u - u*||^2; the operator is linear, so central differences are
exact up to rounding)."""
import numpy as np
import pytest

import oracle as O
from gpu_util import banner, device, rel, to_dev

pytestmark = pytest.mark.gpu
D, N, A = 0, 1, 3


@pytest.fixture(scope="module")
def setup():
    device()
    import focalors_poisson_amd as fpa
    from focalors_poisson_amd.cases import cross
    ctx = fpa.Context(0)
    n = 32
    doms, conn, fields, _ = cross(n, seed=5)
    fs = [f[0] for f in fields]
    rng = np.random.default_rng(11)
    # ghost data on the Dirichlet and Neumann faces of the arms
    faces = [(1, 0), (1, 2), (1, 3), (2, 1), (2, 2), (2, 3), (3, 0), (3, 1), (3, 2), (4, 0), (4, 1), (4, 3)]
    ghost = {k: rng.standard_normal(doms[k[0]][1] if k[1] < 2 else doms[k[0]][0]) for k in faces}
    target = [rng.standard_normal(f.shape) for f in fs]
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=40, tol=1e-14, tol_is_relative=True, max_iter=50)
    return fpa, cat, doms, conn, fs, ghost, target


def loss_and_u(cat, fs, ghost, target):
    import torch
    ts = [to_dev(f) for f in fs]
    cat.solve(ts, {k: to_dev(v) for k, v in ghost.items()})
    torch.cuda.synchronize()
    us = [t.cpu().numpy() for t in ts]
    return 0.5 * sum(float(np.sum((u - t) ** 2)) for u, t in zip(us, target)), us


def test_vjp_matches_finite_differences(setup):
    import torch
    fpa, cat, doms, conn, fs, ghost, target = setup
    L0, us = loss_and_u(cat, fs, ghost, target)
    gu = [to_dev(u - t) for u, t in zip(us, target)]
    gf, gg = cat.vjp(gu, faces=list(ghost))
    torch.cuda.synchronize()
    gf = [g.cpu().numpy() for g in gf]
    gg = {k: v.cpu().numpy() for k, v in gg.items()}
    rng = np.random.default_rng(16)
    h = 1e-6
    for trial in range(8):
        vf = [rng.standard_normal(f.shape) for f in fs]
        vg = {k: rng.standard_normal(v.shape) for k, v in ghost.items()}
        Lp, _ = loss_and_u(cat, [f + h * v for f, v in zip(fs, vf)], {k: ghost[k] + h * vg[k] for k in ghost}, target)
        Lm, _ = loss_and_u(cat, [f - h * v for f, v in zip(fs, vf)], {k: ghost[k] - h * vg[k] for k in ghost}, target)
        fd = (Lp - Lm) / (2 * h)
        an = sum(float(np.sum(g * v)) for g, v in zip(gf, vf)) + sum(float(np.sum(gg[k] * vg[k])) for k in ghost)
        banner("adjoint FD direction %d: fd %.9e analytic %.9e" % (trial, fd, an))
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))


def test_vjp_matches_oracle_gradient(setup):
    import torch
    fpa, cat, doms, conn, fs, ghost, target = setup
    L0, us = loss_and_u(cat, fs, ghost, target)
    gu = [u - t for u, t in zip(us, target)]
    gf, gg = cat.vjp([to_dev(g) for g in gu], faces=list(ghost))
    torch.cuda.synchronize()
    # oracle: lambda = A^-1 dL/du through the converged oracle solve of dL/du / hx^2
    lam = O.reference_solution(doms, conn, [g / (d[2] ** 2) for g, d in zip(gu, doms)])
    ref_f = [d[2] ** 2 * l for d, l in zip(doms, lam)]
    banner("adjoint vs oracle gradient")
    assert rel([g.cpu().numpy() for g in gf], ref_f) < 1e-12
    # dL/dg on each face with data: hx^2 c lambda on the face's line, c the
    # ghost coefficient of the assembly (Dirichlet -1/h^2, Neumann +-1/h)
    lines = {0: lambda x: x[0, :], 1: lambda x: x[-1, :], 2: lambda x: x[:, 0], 3: lambda x: x[:, -1]}
    got, ref = [], []
    for (d, face) in ghost:
        hx, hy = doms[d][2], doms[d][3]
        h = hx if face < 2 else hy
        t = doms[d][4][face]
        c = -1.0 / h ** 2 if t == D else (1.0 / h if face in (0, 2) else -1.0 / h)
        ref.append(hx * hx * c * lines[face](lam[d]))
        got.append(gg[(d, face)].cpu().numpy())
    assert rel(got, ref) < 1e-12


def test_autograd_function(setup):
    import torch
    from focalors_poisson_amd.autograd import composite_solve
    fpa, cat, doms, conn, fs, ghost, target = setup
    ft = [to_dev(f).requires_grad_(True) for f in fs]
    gt = {k: to_dev(v).requires_grad_(True) for k, v in ghost.items()}
    us = composite_solve(cat, ft, gt)
    loss = 0.5 * sum(((u - to_dev(t)) ** 2).sum() for u, t in zip(us, target))
    loss.backward()
    L0, uref = loss_and_u(cat, fs, ghost, target)
    gu = [to_dev(u - t) for u, t in zip(uref, target)]
    gf, gg = cat.vjp(gu, faces=list(ghost))
    banner("autograd wrapper")
    assert abs(float(loss) - L0) < 1e-9 * abs(L0)
    for a, b in zip(ft, gf):
        assert torch.allclose(a.grad, b, rtol=1e-12, atol=1e-12 * float(b.abs().max()))
    for k in ghost:
        assert torch.allclose(gt[k].grad, gg[k], rtol=1e-12, atol=1e-12 * float(gg[k].abs().max()))


def test_autograd_unused_output(setup):
    """A loss that uses only the centre's solution: the other outputs get no
    gradient from autograd (None), which the backward takes as zeros of their
    shape; the gradient equals the vjp with zero dL/du on those domains."""
    import torch
    from focalors_poisson_amd.autograd import composite_solve
    fpa, cat, doms, conn, fs, ghost, target = setup
    ft = [to_dev(f).requires_grad_(True) for f in fs]
    us = composite_solve(cat, ft)
    loss = 0.5 * ((us[0] - to_dev(target[0])) ** 2).sum()
    loss.backward()
    gu = [us[0].detach() - to_dev(target[0])] + [torch.zeros_like(u) for u in us[1:]]
    gf, _ = cat.vjp(gu)
    banner("autograd with unused outputs")
    for a, b in zip(ft, gf):
        assert torch.allclose(a.grad, b, rtol=1e-12, atol=1e-12 * float(b.abs().max()))
