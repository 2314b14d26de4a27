"""The pressure projection of an incompressible flow step (fp_proj_*) on the
device against its numpy restatement (oracle/projection.py, pinned in
tests/test_oracle.py against the assembled composite system): divergence and
gradient kernels, the eager step, and the step replayed from one HIP graph
with the velocity refreshed in place between replays.

Bars: the kernels to FP64 rounding; the pressure to 1e-12 against the
converged reference; the projected velocity's divergence below 4e-15 n^2 of
the provisional one's on an n^2 grid (the divergence left is dt times the
solve's residual, and the scaled 5-point operator has norm 8 / h^2, so a
backward-stable solve leaves a residual about n^2 above rounding); faces
shared by two subdomains equal to rounding."""
import numpy as np
import pytest

import projection as PR
from gpu_util import banner, device, rel, to_dev

pytestmark = pytest.mark.gpu
D, N, P, A = 0, 1, 2, 3


@pytest.fixture(scope="module")
def ctx():
    device()
    import focalors_poisson_amd as fpa
    return fpa.Context(0)


def geometry(kind, n):
    from focalors_poisson_amd.cases import cross_geometry
    if kind == "cross":
        doms, conn, _ = cross_geometry(n)
        return doms, conn
    h = 1.0 / n
    # a channel periodic in y: inlet (Neumann, prescribed velocity) on the
    # left, outlet (Dirichlet pressure) on the right, two subdomains along x
    return [(n, n, h, h, (N, A, P, P)), (n // 2, n, h, h, (A, D, P, P))], [(0, 1, 1)]


def velocity(doms, conn, seed):
    """A random provisional velocity whose shared faces hold one value."""
    rng = np.random.default_rng(seed)
    u = [rng.standard_normal((d[0] + 1, d[1])) for d in doms]
    v = [rng.standard_normal((d[0], d[1] + 1)) for d in doms]
    for a, dr, b in conn:
        if dr == 1:
            u[b][0] = u[a][-1]
        elif dr == 0:
            u[b][-1] = u[a][0]
        elif dr == 3:
            v[b][:, 0] = v[a][:, -1]
        else:
            v[b][:, -1] = v[a][:, 0]
    for k, dd in enumerate(doms):  # periodic y: the two copies of a wrapped face agree
        if dd[4][2] == P:
            v[k][:, -1] = v[k][:, 0]
    return u, v


def ghost(doms, conn, seed):
    import discrete
    rng = np.random.default_rng(seed)
    bcs = discrete.effective_bc(doms, conn)
    return {(d, f): rng.standard_normal(doms[d][1] if f < 2 else doms[d][0])
            for d in range(len(doms)) for f in range(4) if bcs[d][f] == D}


def make(ctx, doms, conn):
    import focalors_poisson_amd as fpa
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=200, tol=1e-13, tol_is_relative=True, max_iter=100)
    return cat, fpa.PressureProjection(cat)


def shared_gap(doms, conn, u, v):
    worst = 0.0
    for a, dr, b in conn:
        if dr == 1:
            worst = max(worst, np.abs(u[a][-1] - u[b][0]).max())
        elif dr == 0:
            worst = max(worst, np.abs(u[a][0] - u[b][-1]).max())
        elif dr == 3:
            worst = max(worst, np.abs(v[a][:, -1] - v[b][:, 0]).max())
        else:
            worst = max(worst, np.abs(v[a][:, 0] - v[b][:, -1]).max())
    return worst


@pytest.mark.parametrize("kind", ["cross", "channel"])
def test_projection_kernels_match_numpy(ctx, kind):
    import torch
    doms, conn = geometry(kind, 32)
    cat, pj = make(ctx, doms, conn)
    u, v = velocity(doms, conn, 1)
    rng = np.random.default_rng(2)
    p = [rng.standard_normal((d[0], d[1])) for d in doms]
    pg = ghost(doms, conn, 3)
    banner("projection kernels %s" % kind)
    du, dv, dp = [to_dev(x) for x in u], [to_dev(x) for x in v], [to_dev(x) for x in p]
    out = [torch.empty(d[0], d[1], dtype=torch.float64, device=device()) for d in doms]
    pj.divergence(du, dv, out, scale=0.5)
    pj.correct(du, dv, dp, {k: to_dev(g) for k, g in pg.items()}, scale=0.25)
    torch.cuda.synchronize()
    assert rel([o.cpu().numpy() for o in out], PR.divergence(doms, u, v, 0.5)) < 1e-15
    uo, vo = PR.correct(doms, conn, u, v, p, pg, 0.25)
    assert rel([x.cpu().numpy() for x in du], uo) < 1e-15
    assert rel([x.cpu().numpy() for x in dv], vo) < 1e-15


def solve_ref(kind):
    if kind == "cross":
        return None  # the direct composite solve
    import oracle as O
    return lambda doms, conn, rhs, g: O.reference_solution(doms, conn, rhs, g)


@pytest.mark.parametrize("kind,n", [("cross", 64), ("channel", 64), ("cross", 256)])
def test_projection_step(ctx, kind, n):
    """One eager step: pressure against the reference, divergence of the
    projected velocity, shared faces."""
    import torch
    doms, conn = geometry(kind, n)
    cat, pj = make(ctx, doms, conn)
    u, v = velocity(doms, conn, 7)
    pg = ghost(doms, conn, 8)
    dt = 0.01
    du, dv = [to_dev(x) for x in u], [to_dev(x) for x in v]
    dp = [torch.empty(d[0], d[1], dtype=torch.float64, device=device()) for d in doms]
    banner("projection step %s %d" % (kind, n))
    info = pj.step(du, dv, dp, dt, {k: to_dev(g) for k, g in pg.items()})
    torch.cuda.synchronize()
    p_ref, u_ref, v_ref = PR.project(doms, conn, u, v, dt, pg, solve_ref(kind))
    got_p = [x.cpu().numpy() for x in dp]
    uo, vo = [x.cpu().numpy() for x in du], [x.cpu().numpy() for x in dv]
    d0 = max(np.abs(x).max() for x in PR.divergence(doms, u, v))
    d1 = max(np.abs(x).max() for x in PR.divergence(doms, uo, vo))
    ep = rel(got_p, p_ref)
    banner("projection step %s %d: pressure rel-L2 %.2e, divergence %.2e of %.2e, solve %s"
           % (kind, n, ep, d1, d0, info))
    assert ep < 1e-12
    assert rel(uo, u_ref) < 1e-12 and rel(vo, v_ref) < 1e-12
    assert d1 < 4e-15 * n * n * d0
    assert shared_gap(doms, conn, uo, vo) <= 1e-14 * max(np.abs(x).max() for x in uo)


def test_projection_graph_replays(ctx):
    """The step captured once and replayed for five new provisional
    velocities (refreshed in place between replays, eager work and waits in
    between): every replay equals an eager step on a copy."""
    import torch
    doms, conn = geometry("cross", 64)
    cat, pj = make(ctx, doms, conn)
    _, pj_eager = make(ctx, doms, conn)
    pg = {k: to_dev(g) for k, g in ghost(doms, conn, 9).items()}
    dt = 0.02
    u0, v0 = velocity(doms, conn, 10)
    du, dv = [to_dev(x) for x in u0], [to_dev(x) for x in v0]
    dp = [torch.zeros(d[0], d[1], dtype=torch.float64, device=device()) for d in doms]
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        pj.capture(du, dv, dp, dt, pg, stream=s)
    torch.cuda.synchronize()
    for it in range(5):
        u, v = velocity(doms, conn, 20 + it)
        for t, x in zip(du, u):
            t.copy_(torch.from_numpy(x))
        for t, x in zip(dv, v):
            t.copy_(torch.from_numpy(x))
        torch.cuda.synchronize()
        banner("projection graph replay %d" % it)
        with torch.cuda.stream(s):
            info = pj.launch(stream=s, wait=True)
        eu, ev = [to_dev(x) for x in u], [to_dev(x) for x in v]
        ep = [torch.empty_like(x) for x in dp]
        pj_eager.step(eu, ev, ep, dt, pg)
        torch.cuda.synchronize()
        assert info["rel_residual"] < 1e-12, info
        assert rel([x.cpu().numpy() for x in dp], [x.cpu().numpy() for x in ep]) < 1e-14
        assert rel([x.cpu().numpy() for x in du], [x.cpu().numpy() for x in eu]) < 1e-14
        d1 = max(np.abs(x).max() for x in PR.divergence(doms, [x.cpu().numpy() for x in du],
                                                        [x.cpu().numpy() for x in dv]))
        d0 = max(np.abs(x).max() for x in PR.divergence(doms, u, v))
        assert d1 < 4e-15 * 64 * 64 * d0


def test_projection_graph_belongs_to_its_projection(ctx):
    """Two projections on one solver: only the one that captured replays
    the graph; a projection made after the capturing one is gone does not
    inherit its graph (the graph refers to the first one's buffers)."""
    import torch
    import focalors_poisson_amd as fpa
    doms, conn = geometry("cross", 32)
    cat, pa = make(ctx, doms, conn)
    pb = fpa.PressureProjection(cat)
    u, v = velocity(doms, conn, 30)
    du, dv = [to_dev(x) for x in u], [to_dev(x) for x in v]
    dp = [torch.zeros(d[0], d[1], dtype=torch.float64, device=device()) for d in doms]
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        pa.capture(du, dv, dp, 0.01, stream=s)
        pa.launch(stream=s, wait=True)
        with pytest.raises(fpa.FPError):
            pb.launch(stream=s)
    torch.cuda.synchronize()
    del pa
    pc = fpa.PressureProjection(cat)
    with pytest.raises(fpa.FPError):
        pc.launch(stream=s)


def test_projection_rejects_bad_input(ctx):
    import focalors_poisson_amd as fpa
    h = 1.0 / 16
    cat = fpa.ConcatPoissonSolver2D(ctx, [(16, 16, h, 2 * h, (D, D, D, D))], [])
    with pytest.raises(fpa.FPError):
        fpa.PressureProjection(cat)
    doms, conn = geometry("cross", 16)
    cat, pj = make(ctx, doms, conn)
    with pytest.raises(fpa.FPError):
        pj.launch()  # nothing captured
