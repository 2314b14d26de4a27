"""Captured composite solve (fp_cat_graph_capture / fp_cat_graph_launch):
every replay, with eager work and synchronisations in between, matches the
eager fp_cat_solve of the same right-hand side, and the oracle."""
import numpy as np
import pytest

import oracle as O
from gpu_util import banner, device, rel, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    device()
    import focalors_poisson_amd as fpa
    return fpa.Context(0)


@pytest.mark.parametrize("n", [32, 128])
def test_cross_graph_replays_match_eager(ctx, n):
    import torch
    import focalors_poisson_amd as fpa
    from focalors_poisson_amd.cases import cross
    banner("cross graph n=%d" % n)
    seeds = [1, 2, 3]
    cases = [cross(n, seed=sd) for sd in seeds]
    doms, conn = cases[0][0], cases[0][1]
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=40, tol=1e-12, tol_is_relative=True, max_iter=100)
    stream = torch.cuda.Stream()
    bound = [torch.empty(f[0].shape, dtype=torch.float64, device=device()) for f in cases[0][2]]
    with torch.cuda.stream(stream):
        cat.graph_capture(bound, stream=stream)
    for k, (_, _, fields, _) in enumerate(cases):
        fs = [f[0] for f in fields]
        eager = [to_dev(f) for f in fs]
        cat.solve(eager)
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            for b, f in zip(bound, fs):
                b.copy_(torch.from_numpy(f).to(device()))
            info = cat.graph_launch(stream=stream)
        torch.cuda.synchronize()
        got = [b.cpu().numpy() for b in bound]
        want = [e.cpu().numpy() for e in eager]
        assert info["status"] == 0 and info["rel_residual"] <= 1e-12, info
        assert rel(got, want) < 1e-13, k
        ref = O.reference_solution(doms, conn, fs)
        assert rel(got, ref) < 1e-12, k


def test_single_domain_graph(ctx):
    import torch
    import focalors_poisson_amd as fpa
    from focalors_poisson_amd.cases import single
    banner("single domain graph")
    doms, conn, fields, _ = single(64)
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn)
    stream = torch.cuda.Stream()
    bound = [torch.empty(fields[0].shape, dtype=torch.float64, device=device())]
    with torch.cuda.stream(stream):
        cat.graph_capture(bound, stream=stream)
        for _ in range(3):
            bound[0].copy_(torch.from_numpy(np.ascontiguousarray(fields[0])).to(device()))
            info = cat.graph_launch(stream=stream)
            torch.cuda.synchronize()
            ref = O.reference_solution(doms, conn, fields)
            assert rel([bound[0].cpu().numpy()], ref) < 1e-12
            assert info["status"] == 0


def test_graph_requires_stream(ctx):
    import focalors_poisson_amd as fpa
    from focalors_poisson_amd.cases import cross
    doms, conn, fields, _ = cross(16, seed=1)
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn)
    with pytest.raises(ValueError):
        cat.graph_capture([to_dev(f[0]) for f in fields])
    with pytest.raises(fpa.FPError):
        cat.graph_launch()


@pytest.mark.parametrize("geom,precision", [("cross", 0), ("cross", 1), ("six", 0)])
def test_solve_into_matches_in_place(ctx, geom, precision):
    """fp_cat_solve_into: the right-hand sides stay unchanged and the output
    equals the in-place solve (two-level path reading them in its transforms;
    the refinement and general-tree paths from a copy)."""
    import os
    import torch
    import focalors_poisson_amd as fpa
    from focalors_poisson_amd.cases import cross
    banner("solve_into %s precision=%d" % (geom, precision))
    if geom == "cross":
        doms, conn, fields, _ = cross(64, seed=7)
        fs = [f[0] for f in fields]
    else:
        g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "concat_six.npz"))
        doms = [(int(a), int(b), 1.0, 1.0, (0, 0, 0, 0)) for a, b in g["sizes"]]
        conn = [tuple(int(v) for v in c) for c in g["conn"]]
        fs = [g["f%d" % k] for k in range(len(doms))]
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=40, tol=1e-13, tol_is_relative=True, precision=precision)
    rhs = [to_dev(f) for f in fs]
    out = [torch.full_like(r, float("nan")) for r in rhs]
    cat.solve_into(rhs, out)
    inplace = [to_dev(f) for f in fs]
    cat.solve(inplace)
    torch.cuda.synchronize()
    assert all(np.array_equal(r.cpu().numpy(), f) for r, f in zip(rhs, fs))
    assert rel([o.cpu().numpy() for o in out], [i.cpu().numpy() for i in inplace]) < 1e-14


def test_graph_capture_into_replays(ctx):
    """Captured out-of-place solve: replays read the bound right-hand sides
    (refilled between replays) and match the eager out-of-place solve."""
    import torch
    import focalors_poisson_amd as fpa
    from focalors_poisson_amd.cases import cross
    banner("graph capture into")
    cases = [cross(64, seed=sd) for sd in (11, 12, 13)]
    doms, conn = cases[0][0], cases[0][1]
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=40, tol=1e-12, tol_is_relative=True, max_iter=100)
    stream = torch.cuda.Stream()
    rhs = [torch.empty(f[0].shape, dtype=torch.float64, device=device()) for f in cases[0][2]]
    out = [torch.empty_like(r) for r in rhs]
    with torch.cuda.stream(stream):
        cat.graph_capture(out, stream=stream, rhs=rhs)
    for k, (_, _, fields, _) in enumerate(cases):
        fs = [f[0] for f in fields]
        eager_out = [torch.empty_like(r) for r in rhs]
        cat.solve_into([to_dev(f) for f in fs], eager_out)
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            for r, f in zip(rhs, fs):
                r.copy_(torch.from_numpy(f).to(device()))
            info = cat.graph_launch(stream=stream)
        torch.cuda.synchronize()
        assert info["status"] == 0, info
        assert rel([o.cpu().numpy() for o in out], [e.cpu().numpy() for e in eager_out]) < 1e-13, k
        assert all(np.array_equal(r.cpu().numpy(), f) for r, f in zip(rhs, fs)), k


def test_tshape_graph_and_multi_rhs(ctx):
    """The T of unequal domains (non-square root, children of 40 and 56 rows
    on the x faces, 48 on y+): graph replays with eager work in between
    match the eager solve and the oracle, and a three-right-hand-side solve
    matches the single ones."""
    import torch
    import focalors_poisson_amd as fpa
    D, N, A = 0, 1, 3
    h = 1.0 / 64
    doms = [(96, 64, h, h, (A, A, D, A)), (40, 64, h, h, (N, A, D, N)), (56, 64, h, h, (A, D, N, D)),
            (96, 48, h, h, (D, N, A, N))]
    conn = [(0, 0, 1), (0, 1, 2), (0, 3, 3)]
    rng = np.random.default_rng(78)
    rhs = [[rng.standard_normal((d[0], d[1])) for d in doms] for _ in range(3)]
    banner("tshape graph and multi rhs")
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=40, tol=1e-12, tol_is_relative=True, max_iter=100)
    stream = torch.cuda.Stream()
    bound = [torch.empty((d[0], d[1]), dtype=torch.float64, device=device()) for d in doms]
    with torch.cuda.stream(stream):
        cat.graph_capture(bound, stream=stream)
    singles = []
    for k, fs in enumerate(rhs):
        eager = [to_dev(f) for f in fs]
        cat.solve(eager)
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            for b, f in zip(bound, fs):
                b.copy_(torch.from_numpy(f).to(device()))
            info = cat.graph_launch(stream=stream)
        torch.cuda.synchronize()
        got = [b.cpu().numpy() for b in bound]
        want = [e.cpu().numpy() for e in eager]
        singles.append(want)
        assert info["status"] == 0 and info["rel_residual"] <= 1e-12, info
        assert rel(got, want) < 1e-13, k
        ref = O.reference_solution(doms, conn, fs)
        assert rel(got, ref) < 1e-12, k
    multi = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=40, tol=1e-12, tol_is_relative=True, max_iter=100)
    ts = [to_dev(np.stack([rhs[r][d] for r in range(3)])) for d in range(len(doms))]
    multi.solve(ts)
    torch.cuda.synchronize()
    for r in range(3):
        assert rel([t[r].cpu().numpy() for t in ts], singles[r]) < 1e-12, r


def test_graph_dropped_when_workspace_grows(ctx):
    """A solve with more right-hand sides than were captured reallocates the
    workspace the graph was recorded on: the graph is destroyed with it, so a
    later replay reports that no solve is captured instead of replaying into
    freed memory; a new capture works again."""
    import torch
    import focalors_poisson_amd as fpa
    from focalors_poisson_amd.cases import cross
    banner("graph dropped when the workspace grows")
    doms, conn, fields, _ = cross(32, seed=21, nrhs=4)
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=40, tol=1e-12, tol_is_relative=True, max_iter=100)
    stream = torch.cuda.Stream()
    bound = [to_dev(f[0]) for f in fields]
    with torch.cuda.stream(stream):
        cat.graph_capture(bound, stream=stream)
    many = [to_dev(f) for f in fields]
    cat.solve(many)
    torch.cuda.synchronize()
    with pytest.raises(fpa.FPError) as e:
        cat.graph_launch(stream=stream)
    assert e.value.status == 1
    with torch.cuda.stream(stream):
        cat.graph_capture(bound, stream=stream)
        for b, f in zip(bound, fields):
            b.copy_(torch.from_numpy(np.ascontiguousarray(f[0])).to(device()))
        info = cat.graph_launch(stream=stream)
    torch.cuda.synchronize()
    assert info["status"] == 0
    assert rel([b.cpu().numpy() for b in bound], [m[0].cpu().numpy() for m in many]) < 1e-13


@pytest.mark.parametrize("n", [64, 256])
def test_refined_graph_replays(ctx, n):
    """FP_F64_REFINE captured: the FP64 solve, the FP64 composite residual,
    one correction solve and the update in one graph; replays match the eager
    refined solve and the converged oracle at 1e-12."""
    import torch
    import focalors_poisson_amd as fpa
    from focalors_poisson_amd.cases import cross
    banner("refined graph n=%d" % n)
    cases = [cross(n, seed=sd) for sd in (31, 32)]
    doms, conn = cases[0][0], cases[0][1]
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=40, tol=1e-12, tol_is_relative=True, max_iter=100,
                                    precision=fpa.FP_F64_REFINE)
    stream = torch.cuda.Stream()
    rhs = [torch.empty(f[0].shape, dtype=torch.float64, device=device()) for f in cases[0][2]]
    out = [torch.empty_like(r) for r in rhs]
    with torch.cuda.stream(stream):
        cat.graph_capture(out, stream=stream, rhs=rhs)
    for k, (_, _, fields, _) in enumerate(cases):
        fs = [f[0] for f in fields]
        eager = [to_dev(f) for f in fs]
        einfo = cat.solve(eager)
        torch.cuda.synchronize()
        assert einfo["refine_steps"] == 2, einfo
        with torch.cuda.stream(stream):
            for r, f in zip(rhs, fs):
                r.copy_(torch.from_numpy(f).to(device()))
            info = cat.graph_launch(stream=stream)
        torch.cuda.synchronize()
        assert info["status"] == 0 and info["refine_steps"] == 2, info
        got = [o.cpu().numpy() for o in out]
        assert rel(got, [e.cpu().numpy() for e in eager]) < 1e-14, k
        assert rel(got, O.reference_solution(doms, conn, fs)) < 1e-12, k
