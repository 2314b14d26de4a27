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
        ref, _, _ = O.Composite(doms, conn, tol=1e-14, tol_is_relative=True, maxiter=400,
                                schur_mode=O.SCHUR_SPECTRAL).solve(fs)
        assert rel(got, ref) < 1e-10, k


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
            ref, _, _ = O.Composite(doms, conn).solve(fields)
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
