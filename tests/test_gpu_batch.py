"""The prepared batch (fp_batch) and estimated planning (fp_ctx_set_planning)
against the CPU oracle (PoissonSolver2D::solve per problem,
pe/poisson/poisson_solver2d.cpp:62-90).  FP64 bar: relative L2 <= 1e-12;
FP32 bar: 2e-5."""
import numpy as np
import pytest

from gpu_util import banner, device, rel, sub_reference, to_dev

pytestmark = pytest.mark.gpu
D, N, P, A = 0, 1, 2, 3


@pytest.fixture(scope="module")
def ctx():
    device()
    import focalors_poisson_amd as fpa
    return fpa.Context(0)


def uneven_problems(ctx, seed, count, f32=False, lo=8, hi=300):
    """Problems of every boundary family, sizes that give smooth, odd-prime
    and Bluestein transform lengths, 1-3 right-hand sides each."""
    import torch
    import focalors_poisson_amd as fpa
    rng = np.random.default_rng(seed)
    fams = [(D, D, D, D), (N, D, N, N), (D, N, P, P), (P, P, D, D), (A, N, D, N), (N, A, N, D), (D, A, A, N)]
    solvers, fields, hosts, specs = [], [], [], []
    for k in range(count):
        bc = fams[k % len(fams)]
        nx = int(rng.integers(lo, hi))
        ny = int(rng.integers(lo, hi)) // 2 * 2
        h = 1.0 / max(nx, ny)
        nrhs = 1 + k % 3
        f = rng.uniform(-1, 1, (nrhs, nx, ny)) * h * h
        solvers.append(fpa.PoissonSolver2D(ctx, nx, ny, h, h, bc))
        fields.append(to_dev(f, torch.float32 if f32 else None))
        hosts.append(f)
        specs.append((h, bc))
    return solvers, fields, hosts, specs


def check(fields, hosts, specs, bar):
    for t, f, (h, bc) in zip(fields, hosts, specs):
        got = t.double().cpu().numpy()
        for r in range(f.shape[0]):
            assert rel(got[r], sub_reference(f[r], h, h, bc)) < bar, (f.shape, bc)


@pytest.mark.parametrize("f32", [False, True])
def test_batch_object_matches_oracle(ctx, f32):
    import torch
    import focalors_poisson_amd as fpa
    solvers, fields, hosts, specs = uneven_problems(ctx, 31 + f32, 45, f32)
    banner("prepared batch of 45, f32=%d" % f32)
    b = fpa.BatchSolve(ctx, solvers, fields)
    b.solve()
    torch.cuda.synchronize()
    check(fields, hosts, specs, 2e-5 if f32 else 1e-12)


def test_batch_object_repeats_and_matches_per_call_batch(ctx):
    """Two solves of the prepared batch (fields refilled between) give the same
    bits, and agree with fp_sub_solve_batched to rounding."""
    import torch
    import focalors_poisson_amd as fpa
    solvers, fields, hosts, specs = uneven_problems(ctx, 43, 30)
    pristine = [f.clone() for f in fields]
    banner("prepared batch repeat")
    b = fpa.BatchSolve(ctx, solvers, fields)
    b.solve()
    first = [f.clone() for f in fields]
    for f, p in zip(fields, pristine):
        f.copy_(p)
    b.solve()
    torch.cuda.synchronize()
    for a, c in zip(first, fields):
        assert torch.equal(a, c)
    other = [p.clone() for p in pristine]
    fpa.solve_batched(ctx, solvers, other)
    torch.cuda.synchronize()
    for a, c in zip(first, other):
        assert rel(a.cpu().numpy(), c.cpu().numpy()) < 1e-13


def test_batch_object_graph_replays(ctx):
    """The prepared batch captured in a graph: every replay (the fields
    refilled eagerly between replays) equals the eager solve."""
    import torch
    import focalors_poisson_amd as fpa
    solvers, fields, hosts, specs = uneven_problems(ctx, 47, 40)
    pristine = [f.clone() for f in fields]
    b = fpa.BatchSolve(ctx, solvers, fields)
    banner("prepared batch eager")
    b.solve()
    torch.cuda.synchronize()
    eager = [f.clone() for f in fields]
    check(fields, hosts, specs, 1e-12)
    banner("prepared batch graph capture")
    for f, p in zip(fields, pristine):
        f.copy_(p)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        g.capture_begin()
        b.solve(stream=s)
        g.capture_end()
    torch.cuda.synchronize()
    for rep in range(3):
        banner("prepared batch graph replay %d" % rep)
        for f, p in zip(fields, pristine):
            f.copy_(p)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, c in zip(eager, fields):
            assert torch.equal(a, c)


def test_estimated_planning_matches_oracle(ctx):
    """Plans built without timing (FP_PLAN_ESTIMATE) solve to the same bar,
    per call and in a prepared batch."""
    import torch
    import focalors_poisson_amd as fpa
    est = fpa.Context(0)
    est.set_planning(fpa.FP_PLAN_ESTIMATE)
    solvers, fields, hosts, specs = uneven_problems(est, 53, 28, hi=700)
    pristine = [f.clone() for f in fields]
    banner("estimated planning, batch of 28")
    fpa.solve_batched(est, solvers, fields)
    torch.cuda.synchronize()
    check(fields, hosts, specs, 1e-12)
    for f, p in zip(fields, pristine):
        f.copy_(p)
    b = fpa.BatchSolve(est, solvers, fields)
    b.solve()
    torch.cuda.synchronize()
    check(fields, hosts, specs, 1e-12)
    with pytest.raises(fpa.FPError):
        est.set_planning(7)


def test_batch_object_rejects_bad_input(ctx):
    import torch
    import focalors_poisson_amd as fpa
    s = fpa.PoissonSolver2D(ctx, 16, 16, 1 / 16, 1 / 16, (D, D, D, D))
    with pytest.raises(ValueError):
        fpa.BatchSolve(ctx, [s], [])
    with pytest.raises(TypeError):
        fpa.BatchSolve(ctx, [s, s], [torch.zeros(256, dtype=torch.float64, device="cuda"),
                                     torch.zeros(256, dtype=torch.float32, device="cuda")])
