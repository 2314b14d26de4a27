"""fp_sub_solve_batched_noalloc (the batched subdomain solve that allocates
nothing: descriptor tables in the library's per-GPU scratch array) against
fp_sub_solve_batched on the same problems: identical results, on two calls in
a row, the second with fewer and smaller problems; and the size limit."""
import ctypes

import numpy as np
import pytest

from gpu_util import banner, device, to_dev

pytestmark = pytest.mark.gpu
D, N, P, A = 0, 1, 2, 3


def problems(ctx, rng, count, lo, hi):
    import focalors_poisson_amd as fpa
    fams = [(D, D, D, D), (N, D, N, N), (D, N, P, P), (A, N, D, N), (P, P, D, N)]
    out = []
    for k in range(count):
        bc = fams[k % len(fams)]
        nx = int(rng.integers(lo, hi))
        ny = int(rng.integers(lo, hi)) // 2 * 2
        h = 1.0 / max(nx, ny)
        nrhs = 1 + k % 3
        f = rng.standard_normal((nrhs, nx, ny)) * h * h
        out.append((fpa.PoissonSolver2D(ctx, nx, ny, h, h, bc), f))
    return out


def test_noalloc_matches_batched_twice():
    import torch
    import focalors_poisson_amd as fpa
    device()
    ctx = fpa.Context(0)
    rng = np.random.default_rng(71)
    for call, (count, lo, hi) in enumerate(((40, 8, 300), (12, 4, 90))):
        probs = problems(ctx, rng, count, lo, hi)
        ref = [to_dev(f) for _, f in probs]
        got = [to_dev(f) for _, f in probs]
        banner("no-alloc batched call %d: %d problems" % (call, count))
        fpa.solve_batched(ctx, [s for s, _ in probs], ref)
        fpa.solve_batched(ctx, [s for s, _ in probs], got, noalloc=True)
        torch.cuda.synchronize()
        for k, (r, g) in enumerate(zip(ref, got)):
            assert torch.equal(r, g), (call, k)


def test_noalloc_rejects_too_many_problems():
    import focalors_poisson_amd as fpa
    from focalors_poisson_amd import _capi
    device()
    ctx = fpa.Context(0)
    n = 4097
    nulls = (ctypes.c_void_p * n)()
    nr = (ctypes.c_int * n)()
    st = _capi.lib().fp_sub_solve_batched_noalloc(ctx._h, n, nulls, nulls, nr, None)
    assert st == _capi.FP_E_INVALID
    assert b"no-allocation" in _capi.lib().fp_last_error()
