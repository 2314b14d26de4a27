"""Transforms longer than one workgroup's LDS (PoissonFFT2D_* at any N, as
the reference's FFTW plans, pe/poisson/poisson_fft2d.cpp:19-57,77-133,
153-262,264-465): the global-memory engine keeps the two FFT buffers of a
row pair in device memory, with Bluestein over a 2,3,5-smooth length for
prime factors above 47.  Coefficients against the closed-form sums
(tests/closed_form.py, angles reduced on integers) and the oracle both
ways, and subdomain and composite solves against the converged oracle and
the assembled sparse system.  FP64 bar: 1e-12.

The reference forms some phases as unreduced angles in double: ND forward
2 N pi (j+1) / (2N+1) (poisson_fft2d.cpp:180-189), PP pi (2m+1) / 4
(:367-465), of order 1e4 at these lengths, whose rounding moves the
coefficients by 1e-12 to 2e-12 relative.  The oracle restates that
rounding; the device transform reduces its angles on integers.  So the
device is held to 1e-12 against the closed form, and to 1e-12 plus the
oracle's own distance from the closed form against the oracle."""
import numpy as np
import pytest

import closed_form
import oracle as O
from gpu_util import banner, device, rel, sub_reference, to_dev

pytestmark = pytest.mark.gpu
D, N, P, A = 0, 1, 2, 3


@pytest.fixture(scope="module")
def ctx():
    device()
    import focalors_poisson_amd as fpa
    return fpa.Context(0)


# FFT length L and why none of the LDS engines holds it in FP64:
#   DN / ND 4096: L = 8193 = 3 x 2731, Bluestein F = 16875 (270 KB)
#   DD 5000:      L = 5001 = 3 x 1667, Bluestein F = 10125 (162 KB + tables)
#   NN 10000:     L = 10000 = 2^4 5^4, plain FFT, one buffer 160 KB + tables
#   PP 12288:     L = 12288 = 2^12 x 3, plain FFT (196 KB)
#   PP 10006:     L = 10006 = 2 x 5003, Bluestein F = 20250
@pytest.mark.parametrize("bcp,ny", [((D, N), 4096), ((N, D), 4096), ((D, D), 5000), ((N, N), 10000),
                                    ((P, P), 12288), ((P, P), 10006)])
def test_long_transform_coefficients(ctx, bcp, ny):
    import torch
    import focalors_poisson_amd as fpa
    rng = np.random.default_rng(ny + 17 * bcp[0] + 3 * bcp[1])
    nx = 5  # two row pairs and a lone row
    f = rng.standard_normal((nx, ny))
    s = fpa.PoissonSolver2D(ctx, nx, ny, 1.0, 1.0, (D, D) + bcp)
    for inv in (False, True):
        t = to_dev(f)
        banner("long transform bc=%s ny=%d inverse=%d" % (bcp, ny, inv))
        s.transform(t, inverse=inv)
        torch.cuda.synchronize()
        got = t.cpu().numpy()
        ref = O.transform(bcp[0], bcp[1], f, inverse=inv)
        err = np.abs(got - ref).max() / max(1.0, np.abs(ref).max())
        msg = "long transform bc=%s ny=%d inverse=%d max error vs the oracle %.2e" % (bcp, ny, inv, err)
        cf = closed_form.transform(bcp[0], bcp[1], f, inverse=inv)
        scale = max(1.0, np.abs(cf).max())
        ecf = np.abs(got - cf).max() / scale
        eor = np.abs(ref - cf).max() / scale
        banner(msg + ", vs the closed form %.2e (oracle vs the closed form %.2e)" % (ecf, eor))
        assert ecf < 1e-12, (bcp, ny, inv, ecf)
        assert err < 1e-12 + eor, (bcp, ny, inv, err, eor)


def test_long_transform_many_row_pairs(ctx):
    """More row pairs than workgroup slots (600 rows: 300 pairs over 256
    slots), so workgroups loop and reuse their scratch slot."""
    import torch
    import focalors_poisson_amd as fpa
    rng = np.random.default_rng(600)
    nx, ny = 600, 4096
    f = rng.standard_normal((nx, ny))
    s = fpa.PoissonSolver2D(ctx, nx, ny, 1.0, 1.0, (D, D, N, D))
    t = to_dev(f)
    banner("long transform 600 rows ND 4096")
    s.transform(t, inverse=False)
    torch.cuda.synchronize()
    got = t.cpu().numpy()
    cf = closed_form.transform(N, D, f)
    scale = max(1.0, np.abs(cf).max())
    ecf = np.abs(got - cf).max() / scale
    err = np.abs(got - O.transform(N, D, f)).max() / scale
    eor = np.abs(O.transform(N, D, f) - cf).max() / scale
    banner("long transform 600 rows: vs the closed form %.2e, vs the oracle %.2e (oracle vs the closed form %.2e)"
           % (ecf, err, eor))
    assert ecf < 1e-12 and err < 1e-12 + eor, (ecf, err, eor)


@pytest.mark.parametrize("bc", [(D, D, D, N), (N, A, N, D), (D, N, D, D)])
def test_long_solves(ctx, bc):
    """Subdomain solves with ny = 4096 quarter-wave (L = 8193) and ny = 5000
    DST-I (L = 5001) y transforms: FP64 against the converged oracle, FP32
    sweeps, and FP32 with FP64 refinement."""
    import torch
    import focalors_poisson_amd as fpa
    ny = 5000 if bc[2:] == (D, D) else 4096
    nx = 48
    rng = np.random.default_rng(nx + ny + bc[1])
    h = 1.0 / ny
    g = rng.standard_normal((nx, ny)) * h * h
    ref = sub_reference(g, h, h, bc)
    s = fpa.PoissonSolver2D(ctx, nx, ny, h, h, bc)
    t = to_dev(g)
    banner("long solve bc=%s ny=%d fp64" % (bc, ny))
    s.solve(t)
    torch.cuda.synchronize()
    e64 = rel(t.cpu().numpy(), ref)
    t32 = to_dev(g, torch.float32)
    banner("long solve bc=%s ny=%d fp32" % (bc, ny))
    s.solve(t32)
    torch.cuda.synchronize()
    e32 = rel(t32.double().cpu().numpy(), ref)
    tr = to_dev(g)
    banner("long solve bc=%s ny=%d fp32+refine" % (bc, ny))
    s.solve_refine(tr, tol=1e-13, max_steps=10)
    torch.cuda.synchronize()
    er = rel(tr.cpu().numpy(), ref)
    banner("long solve bc=%s ny=%d: fp64 %.2e fp32 %.2e fp32+refine %.2e" % (bc, ny, e64, e32, er))
    assert e64 < 1e-12 and e32 < 1e-4 and er < 1e-12, (e64, e32, er)


@pytest.mark.parametrize("precision", [0, 2])
def test_long_composite_channel(ctx, precision):
    """Two domains side by side with Dirichlet / Neumann y faces of 4096
    cells (quarter-wave transforms of L = 8193 in both): the two-level
    composite path with global-memory transforms, FP64 and FP32 sweeps with
    FP64 refinement, against the assembled sparse system."""
    import torch
    import discrete
    import focalors_poisson_amd as fpa
    rng = np.random.default_rng(4099)
    h = 1.0 / 4096
    doms = [(24, 4096, h, h, (D, A, D, N)), (16, 4096, h, h, (A, N, D, N))]
    conn = [(0, 1, 1)]
    fs = [rng.standard_normal((d[0], d[1])) for d in doms]
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=40, tol=1e-13, tol_is_relative=True, max_iter=200,
                                    precision=precision)
    ts = [to_dev(f) for f in fs]
    banner("long composite channel precision=%d" % precision)
    info = cat.solve(ts)
    torch.cuda.synchronize()
    us = [t.cpu().numpy() for t in ts]
    err = rel(us, discrete.solve(doms, conn, fs))
    banner("long composite channel precision=%d rel-L2 vs the sparse system %.2e" % (precision, err))
    assert err < 1e-12, (err, info)


def test_long_multi_rhs(ctx):
    """Three right-hand sides in one call through the global-memory engine
    (ny = 4096, DN in y): each against its own converged oracle solve."""
    import torch
    import focalors_poisson_amd as fpa
    nx, ny, nrhs = 20, 4096, 3
    bc = (N, D, D, N)
    rng = np.random.default_rng(4097)
    h = 1.0 / ny
    g = rng.standard_normal((nrhs, nx, ny)) * h * h
    s = fpa.PoissonSolver2D(ctx, nx, ny, h, h, bc)
    t = to_dev(g)
    banner("long solve, 3 right-hand sides")
    s.solve(t)
    torch.cuda.synchronize()
    got = t.cpu().numpy()
    for r in range(nrhs):
        err = rel(got[r], sub_reference(g[r], h, h, bc))
        assert err < 1e-12, (r, err)


def test_long_composite_graph_replays(ctx):
    """The two-domain channel with 4096-cell Dirichlet/Neumann y faces
    captured into a HIP graph (out of place) and replayed on three
    right-hand sides: every replay equals the eager solve."""
    import torch
    import focalors_poisson_amd as fpa
    h = 1.0 / 4096
    doms = [(24, 4096, h, h, (D, A, D, N)), (16, 4096, h, h, (A, N, D, N))]
    conn = [(0, 1, 1)]
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=40, tol=1e-13, tol_is_relative=True, max_iter=200)
    stream = torch.cuda.Stream()
    rhs = [torch.empty((d[0], d[1]), dtype=torch.float64, device=device()) for d in doms]
    out = [torch.empty_like(r) for r in rhs]
    banner("long composite graph capture")
    with torch.cuda.stream(stream):
        cat.graph_capture(out, stream=stream, rhs=rhs)
    for k in range(3):
        rng = np.random.default_rng(500 + k)
        fs = [rng.standard_normal((d[0], d[1])) for d in doms]
        eager = [torch.empty_like(r) for r in rhs]
        cat.solve_into([to_dev(f) for f in fs], eager)
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            for r, f in zip(rhs, fs):
                r.copy_(torch.from_numpy(f).to(device()))
            info = cat.graph_launch(stream=stream)
        torch.cuda.synchronize()
        banner("long composite graph replay %d status %s" % (k, info["status"]))
        assert info["status"] == 0, info
        assert rel([o.cpu().numpy() for o in out], [e.cpu().numpy() for e in eager]) < 1e-13, k


@pytest.mark.parametrize("ny", [1024, 2048, 4096])
def test_dst1_accuracy_against_closed_form(ctx, ny):
    """DST-I (the Dirichlet pair) by the half-length algorithm, whose odd
    coefficients are running sums of the FFT's outputs (so their error
    grows with the length): both directions against the closed form, max
    error relative to the largest coefficient below 3e-14 up to 4096
    (about 3e-15 at 1024, 8e-15 at 2048, 1.3e-14 at 4096)."""
    import torch
    import focalors_poisson_amd as fpa
    rng = np.random.default_rng(ny + 99)
    f = rng.standard_normal((6, ny))
    s = fpa.PoissonSolver2D(ctx, 6, ny, 1.0, 1.0, (D, D, D, D))
    for inv in (False, True):
        t = to_dev(f)
        s.transform(t, inverse=inv)
        torch.cuda.synchronize()
        cf = closed_form.transform(D, D, f, inverse=inv)
        err = np.abs(t.cpu().numpy() - cf).max() / np.abs(cf).max()
        banner("DST-I ny=%d inverse=%d max error vs the closed form %.2e" % (ny, inv, err))
        assert err < 3e-14, (ny, inv, err)
