"""Subdomain solver parity on the device against the CPU oracle
(PoissonSolver2D::solve, pe/poisson/poisson_solver2d.cpp:62-90).
FP64 bar: relative L2 <= 1e-12.  FP32 bar: relative L2 <= 2e-5."""
import itertools
import os

import numpy as np
import pytest

import oracle as O
from gpu_util import banner, device, rel, sub_reference, to_dev

pytestmark = pytest.mark.gpu
D, N, P, A = 0, 1, 2, 3
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ctx():
    device()
    import focalors_poisson_amd as fpa
    return fpa.Context(0)


def gpu_solve(ctx, f, hx, hy, bc, dtype=None):
    import torch
    import focalors_poisson_amd as fpa
    s = fpa.PoissonSolver2D(ctx, f.shape[-2], f.shape[-1], hx, hy, bc)
    t = to_dev(f, dtype)
    s.solve(t)
    torch.cuda.synchronize()
    return t.double().cpu().numpy()


def test_poisson_3x4_golden(ctx):
    g = np.load(os.path.join(GOLD, "poisson_3x4.npz"))
    hx, hy = float(g["hx"]), float(g["hy"])
    for bc, u in zip(g["bc"], g["u"]):
        banner("poisson_3x4 bc=%s" % (tuple(bc),))
        got = gpu_solve(ctx, g["f"] * hx * hx, hx, hy, tuple(int(b) for b in bc))
        assert rel(got, u) < 1e-12, bc


@pytest.mark.parametrize("bc", [b for b in itertools.product([D, N, A], repeat=4) if b != (N, N, N, N)])
def test_all_boundary_pairs_small(ctx, bc):
    rng = np.random.default_rng(sum(bc) + 7)
    nx, ny, hx, hy = 9, 7, 0.3, 0.5
    f = rng.standard_normal((nx, ny)) * hx * hx
    banner("small bc=%s" % (bc,))
    assert rel(gpu_solve(ctx, f, hx, hy, bc), sub_reference(f, hx, hy, bc)) < 1e-12


SIZES = [(2, 1), (3, 2), (5, 41), (17, 64), (33, 100), (64, 257), (40, 512), (65, 1024), (1024, 37)]


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("bc", [(D, D, D, D), (N, D, N, N), (D, N, D, N), (A, N, N, D)])
def test_sizes(ctx, nx, ny, bc):
    rng = np.random.default_rng(nx * 1000 + ny)
    h = 1.0 / max(nx, ny)
    f = rng.standard_normal((nx, ny)) * h * h
    banner("sizes nx=%d ny=%d bc=%s" % (nx, ny, bc))
    assert rel(gpu_solve(ctx, f, h, h, bc), sub_reference(f, h, h, bc)) < 1e-12


@pytest.mark.parametrize("nx,ny,bc", [(256, 256, (D, D, D, D)), (512, 1024, (D, A, N, N)),
                                      (1024, 512, (N, D, A, N)), (1024, 1024, (A, A, A, A))])
def test_config_shapes(ctx, nx, ny, bc):
    rng = np.random.default_rng(nx + ny)
    h = 1.0 / 1024
    f = rng.standard_normal((nx, ny)) * h * h
    banner("config shape nx=%d ny=%d bc=%s" % (nx, ny, bc))
    assert rel(gpu_solve(ctx, f, h, h, bc), sub_reference(f, h, h, bc)) < 1e-12


def test_all_neumann_singular_mode(ctx):
    rng = np.random.default_rng(11)
    nx, ny = 24, 16
    f = rng.standard_normal((nx, ny))
    f -= f.mean()
    banner("singular NNNN")
    assert rel(gpu_solve(ctx, f, 1.0, 1.0, (N, N, N, N)), O.sub_solve(f, 1.0, 1.0, (N, N, N, N))) < 1e-12


def test_multi_rhs_and_ld(ctx):
    import torch
    import focalors_poisson_amd as fpa
    rng = np.random.default_rng(3)
    nx, ny, nrhs = 48, 30, 5
    f = rng.standard_normal((nrhs, nx, ny))
    s = fpa.PoissonSolver2D(ctx, nx, ny, 1.0 / 48, 1.0 / 48, (D, N, N, D))
    t = to_dev(f)
    banner("multi rhs")
    s.solve(t)
    torch.cuda.synchronize()
    got = t.cpu().numpy()
    for r in range(nrhs):
        assert rel(got[r], sub_reference(f[r], 1.0 / 48, 1.0 / 48, (D, N, N, D))) < 1e-12


@pytest.mark.parametrize("nx,ny,bc", [(64, 64, (D, D, D, D)), (100, 77, (N, D, D, N)), (256, 512, (D, N, N, N))])
def test_fp32_solve(ctx, nx, ny, bc):
    import torch
    rng = np.random.default_rng(5)
    h = 1.0 / max(nx, ny)
    f = rng.standard_normal((nx, ny)) * h * h
    banner("fp32 nx=%d ny=%d" % (nx, ny))
    got = gpu_solve(ctx, f, h, h, bc, dtype=torch.float32)
    assert rel(got, O.sub_solve(f.astype(np.float32).astype(np.float64), h, h, bc)) < 2e-5


def test_batched_uneven(ctx):
    import torch
    import focalors_poisson_amd as fpa
    from focalors_poisson_amd.cases import batched_sizes
    sizes = batched_sizes(24, seed=2, lo=8, hi=96)
    rng = np.random.default_rng(9)
    solvers, fields, hosts = [], [], []
    for nx, ny in sizes:
        h = 1.0 / max(nx, ny)
        solvers.append(fpa.PoissonSolver2D(ctx, nx, ny, h, h, (D, D, D, D)))
        f = rng.uniform(-1, 1, (nx, ny)) * h * h
        hosts.append(f)
        fields.append(to_dev(f))
    banner("batched uneven 24")
    fpa.solve_batched(ctx, solvers, fields)
    torch.cuda.synchronize()
    for (nx, ny), f, t in zip(sizes, hosts, fields):
        h = 1.0 / max(nx, ny)
        assert rel(t.cpu().numpy(), sub_reference(f, h, h, (D, D, D, D))) < 1e-12


@pytest.mark.parametrize("f32", [False, True])
def test_batched_mixed_kinds(ctx, f32):
    """Batches that span every transform kind, several large-prime codelets,
    both tridiagonal chunk lengths (nx above 1024), several right-hand sides
    per problem and more problems than one launch takes."""
    import torch
    import focalors_poisson_amd as fpa
    rng = np.random.default_rng(17 + f32)
    pairs = [(D, D), (N, N), (D, N), (N, D), (A, N), (N, A)]
    sizes = [(40, 22), (1100, 37), (64, 28), (33, 100), (1030, 64), (12, 255), (70, 45), (96, 96),
             (20, 511), (48, 130), (9, 18), (300, 60), (17, 200), (64, 1024)]
    solvers, fields, hosts, specs = [], [], [], []
    for k, (nx, ny) in enumerate(sizes):
        bc = (D if k % 3 else N, A if k % 2 else D) + pairs[k % len(pairs)]
        h = 1.0 / max(nx, ny)
        nrhs = 1 + k % 3
        f = rng.uniform(-1, 1, (nrhs, nx, ny)) * h * h
        solvers.append(fpa.PoissonSolver2D(ctx, nx, ny, h, h, bc))
        fields.append(to_dev(f, torch.float32 if f32 else None))
        hosts.append(f)
        specs.append((h, bc))
    banner("batched mixed kinds f32=%d" % f32)
    fpa.solve_batched(ctx, solvers, fields)
    torch.cuda.synchronize()
    bar = 2e-5 if f32 else 1e-12
    for f, t, (h, bc) in zip(hosts, fields, specs):
        got = t.double().cpu().numpy()
        for r in range(f.shape[0]):
            assert rel(got[r], sub_reference(f[r], h, h, bc)) < bar, (f.shape, bc)


def test_errors_are_status_codes(ctx):
    import focalors_poisson_amd as fpa
    with pytest.raises(fpa.FPError) as e:
        fpa.PoissonSolver2D(ctx, 1, 8, 1.0, 1.0, (D, D, D, D))
    assert e.value.status == 1
    with pytest.raises(fpa.FPError) as e:  # a periodic face without its partner
        fpa.PoissonSolver2D(ctx, 8, 8, 1.0, 1.0, (P, D, D, D))
    assert e.value.status == 1
    with pytest.raises(fpa.FPError) as e:  # periodic y needs an even ny
        fpa.PoissonSolver2D(ctx, 8, 7, 1.0, 1.0, (D, D, P, P))
    assert e.value.status == 1


@pytest.mark.parametrize("bcp,ny", [(b, n) for b in [(D, D), (N, N), (D, N), (N, D)] for n in [5, 64, 100, 256, 1024, 2048]]
                         + [((D, D), 4096), ((N, N), 4096)])
def test_transforms_match_reference_coefficients(ctx, bcp, ny):
    """PoissonFFT2D_*::transform / transform_transpose coefficient for
    coefficient (same signs and mode order as the reference)."""
    import torch
    import focalors_poisson_amd as fpa
    rng = np.random.default_rng(ny)
    nx = 7
    f = rng.standard_normal((nx, ny))
    s = fpa.PoissonSolver2D(ctx, nx, ny, 1.0, 1.0, (D, D) + bcp)
    for inv in (False, True):
        t = to_dev(f)
        banner("transform bc=%s ny=%d inverse=%d" % (bcp, ny, inv))
        s.transform(t, inverse=inv)
        torch.cuda.synchronize()
        ref = O.transform(bcp[0], bcp[1], f, inverse=inv)
        assert np.abs(t.cpu().numpy() - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("bcp,ny", [(b, n) for b in [(D, D), (N, N), (D, N), (N, D)]
                                     for n in [1, 2, 3, 7, 33, 127, 255, 257, 511, 1023, 1025]])
def test_transform_coefficients_edge_lengths(ctx, bcp, ny):
    """The same coefficient check at lengths 1-3 and odd lengths around powers
    of two: the DCT pair's even/odd reordering with odd N, the DST-I pair
    loads with even and odd L = N + 1 (the middle entry of even L), and the
    lengths whose FFT goes to Bluestein (L = 257, 2 x 257 + 1, ...)."""
    import torch
    import focalors_poisson_amd as fpa
    rng = np.random.default_rng(1000 + ny)
    nx = 5
    f = rng.standard_normal((nx, ny))
    s = fpa.PoissonSolver2D(ctx, nx, ny, 1.0, 1.0, (D, D) + bcp)
    for inv in (False, True):
        t = to_dev(f)
        banner("transform edge bc=%s ny=%d inverse=%d" % (bcp, ny, inv))
        s.transform(t, inverse=inv)
        torch.cuda.synchronize()
        ref = O.transform(bcp[0], bcp[1], f, inverse=inv)
        assert np.abs(t.cpu().numpy() - ref).max() < 1e-12 * max(1.0, np.abs(ref).max()), (bcp, ny, inv)


@pytest.mark.parametrize("form", [1, 2, 3])
@pytest.mark.parametrize("bcp,ny", [((D, D), 1024), ((D, N), 512), ((N, D), 512), ((D, D), 1022), ((N, N), 1025),
                                    ((D, D), 40)])
def test_odd_prime_pass_forms(ctx, monkeypatch, form, bcp, ny):
    """Each odd-prime pass form, forced through FP_ODD_FORM (1: matrix
    cores, 2 / 3: VALU pair sums with 5 / 3 output pairs per work item), on
    lengths with prime factors 41 (L = 1025, 41), 31 and 11 (L = 1023, the
    second pass twiddled): coefficients against the oracle both ways, and a
    solve."""
    import torch
    import focalors_poisson_amd as fpa
    monkeypatch.setenv("FP_ODD_FORM", str(form))
    rng = np.random.default_rng(7 * ny + form)
    nx = 6
    f = rng.standard_normal((nx, ny))
    s = fpa.PoissonSolver2D(ctx, nx, ny, 1.0, 1.0, (D, D) + bcp)
    for inv in (False, True):
        t = to_dev(f)
        banner("odd form %d bc=%s ny=%d inverse=%d" % (form, bcp, ny, inv))
        s.transform(t, inverse=inv)
        torch.cuda.synchronize()
        ref = O.transform(bcp[0], bcp[1], f, inverse=inv)
        assert np.abs(t.cpu().numpy() - ref).max() < 1e-12 * max(1.0, np.abs(ref).max()), (form, bcp, ny, inv)
    h = 1.0 / ny
    g = f * h * h
    assert rel(gpu_solve(ctx, g, h, h, (D, D) + bcp), sub_reference(g, h, h, (D, D) + bcp)) < 1e-12


def test_long_prime_47_length(ctx):
    """L = N + 1 = 3055 = 5 x 13 x 47 with 65 radix-47 butterflies: more work
    items than the VALU odd-prime forms fit in one workgroup, so the FP64
    transform takes the matrix-core form and the FP32 one the two-buffer
    engine; coefficients against the oracle, and an FP32+refinement solve."""
    import torch
    import focalors_poisson_amd as fpa
    rng = np.random.default_rng(3054)
    ny = 3054
    f = rng.standard_normal((5, ny))
    s = fpa.PoissonSolver2D(ctx, 5, ny, 1.0, 1.0, (D, D, D, D))
    for inv in (False, True):
        t = to_dev(f)
        banner("transform ny=3054 inverse=%d" % inv)
        s.transform(t, inverse=inv)
        torch.cuda.synchronize()
        ref = O.transform(D, D, f, inverse=inv)
        assert np.abs(t.cpu().numpy() - ref).max() < 1e-12 * max(1.0, np.abs(ref).max()), inv
    h = 1.0 / ny
    g = rng.standard_normal((8, ny)) * h * h
    s2 = fpa.PoissonSolver2D(ctx, 8, ny, h, h, (D, D, D, D))
    t = to_dev(g)
    banner("fp32 refine ny=3054")
    s2.solve_refine(t, tol=1e-13, max_steps=10)
    torch.cuda.synchronize()
    assert rel(t.cpu().numpy(), sub_reference(g, h, h, (D, D, D, D))) < 1e-12


@pytest.mark.parametrize("form", [2, 3])
def test_odd_prime_forms_fp32_refine(ctx, monkeypatch, form):
    """The FP32 transforms of the refinement path in both VALU forms (the
    matrix-core form is FP64 only): 1024 DD rows (L = 1025)."""
    import torch
    import focalors_poisson_amd as fpa
    monkeypatch.setenv("FP_ODD_FORM", str(form))
    rng = np.random.default_rng(90 + form)
    nx, ny = 48, 1024
    h = 1.0 / ny
    f = rng.standard_normal((nx, ny)) * h * h
    s = fpa.PoissonSolver2D(ctx, nx, ny, h, h, (D, D, D, D))
    t = to_dev(f)
    banner("odd form %d fp32 refine" % form)
    s.solve_refine(t, tol=1e-13, max_steps=10)
    torch.cuda.synchronize()
    assert rel(t.cpu().numpy(), sub_reference(f, h, h, (D, D, D, D))) < 1e-12


@pytest.mark.parametrize("nx,ny,bc", [(64, 64, (D, D, D, D)), (200, 131, (N, D, A, N)), (512, 1024, (D, A, N, N)),
                                      (1024, 1024, (A, A, A, A))])
def test_fp32_refine_reaches_fp64_oracle(ctx, nx, ny, bc):
    """FP32 inner solves + FP64 refinement: every step's error vs the FP64
    oracle is printed; the final one meets the FP64 bar of 1e-12 against the
    converged oracle."""
    import torch
    import focalors_poisson_amd as fpa
    rng = np.random.default_rng(nx + 3 * ny)
    h = 1.0 / max(nx, ny)
    f = rng.standard_normal((nx, ny)) * h * h
    ref = sub_reference(f, h, h, bc)
    s = fpa.PoissonSolver2D(ctx, nx, ny, h, h, bc)
    errs = []
    for steps in (1, 2, 3, 6):
        t = to_dev(f)
        info = s.solve_refine(t, tol=0.0, max_steps=steps, allow_not_converged=True)
        torch.cuda.synchronize()
        errs.append(rel(t.cpu().numpy(), ref))
        banner("refine nx=%d ny=%d steps=%d rel err %.3e (correction %.2e)" % (nx, ny, steps, errs[-1],
                                                                            info["refine_residual"]))
    assert errs[0] < 1e-3
    assert errs[1] < errs[0]
    t = to_dev(f)
    info = s.solve_refine(t, tol=1e-13, max_steps=40)
    torch.cuda.synchronize()
    err = rel(t.cpu().numpy(), ref)
    banner("refine nx=%d ny=%d converged in %d steps, rel err %.3e" % (nx, ny, info["refine_steps"], err))
    assert err < 1e-12 and info["refine_residual"] <= 1e-13


# ---- periodic y (PoissonFFT2D_PP, poisson_fft2d.cpp:264-465) ----

@pytest.mark.parametrize("ny", [2, 4, 64, 100, 256, 514, 1024, 2048])
def test_periodic_transform_coefficients(ctx, ny):
    """PoissonFFT2D_PP::transform / transform_transpose coefficient for
    coefficient (mode 0 the Nyquist mode, mode ny-1 the constant); 514 =
    2 x 257 runs by Bluestein."""
    import torch
    import focalors_poisson_amd as fpa
    rng = np.random.default_rng(ny + 5)
    nx = 7
    f = rng.standard_normal((nx, ny))
    s = fpa.PoissonSolver2D(ctx, nx, ny, 1.0, 1.0, (D, D, P, P))
    for inv in (False, True):
        t = to_dev(f)
        banner("periodic transform ny=%d inverse=%d" % (ny, inv))
        s.transform(t, inverse=inv)
        torch.cuda.synchronize()
        ref = O.transform(P, P, f, inverse=inv)
        assert np.abs(t.cpu().numpy() - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("nx,ny,bc", [(9, 8, (D, D, P, P)), (9, 8, (N, D, P, P)), (9, 8, (A, N, P, P)),
                                      (33, 100, (D, N, P, P)), (64, 514, (D, D, P, P)),
                                      (512, 1024, (N, D, P, P)), (256, 256, (D, D, P, P))])
def test_periodic_y_solve(ctx, nx, ny, bc):
    rng = np.random.default_rng(nx + ny)
    h = 1.0 / max(nx, ny)
    f = rng.standard_normal((nx, ny)) * h * h
    banner("periodic y nx=%d ny=%d bc=%s" % (nx, ny, bc))
    assert rel(gpu_solve(ctx, f, h, h, bc), sub_reference(f, h, h, bc)) < 1e-12


def test_periodic_y_singular_gauge(ctx):
    """Neumann x ends with periodic y: the constant mode (last) takes the
    reference's x0 = 0 gauge; compatible right-hand side."""
    rng = np.random.default_rng(41)
    nx, ny, h = 12, 16, 1.0 / 16
    f = rng.standard_normal((nx, ny))
    f -= f.mean()
    f *= h * h
    bc = (N, N, P, P)
    banner("periodic y singular")
    err = rel(gpu_solve(ctx, f, h, h, bc), O.sub_solve(f, h, h, bc))
    banner("periodic y singular rel-L2 %.3e" % err)
    assert err < 1e-12


# ---- periodic x (ChasingMethodBase::chasing_periodic_*, chasing_method_base.cpp:178-285) ----

@pytest.mark.parametrize("nx,ny,bc", [(2, 5, (P, P, D, D)), (3, 4, (P, P, D, N)), (9, 7, (P, P, D, D)),
                                      (9, 7, (P, P, N, D)), (9, 8, (P, P, A, N)), (64, 257, (P, P, D, D)),
                                      (1024, 37, (P, P, D, N)), (512, 1024, (P, P, D, D)), (100, 64, (P, P, N, D)),
                                      (48, 64, (P, P, D, D))])
def test_periodic_x_solve(ctx, nx, ny, bc):
    rng = np.random.default_rng(nx * 7 + ny)
    h = 1.0 / max(nx, ny)
    f = rng.standard_normal((nx, ny)) * h * h
    banner("periodic x nx=%d ny=%d bc=%s" % (nx, ny, bc))
    err = rel(gpu_solve(ctx, f, h, h, bc), sub_reference(f, h, h, bc))
    banner("periodic x rel-L2 %.3e" % err)
    assert err < 1e-12


@pytest.mark.parametrize("nx,ny,bc", [(12, 16, (P, P, P, P)), (12, 9, (P, P, N, N)), (256, 256, (P, P, P, P))])
def test_periodic_x_singular_gauge(ctx, nx, ny, bc):
    """No Dirichlet face: the constant mode of the periodic x axis takes the
    reference's x0 = 0 gauge (compatible right-hand side)."""
    rng = np.random.default_rng(nx + 3 * ny)
    h = 1.0 / max(nx, ny)
    f = rng.standard_normal((nx, ny))
    f -= f.mean()
    f *= h * h
    banner("periodic x singular nx=%d ny=%d bc=%s" % (nx, ny, bc))
    err = rel(gpu_solve(ctx, f, h, h, bc), O.sub_solve(f, h, h, bc))
    banner("periodic x singular rel-L2 %.3e" % err)
    assert err < 1e-12


@pytest.mark.parametrize("bc", [(P, P, D, N), (D, N, P, P)])
def test_periodic_fp32_refine(ctx, bc):
    """FP32 inner solves with the FP64 residual of the periodic operator."""
    import torch
    import focalors_poisson_amd as fpa
    rng = np.random.default_rng(77)
    nx, ny = 256, 128
    h = 1.0 / 256
    f = rng.standard_normal((nx, ny)) * h * h
    banner("periodic fp32 refine bc=%s" % (bc,))
    s = fpa.PoissonSolver2D(ctx, nx, ny, h, h, bc)
    t = to_dev(f)
    info = s.solve_refine(t, tol=1e-13, max_steps=12)
    torch.cuda.synchronize()
    assert rel(t.cpu().numpy(), sub_reference(f, h, h, bc)) < 1e-12, info


def test_batched_sixty_with_periodic(ctx):
    """Sixty problems in one call (device-resident descriptor tables: one
    launch per kernel group), every boundary family including periodic axes."""
    import torch
    import focalors_poisson_amd as fpa
    rng = np.random.default_rng(61)
    fams = [(D, D, D, D), (N, D, N, N), (D, N, P, P), (P, P, D, D), (P, P, N, D), (A, N, D, N), (N, A, P, P)]
    solvers, fields, hosts, specs = [], [], [], []
    for k in range(60):
        bc = fams[k % len(fams)]
        nx = int(rng.integers(8, 300))
        ny = int(rng.integers(4, 300)) // 2 * 2
        h = 1.0 / max(nx, ny)
        f = rng.standard_normal((nx, ny)) * h * h
        solvers.append(fpa.PoissonSolver2D(ctx, nx, ny, h, h, bc))
        fields.append(to_dev(f))
        hosts.append(f)
        specs.append((h, bc))
    banner("batched sixty with periodic")
    fpa.solve_batched(ctx, solvers, fields)
    torch.cuda.synchronize()
    for f, t, (h, bc) in zip(hosts, fields, specs):
        assert rel(t.cpu().numpy(), sub_reference(f, h, h, bc)) < 1e-12, (f.shape, bc)


def test_batched_tables_queued_back_to_back(ctx):
    """Two different batches of twenty problems (descriptor tables in device
    memory, staged through pinned host blocks that are reused once their copy
    has run), called alternately ten times without synchronising, so that
    uploads are queued while earlier ones are still in flight: every field
    of the last call of each batch matches its own oracle solve."""
    import torch
    import focalors_poisson_amd as fpa
    rng = np.random.default_rng(62)
    batches = []
    for b in range(2):
        solvers, hosts = [], []
        for k in range(20):
            nx = int(rng.integers(16, 200))
            ny = int(rng.integers(16, 200))
            h = 1.0 / max(nx, ny)
            bc = (D, D, D, D) if b == 0 else (N, D, D, N)
            solvers.append(fpa.PoissonSolver2D(ctx, nx, ny, h, h, bc))
            hosts.append((rng.standard_normal((nx, ny)) * h * h, h, bc))
        batches.append((solvers, hosts))
    banner("batched tables queued back to back")
    # every call's fields on the device first, so that the calls themselves
    # queue without a host-side copy in between
    sets = [[to_dev(f) for f, _, _ in batches[it % 2][1]] for it in range(10)]
    torch.cuda.synchronize()
    for it in range(10):
        fpa.solve_batched(ctx, batches[it % 2][0], sets[it])
    torch.cuda.synchronize()
    for it in range(10):
        for t, (f, h, bc) in zip(sets[it], batches[it % 2][1]):
            assert rel(t.cpu().numpy(), sub_reference(f, h, h, bc)) < 1e-12, (it, f.shape)


@pytest.mark.parametrize("f32", [False, True])
def test_dirichlet_4096_long_bluestein(ctx, f32):
    """4096 DD rows: L = 4097 = 17 x 241, Bluestein over F = 8640 with 24
    values per thread; the whole 4096^2 solve against the oracle."""
    import torch
    rng = np.random.default_rng(4096)
    n = 4096
    h = 1.0 / n
    f = rng.standard_normal((n, n)) * h * h
    banner("dirichlet 4096 f32=%d" % f32)
    got = gpu_solve(ctx, f, h, h, (D, D, D, D), torch.float32 if f32 else None)
    banner("dirichlet 4096 f32=%d solved on the device" % f32)
    err = rel(got, sub_reference(f, h, h, (D, D, D, D)))
    banner("dirichlet 4096 f32=%d rel-L2 vs the converged oracle %.3e" % (f32, err))
    assert err < (1e-4 if f32 else 1e-12)
