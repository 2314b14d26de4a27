"""Composite solve parity on the device (ConcatPoissonSolver2D::solve,
pe/concat/concat_solver2d.cpp:101-157): golden fixtures from the assembled
sparse system and the converged CPU oracle (GMRES to 1e-13 plus refinement
with the long-double composite residual, O.reference_solution).  Interface
GMRES to relative 1e-13.  Bar: relative L2 <= 1e-12 against the discrete
solution for the FP64 path (FP_F64 up to n = 256, where its rounding floor
is at most 4e-13; FP_F64_REFINE everywhere), and for FP32 sweeps with FP64
iterative refinement (FP_F32_REFINE, refinement target 1e-13)."""
import os

import numpy as np
import pytest

import oracle as O
from gpu_util import banner, device, rel, to_dev

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
D, N, A = 0, 1, 3


@pytest.fixture(scope="module")
def ctx():
    device()
    import focalors_poisson_amd as fpa
    return fpa.Context(0)


BAR = 1e-12


def run(ctx, doms, conn, fields, ghost=None, m=40, tol=1e-13, rel_tol=True, max_iter=200, precision=0,
        allow_not_converged=False, refine_tol=1e-13):
    import torch
    import focalors_poisson_amd as fpa
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=m, tol=tol, tol_is_relative=rel_tol, max_iter=max_iter,
                                    precision=precision)
    if precision == 1:
        cat.set_refinement(tol=refine_tol, max_steps=10)
    ts = [to_dev(f) for f in fields]
    gs = {k: to_dev(v) for k, v in (ghost or {}).items()}
    info = cat.solve(ts, gs, allow_not_converged=allow_not_converged)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in ts], info, cat


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_cross32_golden(ctx, precision):
    g = np.load(os.path.join(GOLD, "cross_32.npz"))
    h = float(g["h"])
    doms = [(int(a), int(b), h, h, tuple(int(t) for t in bc)) for (a, b), bc in zip(g["sizes"], g["bcs"])]
    conn = [tuple(int(v) for v in c) for c in g["conn"]]
    fs = [g["f%d" % k] for k in range(5)]
    ghost = {}
    for k in g.files:
        if k.startswith("g"):
            d, face = k[1:].split("_")
            ghost[(int(d), int(face))] = g[k]
    banner("cross32 golden precision=%d" % precision)
    us, info, cat = run(ctx, doms, conn, fs, ghost, precision=precision)
    assert cat.root == 0
    assert rel(us, [g["u%d" % k] for k in range(5)]) < BAR, info


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_concat_six_golden(ctx, precision):
    g = np.load(os.path.join(GOLD, "concat_six.npz"))
    sizes = [tuple(int(v) for v in s) for s in g["sizes"]]
    doms = [(a, b, 1.0, 1.0, (D, D, D, D)) for a, b in sizes]
    conn = [tuple(int(v) for v in c) for c in g["conn"]]
    fs = [g["f%d" % k] for k in range(6)]
    banner("concat six golden (depth-2 tree) precision=%d" % precision)
    us, info, cat = run(ctx, doms, conn, fs, precision=precision)
    assert rel(us, [g["u%d" % k] for k in range(6)]) < BAR, info


@pytest.mark.parametrize("n,precision", [(16, 0), (64, 0), (128, 0), (256, 0), (64, 2), (256, 2), (256, 1)])
def test_cross_vs_oracle(ctx, n, precision):
    from focalors_poisson_amd.cases import cross
    doms, conn, fields, _ = cross(n, seed=1)
    fs = [f[0] for f in fields]
    banner("cross n=%d precision=%d" % (n, precision))
    us, info, _ = run(ctx, doms, conn, fs, precision=precision)
    ref = O.reference_solution(doms, conn, fs)
    err = rel(us, ref)
    print("cross n=%d precision=%d rel-L2 vs oracle %.3e" % (n, precision, err), info)
    assert err < BAR, info


@pytest.mark.parametrize("n,nrhs", [(64, 1), (256, 1), (128, 3)])
def test_cross_interface_gmres_without_explicit_inverse(ctx, monkeypatch, n, nrhs):
    """The reference's interface method (gmres_solver2d.cpp:123-252): GMRES
    over the Schur system from x0 = C with no explicit-inverse start
    (FP_INTERFACE_PRECOND=0), Krylov steps by the line-sourced operator, to
    relative 1e-13: the converged solution of every right-hand side."""
    from focalors_poisson_amd.cases import cross
    monkeypatch.setenv("FP_INTERFACE_PRECOND", "0")
    doms, conn, fields, _ = cross(n, seed=21, nrhs=nrhs)
    banner("cross interface GMRES n=%d nrhs=%d" % (n, nrhs))
    us, info, cat = run(ctx, doms, conn, fields if nrhs > 1 else [f[0] for f in fields], m=200)
    assert info["iterations"] > 5, info
    for r in range(nrhs):
        fs = [f[r] for f in fields]
        got = [u[r] for u in us] if nrhs > 1 else us
        err = rel(got, O.reference_solution(doms, conn, fs))
        print("cross %d rhs %d: %d Krylov steps, rel-L2 %.2e" % (n, r, info["iterations"], err))
        assert err < BAR, (r, info)
    import torch
    import focalors_poisson_amd as fpa
    s = torch.cuda.Stream()
    with pytest.raises(fpa.FPError, match="explicit-inverse"):  # a captured solve needs the explicit-inverse start
        cat.graph_capture([to_dev(f[0]) for f in fields], stream=s)


def test_cross_multi_rhs(ctx):
    from focalors_poisson_amd.cases import cross
    n, nrhs = 32, 4
    doms, conn, fields, _ = cross(n, seed=100, nrhs=nrhs)
    banner("cross multi rhs")
    us, info, _ = run(ctx, doms, conn, fields)
    for r in range(nrhs):
        ref = O.reference_solution(doms, conn, [f[r] for f in fields])
        assert rel([u[r] for u in us], ref) < BAR


def test_single_domain_through_composite(ctx):
    from focalors_poisson_amd.cases import single
    doms, conn, fields, _ = single(64)
    banner("single domain composite")
    us, info, _ = run(ctx, doms, conn, fields)
    ref = O.reference_solution(doms, conn, fields)
    assert rel(us, ref) < BAR


def test_not_converged_status(ctx):
    import focalors_poisson_amd as fpa
    from focalors_poisson_amd.cases import cross
    doms, conn, fields, _ = cross(32, seed=1)
    banner("not converged status")
    # an unreachable tolerance: the explicit-inverse start is accurate to
    # rounding, so only a tolerance below rounding keeps GMRES from converging
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=2, tol=1e-30, tol_is_relative=True, max_iter=1)
    with pytest.raises(fpa.FPError) as e:
        cat.solve([to_dev(f[0]) for f in fields])
    assert e.value.status == 4


@pytest.mark.parametrize("n", [32, 128, 256])
def test_cross_f32_refine_vs_oracle(ctx, n):
    """FP32 sweeps + FP64 refinement reach the FP64 oracle solution."""
    from focalors_poisson_amd.cases import cross
    doms, conn, fields, _ = cross(n, seed=1)
    fs = [f[0] for f in fields]
    banner("cross f32+refine n=%d" % n)
    us, info, _ = run(ctx, doms, conn, fs, precision=1)
    ref = O.reference_solution(doms, conn, fs)
    assert rel(us, ref) < BAR, info
    assert info["refine_steps"] >= 2 and info["refine_residual"] <= 1e-13, info
    # the estimate is floored at what an FP64 residual resolves
    assert info["refine_residual"] >= 64 * 2.220446049250313e-16, info


def test_cross_multi_rhs_f32_refine(ctx):
    from focalors_poisson_amd.cases import cross
    n, nrhs = 32, 3
    doms, conn, fields, _ = cross(n, seed=200, nrhs=nrhs)
    banner("cross multi rhs f32+refine")
    us, info, _ = run(ctx, doms, conn, fields, precision=1)
    for r in range(nrhs):
        ref = O.reference_solution(doms, conn, [f[r] for f in fields])
        assert rel([u[r] for u in us], ref) < BAR


def test_refine_cap_reports_not_converged(ctx):
    import focalors_poisson_amd as fpa
    from focalors_poisson_amd.cases import cross
    doms, conn, fields, _ = cross(32, seed=1)
    banner("refine cap status")
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=40, tol=1e-13, tol_is_relative=True, precision=1)
    cat.set_refinement(tol=1e-30, max_steps=2)
    info = cat.solve([to_dev(f[0]) for f in fields], allow_not_converged=True)
    assert info["status"] == 4 and info["refine_steps"] == 2
    with pytest.raises(fpa.FPError):
        cat.set_refinement(tol=0.0)


def test_restart_after_start_check(ctx):
    """A tolerance below the explicit-inverse start's residual: the first
    check (taken from the root's final sweep) fails, GMRES iterates and the
    root's final sweep runs again; the result still matches the oracle."""
    from focalors_poisson_amd.cases import cross
    doms, conn, fields, _ = cross(64, seed=3)
    fs = [f[0] for f in fields]
    banner("restart after the start check")
    _, info0, _ = run(ctx, doms, conn, fs)
    assert info0["iterations"] == 0, info0
    us, info, _ = run(ctx, doms, conn, fs, tol=info0["rel_residual"] / 4, max_iter=3,
                      allow_not_converged=True)
    assert info["iterations"] >= 1, info
    ref = O.reference_solution(doms, conn, fs)
    assert rel(us, ref) < BAR, info


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_lshape_vs_oracle(ctx, precision):
    """L of three Dirichlet squares (children on the root's x+ and y+ faces)
    with a multiscale right-hand side, FP64 and FP32+refine."""
    from focalors_poisson_amd.cases import lshape
    doms, conn, fields, _ = lshape(96, seed=4)
    banner("lshape 3x96 precision=%d" % precision)
    us, info, _ = run(ctx, doms, conn, fields, precision=precision)
    ref = O.reference_solution(doms, conn, fields)
    assert rel(us, ref) < BAR, info


def test_cross_multi_rhs_split_gemm(ctx):
    """Eight right-hand sides at n = 256: the explicit-inverse and Schur
    products run as split-k GEMMs (k in 2..8 chunks); every right-hand side
    matches its own one-right-hand-side solve (matrix-vector kernels) and the
    oracle."""
    from focalors_poisson_amd.cases import cross
    n, nrhs = 256, 8
    doms, conn, fields, _ = cross(n, seed=300, nrhs=nrhs)
    banner("cross multi rhs split gemm")
    us, info, _ = run(ctx, doms, conn, fields, tol=1e-12)
    for r in range(nrhs):
        one, _, _ = run(ctx, doms, conn, [f[r] for f in fields], tol=1e-12)
        assert rel([u[r] for u in us], one) < 1e-12, r
    ref = O.reference_solution(doms, conn, [f[0] for f in fields])
    assert rel([u[0] for u in us], ref) < BAR, info


@pytest.mark.parametrize("nrhs", [5, 19, 37, 70])
def test_ragged_multi_rhs_matrix_core_products(ctx, nrhs):
    """Right-hand-side counts that take one, two and four 16-wide tiles of the
    matrix-core products (and two tile groups at 70), on a T-shape whose
    interface lengths (45 rows, 37 and 53 columns) are not multiples of 16:
    every right-hand side against its own one-right-hand-side solve (the
    matrix-vector kernels) and, for three of them, the oracle."""
    rng = np.random.default_rng(500 + nrhs)
    h = 1.0 / 53
    doms = [(37, 45, h, h, (A, A, D, A)), (21, 45, h, h, (N, A, D, N)), (29, 45, h, h, (A, D, N, D)),
            (37, 30, h, h, (D, N, A, N))]
    conn = [(0, 0, 1), (0, 1, 2), (0, 3, 3)]
    fs = [rng.standard_normal((nrhs, d[0], d[1])) for d in doms]
    banner("ragged multi rhs %d" % nrhs)
    us, info, _ = run(ctx, doms, conn, fs, tol=1e-13)
    for r in sorted({0, nrhs // 2, nrhs - 1}):
        one, _, _ = run(ctx, doms, conn, [f[r] for f in fs], tol=1e-13)
        assert rel([u[r] for u in us], one) < 1e-12, r
        ref = O.reference_solution(doms, conn, [f[r] for f in fs])
        assert rel([u[r] for u in us], ref) < BAR, (r, info)


P = 2


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_periodic_y_channel_two_level(ctx, precision):
    """Two domains side by side, both periodic in y (PP transforms): the
    two-level transform-space path with x-lines in the periodic basis."""
    rng = np.random.default_rng(51)
    h = 1.0 / 64
    doms = [(48, 64, h, h, (D, A, P, P)), (40, 64, h, h, (A, N, P, P))]
    conn = [(0, 1, 1)]
    fs = [rng.standard_normal((d[0], d[1])) for d in doms]
    banner("periodic y channel precision=%d" % precision)
    us, info, _ = run(ctx, doms, conn, fs, precision=precision)
    ref = O.reference_solution(doms, conn, fs)
    assert rel(us, ref) < BAR, info
    import discrete
    assert rel(us, discrete.solve(doms, conn, fs)) < BAR


def test_periodic_x_root_with_child(ctx):
    """A root periodic in x with a child on its y+ face (also periodic in x):
    the general sweep with periodic Thomas, full-field interface matvec."""
    rng = np.random.default_rng(52)
    h = 1.0 / 64
    doms = [(64, 32, h, h, (P, P, D, A)), (64, 24, h, h, (P, P, A, D))]
    conn = [(0, 3, 1)]
    fs = [rng.standard_normal((d[0], d[1])) for d in doms]
    banner("periodic x root")
    us, info, _ = run(ctx, doms, conn, fs, m=80)
    import discrete
    assert rel(us, discrete.solve(doms, conn, fs)) < BAR, info
    ref = O.reference_solution(doms, conn, fs)
    assert rel(us, ref) < BAR, info


@pytest.mark.parametrize("shape", ["8,1", "8,4", "16,8", "16,16", "32,8"])
def test_cross_sweep_shapes(ctx, monkeypatch, shape):
    """The composite sweep at forced chunk lengths and modes per workgroup
    (FP_TRI_SHAPE): the same solution as the oracle at each shape."""
    from focalors_poisson_amd.cases import cross
    monkeypatch.setenv("FP_TRI_SHAPE", shape)
    doms, conn, fields, _ = cross(128, seed=5)
    fs = [f[0] for f in fields]
    banner("cross sweep shape %s" % shape)
    us, info, _ = run(ctx, doms, conn, fs)
    ref = O.reference_solution(doms, conn, fs)
    assert rel(us, ref) < BAR, (shape, info)


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_tshape_unequal_children(ctx, precision):
    """A non-square root (96 x 64) with children of different depths on both
    x faces (40 and 56 rows) and a 48-row child on its y+ face, mixed
    Dirichlet/Neumann outer faces: the two-level path against the oracle and
    the assembled sparse system, FP64 and FP32+refine."""
    rng = np.random.default_rng(77)
    h = 1.0 / 64
    doms = [(96, 64, h, h, (A, A, D, A)), (40, 64, h, h, (N, A, D, N)), (56, 64, h, h, (A, D, N, D)),
            (96, 48, h, h, (D, N, A, N))]
    conn = [(0, 0, 1), (0, 1, 2), (0, 3, 3)]
    fs = [rng.standard_normal((d[0], d[1])) for d in doms]
    banner("tshape precision=%d" % precision)
    us, info, _ = run(ctx, doms, conn, fs, precision=precision)
    ref = O.reference_solution(doms, conn, fs)
    assert rel(us, ref) < BAR, info
    import discrete
    assert rel(us, discrete.solve(doms, conn, fs)) < BAR


def test_forced_batch_shape_applies_only_where_it_fits(ctx, monkeypatch):
    """A shared transform shape of 128 threads does not fit the 4096-point
    root transform (its radix-8 passes need 256 threads or more) but fits the
    short child: forced with FP_BATCH_SHAPE, the root keeps its own shape and
    the solve still matches the oracle."""
    monkeypatch.setenv("FP_BATCH_SHAPE", "128,-1")
    rng = np.random.default_rng(91)
    h = 1.0 / 64
    doms = [(32, 4095, h, h, (D, D, D, A)), (32, 48, h, h, (D, D, A, N))]
    conn = [(0, 3, 1)]
    fs = [rng.standard_normal((d[0], d[1])) for d in doms]
    banner("forced batch shape 128 on a 4096-point root")
    us, info, _ = run(ctx, doms, conn, fs)
    ref = O.reference_solution(doms, conn, fs)
    assert rel(us, ref) < BAR, info


def test_single_domain_all_neumann_through_composite(ctx):
    """One all-Neumann domain through ConcatPoissonSolver2D: the sweep gauges
    the zero mode (x_0 = 0, chasing_method_base.cpp:146-154) as the
    oracle does, so the results agree cell for cell."""
    rng = np.random.default_rng(92)
    doms = [(40, 56, 1.0 / 56, 1.0 / 56, (N, N, N, N))]
    f = rng.standard_normal((40, 56))
    f -= f.mean()
    banner("single all-Neumann domain through the composite solver")
    us, info, _ = run(ctx, doms, [], [f])
    ref, _, _ = O.Composite(doms, []).solve([f])
    assert rel(us, ref) < BAR, info


def test_all_neumann_cross(ctx):
    """A cross with no Dirichlet face: the composite operator's null space is
    the constants.  No explicit interface inverse exists; the interface GMRES
    runs unpreconditioned (as the reference's) on a compatible right-hand
    side.  The solution solves the assembled system and equals the oracle's
    up to a constant."""
    import focalors_poisson_amd as fpa
    from focalors_poisson_amd.cases import cross_geometry
    doms, conn, _ = cross_geometry(32)
    doms = [(d[0], d[1], d[2], d[3], tuple(N if t == D else t for t in d[4])) for d in doms]
    rng = np.random.default_rng(93)
    fs = [rng.standard_normal((d[0], d[1])) for d in doms]
    mean = sum(f.sum() for f in fs) / sum(f.size for f in fs)
    fs = [f - mean for f in fs]
    banner("all-Neumann cross (singular composite)")
    us, info, _ = run(ctx, doms, conn, fs, m=80, max_iter=400)
    ocat = O.Composite(doms, conn, m=80, tol=1e-13, tol_is_relative=True, maxiter=400,
                       schur_mode=O.SCHUR_SPECTRAL)
    ref, _, _ = ocat.solve(fs)
    cu = sum(u.sum() for u in us) / sum(u.size for u in us)
    cr = sum(r.sum() for r in ref) / sum(r.size for r in ref)
    assert rel([u - cu for u in us], [r - cr for r in ref]) < 1e-10, info
    res = ocat.residual(us, fs)
    bn = np.sqrt(sum(float(np.sum((f * doms[0][2] ** 2) ** 2)) for f in fs))
    assert np.sqrt(sum(float(np.sum(r * r)) for r in res)) < 1e-10 * bn
    # refinement needs a nonsingular operator
    with pytest.raises(fpa.FPError) as e:
        fpa.ConcatPoissonSolver2D(ctx, doms, conn, precision=fpa.FP_F64_REFINE)
    assert e.value.status == 6


@pytest.mark.parametrize("inverse", [False, True])
def test_composite_transform_matches_each_domain(ctx, inverse):
    """fp_cat_transform (every domain's y transform in the solve's one batched
    launch) equals each domain's own PoissonFFT2D::transform /
    transform_transpose (fp_sub_transform) and the oracle's, coefficient for
    coefficient, on two right-hand sides."""
    import torch
    import focalors_poisson_amd as fpa
    from focalors_poisson_amd.cases import cross
    doms, conn, fields, _ = cross(96, seed=13, nrhs=2)
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=40, tol=1e-12, tol_is_relative=True)
    ts = [to_dev(f) for f in fields]
    banner("composite batched transform inverse=%d" % inverse)
    cat.transform(ts, inverse=inverse)
    torch.cuda.synchronize()
    eff = cat_bcs(doms, conn)
    for d, (dom, f, t) in enumerate(zip(doms, fields, ts)):
        sub = fpa.PoissonSolver2D(ctx, dom[0], dom[1], dom[2], dom[3], eff[d])
        one = to_dev(f)
        sub.transform(one, inverse=inverse)
        torch.cuda.synchronize()
        got = t.cpu().numpy()
        assert np.abs(got - one.cpu().numpy()).max() <= 1e-13 * np.abs(got).max(), d
        for r in range(2):
            ref = O.transform(eff[d][2], eff[d][3], f[r], inverse=inverse)
            assert np.abs(got[r] - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), (d, r)


def cat_bcs(doms, conn):
    bcs = [list(d[4]) for d in doms]
    for a, dr, b in conn:
        bcs[a][dr] = A
        bcs[b][dr ^ 1] = A
    return [tuple(b) for b in bcs]
