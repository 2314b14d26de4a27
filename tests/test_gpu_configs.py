"""The five benchmark workloads of the solve path at their full sizes, on the
device, against converged CPU references that reach the discrete solution to
about FP64 rounding: the oracle's restatement of the reference's algorithm
refined with the long-double composite residual for the single domains, and
the direct solve of oracle/direct.py (closed-form bases in both directions,
exact elimination of the interface, long-double residual refinement) for the
full-size cross and L, where the reference's own algorithm takes minutes.
tests/test_oracle.py pins the direct solve against that algorithm and the
assembled sparse system:

  [0] one 256^2 Dirichlet domain, manufactured right-hand side;
  [1] the cross (centre 1024^2, arms 512 x 1024 / 1024 x 512), FP64, FP64 with
      one refinement step, FP32 sweeps with FP64 refinement;
  [2] 512 uneven Dirichlet problems in one call, FP64 and FP32, and the cross
      with 64 right-hand sides in one call;
  [3] 4096^2 Dirichlet and the L of three 2048^2 domains, FP32 sweeps with FP64
      refinement, error against the oracle after every refinement step;
  [4] the adjoint on the cross with 32 right-hand sides: 16 seeded central
      finite-difference directions (h = 1e-6) and the oracle's gradient with
      respect to the right-hand side and the Dirichlet ghost data.

Bar: relative L2 <= 1e-12 against the oracle for every FP64 result and every
refined result, except the plain FP64 solve of the 1024 cross (<= 2e-12: its
rounding floor is 1.1e-12, see PLAIN_F64_1024_BAR); FP32 solves without
refinement against their own precision."""
import numpy as np
import pytest

import direct
import oracle as O
from gpu_util import banner, device, rel, sub_reference, to_dev

pytestmark = pytest.mark.gpu
D, N, A = 0, 1, 3
BAR = 1e-12


@pytest.fixture(scope="module")
def ctx():
    device()
    import focalors_poisson_amd as fpa
    return fpa.Context(0)


def solve_cat(ctx, doms, conn, fields, precision=0, ghost=None, refine_tol=None, max_steps=10,
              allow_not_converged=False):
    import torch
    import focalors_poisson_amd as fpa
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=200, tol=1e-12, tol_is_relative=True, max_iter=100,
                                    precision=precision)
    if refine_tol is not None:
        cat.set_refinement(tol=refine_tol, max_steps=max_steps)
    ts = [to_dev(f) for f in fields]
    info = cat.solve(ts, {k: to_dev(v) for k, v in (ghost or {}).items()},
                     allow_not_converged=allow_not_converged)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in ts], info


# ---- [0] ----

def test_config0_single_256(ctx):
    from focalors_poisson_amd.cases import single
    doms, conn, fields, _ = single(256, seed=0)
    banner("config 0: single 256^2 Dirichlet")
    us, info = solve_cat(ctx, doms, conn, fields)
    err = rel(us, O.reference_solution(doms, conn, fields))
    banner("config 0 rel-L2 vs oracle %.3e" % err)
    assert err < BAR


# ---- [1] ----

@pytest.fixture(scope="module")
def cross1024():
    from focalors_poisson_amd.cases import cross
    doms, conn, fields, _ = cross(1024, seed=1)
    fs = [f[0] for f in fields]
    banner("config 1: direct reference solution of the 1024 cross")
    return doms, conn, fs, direct.direct_solution(doms, conn, fs)


# The plain FP64 composite solve carries the rounding floor of the interface
# problem, whose condition grows like 1/h: about 1.1e-12 relative on the 1024
# cross (the reference's own algorithm restated in FP64 sits at 4.6e-11 there,
# from its x_diag = -2 + r (lambda - 2) rounding, which this build avoids).  One
# refinement step with the FP64 residual (FP_F64_REFINE) or FP32 sweeps with
# FP64 refinement reach 1e-15.
PLAIN_F64_1024_BAR = 2e-12


@pytest.mark.parametrize("precision", [0, 2, 1])
def test_config1_cross_1024(ctx, cross1024, precision):
    doms, conn, fs, ref = cross1024
    banner("config 1: cross 1024 precision=%d" % precision)
    us, info = solve_cat(ctx, doms, conn, fs, precision=precision, refine_tol=1e-13 if precision == 1 else None)
    err = rel(us, ref)
    banner("config 1 precision=%d rel-L2 vs the direct reference %.3e (%s)" % (precision, err, info))
    assert info["status"] == 0
    assert err < (PLAIN_F64_1024_BAR if precision == 0 else BAR)


# ---- [2] ----

def backward_error(u, b):
    """max |b - A u| / (max |A| |u| + max |b|) of the 5-point Dirichlet
    operator (hx = hy), on the device."""
    import torch
    up = torch.nn.functional.pad(u, (1, 1, 1, 1))
    au = up[:-2, 1:-1] + up[2:, 1:-1] + up[1:-1, :-2] + up[1:-1, 2:] - 4.0 * u
    aa = up[:-2, 1:-1].abs() + up[2:, 1:-1].abs() + up[1:-1, :-2].abs() + up[1:-1, 2:].abs() + 4.0 * u.abs()
    return float((b - au).abs().max() / (aa.max() + b.abs().max()))


def test_config2_batched_512(ctx):
    """512 Dirichlet problems, sides log-uniform in [64, 1024] (seed 2),
    right-hand sides U(-1, 1) (seed 3), in one call in FP64 and in FP32: every
    FP64 solution's backward error, every 16th problem and the largest against
    the oracle, every FP32 solution against the FP64 one."""
    import torch
    import focalors_poisson_amd as fpa
    from focalors_poisson_amd.cases import batched_sizes
    sizes = batched_sizes(512, seed=2)
    rng = np.random.default_rng(3)
    rhs = [rng.uniform(-1, 1, (nx, ny)) for nx, ny in sizes]
    subs = [fpa.PoissonSolver2D(ctx, nx, ny, 1.0 / max(nx, ny), 1.0 / max(nx, ny), (D, D, D, D)) for nx, ny in sizes]
    banner("config 2: 512 uneven problems, FP64")
    u64 = [to_dev(f) for f in rhs]
    fpa.solve_batched(ctx, subs, u64)
    torch.cuda.synchronize()
    worst = max(backward_error(u, to_dev(f)) for u, f in zip(u64, rhs))
    banner("config 2 FP64 worst backward error %.2e" % worst)
    assert worst < 1e-13
    big = max(range(512), key=lambda k: sizes[k][0] * sizes[k][1])
    errs = []
    for k in sorted(set(range(0, 512, 16)) | {big}):
        h = 1.0 / max(sizes[k])
        errs.append(rel(u64[k].cpu().numpy(), sub_reference(rhs[k], h, h, (D, D, D, D))))
    banner("config 2 FP64 worst rel-L2 vs oracle over %d problems %.3e" % (len(errs), max(errs)))
    assert max(errs) < BAR
    banner("config 2: 512 uneven problems, FP32")
    u32 = [to_dev(f, torch.float32) for f in rhs]
    fpa.solve_batched(ctx, subs, u32)
    torch.cuda.synchronize()
    e32 = max(float((a.double() - b).norm() / b.norm()) for a, b in zip(u32, u64))
    banner("config 2 FP32 worst rel-L2 vs FP64 %.3e" % e32)
    assert e32 < 1e-4
    # the same workload as a prepared batch captured in a graph: every replay
    # against the per-call FP64 results
    for dt, ref, bar in ((torch.float64, u64, 1e-13), (torch.float32, u64, 1e-4)):
        banner("config 2: prepared batch, graph replays, %s" % dt)
        src = [to_dev(f, dt) for f in rhs]
        work = [x.clone() for x in src]
        b = fpa.BatchSolve(ctx, subs, work)
        g = torch.cuda.CUDAGraph()
        st = torch.cuda.Stream()
        with torch.cuda.stream(st):
            g.capture_begin()
            b.solve(stream=st)
            g.capture_end()
        torch.cuda.synchronize()
        for rep in range(2):
            for w, x in zip(work, src):
                w.copy_(x)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            e = max(float((w.double() - r.double()).norm() / r.double().norm()) for w, r in zip(work, ref))
            banner("config 2 prepared batch replay %d worst rel-L2 vs per-call FP64 %.3e" % (rep, e))
            assert e < bar
        del g, b


def test_config2_cross_64_rhs(ctx):
    """The cross with 64 seeded right-hand sides (seeds 100..163) in one call:
    right-hand side 0 against the direct reference, every one against its own
    one-right-hand-side solve and by its composite residual."""
    import torch
    import focalors_poisson_amd as fpa
    from focalors_poisson_amd.cases import cross
    doms, conn, fields, _ = cross(1024, seed=100, nrhs=64)
    banner("config 2: cross 1024 x 64 right-hand sides")
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=200, tol=1e-12, tol_is_relative=True, max_iter=100)
    ts = [to_dev(f) for f in fields]
    info = cat.solve(ts)
    torch.cuda.synchronize()
    assert info["status"] == 0, info
    one = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=200, tol=1e-12, tol_is_relative=True, max_iter=100)
    oc = O.Composite(doms, conn, schur_mode=O.SCHUR_SPECTRAL)
    # backward error |b - A u| / (8 |u| + |b|): every row of the scaled
    # operator has absolute row sum at most 8
    nrm = lambda xs: np.sqrt(sum(float(np.sum(x * x)) for x in xs))
    worst_same = worst_res = 0.0
    for r in range(64):
        fs = [f[r] for f in fields]
        us = [t[r].cpu().numpy() for t in ts]
        single = [to_dev(f) for f in fs]
        one.solve(single)
        torch.cuda.synchronize()
        worst_same = max(worst_same, rel(us, [s.cpu().numpy() for s in single]))
        res = oc.residual(us, fs)
        worst_res = max(worst_res, nrm(res) / (8 * nrm(us) + nrm([f * d[2] ** 2 for f, d in zip(fs, doms)])))
        if r == 0:
            err = rel(us, direct.direct_solution(doms, conn, fs))
            banner("config 2 cross x64 rhs 0 rel-L2 vs the direct reference %.3e" % err)
            assert err < PLAIN_F64_1024_BAR
    banner("config 2 cross x64: worst rel diff vs single solves %.2e, worst backward error %.2e"
           % (worst_same, worst_res))
    assert worst_same < 1e-13
    assert worst_res < 1e-13


# ---- [3] ----

def test_config3_dirichlet_4096_refine_steps(ctx):
    """4096^2 Dirichlet, multiscale right-hand side (32 seeded Fourier modes,
    amplitudes ~ |k|^-2): FP32 sweeps with FP64 refinement, error against the
    oracle after each step, and the FP64 solve."""
    import torch
    import focalors_poisson_amd as fpa
    from focalors_poisson_amd.cases import multiscale_rhs
    n = 4096
    h = 1.0 / n
    f = multiscale_rhs(n, n, h, h, 0.0, 0.0, seed=4) * h * h
    banner("config 3: oracle solution of 4096^2")
    ref = sub_reference(f, h, h, (D, D, D, D))
    s = fpa.PoissonSolver2D(ctx, n, n, h, h, (D, D, D, D))
    errs = []
    for steps in range(1, 6):
        t = to_dev(f)
        info = s.solve_refine(t, tol=0.0, max_steps=steps, allow_not_converged=True)
        torch.cuda.synchronize()
        errs.append(rel(t.cpu().numpy(), ref))
        banner("config 3: 4096^2 FP32+refine step %d rel-L2 vs oracle %.3e (correction %.2e)"
               % (steps, errs[-1], info["refine_residual"]))
    assert errs[1] < errs[0] and min(errs) < BAR
    t = to_dev(f)
    info = s.solve_refine(t, tol=1e-13, max_steps=10)
    torch.cuda.synchronize()
    err = rel(t.cpu().numpy(), ref)
    banner("config 3: 4096^2 FP32+refine converged in %d steps, rel-L2 %.3e" % (info["refine_steps"], err))
    assert err < BAR
    t = to_dev(f)
    s.solve(t)
    torch.cuda.synchronize()
    err64 = rel(t.cpu().numpy(), ref)
    banner("config 3: 4096^2 FP64 rel-L2 %.3e" % err64)
    assert err64 < BAR


@pytest.fixture(scope="module")
def lshape2048():
    from focalors_poisson_amd.cases import lshape
    doms, conn, fields, _ = lshape(2048, seed=4)
    banner("config 3: direct reference solution of the L of three 2048^2 domains")
    return doms, conn, fields, direct.direct_solution(doms, conn, fields)


def test_config3_lshape_refine_steps(ctx, lshape2048):
    """FP32 sweeps with FP64 refinement on the L: error after every sweep."""
    doms, conn, fields, ref = lshape2048
    errs = []
    for steps in range(2, 7):
        us, info = solve_cat(ctx, doms, conn, fields, precision=1, refine_tol=1e-30, max_steps=steps,
                             allow_not_converged=True)
        errs.append(rel(us, ref))
        banner("config 3: L 3x2048^2 FP32+refine %d sweeps rel-L2 vs the direct reference %.3e (estimate %.2e)"
               % (steps, errs[-1], info["refine_residual"]))
    assert errs[1] < errs[0] and min(errs) < BAR
    us, info = solve_cat(ctx, doms, conn, fields, precision=1, refine_tol=1e-13)
    err = rel(us, ref)
    banner("config 3: L FP32+refine converged in %d sweeps, rel-L2 %.3e" % (info["refine_steps"], err))
    assert err < BAR


@pytest.mark.parametrize("precision", [0, 2])
def test_config3_lshape_f64(ctx, lshape2048, precision):
    doms, conn, fields, ref = lshape2048
    us, info = solve_cat(ctx, doms, conn, fields, precision=precision)
    err = rel(us, ref)
    banner("config 3: L 3x2048^2 precision=%d rel-L2 vs the direct reference %.3e" % (precision, err))
    assert err < BAR


# ---- [4] ----

def test_config4_adjoint_cross_1024_x32(ctx):
    """Adjoint on the cross, 32 right-hand sides, L = 1/2 ||u - u*||^2 with a
    seeded smooth u*, seeded Dirichlet ghost data: dL/df and dL/dg by
    fp_cat_vjp against central differences along 16 seeded directions
    (h = 1e-6, the loss difference summed as (u+ - u-)/2h . ((u+ + u-)/2 - u*)
    on the device, which has no cancellation) and against the oracle's
    gradient of right-hand side 0."""
    import torch
    import focalors_poisson_amd as fpa
    from focalors_poisson_amd.cases import cross, cross_geometry, smooth_target
    n, nrhs = 1024, 32
    doms, conn, fields, _ = cross(n, seed=1, nrhs=nrhs)
    _, _, offs = cross_geometry(n)
    target = smooth_target(doms, offs, seed=6, nrhs=nrhs)
    rng = np.random.default_rng(12)
    bcs = [d[4] for d in doms]
    faces = [(d, f) for d in range(5) for f in range(4) if bcs[d][f] == D]
    glen = lambda d, f: doms[d][1] if f < 2 else doms[d][0]
    ghost = {k: rng.standard_normal(nrhs * glen(*k)) for k in faces}
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=200, tol=1e-12, tol_is_relative=True, max_iter=100)
    dev = device()
    f0 = [to_dev(f) for f in fields]
    g0 = {k: to_dev(v) for k, v in ghost.items()}
    ut = [to_dev(t) for t in target]

    def forward(fs, gs):
        us = [x.clone() for x in fs]
        info = cat.solve(us, {k: v.clone() for k, v in gs.items()})
        assert info["status"] == 0, info
        return us

    banner("config 4: adjoint cross 1024 x 32: forward and vjp")
    u0 = forward(f0, g0)
    gu = [u - t for u, t in zip(u0, ut)]
    gf, gg = cat.vjp(gu, faces=faces)
    torch.cuda.synchronize()
    gen = torch.Generator(device=dev)
    gen.manual_seed(16)
    h = 1e-6
    worst = 0.0
    for trial in range(16):
        vf = [torch.randn(f.shape, dtype=torch.float64, device=dev, generator=gen) for f in f0]
        vg = {k: torch.randn(v.shape, dtype=torch.float64, device=dev, generator=gen) for k, v in g0.items()}
        up = forward([f + h * v for f, v in zip(f0, vf)], {k: g0[k] + h * vg[k] for k in g0})
        um = forward([f - h * v for f, v in zip(f0, vf)], {k: g0[k] - h * vg[k] for k in g0})
        fd = sum(float(((a - b) / (2 * h) * ((a + b) / 2 - t)).sum()) for a, b, t in zip(up, um, ut))
        an = sum(float((g * v).sum()) for g, v in zip(gf, vf)) + sum(float((gg[k] * vg[k]).sum()) for k in g0)
        worst = max(worst, abs(fd - an) / abs(an))
        banner("config 4 FD direction %d: fd %.12e vjp %.12e" % (trial, fd, an))
    banner("config 4 worst relative FD difference %.2e" % worst)
    assert worst < 1e-6
    # the oracle's gradient of right-hand side 0: lambda = A^-1 dL/du (the
    # assembled operator is symmetric), dL/df = hx^2 lambda, dL/dg = hx^2
    # (-1/h^2) lambda on the Dirichlet face's line
    du = [g[0].cpu().numpy() for g in gu]
    lam = direct.direct_solution(doms, conn, [x / d[2] ** 2 for x, d in zip(du, doms)])
    ref_f = [d[2] ** 2 * l for d, l in zip(doms, lam)]
    err_f = rel([g[0].cpu().numpy() for g in gf], ref_f)
    lines = {0: lambda x: x[0, :], 1: lambda x: x[-1, :], 2: lambda x: x[:, 0], 3: lambda x: x[:, -1]}
    ref_g, got_g = [], []
    for (d, face) in faces:
        hh = doms[d][2] if face < 2 else doms[d][3]
        ref_g.append(doms[d][2] ** 2 * (-1.0 / hh ** 2) * lines[face](lam[d]))
        got_g.append(gg[(d, face)].cpu().numpy()[:glen(d, face)])
    err_g = rel(got_g, ref_g)
    banner("config 4 rel-L2 vs the reference gradient: dL/df %.3e, dL/dg %.3e" % (err_f, err_g))
    assert err_f < BAR and err_g < BAR
