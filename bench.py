#!/usr/bin/env python3
"""Cross-geometry composite Poisson solves per second on one MI355X.

Workload: the cross of 5 subdomains (centre 1024 x 1024, arms 512 x 1024 and
1024 x 512, h = 1/1024, mixed Dirichlet/Neumann outer faces), seeded RHS
sin(2 pi x) cos(pi y) + 1e-3 N(0,1).  One step = one composite solve
(ConcatPoissonSolver2D::solve equivalent): refill the RHS from a resident copy,
boundary assembly, 4 arm solves, interface GMRES on the centre, 4 arm solves.
Inputs are resident in HBM before the timed region.  The GPU solution is
checked against the CPU oracle's solution of the same problem (rel L2).

Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1024, help="centre size (arms n/2)")
    ap.add_argument("--m", type=int, default=200, help="GMRES restart length on the interface")
    ap.add_argument("--tol", type=float, default=1e-12, help="relative GMRES tolerance on the interface")
    ap.add_argument("--oracle-tol", type=float, default=1e-11,
                    help="relative GMRES tolerance of the timed CPU oracle solve")
    ap.add_argument("--refine-tol", type=float, default=1e-11,
                    help="target estimated relative error of the FP32+refine variant")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--extras", action="store_true",
                    help="measure the other workloads instead (cross x 64 right-hand sides, 512 batched uneven "
                         "problems in FP64 and FP32, 4096^2 FP64 vs FP32+refine, L-shape, periodic channel, "
                         "adjoint fwd+bwd ratio); "
                         "prints one JSON line")
    ap.add_argument("--only", default="",
                    help="with --extras: a comma-separated subset of the workloads (cross_x64, batched, dirichlet_4096, "
                         "lshape, periodic, single_256, case_api, adjoint, box_256, cross3d, cross_krylov, projection)")
    ap.add_argument("--forms", default="",
                    help="a comma-separated subset of the solve forms to time (f64, f64_graph, f64_into, "
                         "f64_graph_into, f64_refine_graph_into, f32_refine); default all")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the per-kernel timing on the centre subdomain plus an 8 MiB device copy (the "
                         "calibration kernel of the rocprofv3 counter runs); no composite solve")
    return ap.parse_args()


def time_kernel(fn, reps, torch):
    """Average duration (ms) of fn() over reps launches, HIP events on the
    current stream (the stream the library launches on).  A spin kernel of a
    few milliseconds holds the stream while the host queues the launches, so
    that they run back to back and the average is the kernel's duration, not
    the host's launch interval."""
    fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda._sleep(5_000_000)
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def measured_traffic(kernel, n):
    """HBM bytes per launch of `kernel` at the default size, from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs recorded under profiles/
    (cross1024_traffic.json; null for another kernel or size)."""
    path = os.path.join(ROOT, "profiles", "cross1024_traffic.json")
    if n != 1024 or not os.path.exists(path):
        return None
    with open(path) as fh:
        rec = json.load(fh)
    k = rec.get("kernels", {}).get(kernel)
    return None if k is None else k.get("traffic_bytes")


def centre_kernels(fpa, ctx, doms, torch, reps, nrhs=1):
    """Average durations of the centre subdomain's kernels (1024 x 1024 DD
    for the default cross): forward and inverse y transform (DST-I, two rows
    packed into one complex FFT of length N+1) and the x sweep, each launched
    alone on nrhs fields.
    Algorithmic bytes: one read and one write of the fields (16 B per cell)."""
    nx, ny = doms[0][0], doms[0][1]
    sub = fpa.PoissonSolver2D(ctx, nx, ny, doms[0][2], doms[0][3], (3, 3, 3, 3))
    shape = (nx, ny) if nrhs == 1 else (nrhs, nx, ny)
    buf = torch.randn(*shape, dtype=torch.float64, device=torch.device("cuda:0"))
    fbytes = 8.0 * nx * ny * nrhs
    kern = {}
    for name, fn, nbytes in [
        ("ytransform_fwd", lambda: sub.transform(buf), 2 * fbytes),
        ("ytransform_inv", lambda: sub.transform(buf, inverse=True), 2 * fbytes),
        ("tridiag", lambda: sub.chasing(buf), 2 * fbytes),
    ]:
        t = time_kernel(fn, reps, torch)
        kern[name] = {"ms": round(t, 5), "bytes": nbytes, "gbs": round(nbytes / (t * 1e-3) / 1e9, 1)}
    return kern


def composite_transform(fpa, ctx, doms, conn, torch, reps, nrhs=1):
    """Average duration of the solve's dominant launch: the batched y
    transform of all five domains in one launch (fp_cat_transform, the same
    kernel and shape as the solve's first and last stages), forward and
    inverse.  Algorithmic bytes: one read and one write of every field (16 B
    per cell)."""
    cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=20, tol=1e-12, tol_is_relative=True)
    dev = torch.device("cuda:0")
    bufs = [torch.randn(*((d[0], d[1]) if nrhs == 1 else (nrhs, d[0], d[1])), dtype=torch.float64, device=dev)
            for d in doms]
    nbytes = 16.0 * sum(d[0] * d[1] for d in doms) * nrhs
    kern = {}
    for name, inv in (("cat_ytransform_fwd", False), ("cat_ytransform_inv", True)):
        t = time_kernel(lambda: cat.transform(bufs, inverse=inv), reps, torch)
        kern[name] = {"ms": round(t, 5), "bytes": nbytes, "gbs": round(nbytes / (t * 1e-3) / 1e9, 1)}
    return kern


def copy_kernel(torch, n, reps):
    """A device-to-device copy of n*n doubles (8 MiB at n = 1024): the
    calibration kernel of the rocprofv3 FETCH_SIZE / WRITE_SIZE runs, whose
    algorithmic traffic is known exactly (one read and one write)."""
    src = torch.randn(n * n, dtype=torch.float64, device=torch.device("cuda:0"))
    dst = torch.empty_like(src)
    t = time_kernel(lambda: dst.copy_(src), reps, torch)
    return {"copy_8MiB": {"ms": round(t, 5), "bytes": 16.0 * n * n, "gbs": round(16.0 * n * n / (t * 1e-3) / 1e9, 1)}}


def timed(fn, torch, reps=3):
    """Mean wall time (s) of fn() over reps runs after one warm-up."""
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def flat_fields(torch, pristine):
    """The right-hand sides back to back in one device buffer, and a work
    buffer of the same layout whose per-domain views the solver takes: one
    copy refills every in-place field."""
    src = torch.cat([p.reshape(-1) for p in pristine])
    dst = torch.empty_like(src)
    views, o = [], 0
    for p in pristine:
        views.append(dst[o:o + p.numel()].view(p.shape))
        o += p.numel()
    return src, dst, views


def extras(fpa, ctx, torch, np, n, only=""):
    """The other workloads of the solve path, one measurement each (all of
    them, or the comma-separated subset `only`)."""
    from focalors_poisson_amd import cases
    dev = torch.device("cuda:0")
    out = {}
    sel = set(w for w in only.split(",") if w)
    want = lambda name: not sel or name in sel

    if want('cross_x64'):
        # cross with 64 right-hand sides (seeds 100..163), FP64, one call
        doms, conn, fields, _ = cases.cross(n, seed=100, nrhs=64)
        cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=200, tol=1e-12, tol_is_relative=True, max_iter=100)
        pristine = [torch.from_numpy(f).to(dev) for f in fields]
        work = [torch.empty_like(p) for p in pristine]
        info = {}

        def multi():
            for w, p in zip(work, pristine):
                w.copy_(p)
            info.update(cat.solve(work))

        t = timed(multi, torch)
        one = [p[:1].clone() for p in pristine]
        cat.solve(one)
        torch.cuda.synchronize()
        same = max(float((w[:1] - o).norm() / o.norm()) for w, o in zip(work, one))
        out["cross_x64_rhs"] = {"rhs_solves_per_s": round(64 / t, 1), "ms_per_call": round(t * 1e3, 3),
                                "gmres_rel_residual": info.get("rel_residual"),
                                "rel_diff_rhs0_vs_single_rhs_solve": same}
        # the same 64 right-hand sides with FP32 sweeps and FP64 refinement to
        # an estimated 1e-11, where the sweeps are bandwidth-bound: error of
        # every right-hand side against the FP64 solve
        u64 = [w.clone() for w in work]
        cat32 = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=200, tol=1e-12, tol_is_relative=True, max_iter=100,
                                          precision=fpa.FP_F32_REFINE)
        cat32.set_refinement(tol=1e-11, max_steps=10)
        info32 = {}

        def multi32():
            for w, p in zip(work, pristine):
                w.copy_(p)
            info32.update(cat32.solve(work))

        t32 = timed(multi32, torch)
        worst = max(float((torch.cat([w[r].flatten() for w in work]) - torch.cat([u[r].flatten() for u in u64])).norm()
                          / torch.cat([u[r].flatten() for u in u64]).norm()) for r in range(64))
        out["cross_x64_rhs_f32_refine"] = {"rhs_solves_per_s": round(64 / t32, 1), "ms_per_call": round(t32 * 1e3, 3),
                                           "refine_sweeps": info32.get("refine_steps"),
                                           "refine_error_estimate": info32.get("refine_residual"),
                                           "worst_rel_diff_vs_f64": worst,
                                           "speedup_vs_f64": round(t / t32, 3)}
        del cat, cat32, pristine, work, one, fields, u64

    if want('batched'):
        # 512 independent Dirichlet problems, sides log-uniform in [64, 1024] (seed 2), RHS U(-1, 1) (seed 3).
        # Per-call batched solve (fp_sub_solve_batched) with timed plans, and the prepared batch
        # (fp_batch) replayed from a graph, with timed and with estimated plans.  Algorithmic bytes: one
        # read and one write of every field per phase (transform, sweep, inverse transform).
        sizes = cases.batched_sizes(512, seed=2)
        rng = np.random.default_rng(3)
        rhs = [torch.from_numpy(rng.uniform(-1, 1, (nx, ny))).to(dev) for nx, ny in sizes]
        cells = sum(nx * ny for nx, ny in sizes)
        est = fpa.Context(0)
        est.set_planning(fpa.FP_PLAN_ESTIMATE)
        plans = {}
        for name, c in (("measure", ctx), ("estimate", est)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            plans[name] = [fpa.PoissonSolver2D(c, nx, ny, 1.0 / max(nx, ny), 1.0 / max(nx, ny), (0, 0, 0, 0))
                           for nx, ny in sizes]
            torch.cuda.synchronize()
            out["batched_512_plan_build_s_" + name] = round(time.perf_counter() - t0, 3)
        for prec, dt in (("f64", torch.float64), ("f32", torch.float32)):
            src, dst, wk = flat_fields(torch, [r.to(dt) for r in rhs])
            nbytes = 3 * 2 * cells * (8 if prec == "f64" else 4)

            def batch():
                dst.copy_(src)  # one copy refills all 512 right-hand sides
                fpa.solve_batched(ctx, plans["measure"], wk)

            batch()  # first calls grow the stream-ordered and pinned staging pools
            t = timed(batch, torch, reps=10)
            res = {"per_call_ms": round(t * 1e3, 3)}
            for name in ("measure", "estimate"):
                b = fpa.BatchSolve(ctx, plans[name], wk)
                stream = torch.cuda.Stream()
                g = torch.cuda.CUDAGraph()
                with torch.cuda.stream(stream):
                    g.capture_begin()
                    b.solve(stream=stream)
                    g.capture_end()
                torch.cuda.synchronize()

                def replay():
                    dst.copy_(src)
                    g.replay()

                tg = timed(replay, torch, reps=10)
                # the refill copy alone, to report the solve's own time
                tc = timed(lambda: dst.copy_(src), torch, reps=10)
                res["prepared_graph_ms_" + name] = round(tg * 1e3, 3)
                res["prepared_graph_solve_only_ms_" + name] = round((tg - tc) * 1e3, 3)
                res["prepared_gbs_" + name] = round(nbytes / (tg - tc) / 1e9, 1)
                del g, b
            res.update({"cells": cells, "algorithmic_bytes": nbytes, "refill_copy_ms": round(tc * 1e3, 3)})
            out["batched_512_" + prec] = res
        del plans, rhs

    if want('dirichlet_4096'):
        # 4096^2 Dirichlet, multiscale RHS (seed 4): FP64 vs FP32 + FP64 refinement to 1e-10
        N4 = 4096
        h = 1.0 / N4
        f = cases.multiscale_rhs(N4, N4, h, h, 0.0, 0.0, seed=4) * h * h
        sub = fpa.PoissonSolver2D(ctx, N4, N4, h, h, (0, 0, 0, 0))
        src = torch.from_numpy(f).to(dev)
        u64 = torch.empty_like(src)

        def s64():
            u64.copy_(src)
            sub.solve(u64)

        t64 = timed(s64, torch)
        u32 = torch.empty_like(src)
        rinfo = {}

        def s32():
            u32.copy_(src)
            rinfo.update(sub.solve_refine(u32, tol=1e-10, max_steps=10))

        t32 = timed(s32, torch)
        out["dirichlet_4096"] = {"f64_ms": round(t64 * 1e3, 3), "f32_refine_ms": round(t32 * 1e3, 3),
                                 "refine_steps": rinfo.get("refine_steps"),
                                 "refine_last_correction": rinfo.get("refine_residual"),
                                 "rel_l2_f32_refine_vs_f64": float((u32 - u64).norm() / u64.norm())}
        del sub, src, u64, u32

    if want('lshape'):
        # L-shape of three 2048^2 Dirichlet domains: FP64 vs FP32 + FP64 refinement
        doms, conn, fields, _ = cases.lshape(2048, seed=4)
        src = [torch.from_numpy(f).to(dev) for f in fields]
        res = {}
        for prec, key in ((fpa.FP_F64, "f64"), (fpa.FP_F32_REFINE, "f32_refine")):
            cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=200, tol=1e-12, tol_is_relative=True, max_iter=100,
                                            precision=prec)
            if prec == fpa.FP_F32_REFINE:
                cat.set_refinement(tol=1e-12, max_steps=10)
            wk = [torch.empty_like(x) for x in src]
            info = {}

            def solve_l():
                for w, x in zip(wk, src):
                    w.copy_(x)
                info.update(cat.solve(wk))

            res[key] = (timed(solve_l, torch), dict(info), [w.clone() for w in wk])
            del cat
        d = max(float((a - b).norm() / b.norm()) for a, b in zip(res["f32_refine"][2], res["f64"][2]))
        out["lshape_3x2048"] = {"f64_ms": round(res["f64"][0] * 1e3, 3),
                                "f32_refine_ms": round(res["f32_refine"][0] * 1e3, 3),
                                "refine_steps": res["f32_refine"][1].get("refine_steps"),
                                "rel_l2_f32_refine_vs_f64": d,
                                "gmres_rel_residual_f64": res["f64"][1].get("rel_residual")}
        del res, src

    if want('periodic'):
        # periodic channel: two 1024^2 domains side by side, periodic in y,
        # Dirichlet / Neumann x ends, coupled across x (the periodic transform on
        # the two-level path), RHS N(0,1) (seed 9); FP64 and FP32 + refinement
        hp = 1.0 / 1024
        pdoms = [(1024, 1024, hp, hp, (0, 3, 2, 2)), (1024, 1024, hp, hp, (3, 1, 2, 2))]
        pconn = [(0, 1, 1)]
        prng = np.random.default_rng(9)
        src = [torch.from_numpy(prng.standard_normal((d[0], d[1]))).to(dev) for d in pdoms]
        res = {}
        for prec, key in ((fpa.FP_F64, "f64"), (fpa.FP_F32_REFINE, "f32_refine")):
            cat = fpa.ConcatPoissonSolver2D(ctx, pdoms, pconn, m=200, tol=1e-12, tol_is_relative=True, max_iter=100,
                                            precision=prec)
            if prec == fpa.FP_F32_REFINE:
                cat.set_refinement(tol=1e-12, max_steps=10)
            wk = [torch.empty_like(x) for x in src]
            info = {}

            def solve_p():
                for w, x in zip(wk, src):
                    w.copy_(x)
                info.update(cat.solve(wk))

            res[key] = (timed(solve_p, torch), dict(info), [w.clone() for w in wk])
            del cat
        d = max(float((a - b).norm() / b.norm()) for a, b in zip(res["f32_refine"][2], res["f64"][2]))
        out["periodic_channel_2x1024"] = {"f64_ms": round(res["f64"][0] * 1e3, 3),
                                          "f32_refine_ms": round(res["f32_refine"][0] * 1e3, 3),
                                          "refine_steps": res["f32_refine"][1].get("refine_steps"),
                                          "rel_l2_f32_refine_vs_f64": d,
                                          "gmres_iterations_f64": res["f64"][1].get("iterations"),
                                          "gmres_rel_residual_f64": res["f64"][1].get("rel_residual")}
        del res, src

    if want('single_256'):
        # single 256^2 Dirichlet domain (the reference's own CPU case): solves/s,
        # error vs the oracle, and the oracle's own time on the host
        doms, conn, fields, _ = cases.single(256)
        cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=20, tol=1e-12, tol_is_relative=True)
        src = [torch.from_numpy(f).to(dev) for f in fields]
        wk = [torch.empty_like(x) for x in src]

        def solve_s():
            for w, x in zip(wk, src):
                w.copy_(x)
            cat.solve(wk)

        ts = timed(solve_s, torch, reps=50)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        t0 = time.perf_counter()
        ref, _, _ = O.Composite(doms, conn).solve(fields)
        t_cpu = time.perf_counter() - t0
        u = wk[0].cpu().numpy()
        out["single_256"] = {"solves_per_s": round(1.0 / ts, 1), "ms_per_solve": round(ts * 1e3, 4),
                             "rel_l2_vs_oracle": float(np.linalg.norm(u - ref[0]) / np.linalg.norm(ref[0])),
                             "cpu_oracle_solves_per_s": round(1.0 / t_cpu, 1), "cpu_threads": O.num_threads()}
        del cat, src, wk

    if want('case_api'):
        # the case API (C++ classes over the C ABI) on the same cross: fields kept
        # on the device between solves by field2's mirror, and with a host round
        # trip per solve (PCIe included)
        prog = os.path.join(ROOT, "tests", "cpp", "_build", "case_cross_timing")
        if os.path.exists(prog):
            import subprocess
            r = subprocess.run([prog, str(n), "20"], capture_output=True, text=True, timeout=300)
            w = r.stdout.split()
            if r.returncode == 0 and "device_resident_ms" in w:
                out["case_api_cross"] = {"device_resident_ms": float(w[w.index("device_resident_ms") + 1]),
                                         "host_roundtrip_ms": float(w[w.index("host_roundtrip_ms") + 1])}

    if want('adjoint'):
        # adjoint on the cross, batch of 32: (forward + backward) / forward time
        doms, conn, fields, _ = cases.cross(n, seed=1, nrhs=32)
        cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=200, tol=1e-12, tol_is_relative=True, max_iter=100)
        fs = [torch.from_numpy(f).to(dev) for f in fields]
        wk = [torch.empty_like(x) for x in fs]

        def fwd():
            for w, x in zip(wk, fs):
                w.copy_(x)
            cat.solve(wk)

        def fwd_bwd():
            fwd()
            cat.vjp(wk)

        tf = timed(fwd, torch)
        tfb = timed(fwd_bwd, torch)
        out["adjoint_cross_x32"] = {"forward_ms": round(tf * 1e3, 3), "forward_backward_ms": round(tfb * 1e3, 3),
                                    "ratio": round(tfb / tf, 3)}
    # one 256^3 Dirichlet box (PoissonSolver3D), FP64: solves/s and the
    # algorithmic bandwidth (7 passes of 16 B per cell: z, y transforms both
    # ways, two plane transposes, the x sweep)
    if want('box_256'):
        nb = 256
        hb = 1.0 / nb
        sub3 = fpa.PoissonSolver3D(ctx, nb, nb, nb, hb, hb, hb, (0, 0, 0, 0, 0, 0))
        g = torch.Generator(device=dev)
        g.manual_seed(5)
        src = torch.rand((nb, nb, nb), dtype=torch.float64, device=dev, generator=g) * hb * hb
        u3 = torch.empty_like(src)

        def s3():
            u3.copy_(src)
            sub3.solve(u3)

        t3 = timed(s3, torch, reps=10)
        tc = timed(lambda: u3.copy_(src), torch, reps=10)
        out["box_256_f64"] = {"ms_per_solve": round(t3 * 1e3, 3), "refill_copy_ms": round(tc * 1e3, 3),
                              "solves_per_s": round(1.0 / t3, 1),
                              "gbs_7_passes": round(7 * 16.0 * nb ** 3 / (t3 - tc) / 1e9, 1)}
        del sub3, src, u3

    # a 3-D cross: centre 64^3 with six arms 32 x 64 x 64 (Dirichlet outer
    # faces but Neumann on the far faces of the x arms), one FP64 composite solve
    if want('cross3d'):
        nc, na = 64, 32
        hc = 1.0 / nc
        A3, D3, N3 = 3, 0, 1
        boxes = [(nc, nc, nc, hc, hc, hc, (A3, A3, A3, A3, A3, A3)),
                 (na, nc, nc, hc, hc, hc, (N3, A3, D3, D3, D3, D3)), (na, nc, nc, hc, hc, hc, (A3, N3, D3, D3, D3, D3)),
                 (nc, na, nc, hc, hc, hc, (D3, D3, D3, A3, D3, D3)), (nc, na, nc, hc, hc, hc, (D3, D3, A3, D3, D3, D3)),
                 (nc, nc, na, hc, hc, hc, (D3, D3, D3, D3, D3, A3)), (nc, nc, na, hc, hc, hc, (D3, D3, D3, D3, A3, D3))]
        conn3 = [(0, 0, 1), (0, 1, 2), (0, 2, 3), (0, 3, 4), (0, 4, 5), (0, 5, 6)]
        t0 = time.perf_counter()
        c3 = fpa.ConcatPoissonSolver3D(ctx, boxes, conn3, m=60, tol=1e-12, tol_is_relative=True, max_iter=100)
        torch.cuda.synchronize()
        tb = time.perf_counter() - t0
        g = torch.Generator(device=dev)
        g.manual_seed(6)
        srcs = [torch.randn(b[:3], dtype=torch.float64, device=dev, generator=g) for b in boxes]
        wks = [torch.empty_like(x) for x in srcs]
        inf3 = {}

        def sc3():
            for w, x in zip(wks, srcs):
                w.copy_(x)
            inf3.update(c3.solve(wks))

        tt = timed(sc3, torch, reps=3)
        out["cross3d_64"] = {"ms_per_solve": round(tt * 1e3, 3), "construct_s": round(tb, 2),
                             "gmres_iterations": inf3.get("iterations"), "rel_residual": inf3.get("rel_residual"),
                             "cells": sum(b[0] * b[1] * b[2] for b in boxes)}
        del c3, srcs, wks

    if want('cross_krylov'):
        # the reference's interface method on the device: the cross solved by
        # GMRES(m=200) over the Schur system from x0 = C, no explicit-inverse
        # start (FP_INTERFACE_PRECOND=0 at construction), to relative 1e-12;
        # against the explicit-inverse solve of the same right-hand side
        doms, conn, fields, ghost = cases.cross(n, seed=1)
        src = [torch.from_numpy(f).to(dev) for f in fields]
        gh = {k: torch.from_numpy(g).to(dev) for k, g in ghost.items()} if ghost else None
        res = {}
        for key, pre in (("krylov", "0"), ("explicit_inverse", "1")):
            old = os.environ.get("FP_INTERFACE_PRECOND")
            os.environ["FP_INTERFACE_PRECOND"] = pre
            try:
                cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=200, tol=1e-12, tol_is_relative=True,
                                                max_iter=100)
            finally:
                if old is None:
                    os.environ.pop("FP_INTERFACE_PRECOND")
                else:
                    os.environ["FP_INTERFACE_PRECOND"] = old
            wk = [torch.empty_like(x) for x in src]
            inf = {}

            def solve_k():
                for w, x in zip(wk, src):
                    w.copy_(x)
                inf.update(cat.solve(wk, ghost=gh))

            res[key] = (timed(solve_k, torch, reps=5), dict(inf), [w.clone() for w in wk])
            del cat
        tk, ik, uk = res["krylov"]
        tp, ip, up = res["explicit_inverse"]
        steps = ik.get("iterations") or 0
        diff = math.sqrt(sum(float((a - b).norm()) ** 2 for a, b in zip(uk, up))
                         / sum(float(b.norm()) ** 2 for b in up))
        out["cross_krylov"] = {"ms_per_solve": round(tk * 1e3, 3), "gmres_iterations": steps,
                               "restarts": ik.get("restarts"), "rel_residual": ik.get("rel_residual"),
                               "ms_per_solve_explicit_inverse": round(tp * 1e3, 3),
                               "ms_per_krylov_step": round((tk - tp) * 1e3 / steps, 4) if steps else None,
                               "rel_l2_vs_explicit_inverse_solve": diff}
        del res, src, uk, up

    if want('projection'):
        # the flow solver's pressure projection on the cross (staggered grid,
        # walls Neumann, outlets Dirichlet): divergence, composite solve and
        # velocity correction replayed from one graph per step, the provisional
        # velocity refilled by one device copy per step
        doms, conn, _ = cases.cross_geometry(n)
        cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=200, tol=1e-12, tol_is_relative=True, max_iter=100)
        pj = fpa.PressureProjection(cat)
        g = torch.Generator(device=dev)
        g.manual_seed(11)
        u0 = [torch.randn(d[0] + 1, d[1], dtype=torch.float64, device=dev, generator=g) for d in doms]
        v0 = [torch.randn(d[0], d[1] + 1, dtype=torch.float64, device=dev, generator=g) for d in doms]
        for a_, dr, b_ in conn:  # one value per shared face
            if dr == 1:
                u0[b_][0].copy_(u0[a_][-1])
            elif dr == 0:
                u0[b_][-1].copy_(u0[a_][0])
            elif dr == 3:
                v0[b_][:, 0].copy_(v0[a_][:, -1])
            else:
                v0[b_][:, -1].copy_(v0[a_][:, 0])
        u = [x.clone() for x in u0]
        v = [x.clone() for x in v0]
        p = [torch.zeros(d[0], d[1], dtype=torch.float64, device=dev) for d in doms]
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            pj.capture(u, v, p, 0.01, stream=stream)

            def step():
                for x, y in zip(u + v, u0 + v0):
                    x.copy_(y)
                pj.launch(stream=stream)

            tp = timed(step, torch, reps=20)
        torch.cuda.synchronize()
        div0 = [torch.empty(d[0], d[1], dtype=torch.float64, device=dev) for d in doms]
        div1 = [torch.empty_like(x) for x in div0]
        pj.divergence(u0, v0, div0)
        pj.divergence(u, v, div1)
        torch.cuda.synchronize()
        out["projection_cross"] = {"ms_per_step": round(tp * 1e3, 4),
                                   "divergence_after_rel": max(float(x.abs().max()) for x in div1)
                                   / max(float(x.abs().max()) for x in div0)}
        del pj, cat
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on one GPU: the solve path has no multi-GPU form (the composite solve would shard "
              "by subdomain, exchanging interface lines between ranks). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import focalors_poisson_amd as fpa
    from focalors_poisson_amd.cases import cross

    dev = torch.device("cuda:0")
    n = a.n
    doms, conn, fields, _ = cross(n, seed=1)
    ctx = fpa.Context(0)
    if a.extras:
        print(json.dumps({"extras": extras(fpa, ctx, torch, np, n, a.only)}), flush=True)
        return
    if a.kernels_only:
        kern = centre_kernels(fpa, ctx, doms, torch, a.kernel_reps)
        kern.update(composite_transform(fpa, ctx, doms, conn, torch, a.kernel_reps))
        kern.update(copy_kernel(torch, n, a.kernel_reps))
        print(json.dumps({"kernels": kern}), flush=True)
        return
    cells = sum(d[0] * d[1] for d in doms)
    pristine = [torch.from_numpy(f[0]).to(dev) for f in fields]

    def make(precision):
        t0 = time.perf_counter()
        cat = fpa.ConcatPoissonSolver2D(ctx, doms, conn, m=a.m, tol=a.tol, tol_is_relative=True, max_iter=100,
                                        precision=precision)
        if precision == fpa.FP_F32_REFINE:
            cat.set_refinement(tol=a.refine_tol, max_steps=10)
        torch.cuda.synchronize()
        return cat, time.perf_counter() - t0

    def run_variant(precision, into=False):
        """Eager solves.  into=False: in place (fp_cat_solve), the right-hand
        sides refilled by one device copy per step; into=True: out of place
        (fp_cat_solve_into), the resident right-hand sides read directly."""
        cat, t_build = make(precision)
        src, dst, work = flat_fields(torch, pristine)
        rhs = pristine if into else None  # resident device tensors, read only
        info = {}

        def step():
            if into:
                info.update(cat.solve_into(rhs, work))
            else:
                dst.copy_(src)  # refill the in-place right-hand sides
                info.update(cat.solve(work))

        for _ in range(a.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / a.steps * 1e3
        return {"ms": ms, "build_s": t_build, "info": dict(info), "u": [w.cpu().numpy() for w in work]}

    def run_graph(into=False, precision=fpa.FP_F64):
        """The solve captured once in a HIP graph and replayed per step: in
        place with the right-hand sides refilled by an eager copy on the same
        stream, or (into=True) out of place from the resident right-hand sides."""
        cat, t_build = make(precision)
        stream = torch.cuda.Stream()
        src, dst, work = flat_fields(torch, pristine)
        rhs = pristine if into else None  # resident device tensors, read only
        with torch.cuda.stream(stream):
            cat.graph_capture(work, stream=stream, rhs=rhs)

            def step(wait=False):
                if not into:
                    dst.copy_(src)
                return cat.graph_launch(stream=stream, wait=wait)

            for _ in range(a.warmup):
                step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(a.steps):
                step()
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) / a.steps * 1e3
            info = step(wait=True)  # the residual check of a replay
        torch.cuda.synchronize()
        return {"ms": ms, "build_s": t_build, "info": dict(info), "u": [w.cpu().numpy() for w in work]}

    forms = {
        "f64": lambda: run_variant(fpa.FP_F64),
        "f64_graph": lambda: run_graph(),
        "f64_into": lambda: run_variant(fpa.FP_F64, into=True),
        "f64_graph_into": lambda: run_graph(into=True),
        "f64_refine_graph_into": lambda: run_graph(into=True, precision=fpa.FP_F64_REFINE),
        "f32_refine": lambda: run_variant(fpa.FP_F32_REFINE),
    }
    pick = [f for f in a.forms.split(",") if f] or list(forms)
    for f in pick:
        if f not in forms:
            sys.exit("bench.py: unknown form %r (known: %s)" % (f, ", ".join(forms)))
    runs = {f: forms[f]() for f in pick}

    # ---- the dominant launch of the solve, timed alone: the batched y
    # transform of all five domains (the solve's first and last stages) ----
    kern = composite_transform(fpa, ctx, doms, conn, torch, a.kernel_reps)
    kern.update(centre_kernels(fpa, ctx, doms, torch, a.kernel_reps))
    kern.update(copy_kernel(torch, n, a.kernel_reps))
    dom_name = "cat_ytransform_inv"
    dom = kern[dom_name]
    # the same kernels on 32 right-hand sides: their throughput once launch
    # latency no longer dominates (not the workload above; context only)
    kern32 = composite_transform(fpa, ctx, doms, conn, torch, max(3, a.kernel_reps // 10), nrhs=32)
    kern32.update(centre_kernels(fpa, ctx, doms, torch, max(3, a.kernel_reps // 10), nrhs=32))

    # ---- accuracy: the converged oracle; CPU baseline: the oracle's plain
    # solve (the reference's algorithm) timed on the host cores ----
    cpu = None
    err = {k: None for k in runs}
    cpu_err = None
    if not a.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        fs = [f[0] for f in fields]
        oc = O.Composite(doms, conn, m=20, tol=a.oracle_tol, tol_is_relative=True, maxiter=500,
                         schur_mode=O.SCHUR_SPECTRAL)
        t0 = time.perf_counter()
        base, it, _ = oc.solve(fs)
        t_cpu = time.perf_counter() - t0
        import direct
        # untimed: the direct composite solve refined with the long-double
        # residual (equal to the refined restatement of the reference's
        # algorithm to below FP64 rounding, and a few seconds instead of minutes)
        ref = direct.direct_solution(doms, conn, fs)
        y = np.concatenate([r.ravel() for r in ref])
        for key, v in runs.items():
            x = np.concatenate([u.ravel() for u in v["u"]])
            err[key] = float(np.linalg.norm(x - y) / np.linalg.norm(y))
        xb = np.concatenate([u.ravel() for u in base])
        cpu_err = float(np.linalg.norm(xb - y) / np.linalg.norm(y))
        cpu = {"value": round(1.0 / t_cpu, 4), "unit": "cross solves/s", "cores": O.num_threads(),
               "kind": "port",
               "sample": "one full %dx%d-centre cross solve on the host (construction excluded) by the CPU "
                         "restatement of the reference's algorithm: full-field GMRES(m=20) with dense Schur blocks "
                         "to relative residual %.0e, %d iterations, %.2f s, rel-L2 %.2e vs the converged solution; "
                         "its FFT is this repository's own Stockham/Bluestein code with O(R^2) DFT stages for odd "
                         "prime factors (FFTW, which the reference calls, is not in the image)"
                         % (n, n, a.oracle_tol, it, t_cpu, cpu_err)}

    def variant(v, key):
        i = v["info"]
        return {"solves_per_s": round(1000.0 / v["ms"], 3), "ms_per_step": round(v["ms"], 4),
                "rel_l2_vs_oracle": err[key], "construct_s": round(v["build_s"], 3),
                "gmres_iterations": i.get("iterations"), "rel_residual": i.get("rel_residual"),
                "refine_sweeps": i.get("refine_steps"), "refine_error_estimate": i.get("refine_residual")}

    variants = {k: variant(v, k) for k, v in runs.items()}

    # the headline is the fastest plain FP64 form whose residual check passed
    # and (when the oracle ran) whose error meets the metric's 1e-10 bar
    def passed(k, bar):
        ok = runs[k]["info"].get("status", 0) == 0
        return ok and (err[k] is None or err[k] <= bar)
    plain = [k for k in ("f64", "f64_graph", "f64_into", "f64_graph_into") if k in runs]
    cands = [k for k in plain if passed(k, 1e-10)] or plain or list(runs)
    main_key = min(cands, key=lambda k: runs[k]["ms"])
    main_v = runs[main_key]
    # the fastest form within 1e-12 of the discrete solution
    c12 = [k for k in runs if passed(k, 1e-12) and err[k] is not None]
    k12 = min(c12, key=lambda k: runs[k]["ms"]) if c12 else None
    out = {
        "metric": "cross-geometry solves/s at 1e-10 rel L2: FP64 vs FP32+refine; HBM GB/s",
        "value": round(1000.0 / main_v["ms"], 3),
        "unit": "cross solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(main_v["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "cross %d: centre %dx%d, arms %dx%d / %dx%d, FP64, interface GMRES(m=%d) to "
                               "relative %.0e" % (n, n, n, n // 2, n, n, n // 2, a.m, a.tol),
                   "form": main_key, "cells": cells, "rel_l2_vs_oracle": err[main_key],
                   "fastest_within_1e-12": None if k12 is None else
                   {"form": k12, "solves_per_s": variants[k12]["solves_per_s"], "rel_l2_vs_oracle": err[k12]},
                   "variants": variants},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": dom["gbs"], "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(dom["gbs"] / PEAK_HBM_GBS, 4),
                     "traffic": measured_traffic(dom_name, n)},
        "kernels": kern,
        "kernels_32rhs": kern32,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
