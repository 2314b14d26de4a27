"""MI355X-native Focalors_Poisson solve path.

Python face of the C ABI (include/focalors_poisson.h), mirroring the
reference's solver classes:

  PoissonSolver2D(ctx, nx, ny, hx, hy, bc)   pe/poisson/poisson_solver2d.h:10-34
      .solve(f)                              in place, f = hx^2-scaled RHS
  ConcatPoissonSolver2D(ctx, domains, conn, m, tol, max_iter)
                                             pe/concat/concat_solver2d.h:18-56
      .solve(fields, ghost)                  in place, raw RHS -> solution

Device memory is torch tensors on a HIP device (torch is the allocator and
stream provider only); every computation runs in libfocalors_poisson.so.
"""
import ctypes

from . import _capi
from ._capi import (FP_F32_REFINE, FP_F64, FP_F64_REFINE, FP_PLAN_ESTIMATE, FP_PLAN_MEASURE,  # noqa: F401
                    FPError, SolveInfo)

DIRICHLET, NEUMANN, PERIODIC, ADJACENTED, NULL = 0, 1, 2, 3, 4
XNEG, XPOS, YNEG, YPOS = 0, 1, 2, 3


def _stream_ptr(stream):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


class Context:
    """fp_ctx on one HIP device."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        _capi.check(_capi.lib().fp_ctx_create(int(device), ctypes.byref(self._h)))
        self.device = device

    def set_planning(self, rigor):
        """FP_PLAN_MEASURE (timed launch shapes, the default) or
        FP_PLAN_ESTIMATE (estimated shapes, no timing) for the plans created
        afterwards (fp_ctx_set_planning)."""
        _capi.check(_capi.lib().fp_ctx_set_planning(self._h, int(rigor)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            _capi.lib().fp_ctx_destroy(h)
            self._h = None


def _check_field(t, dtype, n):
    import torch
    if not t.is_cuda:
        raise ValueError("fields must be device tensors")
    if t.dtype != dtype:
        raise TypeError("expected %s field, got %s" % (dtype, t.dtype))
    if not t.is_contiguous():
        raise ValueError("fields must be contiguous")
    if t.numel() % n:
        raise ValueError("field size %d is not a multiple of nx*ny = %d" % (t.numel(), n))
    return t.numel() // n


class PoissonSolver2D:
    """fp_sub: the subdomain fast solver (transform in y, tridiagonal in x)."""

    def __init__(self, ctx, nx, ny, hx, hy, bc):
        self.ctx, self.nx, self.ny, self.hx, self.hy, self.bc = ctx, nx, ny, hx, hy, tuple(bc)
        self._h = ctypes.c_void_p()
        b = (ctypes.c_ubyte * 4)(*bc)
        _capi.check(_capi.lib().fp_sub_create(ctx._h, nx, ny, float(hx), float(hy), b, ctypes.byref(self._h)))

    def solve(self, f, stream=None):
        """In-place solve of one or more hx^2-scaled right-hand sides
        (tensor of shape (..., nx, ny), float64 or float32)."""
        import torch
        n = self.nx * self.ny
        if f.dtype == torch.float64:
            nrhs = _check_field(f, torch.float64, n)
            fn = _capi.lib().fp_sub_solve
        else:
            nrhs = _check_field(f, torch.float32, n)
            fn = _capi.lib().fp_sub_solve_f32
        _capi.check(fn(self._h, ctypes.c_void_p(f.data_ptr()), nrhs, n, _stream_ptr(stream)))
        return f

    def solve_refine(self, f, tol=1e-12, max_steps=30, stream=None, allow_not_converged=False):
        """FP32 inner solves + FP64 residual/iterative refinement, in place on a
        float64 field; returns the solve statistics dict."""
        import torch
        n = self.nx * self.ny
        nrhs = _check_field(f, torch.float64, n)
        info = _capi.SolveInfo()
        allow = (_capi.FP_E_NOT_CONVERGED,) if allow_not_converged else ()
        st = _capi.check(_capi.lib().fp_sub_solve_refine(self._h, ctypes.c_void_p(f.data_ptr()), nrhs, n, float(tol),
                                                         int(max_steps), _stream_ptr(stream), ctypes.byref(info)),
                         allow)
        d = info.as_dict()
        d["status"] = st
        return d

    def transform(self, f, inverse=False, stream=None):
        """PoissonFFT2D::transform / transform_transpose along y, in place (float64)."""
        import torch
        n = self.nx * self.ny
        nrhs = _check_field(f, torch.float64, n)
        _capi.check(_capi.lib().fp_sub_transform(self._h, ctypes.c_void_p(f.data_ptr()), nrhs, n, int(inverse),
                                                 _stream_ptr(stream)))
        return f

    def chasing(self, f, stream=None):
        """ChasingMethod2D::chasing along x on a transformed field, in place (float64)."""
        import torch
        n = self.nx * self.ny
        nrhs = _check_field(f, torch.float64, n)
        _capi.check(_capi.lib().fp_sub_chasing(self._h, ctypes.c_void_p(f.data_ptr()), nrhs, n,
                                               _stream_ptr(stream)))
        return f

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            _capi.lib().fp_sub_destroy(h)
            self._h = None


class PoissonSolver3D:
    """fp_sub3d: the 3-D subdomain fast solver (transforms in z and y,
    tridiagonal in x), pe/poisson/poisson_solver3d.h.  bc: six face types
    (x-, x+, y-, y+, z-, z+)."""

    def __init__(self, ctx, nx, ny, nz, hx, hy, hz, bc):
        self.ctx, self.nx, self.ny, self.nz = ctx, nx, ny, nz
        self.hx, self.hy, self.hz, self.bc = hx, hy, hz, tuple(bc)
        if len(self.bc) != 6:
            raise ValueError("a 3-D subdomain has six faces")
        self._h = ctypes.c_void_p()
        b = (ctypes.c_ubyte * 6)(*bc)
        _capi.check(_capi.lib().fp_sub3d_create(ctx._h, nx, ny, nz, float(hx), float(hy), float(hz), b,
                                                ctypes.byref(self._h)))

    def solve(self, f, stream=None):
        """In-place solve of one or more hx^2-scaled right-hand sides (float64
        tensor of shape (..., nx, ny, nz), z contiguous)."""
        import torch
        n = self.nx * self.ny * self.nz
        nrhs = _check_field(f, torch.float64, n)
        _capi.check(_capi.lib().fp_sub3d_solve(self._h, ctypes.c_void_p(f.data_ptr()), nrhs, n, _stream_ptr(stream)))
        return f

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            _capi.lib().fp_sub3d_destroy(h)
            self._h = None


def solve_batched(ctx, solvers, fields, stream=None, noalloc=False):
    """Many independent subdomain solves in one call (fp_sub_solve_batched;
    noalloc=True: fp_sub_solve_batched_noalloc, FP64 only, which allocates no
    device memory)."""
    import torch
    k = len(solvers)
    if k == 0:
        return
    f64 = fields[0].dtype == torch.float64
    subs = (ctypes.c_void_p * k)(*[s._h.value for s in solvers])
    ptrs = (ctypes.c_void_p * k)(*[f.data_ptr() for f in fields])
    nr = (ctypes.c_int * k)(*[_check_field(f, fields[0].dtype, s.nx * s.ny) for s, f in zip(solvers, fields)])
    if noalloc:
        if not f64:
            raise TypeError("the no-allocation batched solve takes float64 fields")
        fn = _capi.lib().fp_sub_solve_batched_noalloc
    else:
        fn = _capi.lib().fp_sub_solve_batched if f64 else _capi.lib().fp_sub_solve_batched_f32
    _capi.check(fn(ctx._h, k, subs, ptrs, nr, _stream_ptr(stream)))


class BatchSolve:
    """fp_batch: a fixed batch of independent subdomain solves over bound
    fields, prepared once; solve() queues only kernel launches and can be
    captured into a graph (torch.cuda.graph).  The solvers and fields must
    stay alive as long as the batch."""

    def __init__(self, ctx, solvers, fields):
        import torch
        k = len(solvers)
        if k == 0 or len(fields) != k:
            raise ValueError("need one field per solver")
        self.solvers, self.fields = list(solvers), list(fields)
        fp32 = fields[0].dtype == torch.float32
        subs = (ctypes.c_void_p * k)(*[s._h.value for s in solvers])
        ptrs = (ctypes.c_void_p * k)(*[f.data_ptr() for f in fields])
        nr = (ctypes.c_int * k)(*[_check_field(f, fields[0].dtype, s.nx * s.ny) for s, f in zip(solvers, fields)])
        self._h = ctypes.c_void_p()
        _capi.check(_capi.lib().fp_batch_create(ctx._h, k, subs, ptrs, nr, int(fp32), ctypes.byref(self._h)))

    def solve(self, stream=None):
        _capi.check(_capi.lib().fp_batch_solve(self._h, _stream_ptr(stream)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            _capi.lib().fp_batch_destroy(h)
            self._h = None


class ConcatPoissonSolver2D:
    """fp_cat: composite geometry solve (tree sweep + interface GMRES).

    domains: list of (nx, ny, hx, hy, (bc_xneg, bc_xpos, bc_yneg, bc_ypos))
    conn:    list of (a, dir, b) = Geometry2D::connect(a, dir, b)
    """

    def __init__(self, ctx, domains, conn, m=20, tol=1e-8, tol_is_relative=False, max_iter=100,
                 precision=FP_F64):
        self.ctx, self.domains, self.conn = ctx, list(domains), list(conn)
        nd = len(domains)
        self._keep = [
            (ctypes.c_int * nd)(*[d[0] for d in domains]), (ctypes.c_int * nd)(*[d[1] for d in domains]),
            (ctypes.c_double * nd)(*[d[2] for d in domains]), (ctypes.c_double * nd)(*[d[3] for d in domains]),
            (ctypes.c_ubyte * (4 * nd))(*[t for d in domains for t in d[4]]),
            (ctypes.c_int * max(1, 3 * len(conn)))(*[v for c in conn for v in c]),
        ]
        k = self._keep
        desc = _capi.GeometryDesc(nd, k[0], k[1], k[2], k[3], k[4], len(conn), k[5])
        self._h = ctypes.c_void_p()
        _capi.check(_capi.lib().fp_cat_create(ctx._h, ctypes.byref(desc), int(m), float(tol), int(tol_is_relative),
                                              int(max_iter), int(precision), ctypes.byref(self._h)))
        self.root = _capi.lib().fp_cat_root(self._h)

    def set_refinement(self, tol=1e-12, max_steps=10):
        """FP_F32_REFINE / FP_F64_REFINE handles: target estimated relative
        error and sweep cap of the iterative refinement (fp_cat_set_refinement)."""
        _capi.check(_capi.lib().fp_cat_set_refinement(self._h, float(tol), int(max_steps)))

    def _ptrs(self, fields, ghost):
        import torch
        nd = len(self.domains)
        if len(fields) != nd:
            raise ValueError("need one field tensor per domain")
        nrhs = None
        for d, f in enumerate(fields):
            r = _check_field(f, torch.float64, self.domains[d][0] * self.domains[d][1])
            if nrhs is None:
                nrhs = r
            elif r != nrhs:
                raise ValueError("all domains need the same number of right-hand sides")
        fp = (ctypes.c_void_p * nd)(*[f.data_ptr() for f in fields])
        gp = (ctypes.c_void_p * (4 * nd))()
        for (d, face), g in (ghost or {}).items():
            if not g.is_cuda or g.dtype != torch.float64 or not g.is_contiguous():
                raise ValueError("ghost data must be contiguous float64 device tensors")
            flen = self.domains[d][1] if face < 2 else self.domains[d][0]
            if g.numel() != flen * nrhs:
                raise ValueError("ghost data of domain %d face %d needs %d values" % (d, face, flen * nrhs))
            gp[4 * d + face] = g.data_ptr()
        return fp, gp, nrhs

    def solve(self, fields, ghost=None, stream=None, allow_not_converged=False):
        """In-place composite solve; returns the GMRES statistics dict."""
        fp, gp, nrhs = self._ptrs(fields, ghost)
        info = _capi.SolveInfo()
        allow = (_capi.FP_E_NOT_CONVERGED,) if allow_not_converged else ()
        st = _capi.check(_capi.lib().fp_cat_solve(self._h, fp, gp, nrhs, _stream_ptr(stream), ctypes.byref(info)),
                         allow)
        d = info.as_dict()
        d["status"] = st
        return d

    def solve_into(self, rhs, out, ghost=None, stream=None, allow_not_converged=False):
        """Out-of-place composite solve (fp_cat_solve_into): reads the raw
        right-hand sides `rhs` (unchanged), writes the solution to `out`."""
        rp, gp, nrhs = self._ptrs(rhs, ghost)
        op, _, nout = self._ptrs(out, None)
        if nout != nrhs:
            raise ValueError("rhs and out need the same number of right-hand sides")
        info = _capi.SolveInfo()
        allow = (_capi.FP_E_NOT_CONVERGED,) if allow_not_converged else ()
        st = _capi.check(_capi.lib().fp_cat_solve_into(self._h, rp, op, gp, nrhs, _stream_ptr(stream),
                                                       ctypes.byref(info)), allow)
        d = info.as_dict()
        d["status"] = st
        return d

    def graph_capture(self, fields, ghost=None, stream=None, rhs=None):
        """Records the FP64 solve of these fields (and ghost arrays) into a HIP
        graph on `stream` (a non-default torch stream; fp_cat_graph_capture).
        The tensors stay bound to the graph: refill them, then graph_launch().
        rhs given: the replays read the right-hand sides there and write the
        solution to `fields` (fp_cat_graph_capture_into)."""
        fp, gp, nrhs = self._ptrs(fields, ghost)
        if stream is None:
            raise ValueError("graph capture needs an explicit, non-default stream")
        if rhs is None:
            self._graph_keep = (list(fields), dict(ghost or {}), fp, gp)
            _capi.check(_capi.lib().fp_cat_graph_capture(self._h, fp, gp, nrhs, _stream_ptr(stream)))
            return
        rp, _, nr = self._ptrs(rhs, None)
        if nr != nrhs:
            raise ValueError("rhs and fields need the same number of right-hand sides")
        self._graph_keep = (list(fields), list(rhs), dict(ghost or {}), fp, gp, rp)
        _capi.check(_capi.lib().fp_cat_graph_capture_into(self._h, rp, fp, gp, nrhs, _stream_ptr(stream)))

    def graph_launch(self, stream=None, wait=True, allow_not_converged=False):
        """Replays the captured solve.  wait=True: synchronises and returns the
        residual check as a statistics dict; wait=False: returns None at once."""
        if not wait:
            _capi.check(_capi.lib().fp_cat_graph_launch(self._h, _stream_ptr(stream), None))
            return None
        info = _capi.SolveInfo()
        allow = (_capi.FP_E_NOT_CONVERGED,) if allow_not_converged else ()
        st = _capi.check(_capi.lib().fp_cat_graph_launch(self._h, _stream_ptr(stream), ctypes.byref(info)), allow)
        d = info.as_dict()
        d["status"] = st
        return d

    def transform(self, fields, inverse=False, stream=None):
        """The y transforms of every domain in the solve's one batched launch,
        in place (fp_cat_transform)."""
        fp, _, nrhs = self._ptrs(fields, None)
        _capi.check(_capi.lib().fp_cat_transform(self._h, fp, nrhs, int(inverse), _stream_ptr(stream)))
        return fields

    def vjp(self, grad_u, faces=(), stream=None):
        """Adjoint of solve: returns (dL/df per domain, {(d, face): dL/dg}) for
        dL/du = grad_u (one tensor per domain, same shapes as the fields)."""
        import torch
        nd = len(self.domains)
        gu = [g.contiguous() for g in grad_u]
        up, _, nrhs = self._ptrs(gu, None)
        gf = [torch.empty_like(g) for g in gu]
        fp = (ctypes.c_void_p * nd)(*[t.data_ptr() for t in gf])
        gp = (ctypes.c_void_p * (4 * nd))()
        gg = {}
        for d, face in faces:
            flen = self.domains[d][1] if face < 2 else self.domains[d][0]
            t = torch.empty(nrhs * flen, dtype=torch.float64, device=gu[0].device)
            gg[(d, face)] = t
            gp[4 * d + face] = t.data_ptr()
        info = _capi.SolveInfo()
        _capi.check(_capi.lib().fp_cat_vjp(self._h, up, fp, gp, nrhs, _stream_ptr(stream), ctypes.byref(info)))
        return gf, gg

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            _capi.lib().fp_cat_destroy(h)
            self._h = None


class PressureProjection:
    """fp_proj: the pressure projection of an incompressible flow step on a
    staggered grid over the composite geometry of `cat` (a
    ConcatPoissonSolver2D, kept alive here): u = u* - dt G p with
    D G p = D u* / dt.  Per domain d: u[d] the x-face velocities
    ((nx+1) x ny), v[d] the y-face velocities (nx x (ny+1)), p[d] the
    pressure (nx x ny), all contiguous float64 device tensors; pghost: dict
    (domain, face) -> the pressure's ghost-centre values on Dirichlet faces."""

    def __init__(self, cat):
        self.cat = cat
        self._h = ctypes.c_void_p()
        _capi.check(_capi.lib().fp_proj_create(cat._h, ctypes.byref(self._h)))

    def _arr(self, ts, rows, cols, what):
        import torch
        nd = len(self.cat.domains)
        if len(ts) != nd:
            raise ValueError("need one %s tensor per domain" % what)
        for d, t in enumerate(ts):
            nx, ny = self.cat.domains[d][0], self.cat.domains[d][1]
            _check_field(t, torch.float64, (nx + rows) * (ny + cols))
            if t.numel() != (nx + rows) * (ny + cols):
                raise ValueError("%s of domain %d needs %d values" % (what, d, (nx + rows) * (ny + cols)))
        return (ctypes.c_void_p * nd)(*[t.data_ptr() for t in ts])

    def _ghost(self, pghost):
        import torch
        nd = len(self.cat.domains)
        gp = (ctypes.c_void_p * (4 * nd))()
        for (d, face), g in (pghost or {}).items():
            if not g.is_cuda or g.dtype != torch.float64 or not g.is_contiguous():
                raise ValueError("ghost data must be contiguous float64 device tensors")
            flen = self.cat.domains[d][1] if face < 2 else self.cat.domains[d][0]
            if g.numel() != flen:
                raise ValueError("ghost data of domain %d face %d needs %d values" % (d, face, flen))
            gp[4 * d + face] = g.data_ptr()
        return gp

    def divergence(self, u, v, out, scale=1.0, stream=None):
        """out[d] = scale D(u, v) (fp_proj_divergence)."""
        _capi.check(_capi.lib().fp_proj_divergence(self._h, self._arr(u, 1, 0, "u"), self._arr(v, 0, 1, "v"),
                                                   float(scale), self._arr(out, 0, 0, "out"), _stream_ptr(stream)))

    def correct(self, u, v, p, pghost=None, scale=1.0, stream=None):
        """u -= scale Gx p, v -= scale Gy p (fp_proj_correct)."""
        _capi.check(_capi.lib().fp_proj_correct(self._h, self._arr(u, 1, 0, "u"), self._arr(v, 0, 1, "v"),
                                                self._arr(p, 0, 0, "p"), self._ghost(pghost), float(scale),
                                                _stream_ptr(stream)))

    def step(self, u, v, p, dt, pghost=None, stream=None):
        """One projection, in place on u and v, pressure into p
        (fp_proj_step); returns the solve statistics."""
        info = _capi.SolveInfo()
        _capi.check(_capi.lib().fp_proj_step(self._h, self._arr(u, 1, 0, "u"), self._arr(v, 0, 1, "v"),
                                             self._arr(p, 0, 0, "p"), self._ghost(pghost), float(dt),
                                             _stream_ptr(stream), ctypes.byref(info)))
        return info.as_dict()

    def capture(self, u, v, p, dt, pghost=None, stream=None):
        """Records step() on these tensors into a HIP graph (fp_proj_capture;
        a non-default stream)."""
        self._bound = (list(u), list(v), list(p), dict(pghost or {}))
        _capi.check(_capi.lib().fp_proj_capture(self._h, self._arr(u, 1, 0, "u"), self._arr(v, 0, 1, "v"),
                                                self._arr(p, 0, 0, "p"), self._ghost(pghost), float(dt),
                                                _stream_ptr(stream)))

    def launch(self, stream=None, wait=False):
        """Replays the captured step (fp_proj_launch); with wait, returns the
        solve statistics of the replay."""
        info = _capi.SolveInfo()
        _capi.check(_capi.lib().fp_proj_launch(self._h, _stream_ptr(stream), ctypes.byref(info) if wait else None))
        return info.as_dict() if wait else None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            _capi.lib().fp_proj_destroy(h)
            self._h = None


class _Fields:
    """Device pointers inside registered float64 tensors, as tensor views:
    the buffers a domain-solver callback is handed all lie in tensors the
    Python side owns (workspaces, the right-hand side being solved)."""

    def __init__(self, *tensors):
        self.tensors = list(tensors)

    def view(self, ptr, nrhs, nx, ny):
        n = nrhs * nx * ny
        for t in self.tensors:
            base = t.data_ptr()
            if base <= ptr and ptr + 8 * n <= base + 8 * t.numel():
                off = (ptr - base) // 8
                return t.view(-1)[off:off + n].view(nrhs, nx, ny)
        raise ValueError("the domain solver was handed a buffer outside the registered tensors")


class _DomainSolver:
    """A DomainSolver2D for the C ABI (fp_domain_solve_fn): a PoissonSolver2D
    (its fp_sub through fp_sub_domain_solve, no Python in the loop), or a
    callable solve(fields) acting in place on a (nrhs, nx, ny) float64 view
    on the current stream."""

    def __init__(self, solver, fields):
        self.keep = solver
        if isinstance(solver, PoissonSolver2D):
            self.fn = ctypes.cast(_capi.lib().fp_sub_domain_solve, _capi.DOMAIN_FN)
            self.user = solver._h
        else:
            def cb(user, ptr, nx, ny, nrhs, stream):
                try:
                    solver(fields.view(ptr, nrhs, nx, ny))
                    return 0
                except Exception:  # reported to the library as a failed solve
                    import traceback
                    traceback.print_exc()
                    return 1
            self.fn = _capi.DOMAIN_FN(cb)
            self.user = None


class SchurMat2D:
    """fp_schur2d: SchurMat2D_{xneg,xpos,yneg,ypos} (pe/concat/schur_mat2d.h:7-69)
    of a branch domain bnx x bny on face `loc` of the root, built by unit
    solves of branch_solver (a PoissonSolver2D or a callable, see
    GMRESSolver2D)."""

    def __init__(self, ctx, loc, bnx, bny, branch_solver, stream=None, chunk=64):
        import torch
        cn = bny if loc in (XNEG, XPOS) else bnx
        self.loc, self.cn = loc, cn
        self.work = torch.empty(max(1, min(chunk, cn)) * bnx * bny, dtype=torch.float64,
                                device=torch.device("cuda", ctx.device))
        self._solver = _DomainSolver(branch_solver, _Fields(self.work))
        self._h = ctypes.c_void_p()
        _capi.check(_capi.lib().fp_schur2d_create(ctx._h, int(loc), int(bnx), int(bny), self._solver.fn,
                                                  self._solver.user, self.work.data_ptr(), self.work.numel(),
                                                  _stream_ptr(stream), ctypes.byref(self._h)))

    def values(self):
        import numpy as np
        out = np.empty((self.cn, self.cn))
        _capi.check(_capi.lib().fp_schur2d_values(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def apply(self, root, out=None, accumulate=False, stream=None):
        import torch
        if out is None:
            out = torch.zeros_like(root)
        _capi.check(_capi.lib().fp_schur2d_apply(self._h, root.data_ptr(), root.shape[-2], root.shape[-1],
                                                 out.data_ptr(), int(accumulate), _stream_ptr(stream)))
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            _capi.lib().fp_schur2d_destroy(h)
            self._h = None


class GMRESSolver2D:
    """fp_gmres2d: GMRESSolver2D (pe/concat/gmres_solver2d.h:12-57) on an
    nx x ny root field: solves (I - A^-1 S) x = A^-1 b with A^-1 the solver of
    set_solver and S the Schur blocks of schur_mat_construct / add_schur.
    Solvers are PoissonSolver2D objects or callables f(fields) that solve the
    (nrhs, nx, ny) float64 device view in place on the current stream."""

    def __init__(self, ctx, nx, ny, m, tol, max_iter):
        import torch
        self.ctx, self.nx, self.ny = ctx, nx, ny
        n = _capi.lib().fp_gmres2d_workspace(int(nx), int(ny), int(m))
        self.work = torch.empty(n, dtype=torch.float64, device=torch.device("cuda", ctx.device))
        self._fields = _Fields(self.work)
        self._schur, self._solver, self._x0 = [], None, None
        self._h = ctypes.c_void_p()
        _capi.check(_capi.lib().fp_gmres2d_create(ctx._h, int(nx), int(ny), int(m), float(tol), int(max_iter),
                                                  self.work.data_ptr(), n, ctypes.byref(self._h)))

    def set_solver(self, solver):
        self._solver = _DomainSolver(solver, self._fields)
        _capi.check(_capi.lib().fp_gmres2d_set_solver(self._h, self._solver.fn, self._solver.user))

    def add_schur(self, S):
        _capi.check(_capi.lib().fp_gmres2d_add_schur(self._h, S._h))
        self._schur.append(S)

    def schur_mat_construct(self, adjacency, stream=None):
        """adjacency: {loc: (bnx, bny, branch_solver)} (schur_mat_construct,
        gmres_solver2d.cpp:39-91)."""
        for loc, (bnx, bny, solver) in adjacency.items():
            self.add_schur(SchurMat2D(self.ctx, loc, bnx, bny, solver, stream=stream))

    def set_initial_guess(self, x0):
        self._x0 = x0
        _capi.check(_capi.lib().fp_gmres2d_set_initial_guess(self._h, None if x0 is None else x0.data_ptr()))

    def solve(self, b, stream=None, allow_not_converged=False, max_res=1024):
        """Overwrites b (nx x ny float64, on the device) with x; returns the
        solve info with the restart residuals under "resvec"."""
        import numpy as np
        _check_field(b, b.dtype, self.nx * self.ny)
        self._fields.tensors = [self.work, b]
        info = _capi.SolveInfo()
        res = np.zeros(max_res)
        nres = ctypes.c_int(0)
        st = _capi.lib().fp_gmres2d_solve(self._h, b.data_ptr(), _stream_ptr(stream), ctypes.byref(info),
                                          res.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), max_res,
                                          ctypes.byref(nres))
        self._fields.tensors = [self.work]
        _capi.check(st, allow=(_capi.FP_E_NOT_CONVERGED,) if allow_not_converged else ())
        out = info.as_dict()
        out["status"] = st
        out["resvec"] = res[:min(nres.value, max_res)].copy()
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            _capi.lib().fp_gmres2d_destroy(h)
            self._h = None


class ConcatPoissonSolver3D:
    """fp_cat3d: 3-D composite geometry solve (tree sweep + face GMRES),
    pe/concat/concat_solver3d.h.

    boxes: list of (nx, ny, nz, hx, hy, hz, (bc x-, x+, y-, y+, z-, z+))
    conn:  list of (a, dir, b) = Geometry3D::connect(a, dir, b), dir 0..5
    """

    def __init__(self, ctx, boxes, conn, m=20, tol=1e-8, tol_is_relative=False, max_iter=100):
        self.ctx, self.boxes, self.conn = ctx, list(boxes), list(conn)
        nd = len(boxes)
        ci = lambda k: (ctypes.c_int * nd)(*[b[k] for b in boxes])
        cd = lambda k: (ctypes.c_double * nd)(*[b[k] for b in boxes])
        self._keep = [ci(0), ci(1), ci(2), cd(3), cd(4), cd(5),
                      (ctypes.c_ubyte * (6 * nd))(*[t for b in boxes for t in b[6]]),
                      (ctypes.c_int * max(1, 3 * len(conn)))(*[v for c in conn for v in c])]
        k = self._keep
        desc = _capi.Geometry3DDesc(nd, k[0], k[1], k[2], k[3], k[4], k[5], k[6], len(conn), k[7])
        self._h = ctypes.c_void_p()
        _capi.check(_capi.lib().fp_cat3d_create(ctx._h, ctypes.byref(desc), int(m), float(tol), int(tol_is_relative),
                                                int(max_iter), ctypes.byref(self._h)))
        self.root = _capi.lib().fp_cat3d_root(self._h)

    def solve(self, fields, ghost=None, stream=None, allow_not_converged=False):
        """In-place composite solve: fields[d] (float64, shape (..., nx, ny, nz))
        raw right-hand sides -> solutions; ghost: {(box, face): plane data}."""
        import torch
        nd = len(self.boxes)
        nrhs = None
        for d, f in enumerate(fields):
            b = self.boxes[d]
            r = _check_field(f, torch.float64, b[0] * b[1] * b[2])
            if nrhs is None:
                nrhs = r
            elif r != nrhs:
                raise ValueError("all boxes need the same number of right-hand sides")
        fp = (ctypes.c_void_p * nd)(*[f.data_ptr() for f in fields])
        gp = (ctypes.c_void_p * (6 * nd))()
        for (d, face), g in (ghost or {}).items():
            nx, ny, nz = self.boxes[d][:3]
            cnt = (ny * nz, ny * nz, nx * nz, nx * nz, nx * ny, nx * ny)[face]
            if not g.is_cuda or g.dtype != torch.float64 or not g.is_contiguous() or g.numel() != cnt * nrhs:
                raise ValueError("ghost data of box %d face %d: %d contiguous float64 device values" %
                                 (d, face, cnt * nrhs))
            gp[6 * d + face] = g.data_ptr()
        info = _capi.SolveInfo()
        allow = (_capi.FP_E_NOT_CONVERGED,) if allow_not_converged else ()
        st = _capi.check(_capi.lib().fp_cat3d_solve(self._h, fp, gp, nrhs, _stream_ptr(stream), ctypes.byref(info)),
                         allow)
        d = info.as_dict()
        d["status"] = st
        return d

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            _capi.lib().fp_cat3d_destroy(h)
            self._h = None
