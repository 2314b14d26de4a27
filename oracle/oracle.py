"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product package never
imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

DIRICHLET, NEUMANN, PERIODIC, ADJACENTED, NULL = 0, 1, 2, 3, 4
XNEG, XPOS, YNEG, YPOS = 0, 1, 2, 3
SCHUR_UNIT_SOLVES, SCHUR_SPECTRAL = 0, 1

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.or_dft.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.or_r2c.argtypes = [ctypes.c_int, _dp, _dp]
        L.or_plan_count.argtypes = []
        L.or_transform.argtypes = [ctypes.c_int] * 5 + [_dp, _dp]
        L.or_x_diag.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp]
        L.or_chasing.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.or_chasing.restype = None
        L.or_chasing_singular.argtypes = [ctypes.c_int, _dp, _dp]
        L.or_chasing_singular.restype = None
        L.or_chasing_periodic.argtypes = [ctypes.c_double, ctypes.c_int, _dp, _dp]
        L.or_chasing_periodic.restype = None
        L.or_chasing_periodic_singular.argtypes = [ctypes.c_int, _dp, _dp]
        L.or_chasing_periodic_singular.restype = None
        L.or_sub_solve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _ip, _dp]
        L.or_sub_solve3d.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, _ip, _dp]
        L.or_cat_create.restype = ctypes.c_void_p
        L.or_cat_create.argtypes = [ctypes.c_int, _ip, _ip, _dp, _dp, _ip, ctypes.c_int, _ip, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_cat_root.argtypes = [ctypes.c_void_p]
        L.or_cat_tree.argtypes = [ctypes.c_void_p, _ip, _ip, _ip]
        L.or_cat_solve.argtypes = [ctypes.c_void_p, ctypes.POINTER(_dp), ctypes.POINTER(_dp), _ip, _dp]
        L.or_cat_schur.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _ip]
        L.or_cat_residual.argtypes = [ctypes.c_void_p, ctypes.POINTER(_dp), ctypes.POINTER(_dp),
                                      ctypes.POINTER(_dp), ctypes.POINTER(_dp)]
        L.or_cat_solve_refined.argtypes = [ctypes.c_void_p, ctypes.POINTER(_dp), ctypes.POINTER(_dp), ctypes.c_int,
                                           ctypes.c_double, _ip, _dp]
        L.or_cat_destroy.argtypes = [ctypes.c_void_p]
        L.or_cat_destroy.restype = None
        L.or_gmres_mock.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                    _dp, _ip]
        L.or_num_threads.argtypes = []
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def dft(x, sign=-1):
    x = np.asarray(x, dtype=np.complex128)
    xi = np.ascontiguousarray(x.view(np.float64))
    out = np.empty_like(xi)
    assert lib().or_dft(len(x), sign, _p(xi), _p(out)) == 0
    return out.view(np.complex128)


def r2c(x):
    x = _f64(x)
    out = np.empty(2 * (len(x) // 2 + 1))
    assert lib().or_r2c(len(x), _p(x), _p(out)) == 0
    return out.view(np.complex128)


def transform(bc_neg, bc_pos, field, inverse=False):
    f = _f64(np.atleast_2d(field))
    out = np.empty_like(f)
    rc = lib().or_transform(bc_neg, bc_pos, int(inverse), f.shape[0], f.shape[1], _p(f), _p(out))
    assert rc == 0, "unsupported boundary pair"
    return out


def x_diag(ny, hx, hy, bc_neg, bc_pos):
    out = np.empty(ny)
    assert lib().or_x_diag(ny, hx, hy, bc_neg, bc_pos, _p(out)) == 0
    return out


def chasing(diag, f, bc_start=DIRICHLET, bc_end=DIRICHLET):
    f = _f64(f)
    x = np.empty_like(f)
    lib().or_chasing(float(diag), len(f), bc_start, bc_end, _p(f), _p(x))
    return x


def chasing_singular(f):
    f = _f64(f)
    x = np.empty_like(f)
    lib().or_chasing_singular(len(f), _p(f), _p(x))
    return x


def chasing_periodic(diag, f):
    f = _f64(f)
    x = np.empty_like(f)
    lib().or_chasing_periodic(float(diag), len(f), _p(f), _p(x))
    return x


def chasing_periodic_singular(f):
    f = _f64(f)
    x = np.empty_like(f)
    lib().or_chasing_periodic_singular(len(f), _p(f), _p(x))
    return x


def sub_solve(f, hx, hy, bc):
    """PoissonSolver2D::solve of an already hx^2-scaled RHS (returns a new array)."""
    f = _f64(f).copy()
    b = (ctypes.c_int * 4)(*bc)
    assert lib().or_sub_solve(f.shape[0], f.shape[1], hx, hy, b, _p(f)) == 0
    return f


def sub_solve3d(f, hx, hy, hz, bc):
    """PoissonSolver3D::solve of an already hx^2-scaled RHS of shape
    (nx, ny, nz) (returns a new array)."""
    f = _f64(f).copy()
    b = (ctypes.c_int * 6)(*bc)
    assert lib().or_sub_solve3d(f.shape[0], f.shape[1], f.shape[2], hx, hy, hz, b, _p(f)) == 0
    return f


class Composite:
    """ConcatPoissonSolver2D restated on the CPU.

    domains: list of (nx, ny, hx, hy, (bc_xneg, bc_xpos, bc_yneg, bc_ypos))
    conn: list of (a, dir, b) as Geometry2D::connect(a, dir, b).
    """

    def __init__(self, domains, conn, m=20, tol=1e-8, tol_is_relative=False, maxiter=100,
                 schur_mode=SCHUR_UNIT_SOLVES):
        self.domains = domains
        nd = len(domains)
        self.nd = nd
        nx = (ctypes.c_int * nd)(*[d[0] for d in domains])
        ny = (ctypes.c_int * nd)(*[d[1] for d in domains])
        hx = (ctypes.c_double * nd)(*[d[2] for d in domains])
        hy = (ctypes.c_double * nd)(*[d[3] for d in domains])
        bc = (ctypes.c_int * (4 * nd))(*[t for d in domains for t in d[4]])
        flat = [v for c in conn for v in c]
        cc = (ctypes.c_int * max(1, len(flat)))(*flat)
        self.h = lib().or_cat_create(nd, nx, ny, hx, hy, bc, len(conn), cc, m, tol, int(tol_is_relative),
                                     maxiter, schur_mode)
        if not self.h:
            raise ValueError("or_cat_create failed")

    def root(self):
        return lib().or_cat_root(self.h)

    def tree(self):
        lv = (ctypes.c_int * self.nd)()
        pa = (ctypes.c_int * self.nd)()
        pl = (ctypes.c_int * self.nd)()
        lib().or_cat_tree(self.h, lv, pa, pl)
        return list(lv), list(pa), list(pl)

    def schur(self, child):
        cn = ctypes.c_int(0)
        assert lib().or_cat_schur(self.h, child, None, ctypes.byref(cn)) == 0
        out = np.empty((cn.value, cn.value))
        lib().or_cat_schur(self.h, child, _p(out), ctypes.byref(cn))
        return out

    def _buffers(self, buffers):
        keep = []
        bptr = (_dp * (4 * self.nd))()
        for (d, face), g in (buffers or {}).items():
            g = _f64(g)
            keep.append(g)
            bptr[4 * d + face] = _p(g)
        return bptr, keep

    def solve(self, fields, buffers=None, refine=0, tol_refine=1e-10):
        """fields: list of raw RHS arrays; buffers: dict (dom, face) -> ghost data.
        refine > 0: that many refinement steps with the long-double composite
        residual after the solve (or_cat_solve_refined; converged to about FP64
        rounding).  Returns (solutions, gmres_iterations, final_residual)."""
        fs = [_f64(f).copy() for f in fields]
        fptr = (_dp * self.nd)(*[_p(f) for f in fs])
        bptr, keep = self._buffers(buffers)
        it = ctypes.c_int(0)
        res = ctypes.c_double(0)
        if refine > 0:
            rc = lib().or_cat_solve_refined(self.h, fptr, bptr, int(refine), float(tol_refine), ctypes.byref(it),
                                            ctypes.byref(res))
        else:
            rc = lib().or_cat_solve(self.h, fptr, bptr, ctypes.byref(it), ctypes.byref(res))
        assert rc == 0
        return fs, it.value, res.value

    def residual(self, u, fields, buffers=None):
        """hx^2 F - A u of the assembled composite system (long-double sums)."""
        us = [_f64(x) for x in u]
        fs = [_f64(f) for f in fields]
        rs = [np.empty_like(f) for f in fs]
        bptr, keep = self._buffers(buffers)
        assert lib().or_cat_residual(self.h, (_dp * self.nd)(*[_p(x) for x in us]),
                                     (_dp * self.nd)(*[_p(f) for f in fs]), bptr,
                                     (_dp * self.nd)(*[_p(r) for r in rs])) == 0
        return rs

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_cat_destroy(self.h)
            self.h = None


def reference_solution(domains, conn, fields, buffers=None, refine=2, tol=1e-11, tol_refine=1e-7, m=40):
    """The converged FP64 composite solution the parity tests compare with:
    interface GMRES(m) to relative tol, then `refine` steps of refinement with
    the long-double composite residual, whose correction solves run to
    relative tol_refine (each step multiplies the error by about that
    accuracy, so two steps reach FP64 rounding from any tol <= 1e-8)."""
    cat = Composite(domains, conn, m=m, tol=tol, tol_is_relative=True, maxiter=2000, schur_mode=SCHUR_SPECTRAL)
    return cat.solve(fields, buffers, refine=refine, tol_refine=tol_refine)[0]


def gmres_mock(ainv, S, b, m, tol, maxiter):
    n = len(b)
    b = _f64(b).copy()
    resvec = np.zeros(maxiter + 1)
    nres = ctypes.c_int(0)
    rc = lib().or_gmres_mock(n, _p(_f64(ainv)), _p(_f64(S)), _p(b), m, tol, maxiter, _p(resvec),
                             ctypes.byref(nres))
    return b, resvec[:nres.value], rc


def num_threads():
    return lib().or_num_threads()


def plan_count():
    """FFT lengths with a cached plan in the oracle (one plan per length)."""
    return lib().or_plan_count()


# ---- 3-D composite (ConcatPoissonSolver3D), numpy restatement ----

def _face3(loc, shape):
    """Index tuple of face `loc` (x-, x+, y-, y+, z-, z+) of a field3 array
    (field3::xneg_bond_add and the like, base/field/field3.cpp)."""
    nx, ny, nz = shape
    return {0: (0, slice(None), slice(None)), 1: (nx - 1, slice(None), slice(None)),
            2: (slice(None), 0, slice(None)), 3: (slice(None), ny - 1, slice(None)),
            4: (slice(None), slice(None), 0), 5: (slice(None), slice(None), nz - 1)}[loc]


class Composite3D:
    """ConcatPoissonSolver3D restated on the CPU (pe/concat/concat_solver3d.cpp,
    gmres_solver3d.cpp, schur_mat3d.cpp), subdomain solves by or_sub_solve3d.
    For small boxes: the Schur blocks are built by one child solve per face
    cell, as the reference builds them.

    boxes: list of (nx, ny, nz, hx, hy, hz, bc6); conn: (a, dir, b) as
    Geometry3D::connect(a, dir, b), dir 0..5."""

    def __init__(self, boxes, conn, m=20, tol=1e-8, tol_is_relative=False, maxiter=100):
        self.boxes, self.conn = list(boxes), list(conn)
        self.m, self.tol, self.rel, self.maxiter = m, tol, tol_is_relative, maxiter
        nd = len(boxes)
        self.bc = [list(b[6]) for b in boxes]
        self.adj = [[-1] * 6 for _ in range(nd)]
        for a, dr, b in conn:
            self.adj[a][dr], self.adj[b][dr ^ 1] = b, a
            self.bc[a][dr], self.bc[b][dr ^ 1] = ADJACENTED, ADJACENTED
        self._tree()
        self.S = {}
        # construct_solver_map: deepest level first, root last (concat_solver3d.cpp:74-95)
        for lev in reversed(self.levels):
            for d in lev:
                for loc, c in self.children[d]:
                    self.S[c] = self._schur(d, c, loc)
        self.last_iters = 0

    def _tree(self):
        """TreeUtils root choice (geometry_tree.hpp:22-95; lower index on a tie) and BFS."""
        nd = len(self.boxes)
        root = 0
        if nd > 1:
            deg = [sum(1 for x in self.adj[a] if x >= 0) for a in range(nd)]
            q = [a for a in range(nd) if deg[a] <= 1]
            qh, remaining = 0, nd
            while remaining > 2 and qh < len(q):
                cnt = len(q) - qh
                remaining -= cnt
                for _ in range(cnt):
                    leaf = q[qh]
                    qh += 1
                    for nb in self.adj[leaf]:
                        if nb >= 0:
                            deg[nb] -= 1
                            if deg[nb] == 1:
                                q.append(nb)
            if qh < len(q):
                root = q[qh]
                if qh + 1 < len(q) and q[qh + 1] < root:
                    root = q[qh + 1]
        self.root = root
        self.parent, self.parent_loc = [-1] * nd, [-1] * nd
        self.children = [[] for _ in range(nd)]
        level = [-1] * nd
        level[root] = 0
        order = [root]
        for a in order:
            for loc in range(6):
                b = self.adj[a][loc]
                if b < 0 or b == self.parent[a] or level[b] >= 0:
                    continue
                self.children[a].append((loc, b))
                self.parent[b], self.parent_loc[b] = a, loc ^ 1
                level[b] = level[a] + 1
                order.append(b)
        self.levels = [[a for a in order if level[a] == l] for l in range(max(level) + 1)]

    def _pe(self, d, f):
        nx, ny, nz, hx, hy, hz, _ = self.boxes[d]
        return sub_solve3d(f, hx, hy, hz, self.bc[d])

    def _schur(self, d, c, loc):
        """SchurMat3D_*::construct (schur_mat3d.cpp): column t = the child's
        response on its face next to the parent to a unit source there."""
        shape = self.boxes[c][:3]
        fc = _face3(loc ^ 1, shape)
        cells = np.arange(np.prod(shape)).reshape(shape)[fc].ravel()
        S = np.zeros((len(cells), len(cells)))
        for t, cell in enumerate(cells):
            e = np.zeros(shape)
            e.ravel()[cell] = 1.0
            S[:, t] = self._dsolve(c, e).ravel()[cells]
        return S

    def _afun(self, d, x):
        """GMRESSolver3D::Afun (gmres_solver3d.cpp:127-141)."""
        ft = np.zeros_like(x)
        for loc, c in self.children[d]:
            fr = _face3(loc, x.shape)
            ft[fr] += (self.S[c] @ x[fr].ravel()).reshape(x[fr].shape)
        return x - self._pe(d, ft)

    def _gmres(self, d, b):
        """GMRESSolver3D::solve (gmres_solver3d.cpp:163-306): x0 = b, MGS,
        Givens, breakdown 1e-12, tolerance on the residual norm."""
        b = self._pe(d, b)
        tol = self.tol * (np.linalg.norm(b) if self.rel else 1.0)
        x = b.copy()
        m = self.m
        it = 0
        for _ in range(self.maxiter):
            r = b - self._afun(d, x)
            beta = np.linalg.norm(r)
            if beta < tol:
                break
            H = np.zeros((m + 1, m))
            cs, sn, g = np.zeros(m), np.zeros(m), np.zeros(m + 1)
            V = [r / beta]
            g[0] = beta
            j = 0
            while j < m:
                w = self._afun(d, V[j])
                it += 1
                for i in range(j + 1):
                    H[i, j] = np.sum(V[i] * w)
                    w = w - H[i, j] * V[i]
                hj = np.linalg.norm(w)
                if hj < 1e-12:
                    break
                H[j + 1, j] = hj
                V.append(w / hj)
                for i in range(j):
                    t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                    H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                    H[i, j] = t
                den = np.hypot(H[j, j], H[j + 1, j])
                cs[j], sn[j] = H[j, j] / den, H[j + 1, j] / den
                H[j, j] = cs[j] * H[j, j] + sn[j] * H[j + 1, j]
                H[j + 1, j] = 0.0
                g[j], g[j + 1] = cs[j] * g[j], -sn[j] * g[j]
                if abs(g[j + 1]) < tol:
                    j += 1
                    break
                j += 1
            y = np.zeros(j)
            for i in range(j - 1, -1, -1):
                y[i] = (g[i] - H[i, i + 1:j] @ y[i + 1:j]) / H[i, i]
            for i in range(j):
                x = x + y[i] * V[i]
        if d == self.root:
            self.last_iters = it
        return x

    def _dsolve(self, d, f):
        return self._gmres(d, f) if self.children[d] else self._pe(d, f)

    def solve(self, fields, ghost=None):
        """ConcatPoissonSolver3D::solve (concat_solver3d.cpp:97-157) with the
        boundary assembly (:159-247).  fields: raw right-hand sides; ghost:
        dict (box, face) -> plane of ghost data.  Returns the solutions."""
        f = [np.array(x, dtype=np.float64) for x in fields]
        for d, (nx, ny, nz, hx, hy, hz, _) in enumerate(self.boxes):
            h = (hx, hy, hz)
            for face in range(6):
                g = (ghost or {}).get((d, face))
                t = self.bc[d][face]
                if g is None or t not in (DIRICHLET, NEUMANN):
                    continue
                hh = h[face // 2]
                a = -1.0 / hh ** 2 if t == DIRICHLET else (-1.0 / hh if face & 1 else 1.0 / hh)
                fr = _face3(face, f[d].shape)
                f[d][fr] += a * np.asarray(g).reshape(f[d][fr].shape)
            f[d] *= hx * hx
        temp = {}
        for lev in reversed(self.levels[1:]):
            for d in lev:
                temp[d] = f[d].copy()
                for loc, c in self.children[d]:
                    fr, fc = _face3(loc, f[d].shape), _face3(loc ^ 1, f[c].shape)
                    temp[d][fr] -= temp[c][fc]
                    f[d][fr] -= temp[c][fc]
            for d in lev:
                temp[d] = self._dsolve(d, temp[d])
        r = self.root
        for loc, c in self.children[r]:
            f[r][_face3(loc, f[r].shape)] -= temp[c][_face3(loc ^ 1, f[c].shape)]
        f[r] = self._dsolve(r, f[r])
        for lev in self.levels[1:]:
            for d in lev:
                p, loc = self.parent[d], self.parent_loc[d]
                f[d][_face3(loc, f[d].shape)] -= f[p][_face3(loc ^ 1, f[p].shape)]
            for d in lev:
                f[d] = self._dsolve(d, f[d])
        return f
