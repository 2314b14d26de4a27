"""A direct solve of the assembled composite 5-point system for star-shaped
geometries (one root domain, every other domain a leaf attached to it: the
lone domain, the cross, the L, the T), converged to FP64 rounding by
refinement with a long-double residual.

TEST INFRASTRUCTURE ONLY: the parity tests use it as the converged reference
at sizes where the reference's own algorithm (full-field restarted GMRES,
`oracle.reference_solution`) takes minutes.  It shares no code with the
device path or with the reference restatement in fp_oracle.c:

* each subdomain operator is diagonalised in both directions by the closed
  forms of its 1-D three-point operators (sine / cosine / quarter-wave bases,
  long double), so A_d^-1 is two dense basis products and a division;
* the interface problem is eliminated exactly: with z_c the child's line next
  to the root, (I - S G) z = R_c A_c^-1 b_c - S_c R A_R^-1 b_R, where
  S_c = R_c A_c^-1 P_c (child Green's function on its own line) and
  G = R A_R^-1 P (root Green's function between its interface lines), both
  dense and formed spectrally, and the K x K system is solved by LU;
* the residual b - A u of the whole composite system (5-point stencils, face
  couplings with coefficient 1, as concat_solver2d.cpp:110-157 and
  field2.cpp:257-302 define them) is summed in long double and the same
  direct solve applied to it, a few times.

The discrete problem is the one tests/discrete.py assembles as a sparse
matrix (the pinning tests compare the two).
"""
import numpy as np

D, N, P, A = 0, 1, 2, 3
XNEG, XPOS, YNEG, YPOS = 0, 1, 2, 3
LD = np.longdouble


def _basis(n, neg, pos):
    """Orthonormal eigenbasis (rows = modes) and eigenvalues of the 1-D
    operator tridiag(1, -2, 1) of n cells with a Dirichlet-like (ghost 0) or
    Neumann (ghost = end cell) end on each side."""
    dn, dp = neg in (D, A), pos in (D, A)
    if neg == P or pos == P:
        raise NotImplementedError("periodic axes are not covered by the direct reference")
    k = np.arange(n, dtype=LD)[:, None]
    i = np.arange(n, dtype=LD)[None, :]
    pi = LD(np.pi) if np.finfo(LD).eps >= np.finfo(np.float64).eps else np.arccos(LD(-1))
    if dn and dp:
        th = pi * (k + 1) / (n + 1)
        V = np.sin(th * (i + 1))
    elif not dn and not dp:
        th = pi * k / n
        V = np.cos(th * (i + LD(0.5)))
    elif dn:
        th = pi * (2 * k + 1) / (2 * n + 1)
        V = np.sin(th * (i + 1))
    else:
        th = pi * (2 * k + 1) / (2 * n + 1)
        V = np.sin(th * (n - i))
    V /= np.sqrt((V * V).sum(axis=1, keepdims=True))
    lam = -4 * np.sin(th[:, 0] / 2) ** 2
    return V.astype(np.float64), lam


class _Domain:
    def __init__(self, nx, ny, hx, hy, bc):
        self.nx, self.ny, self.hx, self.hy, self.bc = nx, ny, hx, hy, list(bc)
        self.r = hx * hx / (hy * hy)
        self.Px, nu = _basis(nx, bc[XNEG], bc[XPOS])
        self.Py, mu = _basis(ny, bc[YNEG], bc[YPOS])
        lam = nu[:, None] + LD(self.r) * mu[None, :]
        if np.any(lam == 0):
            raise NotImplementedError("singular subdomain operator (no Dirichlet-like face)")
        self.inv = (1 / lam).astype(np.float64)

    def solve(self, s):
        """A_d^-1 s for a field s (nx x ny)."""
        return self.Px.T @ ((self.Px @ s @ self.Py.T) * self.inv) @ self.Py

    def green(self, tgt, src):
        """Dense block R_tgt A^-1 P_src between two lines of this domain; a
        line is (axis, index): ('row', a) = cells (a, :), ('col', c) = (:, c)."""
        (tk, ta), (sk, sb) = tgt, src
        Px, Py, inv = self.Px, self.Py, self.inv
        if tk == "row" and sk == "row":
            g = ((Px[:, ta] * Px[:, sb])[:, None] * inv).sum(axis=0)
            return (Py.T * g) @ Py
        if tk == "col" and sk == "col":
            h = (inv * (Py[:, ta] * Py[:, sb])[None, :]).sum(axis=1)
            return (Px.T * h) @ Px
        if tk == "row" and sk == "col":
            W = Px[:, ta][:, None] * inv                  # (k, m)
            B = W.T @ Px                                  # (m, i)
            return Py.T @ (Py[:, sb][:, None] * B)        # (y, i)
        return self.green(src, tgt).T


def _line(dom, face):
    return {XNEG: ("row", 0), XPOS: ("row", dom.nx - 1), YNEG: ("col", 0), YPOS: ("col", dom.ny - 1)}[face]


def _get(u, line):
    k, a = line
    return u[a, :] if k == "row" else u[:, a]


def _add(u, line, v):
    k, a = line
    if k == "row":
        u[a, :] += v
    else:
        u[:, a] += v


def effective_bc(domains, conn):
    bcs = [list(d[4]) for d in domains]
    for a, dr, b in conn:
        bcs[a][dr] = A
        bcs[b][dr ^ 1] = A
    return bcs


def assembled_rhs(domains, conn, fields, buffers=None):
    """hx^2 (F + ghost terms) per domain: Dirichlet data -g/h^2, Neumann data
    +g/h on XNeg/YNeg and -g/h on XPos/YPos (concat_solver2d.cpp:159-223)."""
    bcs = effective_bc(domains, conn)
    out = []
    for d, (nx, ny, hx, hy, _) in enumerate(domains):
        f = np.array(fields[d], dtype=np.float64).reshape(nx, ny)
        for face in range(4):
            g = (buffers or {}).get((d, face))
            if g is None or bcs[d][face] not in (D, N):
                continue
            g = np.asarray(g, dtype=np.float64)[: (ny if face < 2 else nx)]
            h = hx if face < 2 else hy
            if bcs[d][face] == D:
                a = -1.0 / (h * h)
            else:
                a = 1.0 / h if face in (XNEG, YNEG) else -1.0 / h
            sl = {XNEG: (0, slice(None)), XPOS: (nx - 1, slice(None)),
                  YNEG: (slice(None), 0), YPOS: (slice(None), ny - 1)}[face]
            f[sl] += a * g
        out.append(hx * hx * f)
    return out


def residual(domains, conn, u, b):
    """b - A u of the scaled composite system, summed in long double."""
    bcs = effective_bc(domains, conn)
    res = []
    for d, (nx, ny, hx, hy, _) in enumerate(domains):
        r = LD(hx * hx / (hy * hy))
        x = np.asarray(u[d], dtype=LD)
        au = (-2 - 2 * r) * x
        au[1:, :] += x[:-1, :]
        au[:-1, :] += x[1:, :]
        au[:, 1:] += r * x[:, :-1]
        au[:, :-1] += r * x[:, 1:]
        bc = bcs[d]
        if bc[XNEG] == N:
            au[0, :] += x[0, :]
        if bc[XPOS] == N:
            au[-1, :] += x[-1, :]
        if bc[YNEG] == N:
            au[:, 0] += r * x[:, 0]
        if bc[YPOS] == N:
            au[:, -1] += r * x[:, -1]
        res.append(np.asarray(b[d], dtype=LD) - au)
    for a, dr, bb in conn:
        ua, ub = np.asarray(u[a], dtype=LD), np.asarray(u[bb], dtype=LD)
        la = _line_of(domains[a], dr)
        lb = _line_of(domains[bb], dr ^ 1)
        _sub_ld(res[a], la, _get(ub, lb))
        _sub_ld(res[bb], lb, _get(ua, la))
    return [x.astype(np.float64) for x in res]


def _line_of(dom, face):
    nx, ny = dom[0], dom[1]
    return {XNEG: ("row", 0), XPOS: ("row", nx - 1), YNEG: ("col", 0), YPOS: ("col", ny - 1)}[face]


def _sub_ld(r, line, v):
    k, a = line
    if k == "row":
        r[a, :] -= v
    else:
        r[:, a] -= v


class StarComposite:
    """The direct solver of a star-shaped composite geometry."""

    def __init__(self, domains, conn):
        self.domains, self.conn = list(domains), list(conn)
        nd = len(domains)
        bcs = effective_bc(domains, conn)
        deg = [0] * nd
        for a, _, b in conn:
            deg[a] += 1
            deg[b] += 1
        self.root = int(np.argmax(deg)) if nd > 1 else 0
        R = self.root
        # children: (child, root face, child face)
        self.ch = []
        for a, dr, b in conn:
            if a == R:
                self.ch.append((b, dr, dr ^ 1))
            elif b == R:
                self.ch.append((a, dr ^ 1, dr))
            else:
                raise NotImplementedError("only star-shaped geometries (every connection touches the root)")
        if len(self.ch) != nd - 1:
            raise NotImplementedError("every non-root domain needs exactly one connection to the root")
        self.dom = [_Domain(d[0], d[1], d[2], d[3], bcs[i]) for i, d in enumerate(domains)]
        Rd = self.dom[R]
        self.rl = [_line(Rd, fr) for (_, fr, _) in self.ch]            # root's interface lines
        self.cl = [_line(self.dom[c], fc) for (c, _, fc) in self.ch]   # children's lines next to the root
        self.S = [self.dom[c].green(l, l) for (c, _, _), l in zip(self.ch, self.cl)]
        sizes = [len(S) for S in self.S]
        self.off = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
        K = int(self.off[-1])
        M = np.eye(K)
        for p in range(len(self.ch)):
            for q in range(len(self.ch)):
                G = Rd.green(self.rl[p], self.rl[q])
                M[self.off[p]:self.off[p + 1], self.off[q]:self.off[q + 1]] -= self.S[p] @ G
        self.K = K
        if K:
            import scipy.linalg as sla
            self.lu = sla.lu_factor(M)

    def apply_inverse(self, b):
        """u = A^-1 b for the assembled, scaled right-hand sides b."""
        R, Rd = self.root, self.dom[self.root]
        u = [None] * len(self.domains)
        if self.K == 0:
            u[R] = Rd.solve(b[R])
            return u
        import scipy.linalg as sla
        uR0 = Rd.solve(b[R])
        rhs = np.empty(self.K)
        for p, (c, _, _) in enumerate(self.ch):
            rhs[self.off[p]:self.off[p + 1]] = _get(self.dom[c].solve(b[c]), self.cl[p]) - \
                self.S[p] @ _get(uR0, self.rl[p])
        z = sla.lu_solve(self.lu, rhs)
        fR = np.array(b[R], dtype=np.float64)
        for p in range(len(self.ch)):
            _add(fR, self.rl[p], -z[self.off[p]:self.off[p + 1]])
        u[R] = Rd.solve(fR)
        for p, (c, _, _) in enumerate(self.ch):
            fc = np.array(b[c], dtype=np.float64)
            _add(fc, self.cl[p], -_get(u[R], self.rl[p]))
            u[c] = self.dom[c].solve(fc)
        return u

    def solve(self, fields, buffers=None, refine=3):
        """Composite solution of the raw right-hand sides `fields` with ghost
        data `buffers` ((domain, face) -> values), refined `refine` times with
        the long-double residual."""
        b = assembled_rhs(self.domains, self.conn, fields, buffers)
        u = self.apply_inverse(b)
        for _ in range(refine):
            r = residual(self.domains, self.conn, u, b)
            d = self.apply_inverse(r)
            u = [x + y for x, y in zip(u, d)]
        return u


def direct_solution(domains, conn, fields, buffers=None, refine=3):
    """The converged composite solution (see the module docstring)."""
    return StarComposite(domains, conn).solve(fields, buffers, refine=refine)
