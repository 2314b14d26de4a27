/*
 * fp_oracle.h -- CPU FP64 restatement of the Focalors_Poisson solve path (2-D
 * composite, 2-D and 3-D subdomain solvers).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against (tests/, __graft_entry__.smoke(), bench.py cpu_baseline leg).
 * It is never linked into, loaded by, or called from the product library.
 *
 * Every function restates the reference algorithm in new code; the reference
 * file:line each one follows is given at its definition in fp_oracle.c.
 *
 * Boundary codes are the reference's PDEBoundaryType values
 * (base/location_boundary.h:62-69): Dirichlet 0, Neumann 1, Periodic 2,
 * Adjacented 3, Null 4.  Faces are indexed like LocationType
 * (base/location_boundary.h:8-16): XNegative 0, XPositive 1, YNegative 2,
 * YPositive 3.  Fields are row-major [i][j] with j (the y index) contiguous,
 * as field2 (base/field/field2.cpp:304-314).
 */
#ifndef FP_ORACLE_H
#define FP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_DIRICHLET = 0, OR_NEUMANN = 1, OR_PERIODIC = 2, OR_ADJACENTED = 3, OR_NULL = 4 };
enum { OR_XNEG = 0, OR_XPOS = 1, OR_YNEG = 2, OR_YPOS = 3 };
enum { OR_SCHUR_UNIT_SOLVES = 0, OR_SCHUR_SPECTRAL = 1 };

/* Complex DFT of length n, interleaved re/im; sign -1 forward, +1 backward (unscaled). */
int or_dft(int n, int sign, const double* in_ri, double* out_ri);
/* Number of FFT lengths whose plan is cached (one plan per length, kept for the process). */
int or_plan_count(void);
/* Real-to-complex DFT of length n: n/2+1 interleaved bins (FFTW r2c layout). */
int or_r2c(int n, const double* in, double* out_ri);

/* Row-wise y transform of an nx x ny field for the y-boundary pair (bc_neg, bc_pos).
 * inverse = 0: PoissonFFT2D_*::transform; inverse = 1: transform_transpose. */
int or_transform(int bc_neg, int bc_pos, int inverse, int nx, int ny, const double* in, double* out);

/* Per-mode pivot x_diag[j] = -2 + (hx/hy)^2 (lambda_{j+1} - 2). */
int or_x_diag(int ny, double hx, double hy, int bc_yneg, int bc_ypos, double* out);

/* Thomas on tridiag(1, diag, 1) with the reference's Neumann end pivots. */
void or_chasing(double diag, int n, int bc_start, int bc_end, const double* f, double* x);
/* Singular zero-mode recurrence with x[0] = 0. */
void or_chasing_singular(int n, const double* f, double* x);

/* PoissonSolver2D::solve on an already hx^2-scaled RHS; result written to f. */
int or_sub_solve(int nx, int ny, double hx, double hy, const int* bc4, double* f);

/* PoissonSolver3D::solve on an already hx^2-scaled RHS, field3 layout
 * [i][j][k] (k contiguous); bc6: x-, x+, y-, y+, z-, z+.  Result written to f. */
int or_sub_solve3d(int nx, int ny, int nz, double hx, double hy, double hz, const int* bc6, double* f);

/* Composite geometry (ConcatPoissonSolver2D).  Domains 0..nd-1 with sizes and
 * spacings; bc is nd*4 face types (Adjacented faces are set by the connection
 * list as Variable2D::set_geometry does); conn holds nconn triples
 * (a, dir, b) meaning Geometry2D::connect(a, dir, b). */
void* or_cat_create(int nd, const int* nx, const int* ny, const double* hx, const double* hy,
                    const int* bc, int nconn, const int* conn, int m, double tol, int tol_is_relative,
                    int maxiter, int schur_mode);
int   or_cat_root(void* h);
/* level of each domain in the BFS tree (root = 0) and its parent (-1 for root) */
int   or_cat_tree(void* h, int* level, int* parent, int* parent_loc);
/* fields: nd raw RHS arrays (overwritten with the solution); buf: nd*4 ghost-centre
 * data arrays (NULL allowed for faces without data).  iters/res: last GMRES stats. */
int   or_cat_solve(void* h, double* const* fields, double* const* buf, int* iters, double* res);
/* Residual r_d = hx_d^2 F_d - A u of the assembled composite system, in long
 * double (F_d = raw f_d plus the ghost terms of buf). */
int   or_cat_residual(void* h, double* const* u, double* const* fields, double* const* buf, double* const* r);
/* or_cat_solve followed by `steps` steps of iterative refinement
 * u += A^-1 (hx^2 F - A u) with that long-double residual, the correction
 * solves run with GMRES to relative tol_refine.  Converges to the discrete
 * solution to about FP64 rounding, independently of the GMRES tolerance. */
int   or_cat_solve_refined(void* h, double* const* fields, double* const* buf, int steps, double tol_refine,
                           int* iters, double* res);
/* Copy the dense Schur block of child `child` (cn x cn, row-major value(j,i)). */
int   or_cat_schur(void* h, int child, double* out, int* cn);
void  or_cat_destroy(void* h);

/* GMRES of the reference on an explicit operator (tests of its semantics):
 * solves (I - Ainv S) x = Ainv b with Ainv = diag(ainv) (a mock subdomain
 * solve) and S = dense n x n.  b is overwritten with x. */
int or_gmres_mock(int n, const double* ainv, const double* S, double* b, int m, double tol, int maxiter,
                  double* resvec, int* nres);

int or_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
