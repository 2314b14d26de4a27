/*
 * focalors_poisson.h -- C ABI of the MI355X-native Focalors_Poisson solve path.
 *
 * The entry points replace, one for one, the reference's solver interface for
 * the 2-D composite Poisson solve (reference files, relative to the reference
 * root):
 *
 *   fp_sub_create / fp_sub_solve*   PoissonSolver2D(nx, ny, hx, hy, bc x4) and
 *                                   PoissonSolver2D::solve(field2&)
 *                                   (pe/poisson/poisson_solver2d.h:14-24,
 *                                    pe/poisson/poisson_solver2d.cpp:3-90)
 *   fp_cat_create                   ConcatPoissonSolver2D(Variable2D*) plus
 *                                   set_parameter(m, tol, maxIter)
 *                                   (pe/concat/concat_solver2d.h:18-56,
 *                                    pe/concat/concat_solver2d.cpp:7-99)
 *   fp_cat_solve                    ConcatPoissonSolver2D::solve()
 *                                   (pe/concat/concat_solver2d.cpp:101-223)
 *   fp_cat_vjp                      no reference counterpart: adjoint of
 *                                   fp_cat_solve for gradients
 *
 * Conventions
 *   - Boundary codes are PDEBoundaryType values (base/location_boundary.h:62-69):
 *     Dirichlet 0, Neumann 1, Periodic 2, Adjacented 3, Null 4.  Face order is
 *     LocationType (base/location_boundary.h:8-16): XNegative, XPositive,
 *     YNegative, YPositive.
 *   - Fields are device arrays, row-major [i][j] with j (y) contiguous, as
 *     field2 (base/field/field2.cpp:304-314).  nrhs fields of one domain are
 *     stored back to back, `ld` elements apart (ld >= nx*ny).
 *   - The caller owns every device buffer.  Handles own plans, tables, Schur
 *     blocks and scratch.  Work is enqueued on the given HIP stream (NULL =
 *     the null stream); results are in place.
 *   - No exception crosses the ABI; every call returns an fp_status code.
 *     fp_last_error() gives the message of the last failure on this thread.
 *   - One handle per host thread at a time: a handle holds scratch memory and
 *     is not reentrant (as the reference's solver objects).
 */
#ifndef FOCALORS_POISSON_H
#define FOCALORS_POISSON_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fp_ctx_s* fp_ctx;
typedef struct fp_sub_s* fp_sub;
typedef struct fp_cat_s* fp_cat;
typedef struct fp_sub3d_s* fp_sub3d;
typedef struct fp_cat3d_s* fp_cat3d;

typedef enum fp_status {
    FP_OK               = 0,
    FP_E_INVALID        = 1, /* bad argument: size, boundary pair, geometry */
    FP_E_NOMEM          = 2, /* device or host allocation failed */
    FP_E_HIP            = 3, /* HIP runtime error */
    FP_E_NOT_CONVERGED  = 4, /* GMRES stopped at max_iter above tolerance (results still written) */
    FP_E_NUMERIC        = 5, /* NaN/Inf met in the iteration */
    FP_E_UNSUPPORTED    = 6  /* valid in the reference but not provided here */
} fp_status;

typedef enum fp_precision {
    FP_F64        = 0, /* transforms, tridiagonal sweeps and Krylov work in FP64 */
    FP_F32_REFINE = 1, /* FP32 transforms and sweeps, FP64 residual and iterative refinement */
    FP_F64_REFINE = 2  /* the FP64 solve plus iterative refinement with the FP64 residual
                          (composite handles): removes the FP64 rounding floor of the
                          composite solve, which grows with the grid */
} fp_precision;

typedef struct fp_solve_info {
    int    iterations;     /* GMRES iterations (Arnoldi steps) of the root, max over right-hand sides */
    int    restarts;       /* GMRES restart cycles */
    double residual;       /* final interface residual norm, max over right-hand sides */
    double rel_residual;   /* residual / initial preconditioned right-hand-side norm */
    int    refine_steps;   /* sweeps of an iterative refinement (the first solve included) */
    double refine_residual;/* estimated relative error of the refined solution (never below
                              64 * DBL_EPSILON, what an FP64 residual resolves); fp_sub_solve_refine:
                              the last relative correction */
    double refine_correction; /* last relative correction |u_k - u_(k-1)| / |u_k| */
} fp_solve_info;

/* Geometry of a composite problem: domains and Geometry2D::connect calls
 * (base/domain/geometry2d.cpp:47-72).  bc holds the user-set face types per
 * domain; faces named in a connection become Adjacented, as
 * Variable2D::set_geometry does (base/domain/variable2d.cpp:25-36). */
typedef struct fp_geometry_desc {
    int                  nd;
    const int*           nx;
    const int*           ny;
    const double*        hx;
    const double*        hy;
    const unsigned char* bc;    /* nd*4 face types */
    int                  nconn;
    const int*           conn;  /* nconn*3: (a, dir, b) = connect(a, dir, b) */
} fp_geometry_desc;

/* ---- context ----------------------------------------------------------- */
int         fp_ctx_create(int device, fp_ctx* out);
void        fp_ctx_destroy(fp_ctx ctx);
/* Planning rigor of the plans created from ctx afterwards, the analogue of
 * the FFTW_MEASURE flag the reference passes to every fftw_plan_dft_r2c_1d
 * (pe/poisson/poisson_fft2d.cpp:34,102,177,239): FP_PLAN_MEASURE (the
 * default) times the candidate launch shapes of each new plan's transforms
 * and sweeps on the device; FP_PLAN_ESTIMATE takes estimated shapes without
 * timing (plan construction in milliseconds; solves may be slower). */
#define FP_PLAN_MEASURE 0
#define FP_PLAN_ESTIMATE 1
int         fp_ctx_set_planning(fp_ctx ctx, int rigor);
const char* fp_status_str(int status);
const char* fp_last_error(void);
int         fp_version(void);

/* ---- one subdomain: PoissonSolver2D ------------------------------------ */
/* PoissonSolver2D(nx, ny, hx, hy, bc_xneg, bc_xpos, bc_yneg, bc_ypos),
 * pe/poisson/poisson_solver2d.cpp:3-21.  Any ny, as the reference's FFTW
 * plans: a y transform whose FFT buffer exceeds one workgroup's LDS runs
 * from device memory the handle allocates here (256 slots of two
 * FFT-length complex buffers per precision, e.g. 138 MB in FP64 for a
 * Dirichlet/Neumann pair at ny = 4096). */
int fp_sub_create(fp_ctx ctx, int nx, int ny, double hx, double hy, const unsigned char bc[4], fp_sub* out);
void fp_sub_destroy(fp_sub sub);
/* PoissonSolver2D::solve(field2& f), pe/poisson/poisson_solver2d.cpp:62-90:
 * d_f holds the hx^2-scaled right-hand side on entry and the solution on exit. */
int fp_sub_solve(fp_sub sub, double* d_f, int nrhs, long long ld, void* stream);
/* The solve's inner plugins on their own, for callers that drive the steps
 * themselves: PoissonFFT2D::transform (inverse = 0) / transform_transpose
 * (inverse = 1) along y (pe/poisson/poisson_fft2d.h:8-16), and
 * ChasingMethod2D::chasing along x on the transformed field
 * (pe/poisson/chasing_method2d.h:10-35; no transposes: mode j stays column j). */
int fp_sub_transform(fp_sub sub, double* d_f, int nrhs, long long ld, int inverse, void* stream);
int fp_sub_chasing(fp_sub sub, double* d_f, int nrhs, long long ld, void* stream);
/* The same solve on FP32 fields (FP32 transforms and sweeps). */
int fp_sub_solve_f32(fp_sub sub, float* d_f, int nrhs, long long ld, void* stream);
/* Mixed precision (no reference counterpart): FP32 transforms and sweeps as
 * the inner solver, FP64 residual of the discrete operator and FP64
 * iterative refinement.  Same in/out contract as fp_sub_solve.  Stops when
 * the last FP64 correction is below tol relative to the solution (2-norms,
 * every right-hand side) or after max_steps inner solves; info->refine_steps
 * and info->refine_residual (that relative correction) report the outcome.
 * Returns FP_E_NOT_CONVERGED if max_steps ended it above tol. */
int fp_sub_solve_refine(fp_sub sub, double* d_f, int nrhs, long long ld, double tol, int max_steps, void* stream,
                        fp_solve_info* info);
/* Many independent subdomain solves of different sizes in one call: the
 * reference's loop over independent PoissonSolver2D objects.  d_f[k] holds
 * nrhs[k] fields of subs[k] (ld = nx*ny).  Each phase (forward transform,
 * x sweep, inverse transform) is one launch per group of problems sharing a
 * launch shape, whatever their number; more than 8 problems cannot be
 * recorded into a HIP graph (their descriptors are uploaded per call). */
int fp_sub_solve_batched(fp_ctx ctx, int nprob, const fp_sub* subs, double* const* d_f, const int* nrhs,
                         void* stream);
int fp_sub_solve_batched_f32(fp_ctx ctx, int nprob, const fp_sub* subs, float* const* d_f, const int* nrhs,
                             void* stream);
/* A fixed batch of independent subdomain solves, prepared once (no reference
 * counterpart; the reference loops over PoissonSolver2D objects,
 * pe/poisson/poisson_solver2d.cpp:62-90): the problems, their fields d_f[k]
 * (nrhs[k] fields of subs[k] back to back, FP64, or FP32 when fp32 != 0) and
 * the launch descriptors are bound at creation and the descriptors copied to
 * device memory the batch owns.  fp_batch_solve then queues only kernel
 * launches on the stream (no allocation, no host copy), so it may be
 * captured into a HIP graph and replayed.  Every problem takes 128 threads
 * per transform workgroup where that fits, so that problems of different
 * lengths share launches; within a launch the problems with the longest
 * workgroups come first.  subs and the fields must outlive the batch; calls
 * of one batch on different streams must not overlap. */
typedef struct fp_batch_s* fp_batch;
int  fp_batch_create(fp_ctx ctx, int nprob, const fp_sub* subs, void* const* d_f, const int* nrhs, int fp32,
                     fp_batch* out);
int  fp_batch_solve(fp_batch b, void* stream);
void fp_batch_destroy(fp_batch b);
/* fp_sub_solve_batched without any allocation (no reference counterpart: for
 * callers that cannot allocate device memory, e.g. tight serving loops or
 * harnesses that forbid hipMalloc).  Same arguments, kernels and results; the
 * descriptor tables of batches of more than 8 problems go to a 4 MiB device
 * array of the library (one per GPU, held until the process ends) instead of
 * stream-ordered allocations and pinned staging, copied from pageable host
 * memory, so the host waits for the work queued on the stream before it.
 * At most FP_NOALLOC_MAX_PROBLEMS problems per call (FP_E_INVALID beyond).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (different streams or threads at the same time) race; calls on one stream
 * are ordered. */
#define FP_NOALLOC_MAX_PROBLEMS 4096
int fp_sub_solve_batched_noalloc(fp_ctx ctx, int nprob, const fp_sub* subs, double* const* d_f, const int* nrhs,
                                 void* stream);

/* ---- one 3-D subdomain: PoissonSolver3D -------------------------------- */
/* PoissonSolver3D(nx, ny, nz, hx, hy, hz, bc x6), pe/poisson/poisson_solver3d.cpp:3-85.
 * bc: XNegative, XPositive, YNegative, YPositive, ZNegative, ZPositive
 * (base/location_boundary.h:8-16).  A periodic y or z axis needs an even
 * length. */
int  fp_sub3d_create(fp_ctx ctx, int nx, int ny, int nz, double hx, double hy, double hz, const unsigned char bc[6],
                     fp_sub3d* out);
void fp_sub3d_destroy(fp_sub3d sub);
/* PoissonSolver3D::solve(field3& f), pe/poisson/poisson_solver3d.cpp:87-123:
 * d_f holds nrhs fields ld apart, each field3's layout [i][j][k] with k (z)
 * contiguous (base/field/field3.cpp), the hx^2-scaled right-hand side on
 * entry and the solution on exit. */
int  fp_sub3d_solve(fp_sub3d sub, double* d_f, int nrhs, long long ld, void* stream);

/* ---- 3-D composite geometry: ConcatPoissonSolver3D --------------------- */
/* Boxes and Geometry3D::connect calls (base/domain/geometry3d.cpp); bc holds
 * nd*6 face types, faces named in a connection become Adjacented. */
typedef struct fp_geometry3d_desc {
    int                  nd;
    const int*           nx;
    const int*           ny;
    const int*           nz;
    const double*        hx;
    const double*        hy;
    const double*        hz;
    const unsigned char* bc;    /* nd*6 face types */
    int                  nconn;
    const int*           conn;  /* nconn*3: (a, dir, b) = connect(a, dir, b), dir 0..5 */
} fp_geometry3d_desc;
/* ConcatPoissonSolver3D(Variable3D*) + set_parameter(m, tol, maxIter)
 * (pe/concat/concat_solver3d.cpp:5-95): the tree, the 3-D subdomain plans
 * and the dense Schur blocks of every non-leaf (schur_mat3d.cpp), built on
 * the device.  Geometries without a Dirichlet face return FP_E_UNSUPPORTED. */
int  fp_cat3d_create(fp_ctx ctx, const fp_geometry3d_desc* g, int m, double tol, int tol_is_relative, int max_iter,
                     fp_cat3d* out);
void fp_cat3d_destroy(fp_cat3d cat);
int  fp_cat3d_root(fp_cat3d cat);
/* ConcatPoissonSolver3D::solve() (concat_solver3d.cpp:97-247): d_fields[d]
 * holds nrhs raw right-hand sides of box d (field3 layout, overwritten with
 * the solution); d_ghost[6*d+face] holds nrhs ghost-data planes of that face
 * back to back (face order of its cells: x faces (j, k), y faces (i, k),
 * z faces (i, j), last index fastest), or NULL. */
int  fp_cat3d_solve(fp_cat3d cat, double* const* d_fields, const double* const* d_ghost, int nrhs, void* stream,
                    fp_solve_info* info);

/* ---- composite geometry: ConcatPoissonSolver2D ------------------------- */
/* ConcatPoissonSolver2D(Variable2D*) + set_parameter(m, tol, maxIter):
 * builds the tree (geometry_tree.hpp:22-200), the subdomain plans and the
 * dense Schur blocks of every non-leaf (schur_mat2d.cpp:5-123) on the device.
 * tol_is_relative = 0 keeps the reference's absolute tolerance
 * (gmres_solver2d.cpp:155,221); 1 scales it by the norm of the preconditioned
 * interface right-hand side.  precision: fp_precision. */
int fp_cat_create(fp_ctx ctx, const fp_geometry_desc* g, int m, double tol, int tol_is_relative, int max_iter,
                  int precision, fp_cat* out);
/* FP_F32_REFINE and FP_F64_REFINE handles (no reference counterpart): the
 * refinement runs sweeps (FP32 subdomain solves, or the plain FP64 solve) on
 * the FP64 residual of the assembled composite operator until the estimated
 * relative error of the solution (size of the last correction times the
 * contraction of the iteration, floored at 64 * DBL_EPSILON) is at most tol
 * for every right-hand side, or max_steps sweeps were done (defaults: tol
 * 1e-12; max_steps 10 for FP_F32_REFINE, 2 = one correction for
 * FP_F64_REFINE).  fp_solve_info reports refine_steps (sweeps),
 * refine_residual (that estimate) and refine_correction;
 * FP_E_NOT_CONVERGED if max_steps ended it above tol.  A captured
 * FP_F64_REFINE solve takes exactly max_steps - 1 corrections. */
int fp_cat_set_refinement(fp_cat cat, double tol, int max_steps);
void fp_cat_destroy(fp_cat cat);
int  fp_cat_root(fp_cat cat);
/* PoissonFFT2D::transform (inverse = 0) / transform_transpose (inverse = 1)
 * of every subdomain of the composite, in place, in the one batched launch
 * the solve uses for its first and last stages (pe/poisson/poisson_fft2d.h:8-16
 * applied to each domain of concat_solver2d.cpp's solver_map); d_fields[d]
 * holds nrhs fields of domain d. */
int  fp_cat_transform(fp_cat cat, double* const* d_fields, int nrhs, int inverse, void* stream);
/* ConcatPoissonSolver2D::solve(): d_fields[d] holds nrhs raw right-hand sides
 * of domain d (overwritten with the solution); d_ghost[4*d+face] holds nrhs
 * ghost-centre boundary data arrays of that face back to back (the
 * Variable2D::buffer_map values, variable2d.cpp:345-416), or NULL for none. */
int fp_cat_solve(fp_cat cat, double* const* d_fields, const double* const* d_ghost, int nrhs, void* stream,
                 fp_solve_info* info);
/* Out-of-place form of fp_cat_solve (no reference counterpart: the
 * reference solves in place): reads the raw right-hand sides d_rhs[d], which
 * stay unchanged, and writes the solution to d_out[d] (d_rhs[d] == d_out[d]
 * is fp_cat_solve).  The FP64 two-level path reads them directly in its
 * forward transforms; other paths start from a copy. */
int fp_cat_solve_into(fp_cat cat, const double* const* d_rhs, double* const* d_out, const double* const* d_ghost,
                      int nrhs, void* stream, fp_solve_info* info);
/* Captured composite solve (no reference counterpart; the reference has no
 * device): records the whole FP64 solve of fp_cat_solve on these fields and
 * ghost arrays into a HIP graph (FP_F64 and FP_F64_REFINE handles).  Covered:
 * a root with leaf children (depth two, e.g. the cross) or one domain, with
 * the explicit-inverse interface preconditioner.  The captured solve starts
 * GMRES from that inverse's product and takes no further Krylov steps; its
 * residual check is recorded too.  stream must not be the null stream.
 * Nothing is computed at capture.  The graph is bound to the handle's
 * workspace: a later solve with more right-hand sides than captured
 * reallocates it and destroys the graph (fp_cat_graph_launch then returns
 * FP_E_INVALID, "no captured solve"). */
int fp_cat_graph_capture(fp_cat cat, double* const* d_fields, const double* const* d_ghost, int nrhs, void* stream);
/* fp_cat_graph_capture of fp_cat_solve_into: replays read d_rhs and write
 * d_out (no reference counterpart). */
int fp_cat_graph_capture_into(fp_cat cat, const double* const* d_rhs, double* const* d_out,
                              const double* const* d_ghost, int nrhs, void* stream);
/* Replays the captured solve on the captured buffers (the caller refills
 * them in between).  info != NULL: waits for it and reports the residual
 * check; FP_E_NOT_CONVERGED if the preconditioned start missed the tolerance
 * (fp_cat_solve then iterates further).  info == NULL: returns at once. */
int fp_cat_graph_launch(fp_cat cat, void* stream, fp_solve_info* info);

/* Adjoint of fp_cat_solve for a loss L(u) (no reference counterpart): given
 * dL/du in d_gu[d] (nrhs fields per domain), writes dL/df into d_gf[d] and
 * dL/dg for the ghost data of every face into d_gghost[4*d+face] (zero for
 * Adjacented faces; NULL entries are skipped).  d_gghost may be NULL.
 * Costs one composite solve (the assembled operator is symmetric), which
 * reads d_gu directly, plus one scaling pass over d_gf; allocates nothing
 * beyond the handle's solve workspace. */
int fp_cat_vjp(fp_cat cat, const double* const* d_gu, double* const* d_gf, double* const* d_gghost, int nrhs,
               void* stream, fp_solve_info* info);

/* ---- the composite's building blocks over caller-supplied solvers ------- */
/* DomainSolver2D::solve(field2& f) (pe/concat/domain_solver.h:9-15) as a
 * callback: solve the nrhs fields (nx x ny each, back to back) at d_f in
 * place, queuing the work on stream; return 0, or nonzero to fail the call
 * that invoked it (FP_E_INVALID).  fp_sub_domain_solve is one: user = an
 * fp_sub, its fp_sub_solve on the fields. */
typedef int (*fp_domain_solve_fn)(void* user, double* d_f, int nx, int ny, int nrhs, void* stream);
int fp_sub_domain_solve(void* sub, double* d_f, int nx, int ny, int nrhs, void* stream);
/* SchurMat2D_xneg / _xpos / _yneg / _ypos (pe/concat/schur_mat2d.h:7-69):
 * loc = the face of the root the branch (bnx x bny) sits on; construct
 * (schur_mat2d.cpp) runs one unit solve of the branch per interface point
 * through the callback, at most 64 per call and as many as the caller's
 * workspace d_work holds (work_len doubles, at least bnx*bny; nothing is
 * allocated there beyond the block itself).  values copies value(j, i) to
 * h_value[j*cn+i]; apply is operator* (root line -> the same line of d_out,
 * which is zeroed first unless accumulate). */
typedef struct fp_schur2d_s* fp_schur2d;
int  fp_schur2d_create(fp_ctx ctx, int loc, int bnx, int bny, fp_domain_solve_fn branch, void* user, double* d_work,
                       long long work_len, void* stream, fp_schur2d* out);
int  fp_schur2d_size(fp_schur2d S);
int  fp_schur2d_values(fp_schur2d S, double* h_value);
int  fp_schur2d_apply(fp_schur2d S, const double* d_root, int rnx, int rny, double* d_out, int accumulate,
                      void* stream);
void fp_schur2d_destroy(fp_schur2d S);
/* GMRESSolver2D (pe/concat/gmres_solver2d.h:12-57): solves
 * (I - A^-1 S) x = A^-1 b on an nx x ny root field, A^-1 = the solver set by
 * fp_gmres2d_set_solver (set_solver), S = the sum of the blocks added
 * (schur_mat_construct), with restart length m, absolute tolerance tol and
 * at most max_iter restarts, step by step as gmres_solver2d.cpp:123-252
 * (x0 = A^-1 b unless set_initial_guess).  Its buffers are the caller's
 * workspace d_work of fp_gmres2d_workspace(nx, ny, m) doubles.  solve
 * overwrites d_b with x; info reports iterations (Arnoldi steps), restarts
 * and residual (the last restart residual); h_resvec (capacity cap) gets the
 * restart residuals (the reference's resVec), *nres their count.  Returns
 * FP_E_NOT_CONVERGED (x written all the same) when max_iter restarts ended
 * above tol, where the reference returns silently.  The blocks must outlive
 * the solver. */
typedef struct fp_gmres2d_s* fp_gmres2d;
long long fp_gmres2d_workspace(int nx, int ny, int m);
int  fp_gmres2d_create(fp_ctx ctx, int nx, int ny, int m, double tol, int max_iter, double* d_work,
                       long long work_len, fp_gmres2d* out);
int  fp_gmres2d_set_solver(fp_gmres2d g, fp_domain_solve_fn pe, void* user);
int  fp_gmres2d_add_schur(fp_gmres2d g, fp_schur2d S);
int  fp_gmres2d_set_initial_guess(fp_gmres2d g, const double* d_x0);
int  fp_gmres2d_solve(fp_gmres2d g, double* d_b, void* stream, fp_solve_info* info, double* h_resvec, int cap,
                      int* nres);
void fp_gmres2d_destroy(fp_gmres2d g);

/* ---- the caller side: pressure projection on a staggered grid ---------- */
/* No reference counterpart in this tree: the Focalors Navier–Stokes solver
 * that the Poisson module serves (README_en.md:26) makes each provisional
 * velocity divergence-free by D G p = D u* / dt, u = u* - dt G p, with p at
 * cell centres and the velocity on the faces.  On that grid D G is the
 * composite 5-point operator of `cat` with its face conventions (Neumann:
 * a wall or inflow face, whose normal velocity stays as given; Dirichlet: an
 * outlet with ghost-centre pressure values; Adjacented: the neighbour's
 * cells), so the corrected velocity's divergence vanishes to the solve's
 * accuracy.  Per subdomain d (nx x ny cells): d_u[d] holds the x-faces
 * [i][j], i = 0..nx ((nx+1) x ny), d_v[d] the y-faces [i][j], j = 0..ny
 * (nx x (ny+1)); a face shared by two subdomains is held by both and stays
 * equal in both.  d_p[d] receives the pressure (nx x ny); d_pghost (NULL or
 * 4*nd pointers, NULL entries allowed) holds the pressure's ghost-centre
 * values on Dirichlet faces (ny values on x faces, nx on y faces).  Needs
 * hx == hy in every subdomain; one right-hand side.
 * fp_proj_step: divergence, the composite solve (fp_cat_solve_into) and the
 * correction, eager.  fp_proj_capture records the same three into one HIP
 * graph (for handles fp_cat_graph_capture covers, on a non-null stream);
 * fp_proj_launch replays it (the caller refreshes d_u, d_v and d_pghost in
 * place between replays).  The graph is kept in the fp_cat handle (one per
 * handle: the last projection that captured), so only the projection that
 * recorded it replays it, a solve with more right-hand sides than captured
 * destroys it (as it does the captured solve).  A projection is not used
 * after its fp_cat is destroyed (fp_proj_destroy itself never touches it).
 * fp_proj_divergence (out = scale D(u, v)) and fp_proj_correct
 * (u -= scale Gx p, v -= scale Gy p) are the two steps on their own. */
typedef struct fp_proj_s* fp_proj;
int  fp_proj_create(fp_cat cat, fp_proj* out);
void fp_proj_destroy(fp_proj proj);
int  fp_proj_divergence(fp_proj proj, const double* const* d_u, const double* const* d_v, double scale,
                        double* const* d_out, void* stream);
int  fp_proj_correct(fp_proj proj, double* const* d_u, double* const* d_v, const double* const* d_p,
                     const double* const* d_pghost, double scale, void* stream);
int  fp_proj_step(fp_proj proj, double* const* d_u, double* const* d_v, double* const* d_p,
                  const double* const* d_pghost, double dt, void* stream, fp_solve_info* info);
int  fp_proj_capture(fp_proj proj, double* const* d_u, double* const* d_v, double* const* d_p,
                     const double* const* d_pghost, double dt, void* stream);
int  fp_proj_launch(fp_proj proj, void* stream, fp_solve_info* info);

#ifdef __cplusplus
}
#endif
#endif
