// Mixed-precision refinement of the subdomain solve (fp_refine.hip).
#pragma once

#include "fp_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace fp {

struct RefineArgs {
    int           nx, ny;
    double        ry;                           // hx^2 / hy^2
    int           nxneg, nxpos, nyneg, nypos;   // Neumann faces
    int           perx, pery;                   // periodic axes (first and last lines neighbours)
    const double* u_old;                        // null: zero
    double*       u_new;
    const float*  d;                            // FP32 correction
    const double* b;                            // FP64 right-hand side
    float*        r32;                          // FP32 residual out
    long long     ldu, ldd, ldb;
    double*       part;                         // per rhs, per block: |d|^2, |u|^2
};

struct RefineStats {
    int                 steps      = 0;
    double              correction = 0.0;  // last ||d|| / ||u|| (max over right-hand sides)
    bool                numeric    = false;
    std::vector<double> history;
};

void        launch_to_f32(const double* in, long long ldi, float* out, long long ldo, long long n, int nrhs,
                          hipStream_t s);
void        launch_refine_step(const RefineArgs& a, int nrhs, hipStream_t s);
RefineStats refine_solve(SubPlan& sp, double* f, long long ld, int nrhs, double tol, int max_steps, hipStream_t s);

// FP64 residual of the assembled composite operator on one subdomain:
//   r = b - A u - (sum of the neighbours' adjacent lines)
// A the hx^2-scaled 5-point operator of the subdomain (ghost 0 on Dirichlet
// and Adjacented faces, ghost = cell on Neumann faces); the coupling
// coefficient 1 is the one ConcatPoissonSolver2D::solve applies
// (pe/concat/concat_solver2d.cpp:127-128,140,150).
struct CatResidualArgs {
    const double* u;
    const double* b;
    double*       r;
    long long     ldu, ldb, ldr;
    int           nx, ny;
    double        ry;           // hx^2 / hy^2
    int           neu[4];       // Neumann faces
    int           perx, pery;   // periodic axes
    const double* nb[4];        // neighbour field on Adjacented faces, else null
    int           nbx[4], nby[4];
    long long     ldn[4];
};
void launch_cat_residual(const CatResidualArgs& a, int nrhs, hipStream_t s);
// the residuals of up to 8 domains in one launch
struct CatResidualBatch {
    int             n = 0;
    CatResidualArgs a[8];
};
void launch_cat_residual_batch(const CatResidualBatch& b, int nrhs, hipStream_t s);
// u += d (d may be null: u unchanged) over nx*ny values per right-hand side;
// writes per-block partial sums (|d|^2, |u|^2) to part[(r * nblk + blk) * 2]
constexpr int kNormBlocks = 256;
void launch_update_norms(double* u, long long ldu, const double* d, long long ldd, long long n, int nrhs,
                         double* part, hipStream_t s);
// launch_update_norms of up to 8 domains in one launch
struct NormsJob {
    double*       u;
    long long     ldu;
    const double* d;
    long long     ldd, n;
    double*       part;
};
struct NormsBatch {
    int      n = 0;
    NormsJob j[8];
};
void launch_update_norms_batch(const NormsBatch& b, int nrhs, hipStream_t s);

}  // namespace fp
