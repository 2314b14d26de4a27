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

}  // namespace fp
