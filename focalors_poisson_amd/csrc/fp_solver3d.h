// The 3-D subdomain plan (fp_solver3d.hip), shared with the 3-D composite
// solver (fp_concat3d.hip).
#pragma once

#include "fp_internal.h"

namespace fp {

// Mode pairs (k, j) of the x sweep are numbered m = k ny + j; the pivot gap of
// a pair is the sum of the y and z gaps, each to full precision:
//   x_diag[k][j] = -2 + (hx/hy)^2 lambda_y[j] + (hx/hz)^2 lambda_z[k]
// (poisson_solver3d.cpp:60-70, lambda = cal_lambda's values, <= 0).
struct SubPlan3D {
    int                   nx = 0, ny = 0, nz = 0;
    double                hx = 0, hy = 0, hz = 0;
    int                   bc[6] = {0, 0, 0, 0, 0, 0};
    TransformPlan<double> tz, ty;
    SubPlan               sweep;  // nx rows, ny nz modes
    DevBuf<double>        work;   // the y-contiguous copy [i][k][j] of every right-hand side
    int                   work_nrhs = 0;
    bool                  measure = true;  // time the launch shapes (else the estimated ones)
    void                  build(int nx, int ny, int nz, double hx, double hy, double hz, const int bc[6]);
};

// PoissonSolver3D::solve on nrhs hx^2-scaled right-hand sides ld apart, in
// place (field3 layout [i][j][k], k contiguous).
void sub3d_solve(SubPlan3D& p, double* f, long long ld, int nrhs, hipStream_t s);

}  // namespace fp
