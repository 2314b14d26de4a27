// Mixed precision for the subdomain solve: FP32 transforms and tridiagonal
// sweeps as the inner solver, FP64 residual of the exact discrete operator and
// FP64 iterative refinement.  No reference counterpart (the reference solves
// in FP64 only); the discrete operator is the one PoissonSolver2D inverts
// (pe/poisson/poisson_solver2d.cpp:43-90): per cell
//   u(i-1) - 2u + u(i+1) + r (u(j-1) - 2u + u(j+1)),  r = hx^2 / hy^2,
// ghosts 0 on Dirichlet-like faces and equal to the cell on Neumann faces.
#include "fp_refine.h"

namespace fp {

constexpr int RB = 256;

__device__ __forceinline__ double wsum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// out32 = (float) in over n values of each right-hand side
__global__ void k_to_f32(const double* in, long long ldi, float* out, long long ldo, long long n)
{
    const long long r = blockIdx.y;
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
        out[r * ldo + q] = (float)in[r * ldi + q];
}

// One refinement step, cells in row-major order:
//   u_new = u_old + d                    (u_old == null: zero start)
//   r32   = (float)(b - A u_new)         (neighbours' u_new formed on the fly)
// plus per-block partial sums of |d|^2 and |u_new|^2 for the stopping test.
__global__ __launch_bounds__(RB) void k_refine_step(RefineArgs a)
{
    __shared__ double sd[RB / 64], su[RB / 64];
    const int       nx = a.nx, ny = a.ny;
    const long long n  = (long long)nx * ny;
    const long long r  = blockIdx.y;
    const double*   uo = a.u_old ? a.u_old + r * a.ldu : nullptr;
    const float*    d  = a.d + r * a.ldd;
    const double*   b  = a.b + r * a.ldb;
    double*         un = a.u_new + r * a.ldu;
    float*          rr = a.r32 + r * a.ldd;
    double          pd = 0.0, pu = 0.0;
    for (long long q = blockIdx.x * (long long)RB + threadIdx.x; q < n; q += (long long)gridDim.x * RB) {
        const int i = (int)(q / ny), j = (int)(q - (long long)i * ny);
        auto      U = [&](long long p) { return (uo ? uo[p] : 0.0) + (double)d[p]; };
        const double c  = U(q);
        double       ax = -2.0 * c, ay = -2.0 * c;
        if (i > 0) ax += U(q - ny);
        else if (a.nxneg) ax += c;
        else if (a.perx) ax += U(q + (long long)(nx - 1) * ny);
        if (i < nx - 1) ax += U(q + ny);
        else if (a.nxpos) ax += c;
        else if (a.perx) ax += U(q - (long long)(nx - 1) * ny);
        if (j > 0) ay += U(q - 1);
        else if (a.nyneg) ay += c;
        else if (a.pery) ay += U(q + ny - 1);
        if (j < ny - 1) ay += U(q + 1);
        else if (a.nypos) ay += c;
        else if (a.pery) ay += U(q - (ny - 1));
        const double Au = ax + a.ry * ay;
        un[q]           = c;
        rr[q]           = (float)(b[q] - Au);
        const double dd = (double)d[q];
        pd += dd * dd;
        pu += c * c;
    }
    pd = wsum(pd);
    pu = wsum(pu);
    if ((threadIdx.x & 63) == 0) {
        sd[threadIdx.x >> 6] = pd;
        su[threadIdx.x >> 6] = pu;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double td = 0.0, tu = 0.0;
        for (int w = 0; w < RB / 64; w++) {
            td += sd[w];
            tu += su[w];
        }
        a.part[(r * gridDim.x + blockIdx.x) * 2 + 0] = td;
        a.part[(r * gridDim.x + blockIdx.x) * 2 + 1] = tu;
    }
}

int refine_blocks(long long n) { return (int)std::min<long long>((n + RB - 1) / RB, 1024); }

void launch_to_f32(const double* in, long long ldi, float* out, long long ldo, long long n, int nrhs, hipStream_t s)
{
    const int nb = (int)std::min<long long>((n + RB - 1) / RB, 2048);
    hipLaunchKernelGGL(k_to_f32, dim3(nb, nrhs), dim3(RB), 0, s, in, ldi, out, ldo, n);
    FP_HIP(hipGetLastError());
}

void launch_refine_step(const RefineArgs& a, int nrhs, hipStream_t s)
{
    const long long n = (long long)a.nx * a.ny;
    hipLaunchKernelGGL(k_refine_step, dim3(refine_blocks(n), nrhs), dim3(RB), 0, s, a);
    FP_HIP(hipGetLastError());
}

// FP32 + refinement solve of nrhs fields of one subdomain, in place on f.
// Stops when ||d|| <= tol ||u|| for every right-hand side (the correction of
// the last step is included), or after max_steps inner solves.
RefineStats refine_solve(SubPlan& sp, double* f, long long ld, int nrhs, double tol, int max_steps, hipStream_t s)
{
    FP_REQUIRE(!sp.no_dirichlet, "refinement needs a Dirichlet-like face (the all-Neumann operator is singular)");
    FP_REQUIRE(max_steps >= 1, "max_steps must be positive");
    sp.ensure_f32();
    const long long n  = (long long)sp.nx * sp.ny;
    const int       nb = refine_blocks(n);
    DevBuf<double>  b((size_t)n * nrhs), u2((size_t)n * nrhs), part((size_t)2 * nb * nrhs);
    DevBuf<float>   d((size_t)n * nrhs), r32((size_t)n * nrhs);
    std::vector<double> hp((size_t)2 * nb * nrhs);
    FP_HIP(hipMemcpy2DAsync(b.p, n * sizeof(double), f, ld * sizeof(double), n * sizeof(double), nrhs,
                            hipMemcpyDeviceToDevice, s));
    // step 0: d = A32^-1 (float) b, u = d
    launch_to_f32(b.p, n, d.p, n, n, nrhs, s);
    sub_solve<float>(sp, d.p, n, nrhs, s);
    RefineArgs a{};
    a.nx = sp.nx;
    a.ny = sp.ny;
    a.ry = sp.hx * sp.hx / (sp.hy * sp.hy);
    a.nxneg = sp.bc[XNEG] == NEUMANN;
    a.nxpos = sp.bc[XPOS] == NEUMANN;
    a.nyneg = sp.bc[YNEG] == NEUMANN;
    a.nypos = sp.bc[YPOS] == NEUMANN;
    a.perx  = sp.bc[XNEG] == PERIODIC;
    a.pery  = sp.bc[YNEG] == PERIODIC;
    a.b = b.p;
    a.ldb = n;
    a.d = d.p;
    a.ldd = n;
    a.r32 = r32.p;
    a.ldu = n;
    a.part = part.p;
    DevBuf<double> u1((size_t)n * nrhs);
    double*        ucur = nullptr;  // zero start
    RefineStats    st{};
    for (int step = 1;; step++) {
        double* unew = (ucur == u1.p) ? u2.p : u1.p;  // never in place: neighbours read u_old
        a.u_old      = ucur;
        a.u_new      = unew;
        launch_refine_step(a, nrhs, s);
        ucur = unew;
        FP_HIP(hipMemcpyAsync(hp.data(), part.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        FP_HIP(hipStreamSynchronize(s));
        double worst = 0.0;
        for (int r = 0; r < nrhs; r++) {
            double td = 0.0, tu = 0.0;
            for (int k = 0; k < nb; k++) {
                td += hp[((size_t)r * nb + k) * 2];
                tu += hp[((size_t)r * nb + k) * 2 + 1];
            }
            const double ratio = tu > 0 ? std::sqrt(td / tu) : std::sqrt(td);
            worst              = std::max(worst, ratio);
        }
        st.steps      = step;
        st.correction = worst;
        st.history.push_back(worst);
        if (worst <= tol || step >= max_steps || !std::isfinite(worst)) break;
        // next correction: solve in place on the residual just written, which
        // then serves as the correction while the old one takes the residual
        float* res = a.r32;
        sub_solve<float>(sp, res, n, nrhs, s);
        a.r32 = const_cast<float*>(a.d);
        a.d   = res;
    }
    FP_HIP(hipMemcpy2DAsync(f, ld * sizeof(double), ucur, n * sizeof(double), n * sizeof(double), nrhs,
                            hipMemcpyDeviceToDevice, s));
    FP_HIP(hipStreamSynchronize(s));
    st.numeric = !std::isfinite(st.correction);
    return st;
}

// ---- composite refinement ------------------------------------------------------

__global__ __launch_bounds__(RB) void k_cat_residual(CatResidualBatch bt)
{
    const CatResidualArgs& a = bt.a[blockIdx.z];
    const int       nx = a.nx, ny = a.ny;
    const long long n  = (long long)nx * ny;
    const long long r  = blockIdx.y;
    const double*   u  = a.u + r * a.ldu;
    const double*   b  = a.b + r * a.ldb;
    double*         rr = a.r + r * a.ldr;
    for (long long q = blockIdx.x * (long long)RB + threadIdx.x; q < n; q += (long long)gridDim.x * RB) {
        const int    i = (int)(q / ny), j = (int)(q - (long long)i * ny);
        const double c = u[q];
        double       ax = -2.0 * c, ay = -2.0 * c, cpl = 0.0;
        if (i > 0) ax += u[q - ny];
        else if (a.neu[XNEG]) ax += c;
        else if (a.perx) ax += u[q + (long long)(nx - 1) * ny];
        else if (a.nb[XNEG]) cpl += a.nb[XNEG][r * a.ldn[XNEG] + (long long)(a.nbx[XNEG] - 1) * a.nby[XNEG] + j];
        if (i < nx - 1) ax += u[q + ny];
        else if (a.neu[XPOS]) ax += c;
        else if (a.perx) ax += u[q - (long long)(nx - 1) * ny];
        else if (a.nb[XPOS]) cpl += a.nb[XPOS][r * a.ldn[XPOS] + j];
        if (j > 0) ay += u[q - 1];
        else if (a.neu[YNEG]) ay += c;
        else if (a.pery) ay += u[q + ny - 1];
        else if (a.nb[YNEG]) cpl += a.nb[YNEG][r * a.ldn[YNEG] + (long long)i * a.nby[YNEG] + a.nby[YNEG] - 1];
        if (j < ny - 1) ay += u[q + 1];
        else if (a.neu[YPOS]) ay += c;
        else if (a.pery) ay += u[q - (ny - 1)];
        else if (a.nb[YPOS]) cpl += a.nb[YPOS][r * a.ldn[YPOS] + (long long)i * a.nby[YPOS]];
        rr[q] = b[q] - cpl - (ax + a.ry * ay);
    }
}

void launch_cat_residual_batch(const CatResidualBatch& b, int nrhs, hipStream_t s)
{
    if (b.n <= 0) return;
    long long nmax = 0;
    for (int k = 0; k < b.n; k++) nmax = std::max(nmax, (long long)b.a[k].nx * b.a[k].ny);
    hipLaunchKernelGGL(k_cat_residual, dim3(refine_blocks(nmax), nrhs, b.n), dim3(RB), 0, s, b);
    FP_HIP(hipGetLastError());
}
void launch_cat_residual(const CatResidualArgs& a, int nrhs, hipStream_t s)
{
    CatResidualBatch b{};
    b.n    = 1;
    b.a[0] = a;
    launch_cat_residual_batch(b, nrhs, s);
}

__global__ __launch_bounds__(RB) void k_update_norms(NormsBatch bt)
{
    __shared__ double sd[RB / 64], su[RB / 64];
    const NormsJob&  J    = bt.j[blockIdx.z];
    double*          u    = J.u;
    const double*    d    = J.d;
    const long long  ldu  = J.ldu, ldd = J.ldd, n = J.n;
    double*          part = J.part;
    const long long r  = blockIdx.y;
    double*         ur = u + r * ldu;
    const double*   dr = d ? d + r * ldd : nullptr;
    double          pd = 0.0, pu = 0.0;
    for (long long q = blockIdx.x * (long long)RB + threadIdx.x; q < n; q += (long long)gridDim.x * RB) {
        double v = ur[q];
        if (dr) {
            const double dd = dr[q];
            v += dd;
            ur[q] = v;
            pd += dd * dd;
        }
        pu += v * v;
    }
    pd = wsum(pd);
    pu = wsum(pu);
    if ((threadIdx.x & 63) == 0) {
        sd[threadIdx.x >> 6] = pd;
        su[threadIdx.x >> 6] = pu;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double td = 0.0, tu = 0.0;
        for (int w = 0; w < RB / 64; w++) {
            td += sd[w];
            tu += su[w];
        }
        part[(r * gridDim.x + blockIdx.x) * 2 + 0] = td;
        part[(r * gridDim.x + blockIdx.x) * 2 + 1] = tu;
    }
}

void launch_update_norms_batch(const NormsBatch& b, int nrhs, hipStream_t s)
{
    if (b.n <= 0) return;
    hipLaunchKernelGGL(k_update_norms, dim3(kNormBlocks, nrhs, b.n), dim3(RB), 0, s, b);
    FP_HIP(hipGetLastError());
}
void launch_update_norms(double* u, long long ldu, const double* d, long long ldd, long long n, int nrhs,
                         double* part, hipStream_t s)
{
    NormsBatch b{};
    b.n    = 1;
    b.j[0] = {u, ldu, d, ldd, n, part};
    launch_update_norms_batch(b, nrhs, s);
}

}  // namespace fp
