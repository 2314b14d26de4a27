// Small device kernels of the composite solve: boundary assembly, interface
// line exchange, Schur-block products and the FP64 GMRES state machine over
// interface unknowns.
#include "fp_vec.h"

namespace fp {

constexpr int VB = 256;  // threads per block of the vector kernels

// ---- boundary assembly + h^2 scaling ----------------------------------------
// ConcatPoissonSolver2D::boundary_assembly and the hx^2 scaling
// (pe/concat/concat_solver2d.cpp:110-116,159-223):
//   Dirichlet: f += -g / h^2 on the face line; Neumann: f += +-g / h;
//   then f *= hx^2.
template <typename T>
__global__ void k_assemble(T* f, int nx, int ny, long long ld, T hx2, AsmFaces<T> fc)
{
    const long long n = (long long)nx * ny;
    T*              fr = f + (long long)blockIdx.y * ld;
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(q / ny), j = (int)(q - (long long)i * ny);
        T         v = fr[q];
        if (i == 0 && fc.g[XNEG]) v += fc.a[XNEG] * fc.g[XNEG][(long long)blockIdx.y * ny + j];
        if (i == nx - 1 && fc.g[XPOS]) v += fc.a[XPOS] * fc.g[XPOS][(long long)blockIdx.y * ny + j];
        if (j == 0 && fc.g[YNEG]) v += fc.a[YNEG] * fc.g[YNEG][(long long)blockIdx.y * nx + i];
        if (j == ny - 1 && fc.g[YPOS]) v += fc.a[YPOS] * fc.g[YPOS][(long long)blockIdx.y * nx + i];
        fr[q] = hx2 * v;
    }
}

template <typename T>
void launch_assemble(T* f, int nx, int ny, long long ld, int nrhs, T hx2, const AsmFaces<T>& fc, hipStream_t s)
{
    const long long n   = (long long)nx * ny;
    int             nbx = (int)std::min<long long>((n + VB - 1) / VB, 2048);
    hipLaunchKernelGGL((k_assemble<T>), dim3(nbx, nrhs), dim3(VB), 0, s, f, nx, ny, ld, hx2, fc);
    FP_HIP(hipGetLastError());
}
template void launch_assemble<double>(double*, int, int, long long, int, double, const AsmFaces<double>&, hipStream_t);

// ---- line exchange -----------------------------------------------------------
// field2::bond_add(location, a, neighbour) (base/field/field2.cpp:257-302):
// this side's line += a * the neighbour's adjacent line.
template <typename T>
__global__ void k_bond_add(T* f, int nx, int ny, long long ldf, const T* nb, int nbx, int nby, long long ldn,
                           int loc, T a)
{
    const int t   = blockIdx.x * blockDim.x + threadIdx.x;
    T*        fr  = f + (long long)blockIdx.y * ldf;
    const T*  nbr = nb + (long long)blockIdx.y * ldn;
    if (loc == XNEG && t < ny) fr[t] += a * nbr[(long long)(nbx - 1) * nby + t];
    else if (loc == XPOS && t < ny) fr[(long long)(nx - 1) * ny + t] += a * nbr[t];
    else if (loc == YNEG && t < nx) fr[(long long)t * ny] += a * nbr[(long long)t * nby + nby - 1];
    else if (loc == YPOS && t < nx) fr[(long long)t * ny + ny - 1] += a * nbr[(long long)t * nby];
}

template <typename T>
void launch_bond_add(T* f, int nx, int ny, long long ldf, const T* nb, int nbx, int nby, long long ldn, int loc, T a,
                     int nrhs, hipStream_t s)
{
    const int len = (loc == XNEG || loc == XPOS) ? ny : nx;
    hipLaunchKernelGGL((k_bond_add<T>), dim3((len + VB - 1) / VB, nrhs), dim3(VB), 0, s, f, nx, ny, ldf, nb, nbx,
                       nby, ldn, loc, a);
    FP_HIP(hipGetLastError());
}
template void launch_bond_add<double>(double*, int, int, long long, const double*, int, int, long long, int, double,
                                      int, hipStream_t);

template <typename T>
__global__ void k_fill(T* p, long long n, T v)
{
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
        p[q] = v;
}

template <typename T>
void launch_fill(T* p, long long n, T v, hipStream_t s)
{
    if (n <= 0) return;
    int nb = (int)std::min<long long>((n + VB - 1) / VB, 4096);
    hipLaunchKernelGGL((k_fill<T>), dim3(nb), dim3(VB), 0, s, p, n, v);
    FP_HIP(hipGetLastError());
}
template void launch_fill<double>(double*, long long, double, hipStream_t);
template void launch_fill<float>(float*, long long, float, hipStream_t);

// y = a*x + b*y over n elements (n may span several right-hand sides)
__global__ void k_axpby(const double* x, double* y, long long n, double a, double b)
{
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
        y[q] = a * x[q] + b * y[q];
}
void launch_axpby(const double* x, double* y, long long n, double a, double b, hipStream_t s)
{
    if (n <= 0) return;
    int nb = (int)std::min<long long>((n + VB - 1) / VB, 4096);
    hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(VB), 0, s, x, y, n, a, b);
    FP_HIP(hipGetLastError());
}

// out[r][k] = a * v[r][k] + b * G[r][cell[k]]   (v may be null: a ignored)
__global__ void k_gather(const double* G, long long ldG, const int* cell, int K, const double* v, double* out,
                         long long ldo, double a, double b)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const long long r   = blockIdx.y;
    const double    g   = G[r * ldG + cell[k]];
    const double    val = (v ? a * v[r * ldo + k] : 0.0) + b * g;
    out[r * ldo + k]    = val;
}
void launch_gather(const double* G, long long ldG, const int* cell, int K, const double* v, double* out, long long ldo,
                   double a, double b, int nrhs, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather, dim3((K + VB - 1) / VB, nrhs), dim3(VB), 0, s, G, ldG, cell, K, v, out, ldo, a, b);
    FP_HIP(hipGetLastError());
}

// G[r][ucell[u]] = w[r][k1[u]] + (k2[u] >= 0 ? w[r][k2[u]] : 0): corner cells
// where two interface lines meet receive both contributions, as the
// reference's ft += S_k * x accumulation (gmres_solver2d.cpp:90-104).
__global__ void k_scatter(double* G, long long ldG, const int* ucell, const int* k1, const int* k2, int nu,
                          const double* w, long long ldw)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const long long r = blockIdx.y;
    double          v = w[r * ldw + k1[u]];
    if (k2[u] >= 0) v += w[r * ldw + k2[u]];
    G[r * ldG + ucell[u]] = v;
}
void launch_scatter(double* G, long long ldG, const int* ucell, const int* k1, const int* k2, int nu, const double* w,
                    long long ldw, int nrhs, hipStream_t s)
{
    hipLaunchKernelGGL(k_scatter, dim3((nu + VB - 1) / VB, nrhs), dim3(VB), 0, s, G, ldG, ucell, k1, k2, nu, w, ldw);
    FP_HIP(hipGetLastError());
}

// ---- dense Schur blocks --------------------------------------------------------
// SchurMat2D_*::operator* (schur_mat2d.cpp:18-33 etc.): y_i = sum_j S(i,j) x_j,
// one wavefront per output row; x and y are segments of interface vectors.
__global__ __launch_bounds__(256) void k_schur_gemv(const double* __restrict__ S, int cn, const double* x,
                                                    long long ldx, int xoff, double* y, long long ldy, int yoff)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i    = blockIdx.x * 4 + wave;
    if (i >= cn) return;
    const long long r   = blockIdx.y;
    const double*   Si  = S + (long long)i * cn;
    const double*   xr  = x + r * ldx + xoff;
    double          acc = 0.0;
    for (int j = lane; j < cn; j += 64) acc += Si[j] * xr[j];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) y[r * ldy + yoff + i] = acc;
}
void launch_schur_gemv(const double* S, int cn, const double* x, long long ldx, int xoff, double* y, long long ldy,
                       int yoff, int nrhs, hipStream_t s)
{
    hipLaunchKernelGGL(k_schur_gemv, dim3((cn + 3) / 4, nrhs), dim3(256), 0, s, S, cn, x, ldx, xoff, y, ldy, yoff);
    FP_HIP(hipGetLastError());
}

// S(j, i) = response at line position j to the unit source at position i0+q:
// buf holds `cnt` child fields, field q solved for the unit at i0+q.
__global__ void k_schur_collect(const double* buf, long long ldb, const int* cell, int cn, int i0, double* S)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cn) return;
    const int q            = blockIdx.y;
    S[(long long)j * cn + i0 + q] = buf[(long long)q * ldb + cell[j]];
}
void launch_schur_collect(const double* buf, long long ldb, const int* cell, int cn, int i0, int cnt, double* S,
                          hipStream_t s)
{
    hipLaunchKernelGGL(k_schur_collect, dim3((cn + VB - 1) / VB, cnt), dim3(VB), 0, s, buf, ldb, cell, cn, i0, S);
    FP_HIP(hipGetLastError());
}

// buf field q := unit at cell[i0+q] (buf already zero)
__global__ void k_set_units(double* buf, long long ldb, const int* cell, int i0, int cnt)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= cnt) return;
    buf[(long long)q * ldb + cell[i0 + q]] = 1.0;
}
void launch_set_units(double* buf, long long ldb, const int* cell, int i0, int cnt, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_units, dim3((cnt + VB - 1) / VB), dim3(VB), 0, s, buf, ldb, cell, i0, cnt);
    FP_HIP(hipGetLastError());
}

// ---- GMRES over interface unknowns (FP64) --------------------------------------
// GMRESSolver2D::solve (pe/concat/gmres_solver2d.cpp:123-252) restated on
// interface vectors of length K, nrhs independent systems in lockstep with
// per-system masks.  Orthogonalisation is classical Gram-Schmidt applied
// twice (the reference uses modified Gram-Schmidt; both produce the same
// Hessenberg matrix in exact arithmetic, CGS2 needs one reduction per pass
// instead of one per basis vector).  Reductions are two-stage and
// deterministic: per-block partial sums, then a fixed-order final sum.

// sum of part[0..n) in a fixed order by one block
__device__ double block_sum_fixed(const double* part, int n, double* sh)
{
    double acc = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) acc += part[q];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__device__ double block_reduce(double v, double* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// R = C - BX into V0, partial norms
__global__ __launch_bounds__(VB) void k_gm_resid(GmresDev g)
{
    __shared__ double sh[VB];
    const int         r = blockIdx.y;
    const long long   K = g.K;
    double*           v0 = g.V + (long long)r * K;
    double            acc = 0.0;
    for (long long k = blockIdx.x * (long long)VB + threadIdx.x; k < K; k += (long long)gridDim.x * VB) {
        const double t = g.C[r * K + k] - g.Wv[r * K + k];
        v0[k]          = t;
        acc += t * t;
    }
    acc = block_reduce(acc, sh);
    if (threadIdx.x == 0) g.part[(long long)r * g.nblk + blockIdx.x] = acc;
}

// beta = ||R||, V0 = R / beta, cycle state; convergence at cycle start
// (gmres_solver2d.cpp:148-168)
__global__ __launch_bounds__(VB) void k_gm_cycle_start(GmresDev g)
{
    __shared__ double sh[VB];
    const int         r    = blockIdx.y;
    const double      beta = sqrt(block_sum_fixed(g.part + (long long)r * g.nblk, g.nblk, sh));
    const long long   K    = g.K;
    const bool        conv = beta < g.tolabs[r];
    if (!isfinite(beta)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            g.state[r * 4 + 0] = 0;
            g.state[r * 4 + 2] = 2;  // numeric failure
            g.resid[r]          = beta;
        }
        return;
    }
    if (!conv) {
        double* v0 = g.V + (long long)r * K;
        for (long long k = blockIdx.x * (long long)VB + threadIdx.x; k < K; k += (long long)gridDim.x * VB)
            v0[k] = v0[k] * (1.0 / beta);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        g.resid[r] = beta;
        const int m = g.m;
        for (int q = 0; q <= m; q++) g.g[(long long)r * (m + 1) + q] = 0.0;
        g.g[(long long)r * (m + 1)] = beta;
        g.state[r * 4 + 0]          = conv ? 0 : 1;  // active in this cycle
        g.state[r * 4 + 1]          = 0;             // valid columns
        if (conv) g.state[r * 4 + 2] = 1;            // converged
    }
}

// partial dots of V_i (i <= j) with w: part2[r][i][blk]
__global__ __launch_bounds__(VB) void k_gm_dots(GmresDev g, int j)
{
    __shared__ double sh[VB];
    const int         r = blockIdx.y;
    if (!g.state[r * 4 + 0]) return;
    const long long K   = g.K;
    const long long ldv = (long long)g.nrhs * K;
    for (int i = 0; i <= j; i++) {
        const double* vi  = g.V + i * ldv + (long long)r * K;
        double        acc = 0.0;
        for (long long k = blockIdx.x * (long long)VB + threadIdx.x; k < K; k += (long long)gridDim.x * VB)
            acc += vi[k] * g.Wv[r * K + k];
        acc = block_reduce(acc, sh);
        if (threadIdx.x == 0) g.part2[((long long)r * (g.m + 1) + i) * g.nblk + blockIdx.x] = acc;
    }
}

// h_i = sum part2; w -= sum_i h_i V_i; Hacc += h (pass 1 sets, pass 2 adds);
// partial norms of the new w
__global__ __launch_bounds__(VB) void k_gm_update(GmresDev g, int j, int pass)
{
    __shared__ double sh[VB];
    __shared__ double hs[1024];
    const int         r = blockIdx.y;
    if (!g.state[r * 4 + 0]) return;
    const int m = g.m;
    for (int i = 0; i <= j; i++) {
        const double h = block_sum_fixed(g.part2 + ((long long)r * (m + 1) + i) * g.nblk, g.nblk, sh);
        if (threadIdx.x == 0) hs[i] = h;
    }
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x <= j) {
        double* H = g.Hacc + (long long)r * (m + 1);
        H[threadIdx.x] = (pass == 0 ? 0.0 : H[threadIdx.x]) + hs[threadIdx.x];
    }
    const long long K   = g.K;
    const long long ldv = (long long)g.nrhs * K;
    double          acc = 0.0;
    for (long long k = blockIdx.x * (long long)VB + threadIdx.x; k < K; k += (long long)gridDim.x * VB) {
        double w = g.Wv[r * K + k];
        for (int i = 0; i <= j; i++) w -= hs[i] * g.V[i * ldv + (long long)r * K + k];
        g.Wv[r * K + k] = w;
        acc += w * w;
    }
    acc = block_reduce(acc, sh);
    if (threadIdx.x == 0) g.part[(long long)r * g.nblk + blockIdx.x] = acc;
}

// h_{j+1,j} = ||w||; breakdown check; V_{j+1}; Givens update of column j and
// of g; convergence estimate |g_{j+1}| < tol (gmres_solver2d.cpp:189-225)
__global__ __launch_bounds__(VB) void k_gm_finish(GmresDev g, int j)
{
    __shared__ double sh[VB];
    const int         r = blockIdx.y;
    if (!g.state[r * 4 + 0]) return;
    const int       m   = g.m;
    const double    hn  = sqrt(block_sum_fixed(g.part + (long long)r * g.nblk, g.nblk, sh));
    const long long K   = g.K;
    const long long ldv = (long long)g.nrhs * K;
    if (!isfinite(hn)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            g.state[r * 4 + 0] = 0;
            g.state[r * 4 + 2] = 2;
        }
        return;
    }
    if (hn < 1e-12) {  // breakdown: stop this cycle with j columns
        if (blockIdx.x == 0 && threadIdx.x == 0) g.state[r * 4 + 0] = 0;
        return;
    }
    double* vn = g.V + (j + 1) * ldv + (long long)r * K;
    for (long long k = blockIdx.x * (long long)VB + threadIdx.x; k < K; k += (long long)gridDim.x * VB)
        vn[k] = g.Wv[r * K + k] * (1.0 / hn);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double*       H  = g.H + (long long)r * (m + 1) * m;
        double*       cs = g.cs + (long long)r * m;
        double*       sn = g.sn + (long long)r * m;
        double*       gg = g.g + (long long)r * (m + 1);
        const double* ha = g.Hacc + (long long)r * (m + 1);
        for (int i = 0; i <= j; i++) H[i * m + j] = ha[i];
        H[(j + 1) * m + j] = hn;
        for (int i = 0; i < j; i++) {
            const double t     = cs[i] * H[i * m + j] + sn[i] * H[(i + 1) * m + j];
            H[(i + 1) * m + j] = -sn[i] * H[i * m + j] + cs[i] * H[(i + 1) * m + j];
            H[i * m + j]       = t;
        }
        const double hjj = H[j * m + j], hj1 = H[(j + 1) * m + j];
        const double den = hypot(hjj, hj1);
        cs[j]            = hjj / den;
        sn[j]            = hj1 / den;
        H[j * m + j]       = cs[j] * H[j * m + j] + sn[j] * H[(j + 1) * m + j];
        H[(j + 1) * m + j] = 0.0;
        const double t     = cs[j] * gg[j];
        gg[j + 1]          = -sn[j] * gg[j];
        gg[j]              = t;
        g.state[r * 4 + 1] = j + 1;
        g.resid[r]         = fabs(gg[j + 1]);
        if (fabs(gg[j + 1]) < g.tolabs[r]) g.state[r * 4 + 0] = 0;
        g.state[r * 4 + 3] += 1;  // Arnoldi steps taken
    }
}

// back substitution for y (gmres_solver2d.cpp:228-236), one thread per system
__global__ void k_gm_backsub(GmresDev g)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= g.nrhs) return;
    const int     m  = g.m;
    const int     jr = g.state[r * 4 + 1];
    const double* H  = g.H + (long long)r * (m + 1) * m;
    const double* gg = g.g + (long long)r * (m + 1);
    double*       y  = g.Y + (long long)r * m;
    for (int i = 0; i < m; i++) y[i] = 0.0;
    for (int i = jr - 1; i >= 0; i--) {
        double acc = gg[i];
        for (int k = i + 1; k < jr; k++) acc -= H[i * m + k] * y[k];
        y[i] = acc / H[i * m + i];
    }
}

// x += V(:, 0:jr) y (gmres_solver2d.cpp:238-241)
__global__ __launch_bounds__(VB) void k_gm_update_x(GmresDev g)
{
    const int       r   = blockIdx.y;
    const int       jr  = g.state[r * 4 + 1];
    const long long K   = g.K;
    const long long ldv = (long long)g.nrhs * K;
    const double*   y   = g.Y + (long long)r * g.m;
    for (long long k = blockIdx.x * (long long)VB + threadIdx.x; k < K; k += (long long)gridDim.x * VB) {
        double x = g.X[r * K + k];
        for (int i = 0; i < jr; i++) x += y[i] * g.V[i * ldv + (long long)r * K + k];
        g.X[r * K + k] = x;
    }
}

// tolabs[r] = tol * ||C_r|| (relative) or tol (absolute)
__global__ __launch_bounds__(VB) void k_gm_norm_part(GmresDev g)
{
    __shared__ double sh[VB];
    const int         r   = blockIdx.y;
    const long long   K   = g.K;
    double            acc = 0.0;
    for (long long k = blockIdx.x * (long long)VB + threadIdx.x; k < K; k += (long long)gridDim.x * VB)
        acc += g.C[r * K + k] * g.C[r * K + k];
    acc = block_reduce(acc, sh);
    if (threadIdx.x == 0) g.part[(long long)r * g.nblk + blockIdx.x] = acc;
}
__global__ __launch_bounds__(VB) void k_gm_set_tol(GmresDev g, double tol, int rel)
{
    __shared__ double sh[VB];
    const int         r  = blockIdx.x;
    const double      nc = sqrt(block_sum_fixed(g.part + (long long)r * g.nblk, g.nblk, sh));
    if (threadIdx.x == 0) {
        g.tolabs[r]        = rel ? tol * nc : tol;
        g.resid[r]         = nc;
        g.state[r * 4 + 0] = 0;
        g.state[r * 4 + 1] = 0;
        g.state[r * 4 + 2] = 0;
        g.state[r * 4 + 3] = 0;
    }
}
void gm_init(const GmresDev& g, double tol, int rel, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_norm_part, dim3(g.nblk, g.nrhs), dim3(VB), 0, s, g);
    FP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_gm_set_tol, dim3(g.nrhs), dim3(VB), 0, s, g, tol, rel);
    FP_HIP(hipGetLastError());
}

void gm_resid(const GmresDev& g, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_resid, dim3(g.nblk, g.nrhs), dim3(VB), 0, s, g);
    FP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_gm_cycle_start, dim3(g.nblk, g.nrhs), dim3(VB), 0, s, g);
    FP_HIP(hipGetLastError());
}
void gm_arnoldi(const GmresDev& g, int j, hipStream_t s)
{
    for (int pass = 0; pass < 2; pass++) {
        hipLaunchKernelGGL(k_gm_dots, dim3(g.nblk, g.nrhs), dim3(VB), 0, s, g, j);
        FP_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_gm_update, dim3(g.nblk, g.nrhs), dim3(VB), 0, s, g, j, pass);
        FP_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_gm_finish, dim3(g.nblk, g.nrhs), dim3(VB), 0, s, g, j);
    FP_HIP(hipGetLastError());
}
void gm_finish_cycle(const GmresDev& g, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_backsub, dim3((g.nrhs + 63) / 64), dim3(64), 0, s, g);
    FP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_gm_update_x, dim3(g.nblk, g.nrhs), dim3(VB), 0, s, g);
    FP_HIP(hipGetLastError());
}

}  // namespace fp
