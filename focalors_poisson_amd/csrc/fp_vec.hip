// Small device kernels of the composite solve: boundary assembly, interface
// line exchange, Schur-block products and the FP64 GMRES state machine over
// interface unknowns.
#include "fp_vec.h"

#include <algorithm>
#include <mutex>

namespace fp {

// ---- dense products of many right-hand sides on the matrix cores ----------
// With several right-hand sides the dense products (explicit inverse, Schur
// blocks, basis maps) are matrix-matrix products
//   Y[r][i] = alpha sum_k A[i][k] X[r][k] + beta Y[r][i],   r < nrhs, i < m,
// with A row-major (lda) and each right-hand side's X[r], Y[r] contiguous.
// They run on v_mfma_f64_16x16x4f64 (FP32 operands are widened to FP64:
// exact products, FP64 sums).  Operand layouts of the instruction (gfx950):
// A(i, q) in lane i + 16 q, B(q, j) in lane j + 16 q, D(i, j) in lane
// j + 16 (i % 4), element i / 4.
//
// A wavefront computes 16 rows by up to 16 NTL right-hand sides over one
// run of k; the nw wavefronts of a workgroup take equal runs of k for the
// same outputs and their partial tiles are added in a fixed order through
// one LDS slot (no atomics and no scratch in device memory: results are
// reproducible and the call allocates nothing).  Two row tiles per
// wavefront (half the operand loads of X) measured slower: at 2048
// wavefronts or fewer the loop is bound by its own latency, and the second
// tile halves the wavefronts.  Each
// block of 16 k is four instructions per (row tile, right-hand-side tile),
// and lane (i, q) takes the four consecutive columns 4q .. 4q+3 of the
// block, one per instruction, in both operands: the sum over the block is
// the same, and every lane reads 32 contiguous bytes of each of its rows per
// block (two 16-byte loads where rows are 16-byte aligned).  Every A load
// serves 4 NTL products; the next block's loads are issued before this
// block's products.
typedef double gd4_t __attribute__((ext_vector_type(4)));
constexpr int kGemmMinRhs = 4;
constexpr int kGemmRowTiles = 1;  // 16-row tiles per wavefront
constexpr int kGemmMaxWaves = 8;

// v[s] = p[kk + s] (s < 4) where kk + s < ke and ok, else 0
template <typename T, bool VEC>
__device__ __forceinline__ void gemm_load4(const T* __restrict__ p, int kk, int ke, bool ok, double (&v)[4])
{
    if (VEC && ok && kk + 3 < ke) {
        if constexpr (sizeof(T) == 8) {
            const double2 x0 = *reinterpret_cast<const double2*>(p + kk);
            const double2 x1 = *reinterpret_cast<const double2*>(p + kk + 2);
            v[0] = x0.x;
            v[1] = x0.y;
            v[2] = x1.x;
            v[3] = x1.y;
        } else {
            const float4 x = *reinterpret_cast<const float4*>(p + kk);
            v[0] = x.x;
            v[1] = x.y;
            v[2] = x.z;
            v[3] = x.w;
        }
    } else {
#pragma unroll
        for (int s = 0; s < 4; s++) v[s] = (ok && kk + s < ke) ? (double)p[kk + s] : 0.0;
    }
}

template <typename T, int NTL, bool VEC>
__global__ __launch_bounds__(64 * kGemmMaxWaves) void k_gemm_mfma(const T* __restrict__ A, int m, int k,
                                                                 long long lda, const T* __restrict__ X,
                                                                 long long ldx, T* Y, long long ldy, int nrhs,
                                                                 double alpha, double beta, int kc)
{
    constexpr int MT = kGemmRowTiles;
    extern __shared__ double gred[];  // one slot of 64 lanes x 4 MT NTL doubles
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int il = lane & 15, q = lane >> 4;
    const int i0 = blockIdx.x * 16 * MT;
    const int r0 = blockIdx.y * 16 * NTL;
    const T*  Ar[MT];
    bool      aok[MT];
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
        const int i = i0 + 16 * mt + il;
        aok[mt]     = i < m;
        Ar[mt]      = A + (long long)(aok[mt] ? i : 0) * lda;
    }
    const T* Xr[NTL];
    bool     xok[NTL];
#pragma unroll
    for (int t = 0; t < NTL; t++) {
        const int r = r0 + 16 * t + il;
        xok[t]      = r < nrhs;
        Xr[t]       = X + (long long)(xok[t] ? r : 0) * ldx;
    }
    gd4_t acc[MT][NTL];
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int t = 0; t < NTL; t++) acc[mt][t] = gd4_t{0.0, 0.0, 0.0, 0.0};
    const int kb = w * kc, ke = min(k, kb + kc);
    double    a[MT][4], b[NTL][4];
    auto      load = [&](int k0) {
        const int kk = k0 + 4 * q;
#pragma unroll
        for (int mt = 0; mt < MT; mt++) gemm_load4<T, VEC>(Ar[mt], kk, ke, aok[mt], a[mt]);
#pragma unroll
        for (int t = 0; t < NTL; t++) gemm_load4<T, VEC>(Xr[t], kk, ke, xok[t], b[t]);
    };
    if (kb < ke) load(kb);
    for (int k0 = kb; k0 < ke; k0 += 16) {
        double ca[MT][4], cb[NTL][4];
#pragma unroll
        for (int s = 0; s < 4; s++) {
#pragma unroll
            for (int mt = 0; mt < MT; mt++) ca[mt][s] = a[mt][s];
#pragma unroll
            for (int t = 0; t < NTL; t++) cb[t][s] = b[t][s];
        }
        if (k0 + 16 < ke) load(k0 + 16);
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int mt = 0; mt < MT; mt++)
#pragma unroll
                for (int t = 0; t < NTL; t++)
                    acc[mt][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[mt][s], cb[t][s], acc[mt][t], 0, 0, 0);
    }
    // fixed order: wavefront 0 adds the partial tiles of nw-1, nw-2, .., 1
    for (int h = nw - 1; h >= 1; h--) {
        if (w == h) {
#pragma unroll
            for (int mt = 0; mt < MT; mt++)
#pragma unroll
                for (int t = 0; t < NTL; t++)
#pragma unroll
                    for (int v = 0; v < 4; v++) gred[((mt * NTL + t) * 4 + v) * 64 + lane] = acc[mt][t][v];
        }
        __syncthreads();
        if (w == 0) {
#pragma unroll
            for (int mt = 0; mt < MT; mt++)
#pragma unroll
                for (int t = 0; t < NTL; t++)
#pragma unroll
                    for (int v = 0; v < 4; v++) acc[mt][t][v] += gred[((mt * NTL + t) * 4 + v) * 64 + lane];
        }
        __syncthreads();
    }
    if (w != 0) return;
#pragma unroll
    for (int t = 0; t < NTL; t++) {
        const int r = r0 + 16 * t + il;
        if (r >= nrhs) continue;
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = i0 + 16 * mt + 4 * v + q;
                if (row >= m) continue;
                T&           y   = Y[(long long)r * ldy + row];
                const double val = alpha * acc[mt][t][v] + (beta != 0.0 ? beta * (double)y : 0.0);
                y                = (T)val;
            }
    }
}

template <typename T, int NTL>
static void launch_gemm_mfma(const T* A, int m, int k, long long lda, const T* X, long long ldx, T* Y,
                             long long ldy, int nrhs, T alpha, T beta, hipStream_t s)
{
    const int gx = (m + 16 * kGemmRowTiles - 1) / (16 * kGemmRowTiles), gy = (nrhs + 16 * NTL - 1) / (16 * NTL);
    // wavefronts per workgroup: enough for about 2048 on the device, each
    // with at least 4 blocks of k
    int nw = 1;
    while (nw < kGemmMaxWaves && (long long)gx * gy * nw < 2048 && k / (16 * 2 * nw) >= 4) nw *= 2;
    const int    kc  = ((k + nw - 1) / nw + 15) / 16 * 16;
    const size_t lds = (size_t)64 * 4 * kGemmRowTiles * NTL * sizeof(double);
    // 16-byte loads when every row of A and X starts on a 16-byte boundary
    constexpr long long vl = 16 / sizeof(T);
    const bool vec = ((uintptr_t)A % 16 == 0) && ((uintptr_t)X % 16 == 0) && lda % vl == 0 && ldx % vl == 0;
    if (vec)
        hipLaunchKernelGGL((k_gemm_mfma<T, NTL, true>), dim3(gx, gy), dim3(64 * nw), lds, s, A, m, k, lda, X, ldx, Y,
                           ldy, nrhs, (double)alpha, (double)beta, kc);
    else
        hipLaunchKernelGGL((k_gemm_mfma<T, NTL, false>), dim3(gx, gy), dim3(64 * nw), lds, s, A, m, k, lda, X, ldx,
                           Y, ldy, nrhs, (double)alpha, (double)beta, kc);
    FP_HIP(hipGetLastError());
}

template <typename T>
static void dense_apply(const T* A, int m, int k, int lda, const T* X, long long ldx, T* Y, long long ldy, int nrhs,
                        T alpha, T beta, hipStream_t s)
{
    if (m <= 0 || nrhs <= 0) return;
    if (nrhs <= 16) launch_gemm_mfma<T, 1>(A, m, k, lda, X, ldx, Y, ldy, nrhs, alpha, beta, s);
    else if (nrhs <= 32) launch_gemm_mfma<T, 2>(A, m, k, lda, X, ldx, Y, ldy, nrhs, alpha, beta, s);
    else launch_gemm_mfma<T, 4>(A, m, k, lda, X, ldx, Y, ldy, nrhs, alpha, beta, s);
}

constexpr int VB = 256;  // threads per block of the vector kernels

template <typename T>
__device__ __forceinline__ T wave_reduce_sum(T v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

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

// out[r][t] = a * f[r][line cell t] for one face of an nx x ny field
__global__ void k_face_extract(const double* f, int nx, int ny, long long ld, int face, double a, double* out)
{
    const int       t   = blockIdx.x * blockDim.x + threadIdx.x;
    const long long r   = blockIdx.y;
    const int       len = face < 2 ? ny : nx;
    if (t >= len) return;
    long long c;
    if (face == XNEG) c = t;
    else if (face == XPOS) c = (long long)(nx - 1) * ny + t;
    else if (face == YNEG) c = (long long)t * ny;
    else c = (long long)t * ny + ny - 1;
    out[r * len + t] = a * f[r * ld + c];
}
void launch_face_extract(const double* f, int nx, int ny, long long ld, int face, double a, double* out, int nrhs,
                         hipStream_t s)
{
    const int len = face < 2 ? ny : nx;
    hipLaunchKernelGGL(k_face_extract, dim3((len + 255) / 256, nrhs), dim3(256), 0, s, f, nx, ny, ld, face, a, out);
    FP_HIP(hipGetLastError());
}

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
// SchurMat2D_*::operator* (schur_mat2d.cpp:18-33 etc.): y_i = sum_j S(i,j) x_j.
// All children's Schur products of one root in one launch, written straight
// into the root's line buffer (x-lines as rows 0/1, y-lines as a/b) or, by
// launch_schur_gemv_all, into an interface vector: one wavefront per output
// row, 4 rows per block.
__global__ __launch_bounds__(256) void k_schur_lines(SchurLinesArgs a)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row  = blockIdx.x * 4 + wave;
    if (row >= a.total_rows) return;
    int c = 0;
    while (c + 1 < a.nch && row >= a.ch[c + 1].row0) c++;
    const SchurChild& C  = a.ch[c];
    const int         i  = row - C.row0;
    const long long   r  = blockIdx.y;
    const double*     Si = C.S + (long long)i * C.cn;
    const double*     xr = a.v + r * a.ldv + C.voff;
    // four independent sums: four row loads in flight per wavefront
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int    j  = lane;
    for (; j + 192 < C.cn; j += 256) {
        a0 += Si[j] * xr[j];
        a1 += Si[j + 64] * xr[j + 64];
        a2 += Si[j + 128] * xr[j + 128];
        a3 += Si[j + 192] * xr[j + 192];
    }
    for (; j < C.cn; j += 64) a0 += Si[j] * xr[j];
    const double acc = wave_reduce_sum((a0 + a1) + (a2 + a3));
    if (lane == 0) a.lb[r * a.ldl + C.lboff + i] = acc;
}
void launch_schur_lines(const SchurLinesArgs& a, int nrhs, hipStream_t s)
{
    if (a.total_rows <= 0) return;
    if (nrhs >= kGemmMinRhs) {
        for (int c = 0; c < a.nch; c++) {
            const SchurChild& C = a.ch[c];
            dense_apply(C.S, C.cn, C.cn, C.cn, a.v + C.voff, a.ldv, a.lb + C.lboff, a.ldl, nrhs, 1.0, 0.0, s);
        }
        return;
    }
    hipLaunchKernelGGL(k_schur_lines, dim3((a.total_rows + 3) / 4, nrhs), dim3(256), 0, s, a);
    FP_HIP(hipGetLastError());
}

void launch_schur_gemv_all(const SchurChild* ch, int nch, const double* x, long long ldx, double* y, long long ldy,
                           int nrhs, hipStream_t s)
{
    for (int c0 = 0; c0 < nch; c0 += kMaxSchurChildren) {
        SchurLinesArgs a{};
        int            rows = 0;
        for (int c = c0; c < nch && c < c0 + kMaxSchurChildren; c++) {
            a.ch[a.nch]      = ch[c];
            a.ch[a.nch].row0 = rows;
            rows += ch[c].cn;
            a.nch++;
        }
        a.total_rows = rows;
        a.v          = x;
        a.ldv        = ldx;
        a.lb         = y;
        a.ldl        = ldy;
        launch_schur_lines(a, nrhs, s);
    }
}

// Reductions over the column partials of the line-sourced sweeps: kRedU
// consecutive outputs per 256-thread workgroup (one 128-byte line per
// output run), kRedS slices of the mode blocks, summed in a fixed order.
constexpr int kRedU = 16, kRedS = 256 / kRedU;
template <typename T>
__device__ __forceinline__ T red_slices(const T (*red)[kRedU], int u)
{
    T v[kRedS];
#pragma unroll
    for (int q = 0; q < kRedS; q++) v[q] = red[q][u];
#pragma unroll
    for (int w = 1; w < kRedS; w *= 2)
#pragma unroll
        for (int q = 0; q + w < kRedS; q += 2 * w) v[q] += v[q + w];
    return v[0];
}

// Column sums of the line-sourced sweeps: yl[r][w*nx + i] (w = 0, 1) =
// sum over the mode blocks b of cp[(r*nblk + b)*2nx + w*nx + i]: kRedU
// consecutive outputs per workgroup (coalesced reads), kRedS slices of the
// blocks, in a fixed order (reproducible).  yl entries are ldy apart per
// right-hand side.
// half = 0 or 1: only that column's sums, written at yl[r * ldy + i].
// Up to 8 independent reductions per launch (blockIdx.z).
template <typename T>
__global__ __launch_bounds__(256) void k_colsum(ColsumBatchT<T> b)
{
    __shared__ T         red[kRedS][kRedU];
    const ColsumJobT<T>& J  = b.j[blockIdx.z];
    const int            u  = threadIdx.x % kRedU, sl = threadIdx.x / kRedU;
    const int            o  = blockIdx.x * kRedU + u;  // output index
    const bool           ok = o < (J.half < 0 ? 2 : 1) * J.nx;
    const long long      r  = blockIdx.y;
    T                    acc = T(0);
    if (ok) {
        const T* src = J.cp + r * (long long)J.nblk * 2 * J.nx + o + (J.half > 0 ? J.nx : 0);
#pragma unroll 4
        for (int q = sl; q < J.nblk; q += kRedS) acc += src[(long long)q * 2 * J.nx];
    }
    red[sl][u] = acc;
    __syncthreads();
    if (sl == 0 && ok) J.yl[r * J.ldy + o] = red_slices<T>(red, u);
}
template <typename T>
void launch_colsum_batch(const ColsumBatchT<T>& b, int nrhs, hipStream_t s)
{
    if (b.n <= 0) return;
    int nout = 0;
    for (int q = 0; q < b.n; q++) nout = std::max(nout, (b.j[q].half < 0 ? 2 : 1) * b.j[q].nx);
    hipLaunchKernelGGL((k_colsum<T>), dim3((nout + kRedU - 1) / kRedU, nrhs, b.n), dim3(256), 0, s, b);
    FP_HIP(hipGetLastError());
}
template void launch_colsum_batch<double>(const ColsumBatchT<double>&, int, hipStream_t);
template void launch_colsum_batch<float>(const ColsumBatchT<float>&, int, hipStream_t);
void launch_colsum(const double* cp, int nblk, int nx, double* yl, long long ldy, int nrhs, hipStream_t s, int half)
{
    ColsumBatch b;
    b.n    = 1;
    b.j[0] = {cp, nblk, nx, half, yl, ldy};
    launch_colsum_batch(b, nrhs, s);
}

// Batched strided dense products y[r*ldy + i] = sum_j A[i*n + j] x[r*ldx + j]
// (basis changes of interface rows between subdomains), one wavefront per row.
template <typename T>
__global__ __launch_bounds__(256) void k_gemv_batch(GemvBatchT<T> b)
{
    const GemvJobT<T>& J    = b.j[blockIdx.z];
    const int          wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int          i    = blockIdx.x * 4 + wave;
    if (i >= J.n) return;
    const long long r   = blockIdx.y;
    const T*        Ai  = J.A + (long long)i * J.n;
    const T*        xr  = J.x + r * J.ldx;
    T               acc = T(0);
    for (int q = lane; q < J.n; q += 64) acc += Ai[q] * xr[q];
    acc = wave_reduce_sum(acc);
    if (lane == 0) J.y[r * J.ldy + i] = acc;
}
template <typename T>
void launch_gemv_batch(const GemvBatchT<T>& b, int nrhs, hipStream_t s)
{
    if (b.n <= 0) return;
    if (nrhs >= kGemmMinRhs) {
        for (int q = 0; q < b.n; q++) {
            const GemvJobT<T>& J = b.j[q];
            dense_apply(J.A, J.n, J.n, J.n, J.x, J.ldx, J.y, J.ldy, nrhs, T(1), T(0), s);
        }
        return;
    }
    int nmax = 0;
    for (int q = 0; q < b.n; q++) nmax = std::max(nmax, b.j[q].n);
    hipLaunchKernelGGL((k_gemv_batch<T>), dim3((nmax + 3) / 4, nrhs, b.n), dim3(256), 0, s, b);
    FP_HIP(hipGetLastError());
}
template void launch_gemv_batch<double>(const GemvBatchT<double>&, int, hipStream_t);
template void launch_gemv_batch<float>(const GemvBatchT<float>&, int, hipStream_t);

// Restriction of a line-sourced solution to the interface unknowns: t[k] is
// the x-line coefficient (rows 0 / nx-1 in xu) or the physical y-line value
// of unknown k, the latter the total of the sweep's per-workgroup partial
// column sums (cp, nblk blocks of 2 nx per right-hand side), reduced here
// (kRedU unknowns per workgroup).  mode 0: out = v - t (v == nullptr: the unit
// vector e_{k0 + r}); mode 1: out = t; mode 2: out = v + w - t.
template <typename TL>
__global__ __launch_bounds__(256) void k_lines_assemble(const double* v, double* out, int K, const int* kface,
                                                        const int* kpos, const TL* xu, int ny, const TL* cp, int nblk,
                                                        int nx, int k0, int mode, const double* w)
{
    // kRedU consecutive unknowns per workgroup; on y faces kRedS slices of
    // the mode blocks of the column partials (coalesced reads, fixed order)
    __shared__ double red[kRedS][kRedU];
    const int       u  = threadIdx.x % kRedU, sl = threadIdx.x / kRedU;
    const int       k  = blockIdx.x * kRedU + u;
    const bool      ok = k < K;
    const long long r  = blockIdx.y;
    const int       f = ok ? kface[k] : XNEG, p = ok ? kpos[k] : 0;
    double          t = 0.0;
    if (ok && (f == XNEG || f == XPOS)) {
        if (sl == 0) t = (double)xu[r * 2 * ny + (f == XPOS ? ny : 0) + p];
    } else if (ok) {
        const TL* src = cp + r * (long long)nblk * 2 * nx + (f == YNEG ? 0 : nx) + p;
        TL        acc = TL(0);
#pragma unroll 4
        for (int q = sl; q < nblk; q += kRedS) acc += src[(long long)q * 2 * nx];
        t = (double)acc;
    }
    red[sl][u] = t;
    __syncthreads();
    if (sl != 0 || !ok) return;
    t = red_slices<double>(red, u);
    if (mode == 1) {
        out[r * K + k] = t;
        return;
    }
    if (mode == 2) {
        out[r * K + k] = v[r * K + k] + w[r * K + k] - t;
        return;
    }
    const double vv = v ? v[r * K + k] : (k == k0 + (int)r ? 1.0 : 0.0);
    out[r * K + k]  = vv - t;
}
template <typename TL>
void launch_lines_assemble(const double* v, double* out, int K, const int* kface, const int* kpos, const TL* xu,
                           int ny, const TL* cp, int nblk, int nx, int nrhs, hipStream_t s, int k0, int mode,
                           const double* w)
{
    if (mode == 2 && (!v || !w)) throw Error(FP_E_INVALID, "lines_assemble mode 2 needs v and w");
    hipLaunchKernelGGL((k_lines_assemble<TL>), dim3((K + kRedU - 1) / kRedU, nrhs), dim3(256), 0, s, v, out, K, kface,
                       kpos, xu,
                       ny, cp, nblk, nx, k0, mode, w);
    FP_HIP(hipGetLastError());
}
template void launch_lines_assemble<double>(const double*, double*, int, const int*, const int*, const double*, int,
                                            const double*, int, int, int, hipStream_t, int, int, const double*);
template void launch_lines_assemble<float>(const double*, double*, int, const int*, const int*, const float*, int,
                                           const float*, int, int, int, hipStream_t, int, int, const double*);

// Batched strided line copies: dst[r*ldd + t] = coef * src[r*lds + t*stride],
// t < len, for up to 8 lines (interface lines handed between subdomains).
template <typename T>
__global__ void k_line_copy(LineCopyBatchT<T> b)
{
    const LineCopyT<T>& c = b.c[blockIdx.z];
    const long long     r = blockIdx.y;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < c.len; t += gridDim.x * blockDim.x)
        c.dst[r * c.ldd + t] = c.coef * c.src[r * c.lds + (long long)t * c.stride];
}
template <typename T>
void launch_line_copy(const LineCopyBatchT<T>& b, int nrhs, hipStream_t s)
{
    if (b.n <= 0) return;
    int maxlen = 0;
    for (int q = 0; q < b.n; q++) maxlen = std::max(maxlen, b.c[q].len);
    hipLaunchKernelGGL((k_line_copy<T>), dim3((maxlen + VB - 1) / VB, nrhs, b.n), dim3(VB), 0, s, b);
    FP_HIP(hipGetLastError());
}
template void launch_line_copy<double>(const LineCopyBatchT<double>&, int, hipStream_t);
template void launch_line_copy<float>(const LineCopyBatchT<float>&, int, hipStream_t);

template <typename T>
struct ExchangeBatch {
    ColsumBatchT<T>   cs;
    GemvBatchT<T>     gv;
    LineCopyBatchT<T> lc;
    int               start[25];  // first block of each job: colsums, products, copies
    int               njobs;
};
template <typename T>
__global__ __launch_bounds__(256) void k_exchange(ExchangeBatch<T> b)
{
    int q = 0;
    while (q + 1 < b.njobs && (int)blockIdx.x >= b.start[q + 1]) q++;
    const int       blk  = (int)blockIdx.x - b.start[q];
    const int       wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long r    = blockIdx.y;
    if (q < b.cs.n) {
        // kRedU consecutive outputs per workgroup (consecutive partials of
        // one mode block per load), kRedS slices of the mode blocks
        __shared__ T         red[kRedS][kRedU];
        const ColsumJobT<T>& J  = b.cs.j[q];
        const int            u  = threadIdx.x % kRedU, sl = threadIdx.x / kRedU;
        const int            o  = blk * kRedU + u;
        const bool           ok = o < (J.half < 0 ? 2 : 1) * J.nx;
        T                    acc = T(0);
        if (ok) {
            const T* src = J.cp + r * (long long)J.nblk * 2 * J.nx + o + (J.half > 0 ? J.nx : 0);
#pragma unroll 4
            for (int k = sl; k < J.nblk; k += kRedS) acc += src[(long long)k * 2 * J.nx];
        }
        red[sl][u] = acc;
        __syncthreads();
        if (sl == 0 && ok) J.yl[r * J.ldy + o] = red_slices<T>(red, u);
        return;
    }
    q -= b.cs.n;
    if (q < b.gv.n) {
        const GemvJobT<T>& J = b.gv.j[q];
        const int          i = blk * 4 + wave;
        if (i >= J.n) return;
        const T* Ai = J.A + (long long)i * J.n;
        const T* xr = J.x + r * J.ldx;
        // four independent sums: four row loads in flight per wavefront
        T   a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
        int k  = lane;
        for (; k + 192 < J.n; k += 256) {
            a0 += Ai[k] * xr[k];
            a1 += Ai[k + 64] * xr[k + 64];
            a2 += Ai[k + 128] * xr[k + 128];
            a3 += Ai[k + 192] * xr[k + 192];
        }
        for (; k < J.n; k += 64) a0 += Ai[k] * xr[k];
        const T acc = wave_reduce_sum((a0 + a1) + (a2 + a3));
        if (lane == 0) J.y[r * J.ldy + i] = acc;
        return;
    }
    q -= b.gv.n;
    const LineCopyT<T>& c = b.lc.c[q];
    const int           t = blk * VB + threadIdx.x;
    if (t < c.len) c.dst[r * c.ldd + t] = c.coef * c.src[r * c.lds + (long long)t * c.stride];
}
template <typename T>
void launch_exchange(const ColsumBatchT<T>& cs, const GemvBatchT<T>& gv, const LineCopyBatchT<T>& lc, int nrhs,
                     hipStream_t s)
{
    static_assert(VB == 256, "one block size for every exchange job");
    ExchangeBatch<T> b{};
    b.cs       = cs;
    b.lc       = lc;
    int blocks = 0, q = 0;
    for (int k = 0; k < cs.n; k++, q++) {
        b.start[q] = blocks;
        blocks += ((cs.j[k].half < 0 ? 2 : 1) * cs.j[k].nx + kRedU - 1) / kRedU;
    }
    if (nrhs >= kGemmMinRhs) {
        launch_gemv_batch(gv, nrhs, s);
    } else {
        b.gv = gv;
        for (int k = 0; k < gv.n; k++, q++) {
            b.start[q] = blocks;
            blocks += (gv.j[k].n + 3) / 4;
        }
    }
    for (int k = 0; k < lc.n; k++, q++) {
        b.start[q] = blocks;
        blocks += (lc.c[k].len + VB - 1) / VB;
    }
    b.njobs    = q;
    b.start[q] = blocks;
    if (blocks == 0) return;
    hipLaunchKernelGGL((k_exchange<T>), dim3(blocks, nrhs), dim3(VB), 0, s, b);
    FP_HIP(hipGetLastError());
}
template void launch_exchange<double>(const ColsumBatchT<double>&, const GemvBatchT<double>&,
                                      const LineCopyBatchT<double>&, int, hipStream_t);
template void launch_exchange<float>(const ColsumBatchT<float>&, const GemvBatchT<float>&,
                                     const LineCopyBatchT<float>&, int, hipStream_t);

__global__ void k_to_f32(const double* src, float* dst, long long n)
{
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
        dst[q] = (float)src[q];
}
void launch_to_f32(const double* src, float* dst, long long n, hipStream_t s)
{
    if (n <= 0) return;
    const int nb = (int)std::min<long long>((n + VB - 1) / VB, 1024);
    hipLaunchKernelGGL(k_to_f32, dim3(nb), dim3(VB), 0, s, src, dst, n);
    FP_HIP(hipGetLastError());
}

// out = in^T for an n x n matrix (32x32 tiles through LDS)
__global__ void k_transpose(const double* in, double* out, int n)
{
    __shared__ double t[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    for (int q = threadIdx.y; q < 32; q += blockDim.y) {
        const int r = by + q, c = bx + threadIdx.x;
        if (r < n && c < n) t[q][threadIdx.x] = in[(long long)r * n + c];
    }
    __syncthreads();
    for (int q = threadIdx.y; q < 32; q += blockDim.y) {
        const int r = bx + q, c = by + threadIdx.x;
        if (r < n && c < n) out[(long long)r * n + c] = t[threadIdx.x][q];
    }
}
void launch_transpose(const double* in, double* out, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_transpose, dim3((n + 31) / 32, (n + 31) / 32), dim3(32, 8), 0, s, in, out, n);
    FP_HIP(hipGetLastError());
}

// ---- dense interface preconditioner ---------------------------------------------
// y = alpha * A x + beta * y for a dense n x n row-major A, one wavefront per
// row, nrhs right-hand sides (x, y: nrhs x n)
__global__ __launch_bounds__(256) void k_dense_gemv(const double* __restrict__ A, int n, const double* x, double* y,
                                                    double alpha, double beta)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i    = blockIdx.x * 4 + wave;
    if (i >= n) return;
    const long long r   = blockIdx.y;
    const double*   Ai  = A + (long long)i * n;
    const double*   xr  = x + r * n;
    double          a0 = 0.0, a1 = 0.0;
    int             j  = lane;
    for (; j + 64 < n; j += 128) {
        a0 += Ai[j] * xr[j];
        a1 += Ai[j + 64] * xr[j + 64];
    }
    if (j < n) a0 += Ai[j] * xr[j];
    const double acc = wave_reduce_sum(a0 + a1);
    if (lane == 0) y[r * n + i] = alpha * acc + (beta == 0.0 ? 0.0 : beta * y[r * n + i]);
}
void launch_dense_gemv(const double* A, int n, const double* x, double* y, double alpha, double beta, int nrhs,
                       hipStream_t s)
{
    if (nrhs >= kGemmMinRhs) {
        dense_apply(A, n, n, n, x, n, y, n, nrhs, alpha, beta, s);
        return;
    }
    hipLaunchKernelGGL(k_dense_gemv, dim3((n + 3) / 4, nrhs), dim3(256), 0, s, A, n, x, y, alpha, beta);
    FP_HIP(hipGetLastError());
}

// line buffers of unit interface vectors k0 .. k0+cnt-1: the Schur product of a
// unit vector is one column of that child's block
__global__ void k_lines_from_units(SchurLinesArgs a, int k0, int cnt)
{
    const int q = blockIdx.y;
    const int k = k0 + q;
    int       c = 0;
    while (c + 1 < a.nch && k >= a.ch[c + 1].voff) c++;
    const SchurChild& C = a.ch[c];
    const int         p = k - C.voff;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < C.cn; i += gridDim.x * blockDim.x)
        a.lb[(long long)q * a.ldl + C.lboff + i] = C.S[(long long)i * C.cn + p];
}
void launch_lines_from_units(const SchurLinesArgs& a, int k0, int cnt, hipStream_t s)
{
    hipLaunchKernelGGL(k_lines_from_units, dim3(4, cnt), dim3(256), 0, s, a, k0, cnt);
    FP_HIP(hipGetLastError());
}

// Gauss-Jordan inversion in place with partial pivoting (columns unscrambled at
// the end).  Step k: pivot search, row swap and scaling of row k by one block,
// then the rank-1 elimination of every other row.
__global__ __launch_bounds__(1024) void k_gj_pivot(double* A, int n, int k, int* perm, double* fcol, int* bad)
{
    __shared__ double bv[1024];
    __shared__ int    bi[1024];
    double            best = -1.0;
    int               arg  = k;
    for (int i = k + threadIdx.x; i < n; i += 1024) {
        const double v = fabs(A[(long long)i * n + k]);
        if (v > best) {
            best = v;
            arg  = i;
        }
    }
    bv[threadIdx.x] = best;
    bi[threadIdx.x] = arg;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            const double o = bv[threadIdx.x + w];
            const int    oi = bi[threadIdx.x + w];
            if (o > bv[threadIdx.x] || (o == bv[threadIdx.x] && oi < bi[threadIdx.x])) {
                bv[threadIdx.x] = o;
                bi[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    const int p = bi[0];
    if (threadIdx.x == 0) {
        perm[k] = p;
        if (!(bv[0] > 0.0) || !isfinite(bv[0])) *bad = 1;
    }
    double* rk = A + (long long)k * n;
    double* rp = A + (long long)p * n;
    if (p != k)
        for (int j = threadIdx.x; j < n; j += 1024) {
            const double t = rk[j];
            rk[j]          = rp[j];
            rp[j]          = t;
        }
    __syncthreads();
    const double piv = rk[k];
    __syncthreads();
    const double inv = 1.0 / piv;
    for (int j = threadIdx.x; j < n; j += 1024) rk[j] = (j == k ? 1.0 : rk[j]) * inv;
    for (int i = threadIdx.x; i < n; i += 1024) fcol[i] = (i == k) ? 0.0 : A[(long long)i * n + k];
}

__global__ void k_gj_eliminate(double* A, int n, int k, const double* fcol)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= n || i == k) return;
    const double f  = fcol[i];
    const double rk = A[(long long)k * n + j];
    double*      a  = A + (long long)i * n + j;
    *a              = (j == k ? 0.0 : *a) - f * rk;
}

__global__ void k_gj_unscramble(double* A, int n, const int* perm)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double* row = A + (long long)i * n;
    for (int k = n - 1; k >= 0; k--) {
        const int p = perm[k];
        if (p != k) {
            const double t = row[k];
            row[k]         = row[p];
            row[p]         = t;
        }
    }
}

void gauss_jordan_invert(double* A, int n, hipStream_t s)
{
    DevBuf<int>    perm(n), bad(1);
    DevBuf<double> fcol(n);
    FP_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
    for (int k = 0; k < n; k++) {
        hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(1024), 0, s, A, n, k, perm.p, fcol.p, bad.p);
        hipLaunchKernelGGL(k_gj_eliminate, dim3((n + 255) / 256, n), dim3(256), 0, s, A, n, k, fcol.p);
    }
    FP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_gj_unscramble, dim3((n + 255) / 256), dim3(256), 0, s, A, n, perm.p);
    FP_HIP(hipGetLastError());
    int h_bad = 0;
    FP_HIP(hipMemcpyAsync(&h_bad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    FP_HIP(hipStreamSynchronize(s));
    if (h_bad) throw Error(FP_E_NUMERIC, "interface operator is singular");
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
// twice (the reference uses modified Gram-Schmidt; both give the same
// Hessenberg matrix in exact arithmetic, and CGS2 needs one reduction per
// pass instead of one per basis vector).  All reductions have a fixed order
// (no atomics), so results are reproducible run to run.
//
// Element kernels use one 64-lane wavefront per block and g.nblk = ceil(K/64)
// blocks per system; a dot product is one block per basis vector.

constexpr int GB = 64;
constexpr int GU = 4;  // wavefronts of the basis-vector updates (k_gm_update, k_gm_update_x)
static_assert(GU == 4, "the updates add their four partial sums pairwise");

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// fixed-order sum of n values by one 64-lane block
__device__ double sum_fixed64(const double* p, int n)
{
    double acc = 0.0;
    for (int q = threadIdx.x; q < n; q += GB) acc += p[q];
    return wave_sum(acc);
}

// tolabs[r] = tol * ||C_r|| (relative) or tol (absolute)
__global__ __launch_bounds__(GB) void k_gm_norm_part(GmresDev g)
{
    const int       r   = blockIdx.y;
    const long long K   = g.K;
    const long long k   = blockIdx.x * (long long)GB + threadIdx.x;
    const double    c   = k < K ? g.C[r * K + k] : 0.0;
    const double    acc = wave_sum(c * c);
    if (threadIdx.x == 0) g.part[(long long)r * g.nblk + blockIdx.x] = acc;
}
__global__ __launch_bounds__(GB) void k_gm_set_tol(GmresDev g, double tol, int rel)
{
    const int    r  = blockIdx.x;
    const double nc = sqrt(sum_fixed64(g.part + (long long)r * g.nblk, g.nblk));
    if (threadIdx.x == 0) {
        g.tolabs[r]        = rel ? tol * nc : tol;
        g.resid[r]         = nc;
        g.rnorm[r]         = nc;
        g.state[r * 4 + 0] = 0;
        g.state[r * 4 + 1] = 0;
        g.state[r * 4 + 2] = 0;
        g.state[r * 4 + 3] = 0;
    }
}

// R = C - BX into V0 and partial norms
__global__ __launch_bounds__(GB) void k_gm_resid(GmresDev g)
{
    const int       r = blockIdx.y;
    const long long K = g.K;
    const long long k = blockIdx.x * (long long)GB + threadIdx.x;
    double          t = 0.0;
    if (k < K) {
        t                       = g.C[r * K + k] - g.Wv[r * K + k];
        g.V[(long long)r * K + k] = t;
    }
    const double acc = wave_sum(t * t);
    if (threadIdx.x == 0) g.part[(long long)r * g.nblk + blockIdx.x] = acc;
}

// beta = ||R||, V0 = R / beta; convergence test at cycle start
// (gmres_solver2d.cpp:148-168)
__global__ __launch_bounds__(GB) void k_gm_cycle_start(GmresDev g)
{
    const int       r    = blockIdx.y;
    const double    beta = sqrt(sum_fixed64(g.part + (long long)r * g.nblk, g.nblk));
    const long long K    = g.K;
    if (!isfinite(beta)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            g.state[r * 4 + 0] = 0;
            g.state[r * 4 + 2] = 2;  // numeric failure
            g.resid[r]         = beta;
        }
        return;
    }
    const bool conv = beta < g.tolabs[r];
    const long long k = blockIdx.x * (long long)GB + threadIdx.x;
    if (!conv && k < K) g.V[(long long)r * K + k] *= (1.0 / beta);
    if (blockIdx.x == 0) {
        const int m = g.m;
        for (int q = threadIdx.x; q <= m; q += GB) g.g[(long long)r * (m + 1) + q] = q == 0 ? beta : 0.0;
        if (threadIdx.x == 0) {
            g.resid[r]         = beta;
            g.state[r * 4 + 0] = conv ? 0 : 1;  // active in this cycle
            g.state[r * 4 + 1] = 0;             // valid columns
            if (conv) g.state[r * 4 + 2] = 1;   // converged
        }
    }
}

// Hd[r][i] = V_i . w: one block per (basis vector, segment of nseg), two
// independent sums per thread; the segment sums go to part2 and are added
// in segment order by the update
__global__ __launch_bounds__(256) void k_gm_dots(GmresDev g, int j, int nseg)
{
    __shared__ double sh[4];
    const int r = blockIdx.y, i = blockIdx.x, sg = blockIdx.z;
    if (!g.state[r * 4 + 0]) return;
    const long long K   = g.K;
    const long long len = ((K + nseg - 1) / nseg + 255) / 256 * 256;
    const long long k0 = sg * len, k1 = min(K, k0 + len);
    const double*   vi = g.V + (long long)i * g.nrhs * K + (long long)r * K;
    const double*   w  = g.Wv + (long long)r * K;
    double          a0 = 0.0, a1 = 0.0;
    long long       k  = k0 + threadIdx.x;
    for (; k + 256 < k1; k += 512) {
        a0 += vi[k] * w[k];
        a1 += vi[k + 256] * w[k + 256];
    }
    if (k < k1) a0 += vi[k] * w[k];
    const double acc = wave_sum(a0 + a1);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0)
        g.part2[((long long)r * (g.m + 1) + i) * kDotSeg + sg] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// w -= sum_i h_i V_i; Hacc (= h on pass 0, += h on pass 1); partial norms.
// GB entries of w per workgroup, the basis vectors dealt round-robin to its
// GU wavefronts (a fixed order: wavefront q sums i = q, q + GU, ...; the GU
// partial sums are added pairwise), so each wavefront's dependent chain of
// loads is a GU-th of j + 1.
__global__ __launch_bounds__(GB * GU) void k_gm_update(GmresDev g, int j, int pass, int nseg)
{
    __shared__ double hs[1024];
    __shared__ double red[GU - 1][GB];
    const int r = blockIdx.y;
    if (!g.state[r * 4 + 0]) return;
    const int m = g.m;
    for (int i = threadIdx.x; i <= j; i += GB * GU) {
        const double* ps = g.part2 + ((long long)r * (m + 1) + i) * kDotSeg;
        double        h  = ps[0];
        for (int sg = 1; sg < nseg; sg++) h += ps[sg];
        hs[i] = h;
    }
    __syncthreads();
    if (blockIdx.x == 0) {
        double* H = g.Hacc + (long long)r * (m + 1);
        for (int i = threadIdx.x; i <= j; i += GB * GU) H[i] = (pass == 0 ? 0.0 : H[i]) + hs[i];
    }
    const int       lane = threadIdx.x % GB, wv = threadIdx.x / GB;
    const long long K    = g.K;
    const long long ldv  = (long long)g.nrhs * K;
    const long long k    = blockIdx.x * (long long)GB + lane;
    double          acc  = 0.0;
    if (k < K) {
        const double* vk = g.V + (long long)r * K + k;
        for (int i = wv; i <= j; i += GU) acc += hs[i] * vk[i * ldv];
    }
    if (wv > 0) red[wv - 1][lane] = acc;
    __syncthreads();
    if (wv != 0) return;
    double w = 0.0;
    if (k < K) {
        w               = g.Wv[r * K + k] - ((acc + red[0][lane]) + (red[1][lane] + red[2][lane]));
        g.Wv[r * K + k] = w;
    }
    const double a = wave_sum(w * w);
    if (lane == 0) g.part[(long long)r * g.nblk + blockIdx.x] = a;
}

// h_{j+1,j} = ||w||; breakdown check; V_{j+1}; Givens update of column j and
// of g; convergence estimate |g_{j+1}| < tol (gmres_solver2d.cpp:189-225)
__global__ __launch_bounds__(GB) void k_gm_finish(GmresDev g, int j)
{
    __shared__ double col[1026];
    __shared__ double scs[1024], ssn[1024];
    const int r = blockIdx.y;
    if (!g.state[r * 4 + 0]) return;
    const int       m   = g.m;
    const double    hn  = sqrt(sum_fixed64(g.part + (long long)r * g.nblk, g.nblk));
    const long long K   = g.K;
    const long long ldv = (long long)g.nrhs * K;
    if (!isfinite(hn)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            g.state[r * 4 + 0] = 0;
            g.state[r * 4 + 2] = 2;
        }
        return;
    }
    // h_{j+1,j} < 1e-12 (the reference's breakdown threshold): the Krylov
    // space is invariant.  Column j is kept (the reference drops it, which
    // stalls when the operator is close to the identity, e.g. preconditioned)
    // and the cycle ends with a zero least-squares residual.
    const bool      breakdown = hn < 1e-12;
    const long long k         = blockIdx.x * (long long)GB + threadIdx.x;
    if (!breakdown && k < K) g.V[(j + 1) * ldv + (long long)r * K + k] = g.Wv[r * K + k] * (1.0 / hn);
    if (blockIdx.x != 0) return;
    double*       H  = g.H + (long long)r * (m + 1) * m;
    double*       cs = g.cs + (long long)r * m;
    double*       sn = g.sn + (long long)r * m;
    double*       gg = g.g + (long long)r * (m + 1);
    const double* ha = g.Hacc + (long long)r * (m + 1);
    for (int i = threadIdx.x; i <= j; i += GB) col[i] = ha[i];
    for (int i = threadIdx.x; i < j; i += GB) {
        scs[i] = cs[i];
        ssn[i] = sn[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        col[j + 1] = hn;
        for (int i = 0; i < j; i++) {
            const double t = scs[i] * col[i] + ssn[i] * col[i + 1];
            col[i + 1]     = -ssn[i] * col[i] + scs[i] * col[i + 1];
            col[i]         = t;
        }
        const double hjj = col[j], hj1 = col[j + 1];
        const double den = hypot(hjj, hj1);
        const double c = hjj / den, sg = hj1 / den;
        cs[j]          = c;
        sn[j]          = sg;
        col[j]         = c * hjj + sg * hj1;
        col[j + 1]     = 0.0;
        const double t = c * gg[j];
        gg[j + 1]      = -sg * gg[j];
        gg[j]          = t;
        g.state[r * 4 + 1] = j + 1;
        g.resid[r]         = fabs(gg[j + 1]);
        if (breakdown || fabs(gg[j + 1]) < g.tolabs[r]) g.state[r * 4 + 0] = 0;
        g.state[r * 4 + 3] += 1;  // Arnoldi steps taken
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= j + 1; i += GB) H[i * m + j] = col[i];
}

// back substitution for y (gmres_solver2d.cpp:228-236): blocks of 32 columns
// from the right, the 32x32 diagonal block solved in LDS, the columns above
// updated in parallel.  One block of 256 threads per system.
__global__ __launch_bounds__(256) void k_gm_backsub(GmresDev g)
{
    __shared__ double y[1024];
    __shared__ double D[32][33];
    const int     r  = blockIdx.x;
    const int     m  = g.m;
    const int     jr = g.state[r * 4 + 1];
    const double* H  = g.H + (long long)r * (m + 1) * m;
    const double* gg = g.g + (long long)r * (m + 1);
    for (int i = threadIdx.x; i < m; i += 256) y[i] = i < jr ? gg[i] : 0.0;
    __syncthreads();
    for (int c1 = jr; c1 > 0; c1 -= 32) {
        const int c0 = c1 > 32 ? c1 - 32 : 0;
        const int nb = c1 - c0;
        for (int q = threadIdx.x; q < nb * nb; q += 256) {
            const int a = q / nb, b = q % nb;
            D[a][b] = H[(long long)(c0 + a) * m + c0 + b];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int a = nb - 1; a >= 0; a--) {
                double acc = y[c0 + a];
                for (int b = a + 1; b < nb; b++) acc -= D[a][b] * y[c0 + b];
                y[c0 + a] = acc / D[a][a];
            }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < c0; t += 256) {
            double acc = 0.0;
            for (int b = 0; b < nb; b++) acc += H[(long long)t * m + c0 + b] * y[c0 + b];
            y[t] -= acc;
        }
        __syncthreads();
    }
    double* yo = g.Y + (long long)r * m;
    for (int i = threadIdx.x; i < m; i += 256) yo[i] = y[i];
}

// x += V(:, 0:jr) y (gmres_solver2d.cpp:238-241), the basis vectors dealt
// to the workgroup's wavefronts as in k_gm_update
__global__ __launch_bounds__(GB * GU) void k_gm_update_x(GmresDev g)
{
    __shared__ double ys[1024];
    __shared__ double red[GU - 1][GB];
    const int r  = blockIdx.y;
    const int jr = g.state[r * 4 + 1];
    for (int i = threadIdx.x; i < jr; i += GB * GU) ys[i] = g.Y[(long long)r * g.m + i];
    __syncthreads();
    const int       lane = threadIdx.x % GB, wv = threadIdx.x / GB;
    const long long K    = g.K;
    const long long ldv  = (long long)g.nrhs * K;
    const long long k    = blockIdx.x * (long long)GB + lane;
    double          acc  = 0.0;
    if (k < K) {
        const double* vk = g.V + (long long)r * K + k;
        for (int i = wv; i < jr; i += GU) acc += ys[i] * vk[i * ldv];
    }
    if (wv > 0) red[wv - 1][lane] = acc;
    __syncthreads();
    if (wv != 0 || k >= K) return;
    g.X[r * K + k] += (acc + red[0][lane]) + (red[1][lane] + red[2][lane]);
}

// state, residuals and initial norms to host-mapped memory (read by the host
// after a stream synchronisation; no copy commands)
__global__ void k_gm_publish(GmresDev g, int* hstate, double* hres)
{
    for (int q = threadIdx.x; q < 4 * g.nrhs; q += blockDim.x) hstate[q] = g.state[q];
    for (int q = threadIdx.x; q < g.nrhs; q += blockDim.x) {
        hres[q]          = g.resid[q];
        hres[g.nrhs + q] = g.rnorm[q];
    }
}
void gm_publish(const GmresDev& g, int* hstate, double* hres, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_publish, dim3(1), dim3(256), 0, s, g, hstate, hres);
    FP_HIP(hipGetLastError());
}

// One-workgroup forms of the cycle set-up for interface vectors that one
// workgroup streams quickly (the usual case): a fixed-order block reduction
// replaces the partial-sum kernel pair, and the cycle start also publishes
// the state to host-mapped memory.
constexpr int GW = 1024;

__device__ double block_sum_fixed(double v, double* red)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    v = wave_sum(v);
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += red[q];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(GW) void k_gm_init1(GmresDev g, double tol, int rel)
{
    __shared__ double red[GW / 64];
    const int       r   = blockIdx.x;
    const long long K   = g.K;
    double          acc = 0.0;
    for (long long k = threadIdx.x; k < K; k += GW) {
        const double c = g.C[r * K + k];
        acc += c * c;
    }
    const double nc = sqrt(block_sum_fixed(acc, red));
    if (threadIdx.x == 0) {
        g.tolabs[r]        = rel ? tol * nc : tol;
        g.resid[r]         = nc;
        g.rnorm[r]         = nc;
        g.state[r * 4 + 0] = 0;
        g.state[r * 4 + 1] = 0;
        g.state[r * 4 + 2] = 0;
        g.state[r * 4 + 3] = 0;
    }
}

// R = C - BX, beta = ||R||, V0 = R / beta, convergence test at cycle start
// (gmres_solver2d.cpp:148-168).  FIRST: the first check of a solve, which
// also sets the tolerance from ||C|| and clears the state (k_gm_init1's work,
// in the same pass over C).
template <bool FIRST>
__global__ __launch_bounds__(GW) void k_gm_start1(GmresDev g, double tol, int rel)
{
    __shared__ double red[GW / 64];
    const int       r   = blockIdx.x;
    const long long K   = g.K;
    double          acc = 0.0, accc = 0.0;
    for (long long k = threadIdx.x; k < K; k += GW) {
        const double c = g.C[r * K + k];
        const double t = c - g.Wv[r * K + k];
        g.V[r * K + k] = t;
        acc += t * t;
        if (FIRST) accc += c * c;
    }
    const double beta = sqrt(block_sum_fixed(acc, red));
    double       tola;
    if constexpr (FIRST) {
        const double nc = sqrt(block_sum_fixed(accc, red));
        tola            = rel ? tol * nc : tol;
        if (threadIdx.x == 0) {
            g.tolabs[r]        = tola;
            g.rnorm[r]         = nc;
            g.state[r * 4 + 2] = 0;
            g.state[r * 4 + 3] = 0;
        }
    } else {
        tola = g.tolabs[r];
    }
    if (!isfinite(beta)) {
        if (threadIdx.x == 0) {
            g.state[r * 4 + 0] = 0;
            g.state[r * 4 + 1] = 0;
            g.state[r * 4 + 2] = 2;  // numeric failure
            g.resid[r]         = beta;
        }
    } else {
        const bool conv = beta < tola;
        if (!conv)
            for (long long k = threadIdx.x; k < K; k += GW) g.V[r * K + k] *= (1.0 / beta);
        const int m = g.m;
        for (int q = threadIdx.x; q <= m; q += GW) g.g[(long long)r * (m + 1) + q] = q == 0 ? beta : 0.0;
        if (threadIdx.x == 0) {
            g.resid[r]         = beta;
            g.state[r * 4 + 0] = conv ? 0 : 1;  // active in this cycle
            g.state[r * 4 + 1] = 0;             // valid columns
            if (conv) g.state[r * 4 + 2] = 1;   // converged
        }
    }
    if (threadIdx.x == 0 && g.hstate) {
        for (int q = 0; q < 4; q++) g.hstate[r * 4 + q] = g.state[r * 4 + q];
        g.hres[r]          = g.resid[r];
        g.hres[g.nrhs + r] = g.rnorm[r];
    }
}

bool gm_first(const GmresDev& g, double tol, int rel, hipStream_t s)
{
    if (g.K > (1 << 20)) return false;
    hipLaunchKernelGGL(k_gm_start1<true>, dim3(g.nrhs), dim3(GW), 0, s, g, tol, rel);
    FP_HIP(hipGetLastError());
    return true;
}

void gm_init(const GmresDev& g, double tol, int rel, hipStream_t s)
{
    if (g.K <= (1 << 20)) {
        hipLaunchKernelGGL(k_gm_init1, dim3(g.nrhs), dim3(GW), 0, s, g, tol, rel);
        FP_HIP(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(k_gm_norm_part, dim3(g.nblk, g.nrhs), dim3(GB), 0, s, g);
    FP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_gm_set_tol, dim3(g.nrhs), dim3(GB), 0, s, g, tol, rel);
    FP_HIP(hipGetLastError());
}

// returns true when the state was also published to host-mapped memory
bool gm_resid(const GmresDev& g, hipStream_t s)
{
    if (g.K <= (1 << 20)) {
        hipLaunchKernelGGL(k_gm_start1<false>, dim3(g.nrhs), dim3(GW), 0, s, g, 0.0, 0);
        FP_HIP(hipGetLastError());
        return g.hstate != nullptr;
    }
    hipLaunchKernelGGL(k_gm_resid, dim3(g.nblk, g.nrhs), dim3(GB), 0, s, g);
    FP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_gm_cycle_start, dim3(g.nblk, g.nrhs), dim3(GB), 0, s, g);
    FP_HIP(hipGetLastError());
    return false;
}

void gm_arnoldi(const GmresDev& g, int j, hipStream_t s)
{
    const int nseg = gm_nseg(g.K);
    for (int pass = 0; pass < 2; pass++) {
        hipLaunchKernelGGL(k_gm_dots, dim3(j + 1, g.nrhs, nseg), dim3(256), 0, s, g, j, nseg);
        FP_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_gm_update, dim3(g.nblk, g.nrhs), dim3(GB * GU), 0, s, g, j, pass, nseg);
        FP_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_gm_finish, dim3(g.nblk, g.nrhs), dim3(GB), 0, s, g, j);
    FP_HIP(hipGetLastError());
}

void gm_finish_cycle(const GmresDev& g, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_backsub, dim3(g.nrhs), dim3(256), 0, s, g);
    FP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_gm_update_x, dim3(g.nblk, g.nrhs), dim3(GB * GU), 0, s, g);
    FP_HIP(hipGetLastError());
}

}  // namespace fp
