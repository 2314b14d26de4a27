// Tridiagonal solves along x, one system per y mode, for the subdomain solver.
//
// Reference semantics: ChasingMethod2D::chasing over the transposed buffer
// with chasing_standard_without_precompute per mode and the singular
// recurrence for the last mode of a domain with no Dirichlet-like face
// (pe/poisson/chasing_method2d.cpp:67-88, chasing_method_base.cpp:24-53,146-154).
//
// Layout: the transformed field stays row-major [i][j]; lane = mode j, so each
// row access is coalesced and the reference's two transposes
// (poisson_solver2d.cpp:74,78) disappear.
//
// Parallel Thomas: the LU factor of tridiag(1, b_i, 1) (b_i = the mode's
// diagonal d_j, plus 1 on a Neumann end row) has the reciprocal pivots
//   p_0 = 1 / b_0,   p_i = 1 / (b_i - p_{i-1})       (the reference's c_i).
// With d_j = -2 - delta_j the reference's form loses the small delta_j of the
// lowest modes to the rounding of d_j (about 1e-10 relative error of a
// 4096-row Dirichlet solve): here the pivots are carried as excesses
// e_i = 1 + p_i, which satisfy, with every term non-negative,
//   q_i = delta + e_{i-1},  p_i = -1 / (1 + q_i),  e_i = -q_i p_i   (interior rows)
//   p_i = -1 / (delta + e_{i-1}),  e_i = -(delta + (e_{i-1} - 1)) p_i  (Neumann rows)
// from e_{-1} = 1, and delta_j = (hx/hy)^2 4 sin^2(.) comes to full precision
// from the eigenvalue's closed form.  The same LU factor to rounding, without
// the cancellation.
// The forward sweep y_i = p_i (f_i - y_{i-1}) and the back substitution
// x_i = y_i - p_i x_{i+1} are first-order affine recurrences, so each thread
// owns a chunk of CH rows of one mode, composes its chunk's affine map, the
// chunk maps are chained in LDS, and every thread then replays its chunk from
// the correct start value.  The pivots are not read from memory: each thread
// regenerates its CH pivots from the pivot that precedes its chunk (a table
// of one value per chunk and mode, 1/CH of the field) with the same IEEE
// divisions as the host's recurrence, so they are bit-identical to it.  The
// kernel therefore moves one read and one write of the field and nothing
// else of its size.  The singular zero mode (x_0 = 0,
// x_{i+1} = f_i - x_{i-1} + 2 x_i) is two prefix sums and uses the same chain.
#include <cstdlib>
#include <cstdio>
#include "fp_internal.h"

#include <algorithm>
#include <map>
#include <mutex>

namespace fp {

// Lane permutations of the wavefront by DPP (no LDS traffic): lanes whose
// source lane does not exist, or whose row is masked off, keep `old`.
template <int CTRL, int RM>
__device__ __forceinline__ double dpp(double old, double v)
{
    const unsigned long long o = (unsigned long long)__double_as_longlong(old);
    const unsigned long long x = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)x, CTRL, RM, 0xf, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(x >> 32), CTRL,
                                                             RM, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int CTRL, int RM>
__device__ __forceinline__ float dpp(float old, float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), CTRL, RM, 0xf, false));
}

// x -> A x + B after the map of the earlier lanes (Ap, Bp)
template <typename T>
__device__ __forceinline__ void affine_after(T& A, T& B, T Ap, T Bp)
{
    B = A * Bp + B;
    A = A * Ap;
}

// Inclusive scan of affine maps over the 64 lanes, lane 0 applied first:
// row_shr 1, 2, 4, 8 within rows of 16 lanes, then row_bcast 15 and 31
// across rows (the identity map fills lanes without a source).
template <typename T>
__device__ __forceinline__ void wave_affine_scan(T& A, T& B)
{
    affine_after(A, B, dpp<0x111, 0xf>(T(1), A), dpp<0x111, 0xf>(T(0), B));
    affine_after(A, B, dpp<0x112, 0xf>(T(1), A), dpp<0x112, 0xf>(T(0), B));
    affine_after(A, B, dpp<0x114, 0xf>(T(1), A), dpp<0x114, 0xf>(T(0), B));
    affine_after(A, B, dpp<0x118, 0xf>(T(1), A), dpp<0x118, 0xf>(T(0), B));
    affine_after(A, B, dpp<0x142, 0xa>(T(1), A), dpp<0x142, 0xa>(T(0), B));
    affine_after(A, B, dpp<0x143, 0xc>(T(1), A), dpp<0x143, 0xc>(T(0), B));
}

// Start values of the chunks of each mode: the chunk maps v -> SA v + SB
// (LDS, chunk-major) are chained in application order and the exclusive
// prefix, applied to 0, is written to SY.  backward = 1 chains the chunks from
// the last row upwards (back substitution); mode sing_j always chains
// forwards (the two prefix sums of the singular recurrence).  One wavefront
// per mode: lane l composes q = ceil(P/64) consecutive chunk maps, a log-depth
// DPP scan chains the lanes, and each lane replays its q chunks.
template <typename T>
__device__ void chain_chunks(const T* SA, const T* SB, T* SY, int P, int MB, int j0, int ny, int backward,
                             int sing_j)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    const int q  = (P + 63) >> 6;
    const int p0 = lane * q, p1 = min(P, p0 + q);
    for (int jm = wave; jm < MB; jm += nwave) {
        if (j0 + jm >= ny) continue;  // uniform over the wavefront
        const bool bw = backward && (j0 + jm) != sing_j;
        T          A = T(1), B = T(0);
        for (int p = p0; p < p1; p++) {
            const int c = bw ? P - 1 - p : p;
            const T   a = SA[c * MB + jm];
            B           = a * B + SB[c * MB + jm];
            A           = a * A;
        }
        wave_affine_scan(A, B);
        T v = dpp<0x138, 0xf>(T(0), B);  // wave_shr 1: the exclusive prefix (0 on lane 0)
        for (int p = p0; p < p1; p++) {
            const int c     = bw ? P - 1 - p : p;
            SY[c * MB + jm] = v;
            v               = SA[c * MB + jm] * v + SB[c * MB + jm];
        }
    }
}

// Column partials of one chunk: over the workgroup's MB modes, the sums of
// phi0(j) U(i,j) and phi1(j) U(i,j) for the chunk's CH rows, stored to
// out[i] and out[nx + i].  The MB modes of a chunk are MB neighbouring lanes
// (a quad for MB = 4), reduced by DPP quad permutations (xor 2, then xor 1:
// the order of a shuffle tree, so every lane of the quad holds the same
// bits); rows are taken MB at a time so that lane jl stores row k0 + jl and
// one store instruction covers MB consecutive rows of every chunk.  MB = 8
// takes the 8 lanes of a half-row: the half-row mirror first (lane l and
// 7 - l), then the two quad swaps; each lane adds the same four pair sums,
// two by two, so all 8 hold the same bits.  Other MB reduce through the LDS
// crossbar and lane 0 stores.
template <typename T, int CH>
__device__ __forceinline__ void store_col_partials(const T (&fv)[CH], T p0, T p1, int MB, int jl, int c, int P,
                                                   int i0, int nx, T* out)
{
    if (MB == 4 || MB == 2) {
#pragma unroll
        for (int k0 = 0; k0 < CH; k0 += 4) {
            T a0 = T(0), a1 = T(0), b0 = T(0), b1 = T(0);  // rows k0 + jl and, for MB = 2, k0 + 2 + jl
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int k = k0 + t;
                if (k >= CH) break;
                T s0 = p0 * fv[k], s1 = p1 * fv[k];
                if (MB == 4) {
                    s0 += dpp<0x4E, 0xf>(s0, s0);  // quad_perm [2,3,0,1]: lane ^ 2
                    s1 += dpp<0x4E, 0xf>(s1, s1);
                }
                s0 += dpp<0xB1, 0xf>(s0, s0);  // quad_perm [1,0,3,2]: lane ^ 1
                s1 += dpp<0xB1, 0xf>(s1, s1);
                if (MB == 4) {
                    if (t == jl) {
                        a0 = s0;
                        a1 = s1;
                    }
                } else if ((t & 1) == jl) {
                    if (t < 2) {
                        a0 = s0;
                        a1 = s1;
                    } else {
                        b0 = s0;
                        b1 = s1;
                    }
                }
            }
            if (c < P) {
                const int ka = k0 + jl, kb = k0 + 2 + jl;
                if (ka < CH && i0 + ka < nx) {
                    out[i0 + ka]      = a0;
                    out[nx + i0 + ka] = a1;
                }
                if (MB == 2 && kb < CH && i0 + kb < nx) {
                    out[i0 + kb]      = b0;
                    out[nx + i0 + kb] = b1;
                }
            }
        }
        return;
    }
    if (CH >= 16 && MB == 8) {
        // the 8 lanes of a half-row: the mirror (lane 7 - l) pairs them, the
        // two quad swaps add the four pair sums, so every lane holds the
        // total and lane jl stores row k0 + jl (not compiled for 8-row
        // chunks, whose sweep it would take from 72 to 92 VGPRs)
#pragma unroll
        for (int k0 = 0; k0 < CH; k0 += 8) {
            T a0 = T(0), a1 = T(0);
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const int k = k0 + t;
                if (k >= CH) break;
                T s0 = p0 * fv[k], s1 = p1 * fv[k];
                s0 += dpp<0x141, 0xf>(s0, s0);  // row_half_mirror: lane 7 - l
                s1 += dpp<0x141, 0xf>(s1, s1);
                s0 += dpp<0x4E, 0xf>(s0, s0);   // lane ^ 2
                s1 += dpp<0x4E, 0xf>(s1, s1);
                s0 += dpp<0xB1, 0xf>(s0, s0);   // lane ^ 1
                s1 += dpp<0xB1, 0xf>(s1, s1);
                if (t == jl) {
                    a0 = s0;
                    a1 = s1;
                }
            }
            const int ka = k0 + jl;
            if (c < P && ka < CH && i0 + ka < nx) {
                out[i0 + ka]      = a0;
                out[nx + i0 + ka] = a1;
            }
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < CH; k++) {
        T s0 = p0 * fv[k], s1 = p1 * fv[k];
        for (int off = MB / 2; off > 0; off >>= 1) {
            s0 += __shfl_xor(s0, off, 64);
            s1 += __shfl_xor(s1, off, 64);
        }
        const int i = i0 + k;
        if (jl == 0 && c < P && i < nx) {
            out[i]      = s0;
            out[nx + i] = s1;
        }
    }
}

// The dispatcher deals workgroups to the 8 XCDs round robin.  Renumbering the
// grid so that consecutive logical blocks (neighbouring modes, which share
// cache lines of every row) run on one XCD keeps those lines in one L2.
__device__ __forceinline__ int xcd_block(int b, int nblocks)
{
    if (nblocks < 64 || (nblocks & 7) != 0) return b;
    return (b & 7) * (nblocks >> 3) + (b >> 3);
}

// One step of the excess recurrence (the host's, bit for bit): the
// reciprocal pivot p of the row and the excess e = 1 + p it passes on.
__host__ __device__ __forceinline__ void pivot_step(double delta, bool neumann_row, double& e, double& p)
{
    if (!neumann_row) {
        const double q = delta + e;
        p              = -1.0 / (1.0 + q);
        e              = -q * p;
    } else {
        p = -1.0 / (delta + e);
        e = -(delta + (e - 1.0)) * p;
    }
}

// The CH reciprocal pivots of rows i0 .. i0+CH-1 of one mode from the pivot
// excess before the chunk (1 before row 0), by the host's recurrence.
template <typename T, int CH>
__device__ __forceinline__ void chunk_pivots(T (&pv)[CH], double delta, double e, int i0, int nx, int ns, int ne,
                                             bool act)
{
    double p = 0.0;
#pragma unroll
    for (int k = 0; k < CH; k++) {
        const int i = i0 + k;
        if (act && i < nx) pivot_step(delta, (i == 0 && ns) || (i == nx - 1 && ne), e, p);
        pv[k] = (T)p;
    }
}

// One subdomain's sweep in a batched launch.
template <typename T>
struct TriProb {
    T*            data;
    long long     ld;
    const double* xd;   // delta (ny)
    const double* pst;  // chunk-start pivot excesses, P per mode
    int           nx, ny, P, MB, singular_last, nblk, ns, ne;  // nblk: mode blocks per right-hand side
    // periodic x (cyc): z1 = A_D^-1 e_0 and z2 = A_D^-1 e_{nx-1} per mode, laid
    // out like the field (A_D: the Dirichlet-end tridiagonal of the mode)
    int           cyc;
    const double* z1;
    const double* z2;
};
template <typename T>
struct TriBatch {
    int        nprob;
    int        start[kMaxBatch + 1];
    TriProb<T> p[kMaxBatch];
};

template <int CH>
constexpr int tri_threads()
{
    return CH >= 32 ? 512 : 1024;
}

template <typename T>
__device__ __forceinline__ const TriProb<T>& tri_find(const TriBatch<T>& bt, int blk, int& local)
{
    int q = 0;
    while (q + 1 < bt.nprob && blk >= bt.start[q + 1]) q++;
    local = blk - bt.start[q];
    return bt.p[q];
}
template <typename T>
__device__ __forceinline__ const TriProb<T>& tri_find(const JobTable<TriProb<T>>& bt, int blk, int& local)
{
    return table_find(bt, blk, local);
}

template <typename T, int CH, typename BT>
__global__ __launch_bounds__(tri_threads<CH>()) void k_tridiag(BT bt)
{
    const int         blk = xcd_block((int)blockIdx.x, (int)gridDim.x);
    int               local;
    const TriProb<T>& pr = tri_find<T>(bt, blk, local);
    const int nx = pr.nx, ny = pr.ny, P = pr.P, MB = pr.MB, singular_last = pr.singular_last;
    const int rhs = local / pr.nblk, mblk = local - rhs * pr.nblk;

    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* SA = reinterpret_cast<T*>(smem_raw);
    T* SB = SA + P * MB;
    T* SY = SB + P * MB;

    const int  tid = threadIdx.x;
    const int  jl  = tid % MB;
    const int  c   = tid / MB;
    const int  j   = mblk * MB + jl;
    const bool act = (c < P) && (j < ny);
    T*         f   = pr.data + (long long)rhs * pr.ld;
    const int  i0  = c * CH;

    T fv[CH], pv[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) {
        const int  i  = i0 + k;
        const bool ok = act && i < nx;
        fv[k]         = ok ? f[(long long)i * ny + j] : T(0);
    }
    const bool sing  = singular_last && (j == ny - 1);
    const int  SINGJ = singular_last ? ny - 1 : -1;
    if (!sing) {
        const double d  = act ? pr.xd[j] : 0.0;
        const double p0 = act ? pr.pst[(long long)j * P + c] : 0.0;
        chunk_pivots<T, CH>(pv, d, p0, i0, nx, pr.ns, pr.ne, act);
    }

    // forward: chunk map y_out = A y_in + B
    T A = T(1), B = T(0);
    if (!sing) {
#pragma unroll
        for (int k = 0; k < CH; k++)
            if (i0 + k < nx) {
                A = -pv[k] * A;
                B = pv[k] * (fv[k] - B);
            }
    } else {
#pragma unroll
        for (int k = 0; k < CH; k++)
            if (i0 + k < nx) B += fv[k];
    }
    if (c < P) {
        SA[c * MB + jl] = A;
        SB[c * MB + jl] = B;
    }
    __syncthreads();
    chain_chunks(SA, SB, SY, P, MB, mblk * MB, ny, 0, SINGJ);
    __syncthreads();
    T y = (c < P) ? SY[c * MB + jl] : T(0);
    if (!sing) {
#pragma unroll
        for (int k = 0; k < CH; k++)
            if (i0 + k < nx) {
                y     = pv[k] * (fv[k] - y);
                fv[k] = y;
            }
    } else {
#pragma unroll
        for (int k = 0; k < CH; k++)
            if (i0 + k < nx) {
                y     = y + fv[k];  // d_i = d_{i-1} + f_i
                fv[k] = y;
            }
    }
    __syncthreads();

    // backward: chunk map x_{i0} = A x_{i0+CH} + B  (x_nx = 0)
    A = T(1);
    B = T(0);
    if (!sing) {
#pragma unroll
        for (int k = CH - 1; k >= 0; k--)
            if (i0 + k < nx) {
                B = fv[k] - pv[k] * B;
                A = -pv[k] * A;
            }
    } else {
#pragma unroll
        for (int k = 0; k < CH; k++)
            if (i0 + k < nx) B += fv[k];
    }
    if (c < P) {
        SA[c * MB + jl] = A;
        SB[c * MB + jl] = B;
    }
    __syncthreads();
    chain_chunks(SA, SB, SY, P, MB, mblk * MB, ny, 1, SINGJ);
    __syncthreads();
    T x = (c < P) ? SY[c * MB + jl] : T(0);
    if (!sing) {
#pragma unroll
        for (int k = CH - 1; k >= 0; k--)
            if (i0 + k < nx) {
                x     = fv[k] - pv[k] * x;
                fv[k] = x;
            }
    } else {
#pragma unroll
        for (int k = 0; k < CH; k++)
            if (i0 + k < nx) {
                const T d = fv[k];
                fv[k]     = x;  // x_i = sum_{q<i} d_q
                x += d;
            }
    }
    if (pr.cyc) {
        // periodic x (ChasingMethodBase::chasing_periodic_with_precompute /
        // chasing_periodic_singular, chasing_method_base.cpp:178-285): the
        // sweep above solved the Dirichlet-end system; the mode's first and
        // last values meet through the workgroup's LDS (every chain read of
        // SA / SB ended at the barrier above) and the Sherman-Morrison
        // correction x = y - g1 z1 - g2 z2 follows with the reference's gamma
        // formulas.  The singular mode ran the Neumann-singular recurrence
        // x^N; the periodic one with the same gauge is x^N_i - (i / nx) x^N_{nx-1}.
        if (act && c == 0) SA[jl] = fv[0];
#pragma unroll
        for (int k = 0; k < CH; k++)
            if (act && i0 + k == nx - 1) SB[jl] = fv[k];
        __syncthreads();
        if (act) {
            const T y0 = SA[jl], y1 = SB[jl];
            if (sing) {
#pragma unroll
                for (int k = 0; k < CH; k++)
                    if (i0 + k < nx) fv[k] -= (T)(i0 + k) / (T)nx * y1;
            } else {
                // the 2 x 2 system nearly cancels for the lowest modes: FP64
                const double a0 = pr.z1[j], a1 = pr.z1[(long long)(nx - 1) * ny + j];
                const double b0 = pr.z2[j], b1 = pr.z2[(long long)(nx - 1) * ny + j];
                const double w0 = (double)y0, w1 = (double)y1;
                const double h2 = (w1 / (1.0 + a1) - w0 / a0) / (b1 / (1.0 + a1) - (1.0 + b0) / a0);
                const T      g2 = (T)h2;
                const T      g1 = (T)(w0 / a0 - h2 * ((1.0 + b0) / a0));
#pragma unroll
                for (int k = 0; k < CH; k++) {
                    const long long q = (long long)(i0 + k) * ny + j;
                    if (i0 + k < nx) fv[k] -= g1 * (T)pr.z1[q] + g2 * (T)pr.z2[q];
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < CH; k++) {
        const int i = i0 + k;
        if (act && i < nx) f[(long long)i * ny + j] = fv[k];
    }
}

// Line-sourced variant for the interface operator R A^-1 E of a root whose
// right-hand side is zero except on its interface lines.  The y transform of
// such a field is rank-limited: row i transforms to
//   [i == 0] Xh0(k) + [i == nx-1] Xh1(k) + phi0(k) a_i + phi1(k) b_i
// where Xh0/Xh1 are the transformed x-lines (rows 0 and nx-1), a/b the y-line
// data (columns 0 and ny-1) and phi0/phi1 the transform of the unit vectors
// e_0/e_{ny-1}.  The right-hand side is generated in registers, the sweep is
// the same chained-chunk Thomas, and only what the restriction needs leaves
// the kernel: rows 0 and nx-1 of the solution's coefficients, and the column
// sums sum_k phi0(k) u(i,k), sum_k phi1(k) u(i,k) (= u(i,0), u(i,ny-1) after
// the inverse transform), reduced over this workgroup's modes.
template <typename T, int CH>
__global__ __launch_bounds__(tri_threads<CH>()) void k_tridiag_lines(LinesArgs<T> a, int P, int MB)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* SA = reinterpret_cast<T*>(smem_raw);
    T* SB = SA + P * MB;
    T* SY = SB + P * MB;

    const int  nx = a.nx, ny = a.ny;
    const int  tid = threadIdx.x;
    const int  jl  = tid % MB;
    const int  c   = tid / MB;
    const int  j   = blockIdx.x * MB + jl;
    const bool act = (c < P) && (j < ny);
    const int  r   = blockIdx.y;
    const T*   lb  = a.lines + (long long)r * a.ldl;
    const T*   xh0 = lb;
    const T*   xh1 = lb + ny;
    const T*   av  = lb + 2 * ny;
    const T*   bv  = av + nx;
    const T    p0  = act ? a.phi[j] : T(0);
    const T    p1  = act ? a.phi[ny + j] : T(0);
    const int  i0  = c * CH;

    T fv[CH], pv[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) {
        const int  i  = i0 + k;
        const bool ok = act && i < nx;
        T          v  = T(0);
        if (ok) {
            v = p0 * av[i] + p1 * bv[i];
            if (i == 0) v += xh0[j];
            if (i == nx - 1) v += xh1[j];
        }
        fv[k] = v;
    }
    {
        const double d   = act ? a.xd[j] : 0.0;
        const double pp0 = act ? a.pst[(long long)j * P + c] : 0.0;
        chunk_pivots<T, CH>(pv, d, pp0, i0, nx, a.ns, a.ne, act);
    }
    T A = T(1), B = T(0);
#pragma unroll
    for (int k = 0; k < CH; k++)
        if (i0 + k < nx) {
            A = -pv[k] * A;
            B = pv[k] * (fv[k] - B);
        }
    if (c < P) {
        SA[c * MB + jl] = A;
        SB[c * MB + jl] = B;
    }
    __syncthreads();
    chain_chunks(SA, SB, SY, P, MB, blockIdx.x * MB, ny, 0, -1);
    __syncthreads();
    T y = (c < P) ? SY[c * MB + jl] : T(0);
#pragma unroll
    for (int k = 0; k < CH; k++)
        if (i0 + k < nx) {
            y     = pv[k] * (fv[k] - y);
            fv[k] = y;
        }
    __syncthreads();
    A = T(1);
    B = T(0);
#pragma unroll
    for (int k = CH - 1; k >= 0; k--)
        if (i0 + k < nx) {
            B = fv[k] - pv[k] * B;
            A = -pv[k] * A;
        }
    if (c < P) {
        SA[c * MB + jl] = A;
        SB[c * MB + jl] = B;
    }
    __syncthreads();
    chain_chunks(SA, SB, SY, P, MB, blockIdx.x * MB, ny, 1, -1);
    __syncthreads();
    T x = (c < P) ? SY[c * MB + jl] : T(0);
#pragma unroll
    for (int k = CH - 1; k >= 0; k--)
        if (i0 + k < nx) {
            x     = fv[k] - pv[k] * x;
            fv[k] = x;
        }
    if (a.xrows && act) {
        if (i0 == 0) a.xu[(long long)r * 2 * ny + j] = fv[0];
#pragma unroll
        for (int k = 0; k < CH; k++)
            if (i0 + k == nx - 1) a.xu[(long long)r * 2 * ny + ny + j] = fv[k];
    }
    if (a.ycols) {
        // reduce phi0 u, phi1 u over the MB modes of this chunk (lanes c*MB .. c*MB+MB-1)
        T* out = a.colpart + ((long long)r * gridDim.x + blockIdx.x) * 2 * nx;
        store_col_partials<T, CH>(fv, p0, p1, MB, jl, c, P, i0, nx, out);
    }
}

// General sweep of the composite solve (two-level trees).  The right-hand
// side of mode j, row i is
//   fin ? F(i,j) : 0
//   + sum over the line buffers (coefficient +1 for lbp, -1 for lbm) of
//     [i == 0] X0(j) + [i == nx-1] X1(j) + phi0(j) Y0(i) + phi1(j) Y1(i)
// where X0/X1 are x-lines in this domain's y basis (transform coefficients of
// rows 0 / nx-1) and Y0/Y1 physical y-lines (columns 0 / ny-1), entering
// through phi0/phi1, the transforms of the unit vectors e_0 / e_{ny-1}.
// After the chained-chunk Thomas sweep the solution's coefficients go to any
// of: the field (fout, in place), rows 0 / nx-1 (xo0 / xo1), and the
// per-workgroup partial sums sum_j phi0(j) U(i,j), sum_j phi1(j) U(i,j) of
// this workgroup's modes (cp), whose total over the workgroups is U at
// columns 0 / ny-1 in physical space.  So a solve whose input or output is
// only a few lines never reads or writes the rest of the field.
template <typename T>
struct SweepProb {
    T*            f;
    T*            fo;
    long long     ld;
    const T*      lbp;
    const T*      lbm;
    long long     ldl;
    const T*      phi;
    T*            xo0;
    T*            xo1;
    long long     ldxo;
    T*            cp;
    const double* xd;
    const double* pst;
    int           nx, ny, P, MB, nblk, ns, ne, fin, fout;
};
template <typename T>
struct SweepBatch {
    int          nprob;
    int          start[kMaxBatch + 1];
    SweepProb<T> p[kMaxBatch];
};


template <typename T, int CH>
__global__ __launch_bounds__(tri_threads<CH>()) void k_sweep(SweepBatch<T> bt)
{
    const int blk = xcd_block((int)blockIdx.x, (int)gridDim.x);
    int       q   = 0;
    while (q + 1 < bt.nprob && blk >= bt.start[q + 1]) q++;
    const SweepProb<T>& pr = bt.p[q];
    const int nx = pr.nx, ny = pr.ny, P = pr.P, MB = pr.MB;
    const int local = blk - bt.start[q];
    const int r = local / pr.nblk, mblk = local - r * pr.nblk;

    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* SA = reinterpret_cast<T*>(smem_raw);
    T* SB = SA + P * MB;
    T* SY = SB + P * MB;

    const int  tid = threadIdx.x;
    const int  jl  = tid % MB;
    const int  c   = tid / MB;
    const int  j   = mblk * MB + jl;
    const bool act = (c < P) && (j < ny);
    const int  i0  = c * CH;
    T*         f   = pr.f ? pr.f + (long long)r * pr.ld : nullptr;
    const T*   lp  = pr.lbp ? pr.lbp + (long long)r * pr.ldl : nullptr;
    const T*   lm  = pr.lbm ? pr.lbm + (long long)r * pr.ldl : nullptr;
    const T    p0  = act ? pr.phi[j] : T(0);
    const T    p1  = act ? pr.phi[ny + j] : T(0);

    T fv[CH], pv[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) {
        const int  i  = i0 + k;
        const bool ok = act && i < nx;
        fv[k]         = (ok && pr.fin) ? f[(long long)i * ny + j] : T(0);
    }
    // line sources lb = [x0 (ny) | x1 (ny) | y0 (nx) | y1 (nx)]: the source of
    // mode j at row i is p0 y0[i] + p1 y1[i] (p = phi of the mode on the y
    // faces), plus x0[j] at row 0 and x1[j] at row nx-1.  The y-line terms
    // depend on the row only, so a thread reads its CH consecutive rows of
    // each y line in one run
    auto add_lines = [&](const T* lb, T sg) {
        const T* y0 = lb + 2 * ny + i0;
        const T* y1 = y0 + nx;
        if (act && i0 + CH <= nx) {
            T a0[CH], a1[CH];
#pragma unroll
            for (int k = 0; k < CH; k++) {
                a0[k] = y0[k];
                a1[k] = y1[k];
            }
#pragma unroll
            for (int k = 0; k < CH; k++) fv[k] += sg * (p0 * a0[k] + p1 * a1[k]);
        } else if (act) {
#pragma unroll
            for (int k = 0; k < CH; k++)
                if (i0 + k < nx) fv[k] += sg * (p0 * y0[k] + p1 * y1[k]);
        }
        if (act && i0 == 0) fv[0] += sg * lb[j];
#pragma unroll
        for (int k = 0; k < CH; k++)
            if (act && i0 + k == nx - 1) fv[k] += sg * lb[ny + j];
    };
    if (lp) add_lines(lp, T(1));
    if (lm) add_lines(lm, T(-1));
    {
        const double d   = act ? pr.xd[j] : 0.0;
        const double pp0 = act ? pr.pst[(long long)j * P + c] : 0.0;
        chunk_pivots<T, CH>(pv, d, pp0, i0, nx, pr.ns, pr.ne, act);
    }
    T A = T(1), B = T(0);
#pragma unroll
    for (int k = 0; k < CH; k++)
        if (i0 + k < nx) {
            A = -pv[k] * A;
            B = pv[k] * (fv[k] - B);
        }
    if (c < P) {
        SA[c * MB + jl] = A;
        SB[c * MB + jl] = B;
    }
    __syncthreads();
    chain_chunks(SA, SB, SY, P, MB, mblk * MB, ny, 0, -1);
    __syncthreads();
    T y = (c < P) ? SY[c * MB + jl] : T(0);
#pragma unroll
    for (int k = 0; k < CH; k++)
        if (i0 + k < nx) {
            y     = pv[k] * (fv[k] - y);
            fv[k] = y;
        }
    __syncthreads();
    A = T(1);
    B = T(0);
#pragma unroll
    for (int k = CH - 1; k >= 0; k--)
        if (i0 + k < nx) {
            B = fv[k] - pv[k] * B;
            A = -pv[k] * A;
        }
    if (c < P) {
        SA[c * MB + jl] = A;
        SB[c * MB + jl] = B;
    }
    __syncthreads();
    chain_chunks(SA, SB, SY, P, MB, mblk * MB, ny, 1, -1);
    __syncthreads();
    T x = (c < P) ? SY[c * MB + jl] : T(0);
#pragma unroll
    for (int k = CH - 1; k >= 0; k--)
        if (i0 + k < nx) {
            x     = fv[k] - pv[k] * x;
            fv[k] = x;
        }
    if (act) {
        if (pr.fout) {
            T* fo = pr.fo + (long long)r * pr.ld;
#pragma unroll
            for (int k = 0; k < CH; k++)
                if (i0 + k < nx) fo[(long long)(i0 + k) * ny + j] = fv[k];
        }
        if (pr.xo0 && i0 == 0) pr.xo0[(long long)r * pr.ldxo + j] = fv[0];
        if (pr.xo1) {
#pragma unroll
            for (int k = 0; k < CH; k++)
                if (i0 + k == nx - 1) pr.xo1[(long long)r * pr.ldxo + j] = fv[k];
        }
    }
    if (pr.cp) {
        T* out = pr.cp + ((long long)r * pr.nblk + mblk) * 2 * nx;
        store_col_partials<T, CH>(fv, p0, p1, MB, jl, c, P, i0, nx, out);
    }
}

// ---------------------------------------------------------------------------
// host side

static int tri_max_threads(int ch) { return ch >= 32 ? 512 : 1024; }

// modes per workgroup: the largest power of two up to mb_cap whose
// workgroup (MB x P threads) fits the kernel's thread limit
static int tri_mb(int P, int ch, int mb_cap)
{
    int MB = 1;
    while (MB * 2 * P <= tri_max_threads(ch) && MB < mb_cap) MB *= 2;
    return MB;
}

// modes per workgroup of a composite sweep that writes column partials: at
// least 4 (fewer partials to write and reduce, 32-byte row segments per
// load), whatever the plain sweep's tuned shape
static int sweep_mb(const SubPlan& sp)
{
    return tri_mb(sp.tri_p, sp.tri_ch, std::max(sp.mb_field, 4));
}

template <typename T, int CH>
static void launch_tridiag_group(const TridiagJob<T>* const* jobs, int n, hipStream_t s)
{
    std::vector<TriProb<T>> probs(n);
    std::vector<int>        start(n + 1);
    int                     blocks = 0, threads = 64;
    size_t                  lds = 0;
    for (int q = 0; q < n; q++) {
        const TridiagJob<T>& jb = *jobs[q];
        const SubPlan&       sp = *jb.sp;
        const bool           tb = jb.thru && sp.tri_ch_b == CH;
        const int            P  = tb ? sp.tri_p_b : sp.tri_p;
        const int            MB = tri_mb(P, CH, tb ? sp.mb_b : sp.mb_field);
        if (MB * P > tri_max_threads(CH)) throw Error(FP_E_INVALID, "nx too large for the tridiagonal kernel");
        TriProb<T>& pr   = probs[q];
        pr.data          = jb.data;
        pr.ld            = jb.ld;
        pr.xd            = sp.xd.p;
        pr.pst           = (tb && sp.pst_b.p) ? sp.pst_b.p : sp.pst.p;
        pr.nx            = sp.nx;
        pr.ny            = sp.ny;
        pr.P             = P;
        pr.MB            = MB;
        pr.singular_last = sp.no_dirichlet ? 1 : 0;
        pr.nblk          = (sp.ny + MB - 1) / MB;
        pr.ns            = sp.bc[XNEG] == NEUMANN;
        pr.ne            = sp.bc[XPOS] == NEUMANN;
        pr.cyc           = sp.cyc ? 1 : 0;
        pr.z1            = sp.cyc ? sp.zc.p : nullptr;
        pr.z2            = sp.cyc ? sp.zc.p + (size_t)sp.nx * sp.ny : nullptr;
        start[q]         = blocks;
        const long long nb = (long long)pr.nblk * jb.nrhs;
        if (blocks + nb > (1LL << 30)) throw Error(FP_E_INVALID, "too many tridiagonal workgroups in one launch");
        blocks += (int)nb;
        threads = std::max(threads, (MB * P + 63) / 64 * 64);
        lds     = std::max(lds, 3 * (size_t)P * MB * sizeof(T));
    }
    start[n] = blocks;
    if (blocks == 0) return;
    if (n <= kMaxBatch) {
        TriBatch<T> bt{};
        bt.nprob = n;
        for (int q = 0; q <= n; q++) bt.start[q] = start[q];
        for (int q = 0; q < n; q++) bt.p[q] = probs[q];
        hipLaunchKernelGGL((k_tridiag<T, CH, TriBatch<T>>), dim3(blocks), dim3(threads), lds, s, bt);
        FP_HIP(hipGetLastError());
        return;
    }
    void*                          mem = nullptr;
    const JobTable<TriProb<T>>     t   = table_upload(probs, start, s, mem);
    hipLaunchKernelGGL((k_tridiag<T, CH, JobTable<TriProb<T>>>), dim3(blocks), dim3(threads), lds, s, t);
    const hipError_t e = hipGetLastError();
    table_release(mem, s);
    FP_HIP(e);
}

// jobs are grouped by their plan's chunk length, one launch per group
template <typename T>
void launch_tridiag_batch(const TridiagJob<T>* jobs, int njobs, hipStream_t s)
{
    std::vector<const TridiagJob<T>*> g8, g16, g32;
    for (int k = 0; k < njobs; k++) {
        if (jobs[k].nrhs <= 0) continue;
        const int ch = jobs[k].thru && jobs[k].sp->tri_ch_b ? jobs[k].sp->tri_ch_b : jobs[k].sp->tri_ch;
        (ch == 8 ? g8 : ch == 16 ? g16 : g32).push_back(&jobs[k]);
    }
    if (!g8.empty()) launch_tridiag_group<T, 8>(g8.data(), (int)g8.size(), s);
    if (!g16.empty()) launch_tridiag_group<T, 16>(g16.data(), (int)g16.size(), s);
    if (!g32.empty()) launch_tridiag_group<T, 32>(g32.data(), (int)g32.size(), s);
}

// The launches of launch_tridiag_batch, prepared: one per chunk length, the
// jobs of each in decreasing number of rows (the length of a workgroup's
// chains).
template <typename T>
std::vector<GroupPlan> plan_tridiag_groups(const TridiagJob<T>* jobs, int njobs)
{
    std::map<int, std::vector<const TridiagJob<T>*>> groups;
    for (int k = 0; k < njobs; k++)
        if (jobs[k].nrhs > 0)
            groups[jobs[k].thru && jobs[k].sp->tri_ch_b ? jobs[k].sp->tri_ch_b : jobs[k].sp->tri_ch].push_back(&jobs[k]);
    std::vector<GroupPlan> out;
    for (auto& kv : groups) {
        const int                          CH = kv.first;
        std::vector<const TridiagJob<T>*>& js = kv.second;
        std::stable_sort(js.begin(), js.end(), [](const TridiagJob<T>* a, const TridiagJob<T>* b) {
            return a->sp->nx > b->sp->nx;
        });
        GroupPlan g;
        g.ch      = CH;
        g.nprob   = (int)js.size();
        g.threads = 64;
        std::vector<TriProb<T>> probs(js.size());
        std::vector<int>        start(js.size() + 1);
        long long               blocks = 0;
        for (size_t q = 0; q < js.size(); q++) {
            const TridiagJob<T>& jb = *js[q];
            const SubPlan&       sp = *jb.sp;
            const bool           tb = jb.thru && sp.tri_ch_b == CH;
            const int            P  = tb ? sp.tri_p_b : sp.tri_p;
            const int            MB = tri_mb(P, CH, tb ? sp.mb_b : sp.mb_field);
            if (MB * P > tri_max_threads(CH)) throw Error(FP_E_INVALID, "nx too large for the tridiagonal kernel");
            TriProb<T>& pr   = probs[q];
            pr.data          = jb.data;
            pr.ld            = jb.ld;
            pr.xd            = sp.xd.p;
            pr.pst           = (tb && sp.pst_b.p) ? sp.pst_b.p : sp.pst.p;
            pr.nx            = sp.nx;
            pr.ny            = sp.ny;
            pr.P             = P;
            pr.MB            = MB;
            pr.singular_last = sp.no_dirichlet ? 1 : 0;
            pr.nblk          = (sp.ny + MB - 1) / MB;
            pr.ns            = sp.bc[XNEG] == NEUMANN;
            pr.ne            = sp.bc[XPOS] == NEUMANN;
            pr.cyc           = sp.cyc ? 1 : 0;
            pr.z1            = sp.cyc ? sp.zc.p : nullptr;
            pr.z2            = sp.cyc ? sp.zc.p + (size_t)sp.nx * sp.ny : nullptr;
            start[q]         = (int)blocks;
            blocks += (long long)pr.nblk * jb.nrhs;
            if (blocks > (1LL << 30)) throw Error(FP_E_INVALID, "too many tridiagonal workgroups in one launch");
            g.threads = std::max(g.threads, (MB * P + 63) / 64 * 64);
            g.lds     = std::max(g.lds, 3 * (size_t)P * MB * sizeof(T));
        }
        start[js.size()] = (int)blocks;
        g.blocks         = (int)blocks;
        g.start_off      = (probs.size() * sizeof(TriProb<T>) + 255) & ~(size_t)255;
        g.table.resize(g.start_off + start.size() * sizeof(int));
        std::memcpy(g.table.data(), probs.data(), probs.size() * sizeof(TriProb<T>));
        std::memcpy(g.table.data() + g.start_off, start.data(), start.size() * sizeof(int));
        if (g.blocks > 0) out.push_back(std::move(g));
    }
    return out;
}

template <typename T>
void launch_tridiag_group(const GroupPlan& g, const unsigned char* dtable, hipStream_t s)
{
    JobTable<TriProb<T>> t;
    t.nprob = g.nprob;
    t.p     = reinterpret_cast<const TriProb<T>*>(dtable);
    t.start = reinterpret_cast<const int*>(dtable + g.start_off);
    if (g.ch == 8)
        hipLaunchKernelGGL((k_tridiag<T, 8, JobTable<TriProb<T>>>), dim3(g.blocks), dim3(g.threads), g.lds, s, t);
    else if (g.ch == 16)
        hipLaunchKernelGGL((k_tridiag<T, 16, JobTable<TriProb<T>>>), dim3(g.blocks), dim3(g.threads), g.lds, s, t);
    else if (g.ch == 32)
        hipLaunchKernelGGL((k_tridiag<T, 32, JobTable<TriProb<T>>>), dim3(g.blocks), dim3(g.threads), g.lds, s, t);
    else throw Error(FP_E_INVALID, "unknown chunk length");
    FP_HIP(hipGetLastError());
}
template std::vector<GroupPlan> plan_tridiag_groups<double>(const TridiagJob<double>*, int);
template std::vector<GroupPlan> plan_tridiag_groups<float>(const TridiagJob<float>*, int);
template void launch_tridiag_group<double>(const GroupPlan&, const unsigned char*, hipStream_t);
template void launch_tridiag_group<float>(const GroupPlan&, const unsigned char*, hipStream_t);

template <typename T>
void launch_tridiag(const SubPlan& sp, T* data, long long ld, int nrhs, hipStream_t s)
{
    const TridiagJob<T> j{&sp, data, ld, nrhs};
    launch_tridiag_batch<T>(&j, 1, s);
}

template <typename T, int CH>
static int launch_lines_ch(const SubPlan& sp, const LinesArgs<T>& a, int nrhs, hipStream_t s)
{
    const int    P   = sp.tri_p;
    const int    MB  = tri_mb(P, CH, sp.mb_lines);
    const size_t lds = 3 * (size_t)P * MB * sizeof(T);
    dim3         grid((sp.ny + MB - 1) / MB, nrhs);
    hipLaunchKernelGGL((k_tridiag_lines<T, CH>), grid, dim3((MB * P + 63) / 64 * 64), lds, s, a, P, MB);
    FP_HIP(hipGetLastError());
    return grid.x;
}

template <typename T>
int lines_workgroups(const SubPlan& sp)
{
    const int MB = tri_mb(sp.tri_p, sp.tri_ch, sp.mb_lines);
    return (sp.ny + MB - 1) / MB;
}
template int lines_workgroups<double>(const SubPlan&);

template <typename T>
int launch_tridiag_lines(const SubPlan& sp, const LinesArgs<T>& a, int nrhs, hipStream_t s)
{
    FP_REQUIRE(!sp.no_dirichlet, "line operator needs a Dirichlet-like face");
    FP_REQUIRE(!sp.cyc, "line operator of a periodic x axis");
    LinesArgs<T> b = a;
    b.xd           = sp.xd.p;
    b.pst          = sp.pst.p;
    b.ns           = sp.bc[XNEG] == NEUMANN;
    b.ne           = sp.bc[XPOS] == NEUMANN;
    if (sp.tri_ch == 8) return launch_lines_ch<T, 8>(sp, b, nrhs, s);
    if (sp.tri_ch == 16) return launch_lines_ch<T, 16>(sp, b, nrhs, s);
    return launch_lines_ch<T, 32>(sp, b, nrhs, s);
}
template int launch_tridiag_lines<double>(const SubPlan&, const LinesArgs<double>&, int, hipStream_t);

template <typename T, int CH>
static void launch_sweep_group(const SweepJob<T>* const* jobs, int n, hipStream_t s)
{
    SweepBatch<T> bt{};
    int           blocks = 0, threads = 64;
    size_t        lds = 0;
    bt.nprob          = n;
    for (int q = 0; q < n; q++) {
        const SweepJob<T>& jb = *jobs[q];
        const SubPlan&     sp = *jb.sp;
        FP_REQUIRE(!sp.no_dirichlet, "the composite sweep needs a Dirichlet-like face");
        FP_REQUIRE(!sp.cyc, "the composite sweep of a periodic x axis");
        const int P  = sp.tri_p;
        const int MB = jb.cp ? sweep_mb(sp) : tri_mb(P, CH, sp.mb_field);
        SweepProb<T>& pr = bt.p[q];
        pr.f    = jb.f;
        pr.fo   = jb.fo ? jb.fo : jb.f;
        pr.ld   = jb.ld;
        pr.lbp  = jb.lbp;
        pr.lbm  = jb.lbm;
        pr.ldl  = jb.ldl;
        pr.phi  = jb.phi;
        pr.xo0  = jb.xo0;
        pr.xo1  = jb.xo1;
        pr.ldxo = jb.ldxo;
        pr.cp   = jb.cp;
        pr.xd   = sp.xd.p;
        pr.pst  = sp.pst.p;
        pr.nx   = sp.nx;
        pr.ny   = sp.ny;
        pr.P    = P;
        pr.MB   = MB;
        pr.nblk = (sp.ny + MB - 1) / MB;
        pr.ns   = sp.bc[XNEG] == NEUMANN;
        pr.ne   = sp.bc[XPOS] == NEUMANN;
        pr.fin  = jb.fin;
        pr.fout = jb.fout;
        bt.start[q] = blocks;
        blocks += pr.nblk * jb.nrhs;
        threads = std::max(threads, (MB * P + 63) / 64 * 64);
        lds     = std::max(lds, 3 * (size_t)P * MB * sizeof(T));
    }
    bt.start[n] = blocks;
    if (blocks == 0) return;
    hipLaunchKernelGGL((k_sweep<T, CH>), dim3(blocks), dim3(threads), lds, s, bt);
    FP_HIP(hipGetLastError());
}

template <typename T>
void launch_sweep_batch(const SweepJob<T>* jobs, int njobs, hipStream_t s)
{
    std::vector<const SweepJob<T>*> g8, g16, g32;
    for (int k = 0; k < njobs; k++) {
        if (jobs[k].nrhs <= 0) continue;
        const int ch = jobs[k].sp->tri_ch;
        (ch == 8 ? g8 : ch == 16 ? g16 : g32).push_back(&jobs[k]);
    }
    for (size_t q0 = 0; q0 < g8.size(); q0 += kMaxBatch)
        launch_sweep_group<T, 8>(g8.data() + q0, (int)std::min<size_t>(kMaxBatch, g8.size() - q0), s);
    for (size_t q0 = 0; q0 < g16.size(); q0 += kMaxBatch)
        launch_sweep_group<T, 16>(g16.data() + q0, (int)std::min<size_t>(kMaxBatch, g16.size() - q0), s);
    for (size_t q0 = 0; q0 < g32.size(); q0 += kMaxBatch)
        launch_sweep_group<T, 32>(g32.data() + q0, (int)std::min<size_t>(kMaxBatch, g32.size() - q0), s);
}
template void launch_sweep_batch<double>(const SweepJob<double>*, int, hipStream_t);
template void launch_sweep_batch<float>(const SweepJob<float>*, int, hipStream_t);

int sweep_blocks(const SubPlan& sp)
{
    const int mb = sweep_mb(sp);
    return (sp.ny + mb - 1) / mb;
}

// ---- chunk schedule and its autotuning ----

// Host pivot excesses of every mode (the recurrence above) sampled at the
// chunk starts of chunk length ch: pst[j*P + c] = e_{c*ch - 1} (1 for c = 0).
// Batch shape: chunks of 16 rows and 8 modes per workgroup where the
// workgroup fits (nx <= 2048), else the plan's own shape; a shape forced by
// FP_TRI_SHAPE applies to batches too.  Measured on 512 uneven problems
// against (8, 4): prepared batch 3.58 against 3.75 ms in FP64, 2.06 against
// 2.26 ms in FP32.
void SubPlan::set_batch_chunks()
{
    tri_ch_b = tri_ch;
    tri_p_b  = tri_p;
    mb_b     = mb_field;
    pst_b.release();
    if (cyc || std::getenv("FP_TRI_SHAPE")) return;
    const int ch = 16, mb = 8, P = (nx + ch - 1) / ch;
    if (tri_mb(P, ch, mb) != mb) return;
    mb_b = mb;
    if (ch == tri_ch) return;
    const int ch0 = tri_ch;
    DevBuf<double> keep = std::move(pst);
    set_chunks(ch);  // the table for the batch chunk length, then the plan's own back
    pst_b    = std::move(pst);
    tri_ch_b = ch;
    tri_p_b  = tri_p;
    pst      = std::move(keep);
    tri_ch   = ch0;
    tri_p    = (nx + ch0 - 1) / ch0;
}

void SubPlan::set_chunks(int ch)
{
    tri_ch = ch;
    tri_p  = (nx + ch - 1) / ch;
    std::vector<double> t((size_t)ny * tri_p, 1.0);
    const bool          ns = bc[XNEG] == NEUMANN, ne = bc[XPOS] == NEUMANN;
    for (int j = 0; j < ny; j++) {
        if (no_dirichlet && j == ny - 1) continue;  // singular mode: recurrence, no pivots
        double e = 1.0, p = 0.0;
        for (int i = 0; i < nx; i++) {
            if (i % ch == 0) t[(size_t)j * tri_p + i / ch] = e;
            pivot_step(delta[j], (i == 0 && ns) || (i == nx - 1 && ne), e, p);
        }
    }
    pst.alloc(t.size());
    pst.upload(t.data(), t.size());
}

// Candidate chunk lengths for nx: the chain needs at least one thread per
// chunk (P <= the kernel's thread limit).
static std::vector<int> chunk_candidates(int nx)
{
    std::vector<int> v;
    for (int ch : {8, 16, 32})
        if ((nx + ch - 1) / ch <= tri_max_threads(ch)) v.push_back(ch);
    return v;
}

// Default schedule without timing: the shortest chunk that gives a
// workgroup of 4 modes (the commonest choice of the timed schedule), else
// the longest.
static int default_chunk(int nx)
{
    for (int ch : {8, 16, 32})
        if (4 * ((nx + ch - 1) / ch) <= tri_max_threads(ch)) return ch;
    return 32;
}

struct TriTune {
    int ch, mb;
};

// Chooses (chunk length, modes per workgroup) for the field sweep of this
// plan by timing every candidate on a scratch field of the plan's size (one
// right-hand side).  Choices are recorded per (nx, ny) (tune_store).
void SubPlan::tune_tridiag(bool measure)
{
    // FP_TRI_SHAPE=<chunk>,<modes per workgroup> (tests, measurements): that
    // shape wherever it fits, without timing
    if (const char* env = std::getenv("FP_TRI_SHAPE")) {
        int ch = 0, mb = 0;
        if (std::sscanf(env, "%d,%d", &ch, &mb) == 2)
            for (int c : chunk_candidates(nx))
                if (c == ch && mb > 0 && tri_mb((nx + ch - 1) / ch, ch, mb) == mb) {
                    set_chunks(ch);
                    mb_field = mb;
                    return;
                }
    }
    if (!measure || (long long)nx * ny < (1LL << 16)) {
        set_chunks(default_chunk(nx));
        mb_field = 4;
        return;
    }
    const std::string key = "tri " + std::to_string(nx) + " " + std::to_string(ny);
    {
        std::vector<int> v;
        if (tune_lookup(key, v) && v.size() == 2) {
            set_chunks(v[0]);
            mb_field = v[1];
            return;
        }
    }
    DevBuf<double> buf((size_t)nx * ny);
    FP_HIP(hipMemset(buf.p, 0, (size_t)nx * ny * sizeof(double)));
    hipEvent_t e0, e1;
    FP_HIP(hipEventCreate(&e0));
    FP_HIP(hipEventCreate(&e1));
    std::vector<TriTune> cand;
    for (int ch : chunk_candidates(nx)) {
        const int P = (nx + ch - 1) / ch;
        for (int mb : {1, 2, 4, 8, 16})
            if (tri_mb(P, ch, mb) == mb) cand.push_back({ch, mb});
    }
    // best of three interleaved rounds per candidate: single launches of one
    // right-hand side are latency-bound, and one round alone is noisy
    std::vector<float> best_of(cand.size(), 1e30f);
    for (int round = 0; round < 3; round++)
        for (size_t ci = 0; ci < cand.size(); ci++) {
            set_chunks(cand[ci].ch);
            mb_field = cand[ci].mb;
            launch_tridiag<double>(*this, buf.p, (long long)nx * ny, 1, nullptr);
            FP_HIP(hipEventRecord(e0, nullptr));
            for (int r = 0; r < 5; r++) launch_tridiag<double>(*this, buf.p, (long long)nx * ny, 1, nullptr);
            FP_HIP(hipEventRecord(e1, nullptr));
            FP_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            FP_HIP(hipEventElapsedTime(&ms, e0, e1));
            best_of[ci] = std::min(best_of[ci], ms);
        }
    TriTune best{default_chunk(nx), 4};
    float   best_ms = 1e30f;
    for (size_t ci = 0; ci < cand.size(); ci++)
        if (best_of[ci] < best_ms) {
            best_ms = best_of[ci];
            best    = cand[ci];
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    set_chunks(best.ch);
    mb_field = best.mb;
    tune_store(key, {best.ch, best.mb});
}

template void launch_tridiag<double>(const SubPlan&, double*, long long, int, hipStream_t);
template void launch_tridiag_batch<double>(const TridiagJob<double>*, int, hipStream_t);
template void launch_tridiag_batch<float>(const TridiagJob<float>*, int, hipStream_t);
template void launch_tridiag<float>(const SubPlan&, float*, long long, int, hipStream_t);

}  // namespace fp
