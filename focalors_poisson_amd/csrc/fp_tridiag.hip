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
// Parallel Thomas: the LU factor of tridiag(1, d_j, 1) with the Neumann end
// pivots is precomputed per (row, mode) as the reciprocal pivot p_i (the
// reference's c_i).  The forward sweep y_i = p_i (f_i - y_{i-1}) and the back
// substitution x_i = y_i - p_i x_{i+1} are first-order affine recurrences, so
// each thread owns a chunk of CH rows of one mode, composes its chunk's affine
// map, the chunk maps are chained in LDS, and every thread then replays its
// chunk from the correct start value.  The whole column stays in registers:
// one read of f and p, one write of x.  The singular zero mode (x_0 = 0,
// x_{i+1} = f_i - x_{i-1} + 2 x_i) is two prefix sums and uses the same chain.
#include "fp_internal.h"

#include <algorithm>

namespace fp {

// Start values of the CH-row chunks of each mode: the chunk maps
// v -> SA v + SB (LDS, chunk-major) are chained in application order and the
// exclusive prefix, applied to 0, is written to SY.  backward = 1 chains the
// chunks from the last row upwards (back substitution); mode sing_j always
// chains forwards (the two prefix sums of the singular recurrence).
// With P <= 64 one wavefront scans one mode with a log-depth affine scan;
// longer chains fall back to one thread per mode.
template <typename T>
__device__ void chain_chunks(const T* SA, const T* SB, T* SY, int P, int MB, int j0, int ny, int backward,
                             int sing_j)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    if (P <= 64) {
        for (int jm = wave; jm < MB; jm += nwave) {
            const int j = j0 + jm;
            if (j >= ny) continue;  // uniform over the wavefront
            const bool bw = backward && j != sing_j;
            const int  c  = bw ? P - 1 - lane : lane;
            T          A = T(1), B = T(0);
            if (lane < P) {
                A = SA[c * MB + jm];
                B = SB[c * MB + jm];
            }
            for (int d = 1; d < 64; d <<= 1) {
                const T Ap = __shfl_up(A, d, 64), Bp = __shfl_up(B, d, 64);
                if (lane >= d) {
                    B = A * Bp + B;
                    A = A * Ap;
                }
            }
            T ex = __shfl_up(B, 1, 64);
            if (lane == 0) ex = T(0);
            if (lane < P) SY[c * MB + jm] = ex;
        }
    } else {
        const int jm = threadIdx.x;
        if (jm < MB && j0 + jm < ny) {
            const bool bw = backward && (j0 + jm) != sing_j;
            T          v  = T(0);
            for (int q = 0; q < P; q++) {
                const int c = bw ? P - 1 - q : q;
                SY[c * MB + jm] = v;
                v               = SA[c * MB + jm] * v + SB[c * MB + jm];
            }
        }
    }
}

// One subdomain's sweep in a batched launch.
template <typename T>
struct TriProb {
    T*        data;
    long long ld;
    const T*  pinv;
    int       nx, ny, ldp, P, MB, singular_last, nblk;  // nblk: mode blocks per right-hand side
};
template <typename T>
struct TriBatch {
    int        nprob;
    int        start[kMaxBatch + 1];
    TriProb<T> p[kMaxBatch];
};

template <typename T, int CH>
__global__ __launch_bounds__(1024) void k_tridiag(TriBatch<T> bt)
{
    int q = 0;
    while (q + 1 < bt.nprob && (int)blockIdx.x >= bt.start[q + 1]) q++;
    const TriProb<T>& pr  = bt.p[q];
    const int         nx = pr.nx, ny = pr.ny, ldp = pr.ldp, P = pr.P, MB = pr.MB, singular_last = pr.singular_last;
    const T* __restrict__ pinv = pr.pinv;
    const int local = (int)blockIdx.x - bt.start[q];
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
    // pivots are mode-major and zero-padded to ldp rows: one contiguous,
    // aligned run of CH values per thread
    const T* pp = reinterpret_cast<const T*>(__builtin_assume_aligned(pinv + (long long)(act ? j : 0) * ldp + i0, 16));
#pragma unroll
    for (int k = 0; k < CH; k++) {
        const int  i  = i0 + k;
        const bool ok = act && i < nx;
        fv[k]         = ok ? f[(long long)i * ny + j] : T(0);
        pv[k]         = act ? pp[k] : T(0);
    }
    const bool sing = singular_last && (j == ny - 1);
    const int  SINGJ = singular_last ? ny - 1 : -1;

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
__global__ __launch_bounds__(1024) void k_tridiag_lines(LinesArgs<T> a, int P, int MB)
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
    const int  SINGJ = -1;

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
        const T* pp = reinterpret_cast<const T*>(
            __builtin_assume_aligned(a.pinv + (long long)(act ? j : 0) * a.ldp + i0, 16));
#pragma unroll
        for (int k = 0; k < CH; k++) pv[k] = act ? pp[k] : T(0);
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
    chain_chunks(SA, SB, SY, P, MB, blockIdx.x * MB, ny, 0, SINGJ);
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
}

static int tridiag_ch(int nx) { return nx <= 1024 ? 16 : 32; }

template <typename T, int CH>
static void launch_tridiag_group(const TridiagJob<T>* const* jobs, int n, hipStream_t s)
{
    TriBatch<T> bt{};
    int         blocks = 0, threads = 64;
    size_t      lds = 0;
    bt.nprob        = n;
    for (int q = 0; q < n; q++) {
        const TridiagJob<T>& jb = *jobs[q];
        const SubPlan&       sp = *jb.sp;
        const int            P  = (sp.nx + CH - 1) / CH;
        int                  MB = 1;
        while (MB * 2 * P <= 1024 && MB < sp.mb_field) MB *= 2;
        if (MB * P > 1024) throw Error(FP_E_INVALID, "nx too large for the tridiagonal kernel");
        TriProb<T>& pr   = bt.p[q];
        pr.data          = jb.data;
        pr.ld            = jb.ld;
        pr.pinv          = jb.pinv;
        pr.nx            = sp.nx;
        pr.ny            = sp.ny;
        pr.ldp           = sp.ldp;
        pr.P             = P;
        pr.MB            = MB;
        pr.singular_last = sp.no_dirichlet ? 1 : 0;
        pr.nblk          = (sp.ny + MB - 1) / MB;
        bt.start[q]      = blocks;
        const long long nb = (long long)pr.nblk * jb.nrhs;
        if (blocks + nb > (1LL << 30)) throw Error(FP_E_INVALID, "too many tridiagonal workgroups in one launch");
        blocks += (int)nb;
        threads = std::max(threads, (MB * P + 63) / 64 * 64);
        lds     = std::max(lds, 3 * (size_t)P * MB * sizeof(T));
    }
    bt.start[n] = blocks;
    if (blocks == 0) return;
    hipLaunchKernelGGL((k_tridiag<T, CH>), dim3(blocks), dim3(threads), lds, s, bt);
    FP_HIP(hipGetLastError());
}

// chunk length: 16 rows per thread up to nx = 1024, 32 above (register
// budget); jobs are grouped by it and launched kMaxBatch at a time
template <typename T>
void launch_tridiag_batch(const TridiagJob<T>* jobs, int njobs, hipStream_t s)
{
    std::vector<const TridiagJob<T>*> g16, g32;
    for (int k = 0; k < njobs; k++) {
        if (jobs[k].nrhs <= 0) continue;
        if (jobs[k].sp->nx > 32768) throw Error(FP_E_INVALID, "nx above 32768 is not supported");
        (tridiag_ch(jobs[k].sp->nx) == 16 ? g16 : g32).push_back(&jobs[k]);
    }
    for (size_t q0 = 0; q0 < g16.size(); q0 += kMaxBatch)
        launch_tridiag_group<T, 16>(g16.data() + q0, (int)std::min<size_t>(kMaxBatch, g16.size() - q0), s);
    for (size_t q0 = 0; q0 < g32.size(); q0 += kMaxBatch)
        launch_tridiag_group<T, 32>(g32.data() + q0, (int)std::min<size_t>(kMaxBatch, g32.size() - q0), s);
}

template <typename T>
void launch_tridiag(const SubPlan& sp, const T* pinv, T* data, long long ld, int nrhs, hipStream_t s)
{
    const TridiagJob<T> j{&sp, pinv, data, ld, nrhs};
    launch_tridiag_batch<T>(&j, 1, s);
}

template <typename T, int CH>
static int launch_lines_ch(const SubPlan& sp, const LinesArgs<T>& a, int nrhs, hipStream_t s)
{
    const int P = (sp.nx + CH - 1) / CH;
    int       MB = 1;
    while (MB * 2 * P <= 1024 && MB < sp.mb_lines) MB *= 2;
    if (MB * P > 1024) throw Error(FP_E_INVALID, "nx too large for the tridiagonal kernel");
    const size_t lds = 3 * (size_t)P * MB * sizeof(T);
    dim3         grid((sp.ny + MB - 1) / MB, nrhs);
    hipLaunchKernelGGL((k_tridiag_lines<T, CH>), grid, dim3((MB * P + 63) / 64 * 64), lds, s, a, P, MB);
    FP_HIP(hipGetLastError());
    return grid.x;
}

template <typename T>
int lines_workgroups(const SubPlan& sp)
{
    const int CH = sp.nx <= 1024 ? 16 : 32;
    const int P  = (sp.nx + CH - 1) / CH;
    int       MB = 1;
    while (MB * 2 * P <= 1024 && MB < sp.mb_lines) MB *= 2;
    return (sp.ny + MB - 1) / MB;
}
template int lines_workgroups<double>(const SubPlan&);

template <typename T>
int launch_tridiag_lines(const SubPlan& sp, const LinesArgs<T>& a, int nrhs, hipStream_t s)
{
    FP_REQUIRE(!sp.no_dirichlet, "line operator needs a Dirichlet-like face");
    if (sp.nx <= 1024) return launch_lines_ch<T, 16>(sp, a, nrhs, s);
    if (sp.nx <= 32768) return launch_lines_ch<T, 32>(sp, a, nrhs, s);
    throw Error(FP_E_INVALID, "nx above 32768 is not supported");
}
template int launch_tridiag_lines<double>(const SubPlan&, const LinesArgs<double>&, int, hipStream_t);

template void launch_tridiag<double>(const SubPlan&, const double*, double*, long long, int, hipStream_t);
template void launch_tridiag_batch<double>(const TridiagJob<double>*, int, hipStream_t);
template void launch_tridiag_batch<float>(const TridiagJob<float>*, int, hipStream_t);
template void launch_tridiag<float>(const SubPlan&, const float*, float*, long long, int, hipStream_t);

}  // namespace fp
