// The 3-D subdomain solver: PoissonSolver3D (pe/poisson/poisson_solver3d.cpp,
// poisson_solver3d_base.cpp, chasing_method3d.cpp) on the device.
//
// The reference transforms along z (contiguous rows), transposes so that y is
// contiguous, transforms along y, transposes again so that x is contiguous,
// runs Thomas along x for every (z mode, y mode) pair, and undoes it all
// (poisson_solver3d.cpp:87-123): four full transposes.  Here the field stays
// in field3's layout [i][j][k] (k contiguous) for the z transforms; one
// tiled LDS transpose per x-plane makes y contiguous ([i][k][j]), where the
// y transforms run on rows and the x sweep runs with lane = mode (k, j),
// every row access coalesced, exactly as the 2-D solver's sweep; one
// transpose takes it back.  Seven passes over the field instead of the
// reference's eleven.
#include "fp_solver3d.h"

#include <cmath>
#include <memory>

namespace fp {

static int kind_of(int neg, int pos)
{
    if (neg == PERIODIC && pos == PERIODIC) return T_PP;
    if (neg == PERIODIC || pos == PERIODIC) throw Error(FP_E_INVALID, "a periodic face needs a periodic partner");
    if (neg == NEUMANN && pos == NEUMANN) return T_NN;
    if (dir_like(neg) && dir_like(pos)) return T_DD;
    if (dir_like(neg) && pos == NEUMANN) return T_DN;
    return T_ND;
}

void SubPlan3D::build(int nx_, int ny_, int nz_, double hx_, double hy_, double hz_, const int bc_[6])
{
    FP_REQUIRE(nx_ >= 2 && ny_ >= 1 && nz_ >= 1, "subdomain needs nx >= 2, ny >= 1 and nz >= 1");
    FP_REQUIRE(hx_ > 0 && hy_ > 0 && hz_ > 0 && std::isfinite(hx_) && std::isfinite(hy_) && std::isfinite(hz_),
               "spacings must be positive");
    for (int f = 0; f < 6; f++)
        FP_REQUIRE(bc_[f] == DIRICHLET || bc_[f] == NEUMANN || bc_[f] == ADJACENTED || bc_[f] == PERIODIC,
                   "invalid boundary type");
    FP_REQUIRE((bc_[0] == PERIODIC) == (bc_[1] == PERIODIC), "a periodic face needs a periodic partner");
    FP_REQUIRE(bc_[2] != PERIODIC || ny_ % 2 == 0, "a periodic y axis needs an even ny");
    FP_REQUIRE(bc_[4] != PERIODIC || nz_ % 2 == 0, "a periodic z axis needs an even nz");
    FP_REQUIRE((long long)ny_ * nz_ < (1LL << 31), "too many y-z modes");
    nx = nx_; ny = ny_; nz = nz_; hx = hx_; hy = hy_; hz = hz_;
    for (int f = 0; f < 6; f++) bc[f] = bc_[f];
    const int ky = kind_of(bc[2], bc[3]), kz = kind_of(bc[4], bc[5]);
    // PoissonSolver3DBase::cal_lambda and PoissonSolver3D::init
    // (poisson_solver3d_base.cpp:30-66, poisson_solver3d.cpp:49-85)
    const std::vector<long double> gy = eigen_gaps(ky, ny), gz = eigen_gaps(kz, nz);
    const long double              ry = (long double)hx * hx / ((long double)hy * hy);
    const long double              rz = (long double)hx * hx / ((long double)hz * hz);
    std::vector<double>            dl((size_t)ny * nz);
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++) dl[(size_t)k * ny + j] = (double)(ry * gy[j] + rz * gz[k]);
    sweep.nx = nx;
    sweep.ny = ny * nz;
    sweep.hx = hx;
    sweep.hy = hy;
    sweep.bc[XNEG] = bc[0];
    sweep.bc[XPOS] = bc[1];
    sweep.bc[YNEG] = bc[2];
    sweep.bc[YPOS] = bc[3];
    sweep.kind     = ky;
    // no Dirichlet-like face of the six: the (last z, last y) pair is the
    // zero mode, gauged by the singular recurrence (chasing_method3d.cpp:83-86)
    bool any = false;
    for (int f = 0; f < 6; f++) any = any || dir_like(bc[f]);
    sweep.no_dirichlet = !any;
    FP_HIP(hipGetDevice(&sweep.dev));
    sweep.measure = measure;
    sweep.init_sweep(dl);
    tz.build(kz, nz);
    tune_transform(tz, nx * ny, measure);
    ty.build(ky, ny);
    tune_transform(ty, nx * nz, measure);
}

// Transpose of the a x b block of every x-plane: dst[s][q][p] = src[s][p][q],
// plane s = (right-hand side r, row i): src at r ld_s + i a b, dst at
// r ld_d + i a b.  32 x 32 tiles through LDS (one padding column: no bank
// conflicts on the transposed reads), 32 x 8 threads, 4 rows per thread.
constexpr int kTT = 32;
__global__ __launch_bounds__(256) void k_transpose_planes(const double* __restrict__ src, double* __restrict__ dst,
                                                          int a, int b, int nxp, long long ld_s, long long ld_d,
                                                          int tiles_a, int tiles_b)
{
    __shared__ double tile[kTT][kTT + 1];
    long long         t     = blockIdx.x;
    const int         tb    = (int)(t % tiles_b);
    t /= tiles_b;
    const int       ta    = (int)(t % tiles_a);
    const long long plane = t / tiles_a;
    const int       r     = (int)(plane / nxp), i = (int)(plane % nxp);
    const double*   s     = src + r * ld_s + (long long)i * a * b;
    double*         d     = dst + r * ld_d + (long long)i * a * b;
    const int       x = threadIdx.x & 31, y = threadIdx.x >> 5;
    const int       q = tb * kTT + x;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int p = ta * kTT + y + 8 * k;
        if (p < a && q < b) tile[y + 8 * k][x] = s[(long long)p * b + q];
    }
    __syncthreads();
    const int p2 = ta * kTT + x;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int q2 = tb * kTT + y + 8 * k;
        if (q2 < b && p2 < a) d[(long long)q2 * a + p2] = tile[x][y + 8 * k];
    }
}

static void launch_transpose_planes(const double* src, double* dst, int a, int b, int nx, int nrhs, long long ld_s,
                                    long long ld_d, hipStream_t s)
{
    const int       ta = (a + kTT - 1) / kTT, tb = (b + kTT - 1) / kTT;
    const long long nb = (long long)ta * tb * nx * nrhs;
    FP_REQUIRE(nb < (1LL << 31), "too many transpose tiles in one launch");
    hipLaunchKernelGGL(k_transpose_planes, dim3((unsigned)nb), dim3(256), 0, s, src, dst, a, b, nx, ld_s, ld_d, ta,
                       tb);
    FP_HIP(hipGetLastError());
}

// PoissonSolver3D::solve (poisson_solver3d.cpp:87-123) on nrhs hx^2-scaled
// right-hand sides ld apart, in place.
void sub3d_solve(SubPlan3D& p, double* f, long long ld, int nrhs, hipStream_t s)
{
    if (nrhs <= 0) return;
    const long long n = (long long)p.nx * p.ny * p.nz;
    if (p.work_nrhs < nrhs) {
        p.work.alloc((size_t)n * nrhs);
        p.work_nrhs = nrhs;
    }
    double* w = p.work.p;
    launch_ytransform<double>(p.tz, f, ld, p.nx * p.ny, nrhs, 0, s);
    launch_transpose_planes(f, w, p.ny, p.nz, p.nx, nrhs, ld, n, s);
    launch_ytransform<double>(p.ty, w, n, p.nx * p.nz, nrhs, 0, s);
    launch_tridiag<double>(p.sweep, w, n, nrhs, s);
    launch_ytransform<double>(p.ty, w, n, p.nx * p.nz, nrhs, 1, s);
    launch_transpose_planes(w, f, p.nz, p.ny, p.nx, nrhs, n, ld, s);
    launch_ytransform<double>(p.tz, f, ld, p.nx * p.ny, nrhs, 1, s);
}

}  // namespace fp

struct fp_sub3d_s {
    fp_ctx_s*     ctx = nullptr;
    fp::SubPlan3D plan;
};

extern "C" {

int fp_sub3d_create(fp_ctx ctx, int nx, int ny, int nz, double hx, double hy, double hz, const unsigned char bc[6],
                    fp_sub3d* out)
{
    return fp::guarded([&] {
        FP_REQUIRE(ctx && out && bc, "null argument");
        FP_HIP(hipSetDevice(ctx->device));
        auto s = std::make_unique<fp_sub3d_s>();
        s->ctx = ctx;
        int b[6];
        for (int f = 0; f < 6; f++) b[f] = bc[f];
        s->plan.measure = ctx->measure;
        s->plan.build(nx, ny, nz, hx, hy, hz, b);
        *out = s.release();
    });
}

void fp_sub3d_destroy(fp_sub3d sub) { delete sub; }

int fp_sub3d_solve(fp_sub3d sub, double* d_f, int nrhs, long long ld, void* stream)
{
    return fp::guarded([&] {
        FP_REQUIRE(sub && d_f && nrhs >= 0, "null argument");
        const long long n = (long long)sub->plan.nx * sub->plan.ny * sub->plan.nz;
        FP_REQUIRE(ld >= n || nrhs <= 1, "ld smaller than nx*ny*nz");
        FP_HIP(hipSetDevice(sub->ctx->device));
        fp::sub3d_solve(sub->plan, d_f, nrhs <= 1 ? n : ld, nrhs, (hipStream_t)stream);
    });
}

}  // extern "C"
