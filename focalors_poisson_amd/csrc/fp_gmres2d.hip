// The composite solve's building blocks on their own, over caller-supplied
// domain solvers: SchurMat2D_{xneg,xpos,yneg,ypos}
// (pe/concat/schur_mat2d.h:7-69, schur_mat2d.cpp) and GMRESSolver2D
// (pe/concat/gmres_solver2d.h:12-57, gmres_solver2d.cpp:123-252), with
// the DomainSolver2D plug-in (pe/concat/domain_solver.h:9-15) as a C
// callback that queues its work on the given stream.
//
// Fields live on the device.  Every device buffer comes from a workspace
// the caller passes in (so a caller that owns its memory, e.g. torch, hands
// the same pointers back to its own solver callback), and nothing is
// allocated per call.  The Arnoldi loop follows the reference step by step
// (modified Gram-Schmidt, Givens rotations, the 1e-12 breakdown test, the
// absolute tolerance on |g_{j+1}| and on the restart residual, x0 = A^-1 b
// unless an initial guess is set); its dot products are fixed-order device
// reductions whose values reach the host once per Arnoldi step.
#include "fp_internal.h"

#include <cmath>
#include <memory>

namespace fp {

constexpr int kDotBlocks = 256;

// partial sums of x . y, kDotBlocks blocks of 256 threads, fixed order
__global__ __launch_bounds__(256) void k_dot_part(const double* __restrict__ x, const double* __restrict__ y,
                                                  long long n, double* part)
{
    __shared__ double sh[256];
    double            s = 0.0;
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += (long long)gridDim.x * 256) s += x[q] * y[q];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
// out[0] = the sum of the partials (sqrt of it when root), fixed order
__global__ __launch_bounds__(256) void k_dot_final(const double* part, int nb, double* out, int root)
{
    __shared__ double sh[256];
    double            s = 0.0;
    for (int q = threadIdx.x; q < nb; q += 256) s += part[q];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = root ? sqrt(sh[0]) : sh[0];
}
// y = a y + b x with a, b from host values or device scalars:
// mode 0: y += sb * x;  1: y -= (*db) * x;  2: y = x / (*db);  3: y = x - y;  4: y = sb * x
__global__ void k_vec(double* y, const double* x, long long n, double sb, const double* db, int mode)
{
    const double d = (mode == 1 || mode == 2) ? *db : 0.0;
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
        if (mode == 0) y[q] += sb * x[q];
        else if (mode == 1) y[q] -= d * x[q];
        else if (mode == 2) y[q] = x[q] / d;
        else if (mode == 3) y[q] = x[q] - y[q];
        else y[q] = sb * x[q];
    }
}
static void vec(double* y, const double* x, long long n, double sb, const double* db, int mode, hipStream_t s)
{
    const int nb = (int)std::min<long long>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(k_vec, dim3(std::max(nb, 1)), dim3(256), 0, s, y, x, n, sb, db, mode);
    FP_HIP(hipGetLastError());
}
// *out = x . y (sqrt when root) on the device
static void dot(const double* x, const double* y, long long n, double* part, double* out, int root, hipStream_t s)
{
    hipLaunchKernelGGL(k_dot_part, dim3(kDotBlocks), dim3(256), 0, s, x, y, n, part);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, s, part, kDotBlocks, out, root);
    FP_HIP(hipGetLastError());
}

// The line of a field (nx x ny, row-major) next to face loc, entry t:
// XNEG row 0, XPOS row nx-1, YNEG column 0, YPOS column ny-1.
__host__ __device__ __forceinline__ long long line_cell(int loc, int nx, int ny, int t)
{
    if (loc == XNEG) return t;
    if (loc == XPOS) return (long long)(nx - 1) * ny + t;
    if (loc == YNEG) return (long long)t * ny;
    return (long long)t * ny + ny - 1;
}

// unit fields for columns i0 .. i0+cnt-1 of a Schur block: field q (of
// bnx x bny, zero elsewhere) holds 1 at the branch's line next to the root
// (schur_mat2d.cpp: t_a(bnx-1, i) for XNEG, t_a(0, i) for XPOS, t_a(i, bny-1)
// for YNEG, t_a(i, 0) for YPOS), i = i0 + q
__global__ void k_schur_units(double* f, long long fn, int cnt, int loc, int bnx, int bny, int i0)
{
    const long long tot = fn * cnt;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < tot; p += (long long)gridDim.x * blockDim.x) {
        const int       q    = (int)(p / fn);
        const long long c    = p - (long long)q * fn;
        const int       face = loc ^ 1;  // the branch's face toward the root
        f[p]                 = c == line_cell(face, bnx, bny, i0 + q) ? 1.0 : 0.0;
    }
}
// value(j, i0 + q) = t_a(line j) of unit solve q
__global__ void k_schur_take(const double* f, long long fn, int cnt, int loc, int bnx, int bny, int cn, int i0,
                             double* value)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (j >= cn || q >= cnt) return;
    value[(long long)j * cn + i0 + q] = f[(long long)q * fn + line_cell(loc ^ 1, bnx, bny, j)];
}
// out(line i) += sum_j root(line j) value(i, j) on the root's line next to loc
// (SchurMat2D_*::operator*), one wavefront per output
__global__ __launch_bounds__(256) void k_schur_apply(const double* __restrict__ value, int cn, int loc,
                                                     const double* root, int rnx, int rny, double* out)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i    = blockIdx.x * 4 + wave;
    if (i >= cn) return;
    double s = 0.0;
    for (int j = lane; j < cn; j += 64) s += root[line_cell(loc, rnx, rny, j)] * value[(long long)i * cn + j];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) out[line_cell(loc, rnx, rny, i)] += s;
}

}  // namespace fp

struct fp_schur2d_s {
    fp_ctx_s*           ctx = nullptr;
    int                 loc = 0, bnx = 0, bny = 0, cn = 0;
    fp::DevBuf<double>  value;  // cn x cn, value(j, i) at j * cn + i
};

struct fp_gmres2d_s {
    fp_ctx_s*                 ctx = nullptr;
    int                       nx = 0, ny = 0, m = 0, max_iter = 0;
    double                    tol = 0.0;
    fp_domain_solve_fn        pe  = nullptr;
    void*                     pe_user = nullptr;
    std::vector<fp_schur2d_s*> S;
    const double*             x0 = nullptr;
    // views of the caller's workspace
    double *x = nullptr, *r = nullptr, *w = nullptr, *ft = nullptr, *V = nullptr, *part = nullptr, *dsc = nullptr;
    double* h_state = nullptr;  // pinned: the device scalars of one Arnoldi step
};

namespace {

using fp::Error;

long long gmres_work_len(int nx, int ny, int m)
{
    const long long n = (long long)nx * ny;
    // x, r, w, ft, V (m + 1), partials, device scalars (m + 2)
    return (4 + (long long)m + 1) * n + fp::kDotBlocks + m + 2;
}

void call_solver(fp_domain_solve_fn fn, void* user, double* f, int nx, int ny, int nrhs, hipStream_t s)
{
    const int rc = fn(user, f, nx, ny, nrhs, (void*)s);
    if (rc != 0) throw Error(FP_E_INVALID, "domain solver callback returned " + std::to_string(rc));
}

// GMRESSolver2D::Afun (gmres_solver2d.cpp:93-106): out = x - A^-1 (sum_s S_s x)
void afun(fp_gmres2d_s* g, const double* x, double* out, hipStream_t s)
{
    const long long n = (long long)g->nx * g->ny;
    fp::vec(g->ft, g->ft, n, 0.0, nullptr, 4, s);  // ft = 0
    for (fp_schur2d_s* S : g->S)
        hipLaunchKernelGGL(fp::k_schur_apply, dim3((S->cn + 3) / 4), dim3(256), 0, s, S->value.p, S->cn, S->loc, x,
                           g->nx, g->ny, g->ft);
    FP_HIP(hipGetLastError());
    call_solver(g->pe, g->pe_user, g->ft, g->nx, g->ny, 1, s);
    FP_HIP(hipMemcpyAsync(out, g->ft, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    fp::vec(out, x, n, 0.0, nullptr, 3, s);  // out = x - ft
}

double host_scalar(fp_gmres2d_s* g, const double* d, hipStream_t s)
{
    FP_HIP(hipMemcpyAsync(g->h_state, d, sizeof(double), hipMemcpyDeviceToHost, s));
    FP_HIP(hipStreamSynchronize(s));
    return g->h_state[0];
}

// GMRESSolver2D::solve (gmres_solver2d.cpp:123-252) on d_b, in place.
void gmres_solve(fp_gmres2d_s* g, double* b, hipStream_t s, fp_solve_info* info, double* resvec, int cap, int* nres)
{
    const long long n = (long long)g->nx * g->ny;
    const int       m = g->m;
    call_solver(g->pe, g->pe_user, b, g->nx, g->ny, 1, s);  // b <- A^-1 b
    FP_HIP(hipMemcpyAsync(g->x, g->x0 ? g->x0 : b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), gv(m + 1), y(m);
    int                 nr = 0, total = 0, outer_done = 0;
    double              beta = 0.0;
    bool                conv = false;
    auto                Vj   = [&](int j) { return g->V + (size_t)j * n; };
    for (int outer = 0; outer < g->max_iter; outer++) {
        afun(g, g->x, g->w, s);
        FP_HIP(hipMemcpyAsync(g->r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        fp::vec(g->r, g->w, n, -1.0, nullptr, 0, s);  // r = b - A x
        fp::dot(g->r, g->r, n, g->part, g->dsc, 1, s);
        beta = host_scalar(g, g->dsc, s);
        if (resvec && nr < cap) resvec[nr] = beta;
        nr++;
        outer_done = outer;
        if (beta < g->tol) {
            conv = true;
            break;
        }
        std::fill(H.begin(), H.end(), 0.0);
        std::fill(cs.begin(), cs.end(), 0.0);
        std::fill(sn.begin(), sn.end(), 0.0);
        std::fill(gv.begin(), gv.end(), 0.0);
        fp::vec(Vj(0), g->r, n, 1.0 / beta, nullptr, 4, s);  // V0 = r (1 / beta)
        gv[0] = beta;
        int j = 0;
        for (; j < m; j++) {
            afun(g, Vj(j), g->w, s);
            total++;
            // modified Gram-Schmidt: H(i, j) = V_i . w, then w -= H(i, j) V_i,
            // the scalars staying on the device until the step ends
            for (int i = 0; i <= j; i++) {
                fp::dot(Vj(i), g->w, n, g->part, g->dsc + i, 0, s);
                fp::vec(g->w, Vj(i), n, 0.0, g->dsc + i, 1, s);
            }
            fp::dot(g->w, g->w, n, g->part, g->dsc + j + 1, 1, s);
            FP_HIP(hipMemcpyAsync(g->h_state, g->dsc, sizeof(double) * (j + 2), hipMemcpyDeviceToHost, s));
            FP_HIP(hipStreamSynchronize(s));
            for (int i = 0; i <= j; i++) H[(size_t)i * m + j] = g->h_state[i];
            const double hn = g->h_state[j + 1];
            if (hn < 1e-12) break;
            H[(size_t)(j + 1) * m + j] = hn;
            fp::vec(Vj(j + 1), g->w, n, 0.0, g->dsc + j + 1, 2, s);  // V_{j+1} = w / hn
            for (int i = 0; i < j; i++) {
                const double t            = cs[i] * H[(size_t)i * m + j] + sn[i] * H[(size_t)(i + 1) * m + j];
                H[(size_t)(i + 1) * m + j] = -sn[i] * H[(size_t)i * m + j] + cs[i] * H[(size_t)(i + 1) * m + j];
                H[(size_t)i * m + j]       = t;
            }
            const double hjj = H[(size_t)j * m + j], hj1 = H[(size_t)(j + 1) * m + j];
            const double den = std::hypot(hjj, hj1);
            cs[j]            = hjj / den;
            sn[j]            = hj1 / den;
            H[(size_t)j * m + j]       = cs[j] * H[(size_t)j * m + j] + sn[j] * H[(size_t)(j + 1) * m + j];
            H[(size_t)(j + 1) * m + j] = 0.0;
            const double t             = cs[j] * gv[j];
            gv[j + 1]                  = -sn[j] * gv[j];
            gv[j]                      = t;
            if (std::fabs(gv[j + 1]) < g->tol) {
                j++;
                break;
            }
        }
        std::fill(y.begin(), y.end(), 0.0);
        for (int i = j - 1; i >= 0; i--) {
            y[i] = gv[i];
            for (int k = i + 1; k < j; k++) y[i] -= H[(size_t)i * m + k] * y[k];
            y[i] /= H[(size_t)i * m + i];
        }
        for (int i = 0; i < j; i++) fp::vec(g->x, Vj(i), n, y[i], nullptr, 0, s);  // x += y_i V_i
    }
    FP_HIP(hipMemcpyAsync(b, g->x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    FP_HIP(hipStreamSynchronize(s));
    if (nres) *nres = nr;
    if (info) {
        *info            = fp_solve_info{};
        info->iterations = total;
        info->restarts   = outer_done;
        info->residual   = beta;
    }
    if (!conv) throw Error(FP_E_NOT_CONVERGED, "GMRES returned after max_iter restarts above the tolerance");
}

}  // namespace

extern "C" {

long long fp_gmres2d_workspace(int nx, int ny, int m) { return nx > 0 && ny > 0 && m > 0 ? gmres_work_len(nx, ny, m) : 0; }

int fp_schur2d_create(fp_ctx ctx, int loc, int bnx, int bny, fp_domain_solve_fn branch, void* user, double* d_work,
                      long long work_len, void* stream, fp_schur2d* out)
{
    return fp::guarded([&] {
        FP_REQUIRE(ctx && branch && d_work && out, "null argument");
        FP_REQUIRE(loc >= 0 && loc < 4, "invalid location");
        FP_REQUIRE(bnx >= 1 && bny >= 1, "invalid branch size");
        const long long fn = (long long)bnx * bny;
        FP_REQUIRE(work_len >= fn, "workspace smaller than one branch field");
        FP_HIP(hipSetDevice(ctx->device));
        hipStream_t s = (hipStream_t)stream;
        auto        S = std::make_unique<fp_schur2d_s>();
        S->ctx = ctx;
        S->loc = loc;
        S->bnx = bnx;
        S->bny = bny;
        S->cn  = (loc == fp::XNEG || loc == fp::XPOS) ? bny : bnx;
        S->value.alloc((size_t)S->cn * S->cn);
        // SchurMat2D_*::construct: one unit solve per column, as many per
        // callback as the workspace holds (fields back to back)
        const int chunk = (int)std::min<long long>(S->cn, std::min<long long>(64, work_len / fn));
        for (int i0 = 0; i0 < S->cn; i0 += chunk) {
            const int cnt = std::min(chunk, S->cn - i0);
            const int nb  = (int)std::min<long long>((fn * cnt + 255) / 256, 4096);
            hipLaunchKernelGGL(fp::k_schur_units, dim3(nb), dim3(256), 0, s, d_work, fn, cnt, loc, bnx, bny, i0);
            FP_HIP(hipGetLastError());
            call_solver(branch, user, d_work, bnx, bny, cnt, s);
            hipLaunchKernelGGL(fp::k_schur_take, dim3((S->cn + 255) / 256, cnt), dim3(256), 0, s, d_work, fn, cnt, loc,
                               bnx, bny, S->cn, i0, S->value.p);
            FP_HIP(hipGetLastError());
        }
        FP_HIP(hipStreamSynchronize(s));
        *out = S.release();
    });
}

int fp_schur2d_size(fp_schur2d S) { return S ? S->cn : 0; }

int fp_schur2d_values(fp_schur2d S, double* h_value)
{
    return fp::guarded([&] {
        FP_REQUIRE(S && h_value, "null argument");
        FP_HIP(hipSetDevice(S->ctx->device));
        FP_HIP(hipMemcpy(h_value, S->value.p, sizeof(double) * S->cn * S->cn, hipMemcpyDeviceToHost));
    });
}

int fp_schur2d_apply(fp_schur2d S, const double* d_root, int rnx, int rny, double* d_out, int accumulate, void* stream)
{
    return fp::guarded([&] {
        FP_REQUIRE(S && d_root && d_out, "null argument");
        FP_REQUIRE(rnx >= 1 && rny >= 1, "invalid root size");
        FP_REQUIRE(S->cn == ((S->loc == fp::XNEG || S->loc == fp::XPOS) ? rny : rnx),
                   "the Schur block's interface length differs from the root's line");
        FP_HIP(hipSetDevice(S->ctx->device));
        hipStream_t s = (hipStream_t)stream;
        if (!accumulate) fp::vec(d_out, d_out, (long long)rnx * rny, 0.0, nullptr, 4, s);
        hipLaunchKernelGGL(fp::k_schur_apply, dim3((S->cn + 3) / 4), dim3(256), 0, s, S->value.p, S->cn, S->loc, d_root,
                           rnx, rny, d_out);
        FP_HIP(hipGetLastError());
    });
}

void fp_schur2d_destroy(fp_schur2d S) { delete S; }

int fp_gmres2d_create(fp_ctx ctx, int nx, int ny, int m, double tol, int max_iter, double* d_work, long long work_len,
                      fp_gmres2d* out)
{
    return fp::guarded([&] {
        FP_REQUIRE(ctx && d_work && out, "null argument");
        FP_REQUIRE(nx >= 1 && ny >= 1 && m >= 1 && max_iter >= 1 && tol >= 0.0, "invalid GMRES parameters");
        FP_REQUIRE(work_len >= gmres_work_len(nx, ny, m), "workspace smaller than fp_gmres2d_workspace");
        FP_HIP(hipSetDevice(ctx->device));
        auto g = std::make_unique<fp_gmres2d_s>();
        g->ctx      = ctx;
        g->nx       = nx;
        g->ny       = ny;
        g->m        = m;
        g->tol      = tol;
        g->max_iter = max_iter;
        const long long n = (long long)nx * ny;
        g->x    = d_work;
        g->r    = g->x + n;
        g->w    = g->r + n;
        g->ft   = g->w + n;
        g->V    = g->ft + n;
        g->part = g->V + (m + 1) * n;
        g->dsc  = g->part + fp::kDotBlocks;
        FP_HIP(hipHostMalloc((void**)&g->h_state, sizeof(double) * (m + 2), hipHostMallocDefault));
        *out = g.release();
    });
}

int fp_gmres2d_set_solver(fp_gmres2d g, fp_domain_solve_fn pe, void* user)
{
    return fp::guarded([&] {
        FP_REQUIRE(g && pe, "null argument");
        g->pe      = pe;
        g->pe_user = user;
    });
}

int fp_gmres2d_add_schur(fp_gmres2d g, fp_schur2d S)
{
    return fp::guarded([&] {
        FP_REQUIRE(g && S, "null argument");
        FP_REQUIRE(S->ctx->device == g->ctx->device, "a Schur block of another device");
        FP_REQUIRE(S->cn == ((S->loc == fp::XNEG || S->loc == fp::XPOS) ? g->ny : g->nx),
                   "the Schur block's interface length differs from the root's line");
        g->S.push_back(S);
    });
}

int fp_gmres2d_set_initial_guess(fp_gmres2d g, const double* d_x0)
{
    return fp::guarded([&] {
        FP_REQUIRE(g, "null argument");
        g->x0 = d_x0;
    });
}

int fp_gmres2d_solve(fp_gmres2d g, double* d_b, void* stream, fp_solve_info* info, double* h_resvec, int cap,
                     int* nres)
{
    return fp::guarded([&] {
        FP_REQUIRE(g && d_b, "null argument");
        FP_REQUIRE(g->pe, "no domain solver set");
        FP_REQUIRE(!h_resvec || cap >= 0, "negative residual capacity");
        FP_HIP(hipSetDevice(g->ctx->device));
        gmres_solve(g, d_b, (hipStream_t)stream, info, h_resvec, cap, nres);
    });
}

void fp_gmres2d_destroy(fp_gmres2d g)
{
    if (!g) return;
    if (g->h_state) (void)hipHostFree(g->h_state);
    delete g;
}

int fp_sub_domain_solve(void* sub, double* d_f, int nx, int ny, int nrhs, void* stream)
{
    fp_sub s = static_cast<fp_sub>(sub);
    if (!s || s->plan.nx != nx || s->plan.ny != ny) {
        fp::last_error_slot() = "fp_sub_domain_solve: the field size differs from the subdomain's";
        return FP_E_INVALID;
    }
    return fp_sub_solve(s, d_f, nrhs, (long long)nx * ny, stream);
}

}  // extern "C"
