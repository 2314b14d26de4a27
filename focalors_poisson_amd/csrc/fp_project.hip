// Pressure projection on a staggered (MAC) grid over a composite geometry:
// the caller side of the solve path.  An incompressible Navier–Stokes step
// (the Focalors solver the Poisson module serves) takes a provisional
// velocity u* and makes it divergence-free,
//   D G p = D u* / dt,   u = u* - dt G p,
// with p at cell centres and the velocity components on the cell faces.  On
// this grid D G is exactly the reference's scaled 5-point operator with its
// face conventions (a wall face: Neumann, ghost = the adjacent cell; an
// outlet: Dirichlet, the ghost-centre value; a face between two subdomains:
// the neighbour's adjacent cell with coefficient 1), so the composite solve
// of D u*/dt gives a velocity whose discrete divergence vanishes to the
// solve's accuracy.  The solver object is reused every step, and the whole
// step (divergence, solve, correction) replays from one HIP graph.
//
// Layout per subdomain d (nx x ny cells, row-major as field2): u holds the
// x-faces [i][j], i = 0..nx (nx+1 rows of ny), v the y-faces [i][j],
// j = 0..ny (nx rows of ny+1).  A face shared by two subdomains is held by
// both (the last x-face row of the left domain and the first of the right
// one, and likewise in y); the correction changes both copies by the same
// amount, so they stay equal.  Boundary faces that are Neumann for the
// pressure (walls, inflow) keep their prescribed normal velocity.
#include "fp_internal.h"

#include <atomic>
#include <cmath>
#include <memory>

namespace {

using fp::ADJACENTED;
using fp::guarded;
using fp::DIRICHLET;
using fp::PERIODIC;
using fp::XNEG;
using fp::XPOS;
using fp::YNEG;
using fp::YPOS;

// out[i][j] = s ((u[i+1][j] - u[i][j]) / hx + (v[i][j+1] - v[i][j]) / hy)
__global__ __launch_bounds__(256) void k_divergence(const double* __restrict__ u, const double* __restrict__ v,
                                                    double* __restrict__ out, int nx, int ny, double sx, double sy)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= (long long)nx * ny) return;
    const int i = (int)(c / ny), j = (int)(c - (long long)i * ny);
    const double du = u[(long long)(i + 1) * ny + j] - u[(long long)i * ny + j];
    const double dv = v[(long long)i * (ny + 1) + j + 1] - v[(long long)i * (ny + 1) + j];
    out[c]          = sx * du + sy * dv;
}

// What lies beyond one face of a subdomain for the pressure gradient: no
// correction (Neumann), a ghost-centre value (Dirichlet; 0 without data, and
// for an Adjacented face without a neighbour), or a line of cells (the
// neighbour's adjacent line, or the opposite line of a periodic pair) read
// with a stride.
struct FaceOut {
    int           kind;  // 0 none, 1 ghost values, 2 cell line
    const double* p;     // ghost array (kind 1, may be null: zeros) or the first cell of the line (kind 2)
    long long     stride;
};

// x-faces: u[i][j] -= s (p[i][j] - p[i-1][j]) for 0 < i < nx, and on the two
// boundary rows with the value beyond the face in place of the missing cell
__global__ __launch_bounds__(256) void k_correct_u(double* __restrict__ u, const double* __restrict__ p, int nx,
                                                   int ny, double s, FaceOut lo, FaceOut hi)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= (long long)(nx + 1) * ny) return;
    const int i = (int)(c / ny), j = (int)(c - (long long)i * ny);
    double    pl, pr;
    if (i > 0 && i < nx) {
        pl = p[(long long)(i - 1) * ny + j];
        pr = p[(long long)i * ny + j];
    } else {
        const FaceOut& f = i == 0 ? lo : hi;
        if (f.kind == 0) return;
        const double out = f.kind == 1 ? (f.p ? f.p[j] : 0.0) : f.p[(long long)j * f.stride];
        const double in  = p[(long long)(i == 0 ? 0 : nx - 1) * ny + j];
        pl               = i == 0 ? out : in;
        pr               = i == 0 ? in : out;
    }
    u[c] -= s * (pr - pl);
}

// y-faces: v[i][j] -= s (p[i][j] - p[i][j-1]) for 0 < j < ny, boundary
// columns as above
__global__ __launch_bounds__(256) void k_correct_v(double* __restrict__ v, const double* __restrict__ p, int nx,
                                                   int ny, double s, FaceOut lo, FaceOut hi)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= (long long)nx * (ny + 1)) return;
    const int i = (int)(c / (ny + 1)), j = (int)(c - (long long)i * (ny + 1));
    double    pb, pt;
    if (j > 0 && j < ny) {
        pb = p[(long long)i * ny + j - 1];
        pt = p[(long long)i * ny + j];
    } else {
        const FaceOut& f = j == 0 ? lo : hi;
        if (f.kind == 0) return;
        const double out = f.kind == 1 ? (f.p ? f.p[i] : 0.0) : f.p[(long long)i * f.stride];
        const double in  = p[(long long)i * ny + (j == 0 ? 0 : ny - 1)];
        pb               = j == 0 ? out : in;
        pt               = j == 0 ? in : out;
    }
    v[c] -= s * (pt - pb);
}

int blocks_of(long long n) { return (int)((n + 255) / 256); }

}  // namespace

struct fp_proj_s {
    fp_cat                          cat = nullptr;
    fp::CatView                     g{};
    std::vector<fp::DevBuf<double>> rhs;  // D u* / dt per subdomain, the solve's right-hand side
    // the id under which the handle's projection slot holds this
    // projection's graph: unique per projection, so that no other projection
    // (one made later at the same address included) replays it
    unsigned long long id = 0;
};

static std::atomic<unsigned long long> g_next_proj_id{1};

namespace {

void divergence(fp_proj_s* pj, const double* const* d_u, const double* const* d_v, double scale,
                double* const* d_out, hipStream_t s)
{
    const fp::CatView& g = pj->g;
    for (int d = 0; d < g.nd; d++) {
        const long long n = (long long)g.nx[d] * g.ny[d];
        hipLaunchKernelGGL(k_divergence, dim3(blocks_of(n)), dim3(256), 0, s, d_u[d], d_v[d], d_out[d], g.nx[d],
                           g.ny[d], scale / g.hx[d], scale / g.hy[d]);
        FP_HIP(hipGetLastError());
    }
}

// the pressure beyond face `face` of subdomain d
FaceOut face_out(const fp::CatView& g, int d, int face, const double* const* d_p, const double* const* d_pghost)
{
    const int nx = g.nx[d], ny = g.ny[d], t = g.bc[4 * d + face];
    if (t == DIRICHLET) return {1, d_pghost ? d_pghost[4 * d + face] : nullptr, 1};
    if (t == PERIODIC) {
        // the opposite line of the same subdomain
        if (face == XNEG) return {2, d_p[d] + (long long)(nx - 1) * ny, 1};
        if (face == XPOS) return {2, d_p[d], 1};
        if (face == YNEG) return {2, d_p[d] + (ny - 1), ny};
        return {2, d_p[d], ny};
    }
    // an Adjacented face without a neighbour: ghost 0, as the operator has it
    if (t == ADJACENTED && g.adj[4 * d + face] < 0) return {1, nullptr, 1};
    if (t == ADJACENTED) {
        const int e = g.adj[4 * d + face], ex = g.nx[e], ey = g.ny[e];
        // the neighbour's line next to this face
        if (face == XNEG) return {2, d_p[e] + (long long)(ex - 1) * ey, 1};
        if (face == XPOS) return {2, d_p[e], 1};
        if (face == YNEG) return {2, d_p[e] + (ey - 1), ey};
        return {2, d_p[e], ey};
    }
    return {0, nullptr, 0};  // Neumann (walls, inflow)
}

void correct(fp_proj_s* pj, double* const* d_u, double* const* d_v, const double* const* d_p,
             const double* const* d_pghost, double scale, hipStream_t s)
{
    const fp::CatView& g = pj->g;
    for (int d = 0; d < g.nd; d++) {
        const int nx = g.nx[d], ny = g.ny[d];
        hipLaunchKernelGGL(k_correct_u, dim3(blocks_of((long long)(nx + 1) * ny)), dim3(256), 0, s, d_u[d], d_p[d],
                           nx, ny, scale / g.hx[d], face_out(g, d, XNEG, d_p, d_pghost),
                           face_out(g, d, XPOS, d_p, d_pghost));
        FP_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_correct_v, dim3(blocks_of((long long)nx * (ny + 1))), dim3(256), 0, s, d_v[d], d_p[d],
                           nx, ny, scale / g.hy[d], face_out(g, d, YNEG, d_p, d_pghost),
                           face_out(g, d, YPOS, d_p, d_pghost));
        FP_HIP(hipGetLastError());
    }
}

// ghost data of the solve: the pressure's Dirichlet values, nothing elsewhere
std::vector<const double*> solve_ghost(const fp::CatView& g, const double* const* d_pghost)
{
    std::vector<const double*> gh(4 * g.nd, nullptr);
    if (d_pghost)
        for (int k = 0; k < 4 * g.nd; k++)
            if (g.bc[k] == DIRICHLET) gh[k] = d_pghost[k];
    return gh;
}

void check_fields(const fp_proj_s* pj, double* const* a, double* const* b, double* const* c)
{
    FP_REQUIRE(a && b && c, "null argument");
    for (int d = 0; d < pj->g.nd; d++) FP_REQUIRE(a[d] && b[d] && c[d], "null field");
}

}  // namespace

extern "C" {

int fp_proj_create(fp_cat cat, fp_proj* out)
{
    return guarded([&] {
        FP_REQUIRE(cat && out, "null argument");
        auto pj = std::make_unique<fp_proj_s>();
        pj->cat = cat;
        pj->id  = g_next_proj_id.fetch_add(1);
        pj->g   = fp::cat_view(cat);
        const fp::CatView& g = pj->g;
        for (int d = 0; d < g.nd; d++) {
            // the interface coupling of the composite operator has coefficient
            // 1 on every face, which is D G only when the cells are square
            FP_REQUIRE(std::abs(g.hx[d] - g.hy[d]) <= 1e-12 * g.hx[d], "the projection needs hx == hy in every subdomain");
            for (int f = 0; f < 4; f++)
                FP_REQUIRE(g.bc[4 * d + f] != fp::BC_NULL, "the projection needs a boundary type on every face");
        }
        FP_HIP(hipSetDevice(g.device));
        pj->rhs.resize(g.nd);
        for (int d = 0; d < g.nd; d++) pj->rhs[d].alloc((size_t)g.nx[d] * g.ny[d]);
        *out = pj.release();
    });
}

void fp_proj_destroy(fp_proj pj) { delete pj; }

int fp_proj_divergence(fp_proj pj, const double* const* d_u, const double* const* d_v, double scale,
                       double* const* d_out, void* stream)
{
    return guarded([&] {
        FP_REQUIRE(pj && d_u && d_v && d_out, "null argument");
        for (int d = 0; d < pj->g.nd; d++) FP_REQUIRE(d_u[d] && d_v[d] && d_out[d], "null field");
        FP_HIP(hipSetDevice(pj->g.device));
        divergence(pj, d_u, d_v, scale, d_out, (hipStream_t)stream);
    });
}

int fp_proj_correct(fp_proj pj, double* const* d_u, double* const* d_v, const double* const* d_p,
                    const double* const* d_pghost, double scale, void* stream)
{
    return guarded([&] {
        FP_REQUIRE(pj && d_p, "null argument");
        check_fields(pj, d_u, d_v, const_cast<double* const*>(d_p));
        FP_HIP(hipSetDevice(pj->g.device));
        correct(pj, d_u, d_v, d_p, d_pghost, scale, (hipStream_t)stream);
    });
}

int fp_proj_step(fp_proj pj, double* const* d_u, double* const* d_v, double* const* d_p,
                 const double* const* d_pghost, double dt, void* stream, fp_solve_info* info)
{
    int st = guarded([&] {
        FP_REQUIRE(pj && dt > 0.0, "null argument or dt <= 0");
        check_fields(pj, d_u, d_v, d_p);
        FP_HIP(hipSetDevice(pj->g.device));
        hipStream_t         s = (hipStream_t)stream;
        std::vector<double*> rhs(pj->g.nd);
        for (int d = 0; d < pj->g.nd; d++) rhs[d] = pj->rhs[d].p;
        divergence(pj, d_u, d_v, 1.0 / dt, rhs.data(), s);
        const std::vector<const double*> gh = solve_ghost(pj->g, d_pghost);
        const int rc = fp_cat_solve_into(pj->cat, rhs.data(), d_p, gh.data(), 1, stream, info);
        if (rc != FP_OK && rc != FP_E_NOT_CONVERGED) throw fp::Error(rc, fp_last_error());
        correct(pj, d_u, d_v, d_p, d_pghost, dt, s);
        if (rc == FP_E_NOT_CONVERGED) throw fp::Error(rc, "pressure solve: GMRES did not converge (velocity corrected with the last iterate)");
    });
    return st;
}

int fp_proj_capture(fp_proj pj, double* const* d_u, double* const* d_v, double* const* d_p,
                    const double* const* d_pghost, double dt, void* stream)
{
    return guarded([&] {
        FP_REQUIRE(pj && dt > 0.0, "null argument or dt <= 0");
        check_fields(pj, d_u, d_v, d_p);
        std::vector<double*>       rhs(pj->g.nd);
        std::vector<const double*> crhs(pj->g.nd);
        for (int d = 0; d < pj->g.nd; d++) crhs[d] = rhs[d] = pj->rhs[d].p;
        const std::vector<const double*> gh = solve_ghost(pj->g, d_pghost);
        // the ghost pointers are read when the projection's kernels are
        // recorded: keep the caller's array of them alive only for this call
        std::vector<const double*> pg;
        if (d_pghost) pg.assign(d_pghost, d_pghost + 4 * pj->g.nd);
        std::vector<double*> u(d_u, d_u + pj->g.nd), v(d_v, d_v + pj->g.nd), p(d_p, d_p + pj->g.nd);
        std::vector<const double*> cp(p.begin(), p.end());
        fp::cat_capture(
            pj->cat, crhs.data(), p.data(), gh.data(), 1, (hipStream_t)stream,
            [&](hipStream_t s) { divergence(pj, u.data(), v.data(), 1.0 / dt, rhs.data(), s); },
            [&](hipStream_t s) { correct(pj, u.data(), v.data(), cp.data(), d_pghost ? pg.data() : nullptr, dt, s); },
            pj->id);
    });
}

int fp_proj_launch(fp_proj pj, void* stream, fp_solve_info* info)
{
    if (!pj) return guarded([] { FP_REQUIRE(false, "null argument"); });
    return fp::cat_replay(pj->cat, pj->id, stream, info);
}

}  // extern "C"
