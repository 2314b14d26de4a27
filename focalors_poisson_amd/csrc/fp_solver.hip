// Host orchestration of the MI355X solve path and its C ABI.
//
//   SubPlan        PoissonSolver2D state (pe/poisson/poisson_solver2d.cpp:37-60)
//   sub_solve      PoissonSolver2D::solve (poisson_solver2d.cpp:62-90)
//   Node / Cat     ConcatPoissonSolver2D + GMRESSolver2D + SchurMat2D_*
//                  (pe/concat/concat_solver2d.cpp, gmres_solver2d.cpp,
//                  schur_mat2d.cpp), with the Krylov iteration moved from
//                  full root fields to the interface lines of the root.
#include "fp_refine.h"
#include "fp_vec.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

namespace fp {

// ---------------------------------------------------------------------------
// subdomain plan

static int transform_kind(int neg, int pos)
{
    if (neg == PERIODIC && pos == PERIODIC) return T_PP;
    if (neg == PERIODIC || pos == PERIODIC) throw Error(FP_E_INVALID, "a periodic face needs a periodic partner");
    if (neg == BC_NULL || pos == BC_NULL || neg < 0 || neg > 4 || pos < 0 || pos > 4)
        throw Error(FP_E_INVALID, "invalid boundary type");
    if (neg == NEUMANN && pos == NEUMANN) return T_NN;
    if (dir_like(neg) && dir_like(pos)) return T_DD;
    if (dir_like(neg) && pos == NEUMANN) return T_DN;
    return T_ND;
}

// The eigenvalue gaps 2 - lambda_i (i = 1..N, reference order) of the
// transform kind.  The reference forms x_diag = -2 + r (lambda - 2) with
// lambda a cosine, which cancels to a few digits for the lowest modes
// (poisson_solver2d_base.cpp:28-62, poisson_solver3d_base.cpp:30-66); here
// 2 - lambda = 4 sin^2(a) with the half angle a of each kind, in long double:
//   DD  lambda = 2 cos(pi i/(N+1))          a = pi i / (2 (N+1))
//   NN  lambda = -2 cos(pi i/N)             a = pi (N - i) / (2N)
//   DN, ND  lambda = -2 cos(2 pi i/(2N+1))  a = pi (2N+1-2i) / (2 (2N+1))
//   PP  lambda = -2 cos(2 pi floor(i/2)/N)  a = pi (N - 2 floor(i/2)) / (2N)
// so that x_diag = -2 - delta with delta = r 4 sin^2(a) >= 0.
std::vector<long double> eigen_gaps(int kind, int n)
{
    const long double        PI = 3.141592653589793238462643383279502884L;
    std::vector<long double> g(n);
    for (int j = 0; j < n; j++) {
        const long double i = j + 1, N = n;
        long double       a;
        if (kind == T_PP) a = PI * (N - 2.0L * (long double)((j + 1) / 2)) / (2.0L * N);
        else if (kind == T_NN) a = PI * (N - i) / (2.0L * N);
        else if (kind == T_DD) a = PI * i / (2.0L * (N + 1.0L));
        else a = PI * (2.0L * N + 1.0L - 2.0L * i) / (2.0L * (2.0L * N + 1.0L));
        const long double sa = std::sin(a);
        g[j]                 = 4.0L * sa * sa;
    }
    return g;
}

// The x-sweep part of a plan: nx rows, ny modes (the y modes of a 2-D
// subdomain or the (z, y) mode pairs of a 3-D one) with per-mode gap delta;
// bc[XNEG], bc[XPOS], no_dirichlet already set.
void SubPlan::init_sweep(const std::vector<double>& dl)
{
    delta = dl;
    x_diag.resize(ny);
    for (int j = 0; j < ny; j++) x_diag[j] = -2.0 - delta[j];
    xd.alloc(ny);
    xd.upload(delta.data(), ny);
    // periodic x (ChasingMethod2D::init, chasing_method2d.cpp:21-33 with
    // chasing_periodic_precompute, chasing_method_base.cpp:232-258): the two
    // correction vectors of every mode, Dirichlet-end Thomas in long double
    cyc = bc[XNEG] == PERIODIC;
    if (cyc) {
        std::vector<double>      z(2 * (size_t)nx * ny);
        std::vector<long double> cp(nx), y(nx);
        for (int j = 0; j < ny; j++) {
            const long double d = -2.0L - (long double)delta[j];
            for (int e = 0; e < 2; e++) {
                // forward elimination of tridiag(1, d, 1) with the unit right-hand side e_0 / e_{nx-1}
                cp[0] = 1.0L / d;
                y[0]  = (e == 0 ? 1.0L : 0.0L) / d;
                for (int i = 1; i < nx; i++) {
                    const long double den = d - cp[i - 1];
                    cp[i]                 = 1.0L / den;
                    y[i]                  = ((e == 1 && i == nx - 1 ? 1.0L : 0.0L) - y[i - 1]) / den;
                }
                for (int i = nx - 2; i >= 0; i--) y[i] -= cp[i] * y[i + 1];
                for (int i = 0; i < nx; i++) z[(size_t)e * nx * ny + (size_t)i * ny + j] = (double)y[i];
            }
        }
        zc.alloc(z.size());
        zc.upload(z.data(), z.size());
    }
    // reciprocal Thomas pivots (the reference's c vector,
    // chasing_method_base.cpp:33-46) are regenerated inside the sweep from
    // one stored pivot excess per row chunk; the chunk schedule is tuned here
    tune_tridiag(measure);
    set_batch_chunks();
}

void SubPlan::build(int nx_, int ny_, double hx_, double hy_, const int bc_[4])
{
    FP_REQUIRE(nx_ >= 2 && ny_ >= 1, "subdomain needs nx >= 2 and ny >= 1");
    FP_REQUIRE(hx_ > 0 && hy_ > 0 && std::isfinite(hx_) && std::isfinite(hy_), "spacings must be positive");
    for (int f = 0; f < 4; f++)
        FP_REQUIRE(bc_[f] == DIRICHLET || bc_[f] == NEUMANN || bc_[f] == ADJACENTED || bc_[f] == PERIODIC,
                   "invalid boundary type");
    FP_REQUIRE((bc_[XNEG] == PERIODIC) == (bc_[XPOS] == PERIODIC), "a periodic face needs a periodic partner");
    FP_REQUIRE(bc_[YNEG] != PERIODIC || ny_ % 2 == 0, "a periodic y axis needs an even ny");
    nx = nx_; ny = ny_; hx = hx_; hy = hy_;
    for (int f = 0; f < 4; f++) bc[f] = bc_[f];
    kind         = transform_kind(bc[YNEG], bc[YPOS]);
    no_dirichlet = !(dir_like(bc[0]) || dir_like(bc[1]) || dir_like(bc[2]) || dir_like(bc[3]));
    FP_HIP(hipGetDevice(&dev));
    // PoissonSolver2DBase::cal_lambda + PoissonSolver2D::init
    // (poisson_solver2d_base.cpp:28-62, poisson_solver2d.cpp:43-51)
    const std::vector<long double> g = eigen_gaps(kind, ny);
    const long double              r = (long double)hx * hx / ((long double)hy * hy);
    std::vector<double>            dl(ny);
    for (int j = 0; j < ny; j++) dl[j] = (double)(r * g[j]);
    init_sweep(dl);
    t64.build(kind, ny);
    tune_transform(t64, nx, measure);
}

void SubPlan::ensure_f32()
{
    if (t32.L) return;
    t32.build(kind, ny);
    tune_transform(t32, nx, measure);
}

template <typename T>
static const TransformPlan<T>& tplan(const SubPlan& sp);
template <>
const TransformPlan<double>& tplan<double>(const SubPlan& sp) { return sp.t64; }
template <>
const TransformPlan<float>& tplan<float>(const SubPlan& sp) { return sp.t32; }

template <typename T>
void sub_solve(SubPlan& sp, T* data, long long ld, int nrhs, hipStream_t s)
{
    if (nrhs <= 0) return;
    if constexpr (sizeof(T) == 4) sp.ensure_f32();
    launch_ytransform<T>(tplan<T>(sp), data, ld, sp.nx, nrhs, 0, s);
    launch_tridiag<T>(sp, data, ld, nrhs, s);
    launch_ytransform<T>(tplan<T>(sp), data, ld, sp.nx, nrhs, 1, s);
}
template void sub_solve<double>(SubPlan&, double*, long long, int, hipStream_t);
template void sub_solve<float>(SubPlan&, float*, long long, int, hipStream_t);

template <typename T>
void sub_solve_batch(SubPlan* const* plans, T* const* data, const long long* ld, const int* nrhs, int n,
                     hipStream_t s, bool thru)
{
    std::vector<TransformJob<T>> tj;
    std::vector<TridiagJob<T>>   rj;
    tj.reserve(n);
    rj.reserve(n);
    for (int k = 0; k < n; k++) {
        if (nrhs[k] <= 0) continue;
        SubPlan& sp = *plans[k];
        if constexpr (sizeof(T) == 4) sp.ensure_f32();
        TransformJob<T> j{};
        j.plan  = &tplan<T>(sp);
        j.data  = data[k];
        j.ld    = ld[k];
        j.nx    = sp.nx;
        j.nrhs  = nrhs[k];
        j.dst64 = sizeof(T) == 8;
        tj.push_back(j);
        rj.push_back({&sp, data[k], ld[k], nrhs[k], thru ? 1 : 0});
    }
    if (tj.empty()) return;
    launch_ytransform_batch<T>(tj.data(), (int)tj.size(), 0, s);
    launch_tridiag_batch<T>(rj.data(), (int)rj.size(), s);
    launch_ytransform_batch<T>(tj.data(), (int)tj.size(), 1, s);
}
template void sub_solve_batch<double>(SubPlan* const*, double* const*, const long long*, const int*, int, hipStream_t,
                                      bool);
template void sub_solve_batch<float>(SubPlan* const*, float* const*, const long long*, const int*, int, hipStream_t,
                                     bool);

// FP32 transforms and sweeps on FP64 fields: the forward transform reads the
// FP64 field and writes FP32 coefficients to scratch[k], the sweep runs there,
// and the inverse transform writes the FP64 field again (32 bytes per cell
// instead of 48, no separate conversion passes).
static void sub_solve_batch_mixed(SubPlan* const* plans, double* const* data, const long long* ld,
                                  float* const* scratch, const int* nrhs, int n, hipStream_t s)
{
    std::vector<TransformJob<float>> fw, iv;
    std::vector<TridiagJob<float>>   rj;
    for (int k = 0; k < n; k++) {
        if (nrhs[k] <= 0) continue;
        SubPlan& sp = *plans[k];
        sp.ensure_f32();
        const long long nn = (long long)sp.nx * sp.ny;
        TransformJob<float> a{};
        a.plan  = &sp.t32;
        a.data  = scratch[k];
        a.ld    = nn;
        a.nx    = sp.nx;
        a.nrhs  = nrhs[k];
        a.src   = data[k];
        a.lds   = ld[k];
        a.src64 = 1;
        a.dst64 = 0;
        fw.push_back(a);
        TransformJob<float> b{};
        b.plan  = &sp.t32;
        b.data  = data[k];
        b.ld    = ld[k];
        b.nx    = sp.nx;
        b.nrhs  = nrhs[k];
        b.src   = scratch[k];
        b.lds   = nn;
        b.src64 = 0;
        b.dst64 = 1;
        iv.push_back(b);
        rj.push_back({&sp, scratch[k], nn, nrhs[k]});
    }
    if (fw.empty()) return;
    launch_ytransform_batch<float>(fw.data(), (int)fw.size(), 0, s);
    launch_tridiag_batch<float>(rj.data(), (int)rj.size(), s);
    launch_ytransform_batch<float>(iv.data(), (int)iv.size(), 1, s);
}

// ---------------------------------------------------------------------------
// composite solver

struct ChildLink {
    int            dom = -1, loc = 0, cn = 0, off = 0;
    DevBuf<double> S;      // cn x cn, value(i, j) row-major
    DevBuf<double> St;     // x-line child: Phi S Phi^T in the parent's y basis
    DevBuf<int>    ccell;  // child's adjacent line cells
    // x-line children whose y basis differs from the parent's: Q = T_parent T_child^T
    // maps the child's transform coefficients of a row to the parent's; Qt = Q^T
    // the other way (the transforms are orthonormal)
    DevBuf<double> Q, Qt;
    DevBuf<float>  Q32, Qt32;  // FP32 copies for the FP32 sweeps of a refinement
};

struct Node {
    int                    dom = -1;
    SubPlan                plan;
    std::vector<ChildLink> ch;
    int                    K = 0;
    DevBuf<int>            cell, ucell, uk1, uk2;
    int                    nu = 0;
    // GMRES workspace, sized for ws_nrhs systems
    int            ws_nrhs = 0, ws_m = 0, nblk = 1;
    DevBuf<double> V, Wv, Tv, X, C, H, Hacc, cs, sn, g, Y, part, part2, tolabs, resid, rnorm, G;
    // structured interface operator R A^-1 E of this node's subdomain
    int            face_off[4] = {-1, -1, -1, -1};
    int            has_x = 0, has_y = 0, nwg = 0, nxl = 0;  // nxl: x-line segments (first in the vector)
    bool           spectral_x = false;  // x-line unknowns held as y-transform coefficients
    bool           precond    = false;  // right preconditioner: explicit inverse of the interface operator
    DevBuf<double> Minv, Zv;
    long long      ldl = 0;
    DevBuf<double> phi, LB, XU, CP;
    DevBuf<float>  phi32;  // FP32 copy of phi for the FP32 sweeps of a refinement
    DevBuf<int>    kface, kpos;
    DevBuf<int>    state;
    int*           h_state = nullptr;
    double*        h_resid = nullptr;  // nrhs residual norms, then nrhs initial norms
    int*           dh_state = nullptr; // the same host memory as the device addresses it
    double*        dh_resid = nullptr;
    ~Node()
    {
        if (h_state) (void)hipHostFree(h_state);
        if (h_resid) (void)hipHostFree(h_resid);
    }
};

// Workspace of the two-level composite sweep (cat_sweep_fast_t) in one
// precision, per child link of the root in link order.
template <typename T>
struct FastLines {
    int                    nrhs = 0;
    DevBuf<T>              LBc;  // root: children's adjacent lines in the root's basis (packed lines)
    std::vector<DevBuf<T>> LBr;  // the root's adjacent line in the child's basis (packed lines)
    std::vector<DevBuf<T>> CPc;  // partial column sums of the child's upward sweep
    std::vector<DevBuf<T>> Xs;   // the child's solved adjacent row (child's basis)
    // FP32 only (FP64 uses the root node's own): the root's rows, column
    // partial sums and Schur lines
    DevBuf<T> XU, CP, LB;
};

}  // namespace fp

// Device scratch of fp_sub_solve_batched_noalloc: the descriptor tables of
// one call (about 100 bytes per problem and launch) are carved from this
// array instead of being allocated.  Each GPU the library runs on holds its
// own copy until the process ends; host code reaches it only through the
// address hipGetSymbolAddress returns on that device (fp_ctx::scratch).
constexpr size_t kNoAllocScratchBytes = 4u << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];


struct fp_cat_s {
    fp_ctx_s*                              ctx = nullptr;
    int                                    nd  = 0;
    std::vector<int>                       nx, ny, bc, adj, child, parent, parent_loc, level;
    std::vector<double>                    hx, hy;
    int                                    root = 0, nlev = 0;
    std::vector<std::vector<int>>          levels;
    std::vector<std::unique_ptr<fp::Node>> nodes;
    std::vector<fp::DevBuf<double>>        temp;
    int                                    temp_nrhs = 0;
    int                                    m = 20, tol_rel = 0, maxiter = 100, precision = FP_F64;
    double                                 tol = 1e-8;
    int                                    check_every = 4;
    // FP32 + refinement (precision FP_F32_REFINE)
    bool                                   mixed = false;  // sweep solves in FP32 arithmetic
    double                                 refine_tol = 1e-12;
    int                                    refine_max = 10;
    int                                    ws32_nrhs  = 0;
    std::vector<fp::DevBuf<float>>         s32;            // per domain FP32 scratch
    std::vector<fp::DevBuf<double>>        rb, rw;         // per domain: assembled RHS, residual / correction
    fp::DevBuf<double>                     rpart;
    double*                                h_rpart = nullptr;
    int                                    last_refine_steps = 0;
    double                                 last_refine_est   = 0.0;
    double                                 last_refine_corr  = 0.0;
    ~fp_cat_s()
    {
        if (h_rpart) (void)hipHostFree(h_rpart);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (pexec) (void)hipGraphExecDestroy(pexec);
    }
    int                                    precond_kind = 1;  // 0: none, 1: dense inverse of the interface operator
    // launch shape {threads per workgroup, odd-prime form} shared by the
    // batched transforms of all domains ({0, -1}: each plan's own; threads
    // -1: not chosen yet)
    std::array<int, 2>                     yt_shape64{-1, -1}, yt_shape32{-1, -1};
    fp::FastLines<double>                  fast;
    fp::FastLines<float>                   fast32;
    // captured solve (fp_cat_graph_capture): replays the two-level sweep with
    // the GMRES start from the explicit inverse and no further Krylov steps
    bool                                   graph_mode = false;
    // inner sweeps of a refinement: the interface solve is the explicit
    // inverse's product alone (no residual check, no host wait); the outer
    // FP64 residual decides convergence
    bool                                   no_wait    = false;
    hipGraphExec_t                         gexec      = nullptr;
    // the same solve recorded between a pressure projection's divergence and
    // velocity correction (fp_proj_capture); bound to the same workspace
    hipGraphExec_t                         pexec      = nullptr;
    unsigned long long                     pexec_id   = 0;  // the projection that recorded pexec
    int                                    gnrhs      = 0;
    int                                    gsteps     = 0;  // refinement sweeps recorded in the graph
    // statistics of the last solve
    int    last_steps = 0, last_restarts = 0, last_status = 0;
    double last_res = 0.0, last_rel = 0.0;
};

namespace fp {

static long long dom_n(const fp_cat_s* c, int d) { return (long long)c->nx[d] * c->ny[d]; }

// The subdomain solves of a sweep on domains doms[0..n): one batch per solve
// phase, in FP64, or in FP32 arithmetic on the FP64 fields while a
// refinement runs its inner sweeps.
static void dom_solve_batch(fp_cat_s* c, const int* doms, double* const* data, int n, int nrhs, hipStream_t s)
{
    std::vector<SubPlan*>  plans(n);
    std::vector<long long> ld(n);
    std::vector<int>       nr(n, nrhs);
    std::vector<float*>    sc(n);
    for (int k = 0; k < n; k++) {
        plans[k] = &c->nodes[doms[k]]->plan;
        ld[k]    = dom_n(c, doms[k]);
        sc[k]    = c->mixed ? c->s32[doms[k]].p : nullptr;
    }
    if (c->mixed) sub_solve_batch_mixed(plans.data(), data, ld.data(), sc.data(), nr.data(), n, s);
    else sub_solve_batch<double>(plans.data(), data, ld.data(), nr.data(), n, s);
}
static void dom_solve(fp_cat_s* c, int d, double* data, int nrhs, hipStream_t s)
{
    dom_solve_batch(c, &d, &data, 1, nrhs, s);
}

// root-side line cells of the face `loc` of an nx x ny field
static std::vector<int> line_cells(int loc, int nx, int ny)
{
    std::vector<int> v;
    if (loc == XNEG) for (int j = 0; j < ny; j++) v.push_back(j);
    else if (loc == XPOS) for (int j = 0; j < ny; j++) v.push_back((nx - 1) * ny + j);
    else if (loc == YNEG) for (int i = 0; i < nx; i++) v.push_back(i * ny);
    else for (int i = 0; i < nx; i++) v.push_back(i * ny + ny - 1);
    return v;
}

// child-side adjacent line of a child sitting at the parent's side `loc`
// (SchurMat2D_{xneg,xpos,yneg,ypos}::construct, schur_mat2d.cpp:5-106)
static std::vector<int> child_cells(int loc, int bnx, int bny)
{
    std::vector<int> v;
    if (loc == XNEG) for (int j = 0; j < bny; j++) v.push_back((bnx - 1) * bny + j);
    else if (loc == XPOS) for (int j = 0; j < bny; j++) v.push_back(j);
    else if (loc == YNEG) for (int i = 0; i < bnx; i++) v.push_back(i * bny + bny - 1);
    else for (int i = 0; i < bnx; i++) v.push_back(i * bny);
    return v;
}

// two subdomains' y transforms agree (same length and boundary pair)
static bool same_basis(const SubPlan& a, const SubPlan& b)
{
    return a.kind == b.kind && a.ny == b.ny;
}

// A captured solve refers to the workspace it was recorded on: every
// reallocation of that workspace (a later solve with more right-hand sides)
// destroys the graph first, so that fp_cat_graph_launch reports that no
// solve is captured instead of replaying into freed memory.
static void drop_graph(fp_cat_s* c)
{
    if (!c->gexec && !c->pexec) return;
    FP_HIP(hipDeviceSynchronize());
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->pexec) (void)hipGraphExecDestroy(c->pexec);
    c->gexec = c->pexec = nullptr;
    c->gnrhs = 0;
}

static void ensure_ws(fp_cat_s* c, Node& nd, int nrhs)
{
    if (nd.ws_nrhs >= nrhs && nd.ws_m == c->m) return;
    drop_graph(c);
    const int       K = nd.K, m = c->m;
    const long long n = dom_n(c, nd.dom);
    nd.nblk           = (K + 63) / 64;
    nd.V.alloc((size_t)(m + 1) * nrhs * K);
    nd.Wv.alloc((size_t)nrhs * K);
    nd.Tv.alloc((size_t)nrhs * K);
    nd.X.alloc((size_t)nrhs * K);
    nd.C.alloc((size_t)nrhs * K);
    nd.H.alloc((size_t)nrhs * (m + 1) * m);
    nd.Hacc.alloc((size_t)nrhs * (m + 1));
    nd.cs.alloc((size_t)nrhs * m);
    nd.sn.alloc((size_t)nrhs * m);
    nd.g.alloc((size_t)nrhs * (m + 1));
    nd.Y.alloc((size_t)nrhs * m);
    nd.part.alloc((size_t)nrhs * nd.nblk);
    nd.part2.alloc((size_t)nrhs * (m + 1) * kDotSeg);
    nd.tolabs.alloc(nrhs);
    nd.resid.alloc(nrhs);
    nd.rnorm.alloc(nrhs);
    nd.state.alloc((size_t)nrhs * 4);
    nd.G.alloc((size_t)nrhs * n);
    nd.LB.alloc((size_t)nrhs * nd.ldl);
    launch_fill<double>(nd.LB.p, (long long)nrhs * nd.ldl, 0.0, nullptr);
    nd.XU.alloc((size_t)nrhs * 2 * c->ny[nd.dom]);
    nd.Zv.alloc((size_t)nrhs * K);
    nd.CP.alloc((size_t)nrhs * std::max(nd.nwg, sweep_blocks(nd.plan)) * 2 * c->nx[nd.dom]);
    FP_HIP(hipDeviceSynchronize());
    if (nd.h_state) (void)hipHostFree(nd.h_state);
    if (nd.h_resid) (void)hipHostFree(nd.h_resid);
    FP_HIP(hipHostMalloc((void**)&nd.h_state, sizeof(int) * 4 * nrhs, hipHostMallocMapped | hipHostMallocCoherent));
    FP_HIP(hipHostMalloc((void**)&nd.h_resid, sizeof(double) * 2 * nrhs, hipHostMallocMapped | hipHostMallocCoherent));
    FP_HIP(hipHostGetDevicePointer((void**)&nd.dh_state, nd.h_state, 0));
    FP_HIP(hipHostGetDevicePointer((void**)&nd.dh_resid, nd.h_resid, 0));
    nd.ws_nrhs = nrhs;
    nd.ws_m    = m;
}

static GmresDev gdev(fp_cat_s* c, Node& nd, int nrhs)
{
    GmresDev g{};
    g.K = nd.K; g.nrhs = nrhs; g.m = c->m; g.nblk = nd.nblk;
    g.V = nd.V.p; g.Wv = nd.Wv.p; g.X = nd.X.p; g.C = nd.C.p; g.H = nd.H.p; g.Hacc = nd.Hacc.p;
    g.cs = nd.cs.p; g.sn = nd.sn.p; g.g = nd.g.p; g.Y = nd.Y.p; g.part = nd.part.p; g.part2 = nd.part2.p;
    g.tolabs = nd.tolabs.p; g.resid = nd.resid.p; g.rnorm = nd.rnorm.p; g.state = nd.state.p;
    g.hstate = nd.dh_state; g.hres = nd.dh_resid;
    return g;
}

// out = v - R A^-1 E S v  (GMRESSolver2D::Afun on interface vectors), with
// R A^-1 E applied through the line-sourced sweep: no full root field is
// formed, transformed or read.
static SchurLinesArgs schur_lines_args(fp_cat_s* c, Node& nd, const double* v)
{
    const int      nx = c->nx[nd.dom], ny = c->ny[nd.dom];
    SchurLinesArgs ga{};
    int            rows = 0;
    for (auto& L : nd.ch) {
        SchurChild& sc = ga.ch[ga.nch++];
        sc.S    = (nd.spectral_x && L.St.p) ? L.St.p : L.S.p;
        sc.cn   = L.cn;
        sc.voff = L.off;
        sc.lboff = L.loc == XNEG ? 0 : L.loc == XPOS ? ny : L.loc == YNEG ? 2 * ny : 2 * ny + nx;
        sc.row0 = rows;
        rows += L.cn;
    }
    ga.total_rows = rows;
    ga.v          = v;
    ga.ldv        = nd.K;
    ga.lb         = nd.LB.p;
    ga.ldl        = nd.ldl;
    return ga;
}

// R A^-1 E on the line buffer LB (filled by the caller), then out = v - that
// (v == nullptr: unit vectors e_{k0+r})
static void lines_solve_assemble(fp_cat_s* c, Node& nd, const double* v, double* out, int nrhs, hipStream_t s,
                                 int k0 = 0)
{
    const int nx = c->nx[nd.dom], ny = c->ny[nd.dom];
    if (nd.has_x && !nd.spectral_x) launch_ytransform<double>(nd.plan.t64, nd.LB.p, nd.ldl, 2, nrhs, 0, s);
    LinesArgs<double> la{};
    la.lines   = nd.LB.p;
    la.ldl     = nd.ldl;
    la.phi     = nd.phi.p;
    la.xu      = nd.XU.p;
    la.colpart = nd.CP.p;
    la.nx      = nx;
    la.ny      = ny;
    la.xrows   = nd.has_x;
    la.ycols   = nd.has_y;
    launch_tridiag_lines<double>(nd.plan, la, nrhs, s);
    if (nd.has_x && !nd.spectral_x) launch_ytransform<double>(nd.plan.t64, nd.XU.p, 2LL * ny, 2, nrhs, 1, s);
    launch_lines_assemble(v, out, nd.K, nd.kface.p, nd.kpos.p, nd.XU.p, ny, nd.CP.p, nd.nwg, nx, nrhs, s, k0);
}

static void matvec_lines(fp_cat_s* c, Node& nd, const double* v, double* out, int nrhs, hipStream_t s)
{
    launch_schur_lines(schur_lines_args(c, nd, v), nrhs, s);
    lines_solve_assemble(c, nd, v, out, nrhs, s);
}

static bool full_field_matvec()
{
    static const bool v = [] {
        const char* e = std::getenv("FP_MATVEC");
        return e && std::strcmp(e, "full") == 0;
    }();
    return v;
}

// The full-field matvec for this node: when asked for, and always for a
// periodic x axis (the line-sourced operator runs Dirichlet-end sweeps only).
static bool node_full_matvec(const Node& nd) { return full_field_matvec() || nd.plan.cyc; }

// Tv = S v on every child's segment of the interface vector, one launch
static void schur_products(Node& nd, const double* v, int nrhs, hipStream_t s)
{
    std::vector<SchurChild> ch;
    for (auto& L : nd.ch) ch.push_back({L.S.p, L.cn, L.off, L.off, 0});
    launch_schur_gemv_all(ch.data(), (int)ch.size(), v, nd.K, nd.Tv.p, nd.K, nrhs, s);
}

static void matvec(fp_cat_s* c, Node& nd, const double* v, double* out, int nrhs, hipStream_t s)
{
    if (!node_full_matvec(nd)) {
        matvec_lines(c, nd, v, out, nrhs, s);
        return;
    }
    const long long n = dom_n(c, nd.dom);
    schur_products(nd, v, nrhs, s);
    launch_fill<double>(nd.G.p, n * nrhs, 0.0, s);
    launch_scatter(nd.G.p, n, nd.ucell.p, nd.uk1.p, nd.uk2.p, nd.nu, nd.Tv.p, nd.K, nrhs, s);
    sub_solve<double>(nd.plan, nd.G.p, n, nrhs, s);
    launch_gather(nd.G.p, n, nd.cell.p, nd.K, v, out, nd.K, 1.0, -1.0, nrhs, s);
}

// w = B M v with M the explicit inverse of the interface operator B
static void op(fp_cat_s* c, Node& nd, const double* v, double* out, int nrhs, hipStream_t s)
{
    if (!nd.precond) {
        matvec(c, nd, v, out, nrhs, s);
        return;
    }
    launch_dense_gemv(nd.Minv.p, nd.K, v, nd.Zv.p, 1.0, 0.0, nrhs, s);
    matvec(c, nd, nd.Zv.p, out, nrhs, s);
}

static void ensure_ws(fp_cat_s* c, Node& nd, int nrhs);

// B = I - R A^-1 E S column by column (unit interface vectors through the
// line-sourced sweep, 256 at a time), then inverted in place on the device.
static void build_interface_inverse(fp_cat_s* c, Node& nd, hipStream_t s)
{
    const int K = nd.K;
    if (!nd.spectral_x || K == 0) return;
    const int chunk = std::min(K, 256);
    ensure_ws(c, nd, chunk);
    DevBuf<double> Bt((size_t)K * K);
    SchurLinesArgs ga = schur_lines_args(c, nd, nullptr);
    for (int k0 = 0; k0 < K; k0 += chunk) {
        const int cnt = std::min(chunk, K - k0);
        launch_fill<double>(nd.LB.p, (long long)cnt * nd.ldl, 0.0, s);
        launch_lines_from_units(ga, k0, cnt, s);
        lines_solve_assemble(c, nd, nullptr, Bt.p + (long long)k0 * K, cnt, s, k0);
    }
    launch_fill<double>(nd.LB.p, (long long)nd.ws_nrhs * nd.ldl, 0.0, s);
    nd.Minv.alloc((size_t)K * K);
    launch_transpose(Bt.p, nd.Minv.p, K, s);
    gauss_jordan_invert(nd.Minv.p, K, s);
    nd.precond = true;
}

static void read_state(Node& nd, int nrhs, hipStream_t s)
{
    GmresDev g{};
    g.nrhs  = nrhs;
    g.state = nd.state.p;
    g.resid = nd.resid.p;
    g.rnorm = nd.rnorm.p;
    gm_publish(g, nd.dh_state, nd.dh_resid, s);
    FP_HIP(hipStreamSynchronize(s));
}

// GMRESSolver2D::solve (gmres_solver2d.cpp:123-252) on the interface
// unknowns of node nd: C holds the right-hand side R A^-1 b' (x-line unknowns
// as y-transform coefficients when spectral_x); X receives the solution y of
// (I - R A^-1 E S) y = C.  With the explicit-inverse preconditioner the
// iteration starts from x0 = M C (the reference starts from its right-hand
// side); the first residual check then usually ends it.
static void gmres_stats(fp_cat_s* c, Node& nd, int nrhs, int restarts);

// x0 = M C with the explicit-inverse preconditioner, else x0 = C; the
// tolerance is set here unless the first residual check sets it (fused: the
// interface fits one workgroup, see gm_first).
static GmresDev gmres_begin(fp_cat_s* c, Node& nd, int nrhs, hipStream_t s, bool& fused)
{
    const int K = nd.K;
    if (nd.precond) launch_dense_gemv(nd.Minv.p, K, nd.C.p, nd.X.p, 1.0, 0.0, nrhs, s);
    else FP_HIP(hipMemcpyAsync(nd.X.p, nd.C.p, sizeof(double) * K * nrhs, hipMemcpyDeviceToDevice, s));
    const GmresDev g = gdev(c, nd, nrhs);
    fused            = g.hstate && g.K <= (1 << 20);
    if (!fused) gm_init(g, c->tol, c->tol_rel, s);
    return g;
}

// The restarted iteration from the current X.  first_pending: no residual
// check has run yet (the first one sets the tolerance when fused).  Returns
// true when the node's line buffer LB holds E S X for the final X (the last
// operator application was the residual check of X by the line-sourced
// matvec), so that the caller need not form it again.
static bool gmres_loop(fp_cat_s* c, Node& nd, int nrhs, hipStream_t s, const GmresDev& g, bool fused,
                       bool first_pending)
{
    const int  K = nd.K, m = c->m;
    const bool lines = !node_full_matvec(nd) && (!nd.has_x || nd.spectral_x);
    int        restarts = 0;
    bool       fresh    = false;  // host copy of the state is current
    bool       lb_x     = false;
    for (int outer = 0; outer < c->maxiter; outer++) {
        matvec(c, nd, nd.X.p, nd.Wv.p, nrhs, s);
        const bool pub = (first_pending && outer == 0 && fused) ? gm_first(g, c->tol, c->tol_rel, s) : gm_resid(g, s);
        if (pub) FP_HIP(hipStreamSynchronize(s));
        else read_state(nd, nrhs, s);
        fresh         = true;
        bool all_done = true;
        for (int r = 0; r < nrhs; r++)
            if (nd.h_state[4 * r + 2] == 0) all_done = false;
        if (all_done) {
            lb_x = lines;
            break;
        }
        const long long ldv = (long long)nrhs * K;
        for (int j = 0; j < m; j++) {
            op(c, nd, nd.V.p + j * ldv, nd.Wv.p, nrhs, s);
            gm_arnoldi(g, j, s);
            // the first step is checked at once (an exact preconditioner
            // converges there), later ones every check_every steps
            if (j == 0 || (j + 1) % c->check_every == 0 || j == m - 1) {
                read_state(nd, nrhs, s);
                bool any = false;
                for (int r = 0; r < nrhs; r++)
                    if (nd.h_state[4 * r + 0]) any = true;
                if (!any) break;
            }
        }
        if (!nd.precond) gm_finish_cycle(g, s);
        else {
            // x += M (V y)
            launch_fill<double>(nd.Zv.p, (long long)K * nrhs, 0.0, s);
            GmresDev gz = g;
            gz.X        = nd.Zv.p;
            gm_finish_cycle(gz, s);
            launch_dense_gemv(nd.Minv.p, K, nd.Zv.p, nd.X.p, 1.0, 1.0, nrhs, s);
        }
        fresh = false;
        restarts++;
    }
    if (!fresh) read_state(nd, nrhs, s);
    gmres_stats(c, nd, nrhs, restarts);
    return lb_x;
}

// Returns true when LB holds E S X for the final X (see gmres_loop).
static bool gmres_iterate(fp_cat_s* c, Node& nd, int nrhs, hipStream_t s)
{
    bool           fused;
    const GmresDev g = gmres_begin(c, nd, nrhs, s, fused);
    if (c->no_wait && nd.precond) return false;
    if (c->graph_mode) {
        // captured: the residual check at the start is published to host
        // memory and read after the replay; no host synchronisation here
        matvec(c, nd, nd.X.p, nd.Wv.p, nrhs, s);
        FP_REQUIRE(fused && gm_first(g, c->tol, c->tol_rel, s), "graph capture needs the one-workgroup GMRES start");
        return !node_full_matvec(nd) && (!nd.has_x || nd.spectral_x);
    }
    return gmres_loop(c, nd, nrhs, s, g, fused, true);
}

// Statistics of the last GMRES of node nd from the host copy of its state.
static void gmres_stats(fp_cat_s* c, Node& nd, int nrhs, int restarts)
{
    int    steps = 0, status = 0;
    double res = 0.0, rel = 0.0;
    for (int r = 0; r < nrhs; r++) {
        const double r0 = nd.h_resid[nrhs + r];
        steps = std::max(steps, nd.h_state[4 * r + 3]);
        res   = std::max(res, nd.h_resid[r]);
        rel   = std::max(rel, r0 > 0 ? nd.h_resid[r] / r0 : nd.h_resid[r]);
        if (nd.h_state[4 * r + 2] == 2) status = FP_E_NUMERIC;
        else if (nd.h_state[4 * r + 2] != 1 && status == 0) status = FP_E_NOT_CONVERGED;
    }
    if (nd.dom == c->root) {
        c->last_steps    = steps;
        c->last_restarts = restarts;
        c->last_res      = res;
        c->last_rel      = rel;
        c->last_status   = status;
    }
}

// General tree node: b <- A^-1 b; C = R b; GMRES; then
// u = A^-1 b + A^-1 E S y (full-field form).
static void gmres_solve(fp_cat_s* c, Node& nd, double* b, int nrhs, hipStream_t s)
{
    ensure_ws(c, nd, nrhs);
    const long long n = dom_n(c, nd.dom);
    const int       K = nd.K;
    dom_solve(c, nd.dom, b, nrhs, s);
    launch_gather(b, n, nd.cell.p, K, nullptr, nd.C.p, K, 0.0, 1.0, nrhs, s);
    // x-line unknowns to the root's y basis (orthogonal: GMRES norms unchanged)
    if (nd.spectral_x) launch_ytransform<double>(nd.plan.t64, nd.C.p, K, nd.nxl, nrhs, 0, s);
    gmres_iterate(c, nd, nrhs, s);
    if (nd.spectral_x) launch_ytransform<double>(nd.plan.t64, nd.X.p, K, nd.nxl, nrhs, 1, s);
    // u = z0 + A^-1 E S y
    schur_products(nd, nd.X.p, nrhs, s);
    launch_fill<double>(nd.G.p, n * nrhs, 0.0, s);
    launch_scatter(nd.G.p, n, nd.ucell.p, nd.uk1.p, nd.uk2.p, nd.nu, nd.Tv.p, K, nrhs, s);
    dom_solve(c, nd.dom, nd.G.p, nrhs, s);
    launch_axpby(nd.G.p, b, n * nrhs, 1.0, 1.0, s);
}

// The solves of one tree level: the leaves of the level in one batch (one
// launch per solve phase for all of them), then each non-leaf's GMRES.
static void level_solve(fp_cat_s* c, const std::vector<int>& doms, double* const* f, int nrhs, hipStream_t s)
{
    std::vector<int>     leaves;
    std::vector<double*> data;
    for (int d : doms)
        if (c->nodes[d]->ch.empty()) {
            leaves.push_back(d);
            data.push_back(f[d]);
        }
    if (!leaves.empty()) dom_solve_batch(c, leaves.data(), data.data(), (int)leaves.size(), nrhs, s);
    for (int d : doms)
        if (!c->nodes[d]->ch.empty()) gmres_solve(c, *c->nodes[d], f[d], nrhs, s);
}

static void node_solve(fp_cat_s* c, int d, double* f, int nrhs, hipStream_t s)
{
    if (c->nodes[d]->ch.empty()) dom_solve(c, d, f, nrhs, s);
    else gmres_solve(c, *c->nodes[d], f, nrhs, s);
}

// GMRESSolver2D::schur_mat_construct (gmres_solver2d.cpp:39-89): S of each
// child from unit sources on the child's adjacent line, solved in batches.
static void build_schur(fp_cat_s* c, Node& nd, ChildLink& L, hipStream_t s)
{
    const int       ch = L.dom, cn = L.cn;
    const long long nc = dom_n(c, ch);
    L.S.alloc((size_t)cn * cn);
    const long long budget = 1LL << 28;  // doubles of scratch (2 GiB)
    int             chunk  = (int)std::max<long long>(1, std::min<long long>(cn, budget / nc));
    DevBuf<double>  buf((size_t)chunk * nc);
    for (int i0 = 0; i0 < cn; i0 += chunk) {
        const int cnt = std::min(chunk, cn - i0);
        launch_fill<double>(buf.p, nc * cnt, 0.0, s);
        launch_set_units(buf.p, nc, L.ccell.p, i0, cnt, s);
        node_solve(c, ch, buf.p, cnt, s);
        launch_schur_collect(buf.p, nc, L.ccell.p, cn, i0, cnt, L.S.p, s);
    }
    FP_HIP(hipStreamSynchronize(s));
}

static void build_node(fp_cat_s* c, int d, hipStream_t s)
{
    auto nd = std::make_unique<Node>();
    nd->dom = d;
    int bc[4];
    for (int f = 0; f < 4; f++) bc[f] = c->bc[4 * d + f];
    nd->plan.measure = c->ctx->measure;
    nd->plan.build(c->nx[d], c->ny[d], c->hx[d], c->hy[d], bc);
    std::vector<int>                 cells;
    std::map<int, std::vector<int>>  users;
    for (int loc = 0; loc < 4; loc++) {
        const int chd = c->child[4 * d + loc];
        if (chd < 0) continue;
        ChildLink L;
        L.dom = chd;
        L.loc = loc;
        L.cn  = (loc == XNEG || loc == XPOS) ? c->ny[chd] : c->nx[chd];
        L.off = (int)cells.size();
        std::vector<int> rc = line_cells(loc, c->nx[d], c->ny[d]);
        FP_REQUIRE((int)rc.size() == L.cn, "connected faces must have the same number of cells");
        for (int t = 0; t < L.cn; t++) {
            users[rc[t]].push_back((int)cells.size());
            cells.push_back(rc[t]);
        }
        std::vector<int> cc = child_cells(loc, c->nx[chd], c->ny[chd]);
        L.ccell.alloc(cc.size());
        L.ccell.upload(cc.data(), cc.size());
        nd->ch.push_back(std::move(L));
    }
    nd->K = (int)cells.size();
    {
        // phi0 / phi1: transform of the unit vectors e_0 and e_{ny-1}; ldl:
        // packed line buffers [x0 (ny) | x1 (ny) | y0 (nx) | y1 (nx)]
        const int ny = c->ny[d];
        nd->ldl      = 2LL * ny + 2LL * c->nx[d];
        std::vector<double> e(2 * (size_t)ny, 0.0);
        e[0]             = 1.0;
        e[ny + (ny - 1)] = 1.0;
        nd->phi.alloc(e.size());
        nd->phi.upload(e.data(), e.size());
        launch_ytransform<double>(nd->plan.t64, nd->phi.p, ny, 2, 1, 0, s);
        FP_HIP(hipStreamSynchronize(s));
    }
    if (nd->K) {
        std::vector<int> kf, kp;
        for (auto& L : nd->ch) {
            nd->face_off[L.loc] = L.off;
            if (L.loc == XNEG || L.loc == XPOS) nd->has_x = 1;
            else nd->has_y = 1;
            for (int t = 0; t < L.cn; t++) {
                kf.push_back(L.loc);
                kp.push_back(t);
            }
        }
        nd->kface.alloc(kf.size()); nd->kface.upload(kf.data(), kf.size());
        nd->kpos.alloc(kp.size()); nd->kpos.upload(kp.data(), kp.size());
        nd->nwg = lines_workgroups<double>(nd->plan);
        nd->cell.alloc(cells.size());
        nd->cell.upload(cells.data(), cells.size());
        std::vector<int> uc, k1, k2;
        for (auto& kv : users) {
            uc.push_back(kv.first);
            k1.push_back(kv.second[0]);
            k2.push_back(kv.second.size() > 1 ? kv.second[1] : -1);
        }
        nd->nu = (int)uc.size();
        nd->ucell.alloc(uc.size()); nd->ucell.upload(uc.data(), uc.size());
        nd->uk1.alloc(k1.size()); nd->uk1.upload(k1.data(), k1.size());
        nd->uk2.alloc(k2.size()); nd->uk2.upload(k2.data(), k2.size());
    }
    c->nodes[d] = std::move(nd);
    Node& N = *c->nodes[d];
    for (auto& L : N.ch) build_schur(c, N, L, s);
    N.spectral_x = N.has_x && !node_full_matvec(N);
    for (auto& L : N.ch) {
        if (L.loc != XNEG && L.loc != XPOS) continue;
        N.nxl++;
        if (!N.spectral_x) continue;
        // St = Phi S Phi^T with the row transform R(M) = M Phi^T:
        // R((R(S^T))^T) = (S^T Phi^T)^T Phi^T = Phi S Phi^T
        const int      cn = L.cn;
        DevBuf<double> t1((size_t)cn * cn);
        L.St.alloc((size_t)cn * cn);
        launch_transpose(L.S.p, t1.p, cn, s);
        launch_ytransform<double>(N.plan.t64, t1.p, cn, cn, 1, 0, s);
        launch_transpose(t1.p, L.St.p, cn, s);
        launch_ytransform<double>(N.plan.t64, L.St.p, cn, cn, 1, 0, s);
        FP_HIP(hipStreamSynchronize(s));
    }
    for (auto& L : N.ch) {
        if (L.loc != XNEG && L.loc != XPOS) continue;
        const Node& C = *c->nodes[L.dom];
        if (same_basis(C.plan, N.plan)) continue;
        // rows e_k -> child inverse transform -> parent forward transform: row k
        // of the result is Q e_k, i.e. the result is Q^T (row-major)
        const int           n = L.cn;
        std::vector<double> eye((size_t)n * n, 0.0);
        for (int k = 0; k < n; k++) eye[(size_t)k * n + k] = 1.0;
        L.Qt.alloc(eye.size());
        L.Qt.upload(eye.data(), eye.size());
        launch_ytransform<double>(C.plan.t64, L.Qt.p, n, n, 1, 1, s);
        launch_ytransform<double>(N.plan.t64, L.Qt.p, n, n, 1, 0, s);
        L.Q.alloc(eye.size());
        launch_transpose(L.Qt.p, L.Q.p, n, s);
        FP_HIP(hipStreamSynchronize(s));
    }
    if (c->precond_kind == 1 && N.K > 0 && N.K <= 16384 && !N.plan.cyc) build_interface_inverse(c, N, s);
}

// TreeUtils::findOptimalRoot / buildTreeMapFromRoot / BFS
// (base/domain/geometry_tree.hpp:22-200).  A two-centre tie is broken by the
// lower domain index (the reference compares Domain2DUniform addresses).
static void build_tree(fp_cat_s* c)
{
    const int nd = c->nd;
    c->root      = 0;
    if (nd > 1) {
        std::vector<int> deg(nd, 0), q;
        for (int a = 0; a < nd; a++)
            for (int l = 0; l < 4; l++)
                if (c->adj[4 * a + l] >= 0) deg[a]++;
        for (int a = 0; a < nd; a++)
            if (deg[a] <= 1) q.push_back(a);
        size_t qh = 0;
        int    remaining = nd;
        while (remaining > 2) {
            const int cnt = (int)(q.size() - qh);
            if (cnt == 0) break;
            remaining -= cnt;
            for (int i = 0; i < cnt; i++) {
                const int leaf = q[qh++];
                for (int l = 0; l < 4; l++) {
                    const int nb = c->adj[4 * leaf + l];
                    if (nb >= 0 && --deg[nb] == 1) q.push_back(nb);
                }
            }
        }
        if (qh < q.size()) {
            c->root = q[qh];
            if (qh + 1 < q.size() && q[qh + 1] < c->root) c->root = q[qh + 1];
        }
    }
    c->child.assign(4 * nd, -1);
    c->parent.assign(nd, -1);
    c->parent_loc.assign(nd, -1);
    c->level.assign(nd, -1);
    std::vector<int> order{c->root};
    c->level[c->root] = 0;
    for (size_t h = 0; h < order.size(); h++) {
        const int a = order[h];
        for (int l = 0; l < 4; l++) {
            const int b = c->adj[4 * a + l];
            if (b < 0 || b == c->parent[a] || c->level[b] >= 0) continue;
            c->child[4 * a + l] = b;
            c->parent[b]        = a;
            c->parent_loc[b]    = l ^ 1;
            c->level[b]         = c->level[a] + 1;
            order.push_back(b);
        }
    }
    FP_REQUIRE((int)order.size() == nd, "geometry is not single connected");
    c->nlev = 0;
    for (int i = 0; i < nd; i++) c->nlev = std::max(c->nlev, c->level[i] + 1);
    c->levels.assign(c->nlev, {});
    for (int a : order) c->levels[c->level[a]].push_back(a);
}

// Boundary assembly and hx^2 scaling of every domain, in place
// (pe/concat/concat_solver2d.cpp:110-116,159-223).
// Ghost-data coefficient of each face of domain d: Dirichlet -1/h^2, Neumann
// +1/h (negative faces) or -1/h (positive faces); none for Adjacented faces
// or missing data.
static void face_terms(const fp_cat_s* c, int d, const double* const* ghost, const double* g[4], double a[4])
{
    const double hx = c->hx[d], hy = c->hy[d];
    for (int face = 0; face < 4; face++) {
        const int t = c->bc[4 * d + face];
        g[face]     = nullptr;
        a[face]     = 0.0;
        const double* gp = ghost ? ghost[4 * d + face] : nullptr;
        if (!gp) continue;
        if (t == DIRICHLET) a[face] = face < 2 ? -1.0 / hx / hx : -1.0 / hy / hy;
        else if (t == NEUMANN) {
            const double h = face < 2 ? hx : hy;
            a[face]        = (face == XNEG || face == YNEG) ? 1.0 / h : -1.0 / h;
        } else continue;
        g[face] = gp;
    }
}

static void cat_assemble(fp_cat_s* c, double* const* f, const double* const* ghost, int nrhs, hipStream_t s)
{
    for (int d = 0; d < c->nd; d++) {
        AsmFaces<double> fc{};
        face_terms(c, d, ghost, fc.g, fc.a);
        launch_assemble<double>(f[d], c->nx[d], c->ny[d], dom_n(c, d), nrhs, c->hx[d] * c->hx[d], fc, s);
    }
}

// The tree sweep of ConcatPoissonSolver2D::solve (concat_solver2d.cpp:118-157)
// on assembled, scaled right-hand sides: u = A^-1 b in place.
static void cat_sweep(fp_cat_s* c, double* const* f, int nrhs, hipStream_t s)
{
    const int nd = c->nd;
    if (c->temp_nrhs < nrhs) {
        c->temp.clear();
        c->temp.resize(nd);
        for (int d = 0; d < nd; d++)
            if (d != c->root) c->temp[d].alloc((size_t)dom_n(c, d) * nrhs);
        c->temp_nrhs = nrhs;
    }
    for (int l = c->nlev - 1; l >= 1; l--) {
        for (int d : c->levels[l]) {
            const long long n = dom_n(c, d);
            FP_HIP(hipMemcpyAsync(c->temp[d].p, f[d], sizeof(double) * n * nrhs, hipMemcpyDeviceToDevice, s));
            for (int loc = 0; loc < 4; loc++) {
                const int ch = c->child[4 * d + loc];
                if (ch < 0) continue;
                launch_bond_add<double>(c->temp[d].p, c->nx[d], c->ny[d], n, c->temp[ch].p, c->nx[ch], c->ny[ch],
                                        dom_n(c, ch), loc, -1.0, nrhs, s);
                launch_bond_add<double>(f[d], c->nx[d], c->ny[d], n, c->temp[ch].p, c->nx[ch], c->ny[ch],
                                        dom_n(c, ch), loc, -1.0, nrhs, s);
            }
        }
        std::vector<double*> tp(c->nd, nullptr);
        for (int d : c->levels[l]) tp[d] = c->temp[d].p;
        level_solve(c, c->levels[l], tp.data(), nrhs, s);
    }
    const int r = c->root;
    for (int loc = 0; loc < 4; loc++) {
        const int ch = c->child[4 * r + loc];
        if (ch < 0) continue;
        launch_bond_add<double>(f[r], c->nx[r], c->ny[r], dom_n(c, r), c->temp[ch].p, c->nx[ch], c->ny[ch],
                                dom_n(c, ch), loc, -1.0, nrhs, s);
    }
    c->last_steps = c->last_restarts = c->last_status = 0;
    c->last_res = c->last_rel = 0.0;
    node_solve(c, r, f[r], nrhs, s);
    for (int l = 1; l < c->nlev; l++) {
        for (int d : c->levels[l]) {
            const int p = c->parent[d];
            launch_bond_add<double>(f[d], c->nx[d], c->ny[d], dom_n(c, d), f[p], c->nx[p], c->ny[p], dom_n(c, p),
                                    c->parent_loc[d], -1.0, nrhs, s);
        }
        level_solve(c, c->levels[l], f, nrhs, s);
    }
}

static void ensure_refine_ws(fp_cat_s* c, int nrhs)
{
    if (c->ws32_nrhs >= nrhs) return;
    drop_graph(c);
    const int nd = c->nd;
    c->s32.clear();
    c->rb.clear();
    c->rw.clear();
    c->s32.resize(nd);
    c->rb.resize(nd);
    c->rw.resize(nd);
    for (int d = 0; d < nd; d++) {
        const size_t n = (size_t)dom_n(c, d) * nrhs;
        c->s32[d].alloc(n);
        c->rb[d].alloc(n);
        c->rw[d].alloc(n);
    }
    c->rpart.alloc((size_t)2 * 2 * nd * nrhs * kNormBlocks);
    if (c->h_rpart) (void)hipHostFree(c->h_rpart);
    c->h_rpart = nullptr;
    FP_HIP(hipHostMalloc((void**)&c->h_rpart, sizeof(double) * c->rpart.n));
    c->ws32_nrhs = nrhs;
}

// r_d = b_d - A_d u_d - (neighbour lines) for every domain
static void cat_residual(fp_cat_s* c, double* const* u, int nrhs, hipStream_t s)
{
    CatResidualBatch bt;
    for (int d = 0; d < c->nd; d++) {
        CatResidualArgs a{};
        const long long n = dom_n(c, d);
        a.u   = u[d];
        a.b   = c->rb[d].p;
        a.r   = c->rw[d].p;
        a.ldu = a.ldb = a.ldr = n;
        a.nx  = c->nx[d];
        a.ny  = c->ny[d];
        a.ry  = c->hx[d] * c->hx[d] / (c->hy[d] * c->hy[d]);
        a.perx = c->bc[4 * d + XNEG] == PERIODIC;
        a.pery = c->bc[4 * d + YNEG] == PERIODIC;
        for (int face = 0; face < 4; face++) {
            a.neu[face] = c->bc[4 * d + face] == NEUMANN;
            const int nb = c->adj[4 * d + face];
            a.nb[face]   = nb >= 0 ? u[nb] : nullptr;
            a.nbx[face]  = nb >= 0 ? c->nx[nb] : 0;
            a.nby[face]  = nb >= 0 ? c->ny[nb] : 0;
            a.ldn[face]  = nb >= 0 ? dom_n(c, nb) : 0;
        }
        bt.a[bt.n++] = a;
        if (bt.n == 8 || d == c->nd - 1) {
            launch_cat_residual_batch(bt, nrhs, s);
            bt.n = 0;
        }
    }
}

// Iterative refinement of the composite solve (no reference counterpart):
//   u_1 = S(b);  u_{k+1} = u_k + S(b - A u_k)
// with the FP64 residual of the assembled composite operator.  S is the tree
// sweep: FP_F32_REFINE runs its subdomain solves in FP32 arithmetic (the
// interface GMRES, Schur products and the residual stay FP64); FP_F64_REFINE
// runs the plain FP64 sweep, whose first solve keeps its GMRES residual check
// and whose correction solves start the interface solve from the explicit
// inverse without a check.  The FP64 composite solve has a rounding floor
// that grows with the grid (about 4e-11 relative on the 1024 cross, for the
// reference's algorithm as for this one); one correction with the FP64
// residual takes it to about 1e-14.
//
// Error estimate of u_k, with d_k = u_k - u_{k-1} and ratio_k = |d_k| / |u_k|:
// d_2 is about the error of u_1, so ratio_2 estimates the inner solve's
// relative accuracy, which is also the contraction of the iteration;
// from k = 3 on the contraction is measured, rho = |d_k| / |d_{k-1}|.  The
// estimate rho / (1 - rho) ratio_k (rho = ratio_2 at k = 2 for FP64 inner
// solves; FP32 ones wait for the measured contraction) is never below
// kRefineFloor, the accuracy an FP64 residual allows, nor used when rho >= 1
// (then ratio_k itself).  The loop stops when the estimate is at most tol for
// every right-hand side, or after refine_max sweeps.
constexpr double kRefineFloor = 64.0 * 2.220446049250313e-16;

static bool fast_shape_ok(const fp_cat_s* c);
template <typename T>
static void cat_sweep_fast_t(fp_cat_s* c, double* const* f, const double* const* ghost, int nrhs, hipStream_t s,
                             const double* const* rhs = nullptr, bool assembled = false);

// u += d (d null: u unchanged) for every domain, with the per-block norms
// |d|^2, |u|^2 into norm slot `slot`
static void refine_norms(fp_cat_s* c, double* const* u, const double* const* dlt, size_t slot, int nrhs,
                         hipStream_t s)
{
    const int    nd   = c->nd;
    const size_t half = (size_t)2 * nd * nrhs * kNormBlocks;  // doubles per norm set
    NormsBatch   nb;
    for (int d = 0; d < nd; d++) {
        nb.j[nb.n++] = {u[d], dom_n(c, d), dlt ? dlt[d] : nullptr, dom_n(c, d), dom_n(c, d),
                        c->rpart.p + slot * half + (size_t)2 * d * nrhs * kNormBlocks};
        if (nb.n == 8 || d == nd - 1) {
            launch_update_norms_batch(nb, nrhs, s);
            nb.n = 0;
        }
    }
}

// The refinement on assembled, hx^2-scaled right-hand sides f (solution in
// place).  capture: the stream is being recorded into a graph, so the loop
// takes exactly refine_max - 1 corrections with no host reads.
static void cat_refine(fp_cat_s* c, double* const* f, int nrhs, hipStream_t s, bool capture = false)
{
    const int  nd  = c->nd;
    const bool f32 = c->precision == FP_F32_REFINE;
    ensure_refine_ws(c, nrhs);
    // the inner sweep: the two-level sweep in transform space where the tree
    // allows it, else the general sweep (FP32 subdomain solves when c->mixed)
    const bool fast = fast_shape_ok(c);
    int        gm_steps = 0, gm_restarts = 0, gm_status = 0;
    double     gm_res = 0.0, gm_rel = 0.0;
    auto       sweep = [&](double* const* u, bool first) {
        if (fast && f32) cat_sweep_fast_t<float>(c, u, nullptr, nrhs, s);
        else if (fast) {
            c->no_wait = !first;
            cat_sweep_fast_t<double>(c, u, nullptr, nrhs, s, nullptr, true);
            c->no_wait = false;
        } else cat_sweep(c, u, nrhs, s);
        if (first) {
            gm_steps    = c->last_steps;
            gm_restarts = c->last_restarts;
            gm_status   = c->last_status;
            gm_res      = c->last_res;
            gm_rel      = c->last_rel;
        }
    };
    for (int d = 0; d < nd; d++)
        FP_HIP(hipMemcpyAsync(c->rb[d].p, f[d], sizeof(double) * dom_n(c, d) * nrhs, hipMemcpyDeviceToDevice, s));
    const size_t half = (size_t)2 * nd * nrhs * kNormBlocks;
    auto         sums = [&](size_t slot, int r, double& dd, double& uu) {
        dd = uu = 0.0;
        for (int d = 0; d < nd; d++) {
            const double* p = c->h_rpart + slot * half + (size_t)2 * d * nrhs * kNormBlocks + (size_t)2 * r * kNormBlocks;
            for (int k = 0; k < kNormBlocks; k++) {
                dd += p[2 * k];
                uu += p[2 * k + 1];
            }
        }
    };
    std::vector<double*>       w(nd);
    std::vector<const double*> wc(nd);
    for (int d = 0; d < nd; d++) wc[d] = w[d] = c->rw[d].p;
    c->mixed   = f32;
    c->no_wait = fast && f32;
    std::vector<double> prev(nrhs, 0.0), first_ratio(nrhs, 0.0);
    int                 steps = 0;
    double              est = INFINITY, corr = INFINITY;
    auto                restore = [&] {
        c->mixed   = false;
        c->no_wait = false;
    };
    try {
        sweep(f, true);  // u_1 = d_1
        steps = 1;
        if (!capture) refine_norms(c, f, nullptr, 0, nrhs, s);  // |u_1|^2 in slot 0
        for (int k = 2; k <= c->refine_max; k++) {
            cat_residual(c, f, nrhs, s);
            sweep(w.data(), false);
            refine_norms(c, f, wc.data(), 1, nrhs, s);  // u += d, |d|^2, |u|^2 in slot 1
            steps = k;
            if (capture) continue;
            FP_HIP(hipMemcpyAsync(c->h_rpart, c->rpart.p, sizeof(double) * 2 * half, hipMemcpyDeviceToHost, s));
            FP_HIP(hipStreamSynchronize(s));
            double worst = 0.0, wcorr = 0.0;
            bool   bad   = false;
            for (int r = 0; r < nrhs; r++) {
                double dd, uu, d0, u0;
                if (k == 2) {
                    sums(0, r, d0, u0);
                    prev[r] = std::sqrt(u0);  // |d_1| = |u_1|
                }
                sums(1, r, dd, uu);
                const double dn = std::sqrt(dd), un = std::sqrt(uu);
                const double ratio = un > 0 ? dn / un : dn;
                if (k == 2) first_ratio[r] = ratio;
                // d_1 = u_1 is the whole solution: the contraction is first
                // measured as |d_3| / |d_2|; at k = 2 an FP64 inner solve's own
                // relative accuracy (ratio_2) stands in for it
                double rho = 1.0;
                if (k > 2 && prev[r] > 0) rho = dn / prev[r];
                else if (k == 2 && !f32) rho = ratio;
                double e = rho < 1.0 ? rho / (1.0 - rho) * ratio : ratio;
                e        = std::max(e, kRefineFloor);
                if (!std::isfinite(e) || !std::isfinite(ratio)) bad = true;
                worst   = std::max(worst, e);
                wcorr   = std::max(wcorr, ratio);
                prev[r] = dn;
            }
            est  = bad ? INFINITY : worst;
            corr = bad ? INFINITY : wcorr;
            if (bad || est <= c->refine_tol) break;
        }
    } catch (...) {
        restore();
        throw;
    }
    restore();
    c->last_steps        = gm_steps;
    c->last_restarts     = gm_restarts;
    c->last_res          = gm_res;
    c->last_rel          = gm_rel;
    c->last_status       = gm_status;
    c->last_refine_steps = steps;
    c->last_refine_est   = capture ? 0.0 : est;
    c->last_refine_corr  = capture ? 0.0 : corr;
    if (capture) return;
    if (!std::isfinite(est)) c->last_status = FP_E_NUMERIC;
    else if (est > c->refine_tol && c->last_status == 0) c->last_status = FP_E_NOT_CONVERGED;
}

// ---- two-level composite sweep in transform space ----
//
// For a root with leaf children (every geometry of depth two: the cross, an
// L, a T, one lone domain) the tree sweep of concat_solver2d.cpp:118-157 is
// run on transform coefficients, and every interface coupling enters as a
// line source of the x sweep (k_sweep) instead of a full-field operation:
//   1. every domain's raw right-hand side is assembled, scaled and
//      transformed in place (one batched launch, assembly fused into the
//      load): F_d;
//   2. the children's upward solves A_c^-1 F_c are swept with only the line
//      next to the root written out (x-lines: the solved row, taken to the
//      root's basis by the dense orthogonal map Q unless the bases agree;
//      y-lines: column sums);
//   3. the root's sweep of F_r minus those lines writes only its interface
//      lines: the GMRES right-hand side R A_r^-1 b';
//   4. after GMRES, the root's sweep of F_r - child lines + E S y writes the
//      whole coefficient field and its interface rows and column sums;
//   5. the root's line next to each child (in the child's basis for x-lines)
//      is subtracted in the children's final sweeps, which write their
//      fields; one batched launch transforms every domain back.
// Every field is read and written in place: two transforms and two sweeps per
// domain, no copies, fills or full-field scatters.

static bool fast_shape_ok(const fp_cat_s* c)
{
    if (c->nlev > 2) return false;
    for (int d = 0; d < c->nd; d++)
        if (c->nodes[d]->plan.cyc) return false;  // the line sweeps are Dirichlet-end only
    const Node& R = *c->nodes[c->root];
    return R.ch.empty() || !R.has_x || R.spectral_x;
}
static bool fast_ok(const fp_cat_s* c) { return !c->mixed && fast_shape_ok(c); }

static void alloc_fast_lines(fp_cat_s* c, FastLines<double>& w, int nrhs)
{
    Node&        R  = *c->nodes[c->root];
    const size_t nc = R.ch.size();
    w.LBc.alloc((size_t)nrhs * R.ldl);
    launch_fill<double>(w.LBc.p, (long long)nrhs * R.ldl, 0.0, nullptr);
    w.LBr.clear();
    w.CPc.clear();
    w.Xs.clear();
    w.LBr.resize(nc);
    w.CPc.resize(nc);
    w.Xs.resize(nc);
    for (size_t k = 0; k < nc; k++) {
        const Node& C = *c->nodes[R.ch[k].dom];
        w.LBr[k].alloc((size_t)nrhs * C.ldl);
        launch_fill<double>(w.LBr[k].p, (long long)nrhs * C.ldl, 0.0, nullptr);
        w.CPc[k].alloc((size_t)nrhs * sweep_blocks(C.plan) * 2 * c->nx[C.dom]);
        w.Xs[k].alloc((size_t)nrhs * c->ny[C.dom]);
    }
}
static void alloc_fast_lines(fp_cat_s* c, FastLines<float>& w, int nrhs)
{
    Node&        R  = *c->nodes[c->root];
    const size_t nc = R.ch.size();
    // zero-filled through an FP64 fill of the same bytes
    auto zero = [](DevBuf<float>& b, size_t n) {
        b.alloc(n + 1);
        launch_fill<double>(reinterpret_cast<double*>(b.p), (long long)(n + 1) / 2, 0.0, nullptr);
    };
    zero(w.LBc, (size_t)nrhs * R.ldl);
    w.LBr.clear();
    w.CPc.clear();
    w.Xs.clear();
    w.LBr.resize(nc);
    w.CPc.resize(nc);
    w.Xs.resize(nc);
    for (size_t k = 0; k < nc; k++) {
        const Node& C = *c->nodes[R.ch[k].dom];
        zero(w.LBr[k], (size_t)nrhs * C.ldl);
        w.CPc[k].alloc((size_t)nrhs * sweep_blocks(C.plan) * 2 * c->nx[C.dom]);
        w.Xs[k].alloc((size_t)nrhs * c->ny[C.dom]);
    }
    zero(w.LB, (size_t)nrhs * R.ldl);
    w.XU.alloc((size_t)nrhs * 2 * c->ny[R.dom]);
    w.CP.alloc((size_t)nrhs * sweep_blocks(R.plan) * 2 * c->nx[R.dom]);
    // FP32 copies of the phi tables and the orthogonal maps (nrhs-independent)
    for (int d = 0; d < c->nd; d++) {
        Node& N = *c->nodes[d];
        N.plan.ensure_f32();
        if (!N.phi32.p && N.phi.p) {
            N.phi32.alloc(N.phi.n);
            launch_to_f32(N.phi.p, N.phi32.p, (long long)N.phi.n, nullptr);
        }
    }
    for (ChildLink& L : R.ch)
        if (L.Q.p && !L.Q32.p) {
            L.Q32.alloc(L.Q.n);
            L.Qt32.alloc(L.Qt.n);
            launch_to_f32(L.Q.p, L.Q32.p, (long long)L.Q.n, nullptr);
            launch_to_f32(L.Qt.p, L.Qt32.p, (long long)L.Qt.n, nullptr);
        }
}

template <typename T>
static FastLines<T>& fast_lines(fp_cat_s* c);
template <>
FastLines<double>& fast_lines<double>(fp_cat_s* c) { return c->fast; }
template <>
FastLines<float>& fast_lines<float>(fp_cat_s* c) { return c->fast32; }

template <typename T>
static void ensure_fast_ws(fp_cat_s* c, int nrhs)
{
    FastLines<T>& w = fast_lines<T>(c);
    if (w.nrhs >= nrhs) return;
    drop_graph(c);
    alloc_fast_lines(c, w, nrhs);
    FP_HIP(hipDeviceSynchronize());
    w.nrhs = nrhs;
}

// The buffers one precision of the two-level sweep works on.  FP64: the
// fields themselves (transformed in place) and the nodes' own line buffers;
// FP32 (the inner solves of a refinement, fields already assembled): per
// domain FP32 coefficient scratch between an FP64-reading forward transform
// and an FP64-writing inverse, and FP32 copies of every line buffer, phi table
// and orthogonal map.
template <typename T>
struct FastView {
    std::vector<T*>       fld;
    std::vector<const T*> phi;
    T *                   XU = nullptr, *CP = nullptr;
    const T*              LB = nullptr;
    std::vector<const T*> Q, Qt;
};
static FastView<double> fast_view(fp_cat_s* c, double* const* f, double*)
{
    FastView<double> v;
    Node&            R = *c->nodes[c->root];
    for (int d = 0; d < c->nd; d++) {
        v.fld.push_back(f[d]);
        v.phi.push_back(c->nodes[d]->phi.p);
    }
    v.XU = R.XU.p;
    v.CP = R.CP.p;
    v.LB = R.LB.p;
    for (const ChildLink& L : R.ch) {
        v.Q.push_back(L.Q.p);
        v.Qt.push_back(L.Qt.p);
    }
    return v;
}
static FastView<float> fast_view(fp_cat_s* c, double* const*, float*)
{
    FastView<float> v;
    Node&           R = *c->nodes[c->root];
    for (int d = 0; d < c->nd; d++) {
        v.fld.push_back(c->s32[d].p);
        v.phi.push_back(c->nodes[d]->phi32.p);
    }
    v.XU = c->fast32.XU.p;
    v.CP = c->fast32.CP.p;
    v.LB = c->fast32.LB.p;
    for (const ChildLink& L : R.ch) {
        v.Q.push_back(L.Q32.p);
        v.Qt.push_back(L.Qt32.p);
    }
    return v;
}

// The launch shape shared by the batched transforms of every domain, so that
// each direction is one launch whose domains run side by side: timed once per
// solver and precision on zeroed scratch fields.
template <typename T>
static std::array<int, 2> batch_shape(fp_cat_s* c)
{
    std::array<int, 2>& v = sizeof(T) == 8 ? c->yt_shape64 : c->yt_shape32;
    if (v[0] >= 0) return v;
    if (c->nd < 2) return v = {0, -1};
    // FP_BATCH_SHAPE=<threads>,<form> forces the shared shape (tests): it is
    // still applied only to the domains it fits
    if (const char* e = std::getenv("FP_BATCH_SHAPE")) {
        int nt = 0, form = -1;
        if (std::sscanf(e, "%d,%d", &nt, &form) == 2) return v = {nt, form};
    }
    std::vector<DevBuf<T>>       buf(c->nd);
    std::vector<TransformJob<T>> jobs(c->nd);
    for (int d = 0; d < c->nd; d++) {
        SubPlan& sp = c->nodes[d]->plan;
        if constexpr (sizeof(T) == 4) sp.ensure_f32();
        const long long n = dom_n(c, d);
        buf[d].alloc((size_t)n);
        FP_HIP(hipMemset(buf[d].p, 0, sizeof(T) * n));
        jobs[d]       = {};
        jobs[d].plan  = &tplan<T>(sp);
        jobs[d].data  = buf[d].p;
        jobs[d].ld    = n;
        jobs[d].nx    = c->nx[d];
        jobs[d].nrhs  = 1;
        jobs[d].dst64 = sizeof(T) == 8;
    }
    v = tune_batch_shape<T>(jobs, nullptr);
    FP_HIP(hipDeviceSynchronize());
    return v;
}

// rhs (FP64 only; null: f holds them): raw right-hand sides read by the
// forward transforms instead of f, which then only receives the solution.
// assembled (FP64 only): f already holds assembled, hx^2-scaled right-hand
// sides (the corrections of a refinement), so the forward transforms read
// them as they are.
template <typename T>
static void cat_sweep_fast_t(fp_cat_s* c, double* const* f, const double* const* ghost, int nrhs, hipStream_t s,
                             const double* const* rhs, bool assembled)
{
    constexpr bool f64 = sizeof(T) == 8;
    const int      nd = c->nd, root = c->root;
    Node&          R  = *c->nodes[root];
    c->last_steps = c->last_restarts = c->last_status = 0;
    c->last_res = c->last_rel = 0.0;
    if (!R.ch.empty()) {
        ensure_ws(c, R, nrhs);
        ensure_fast_ws<T>(c, nrhs);
    } else if constexpr (!f64)
        R.plan.ensure_f32();
    const FastView<T> v  = fast_view(c, f, (T*)nullptr);
    const auto        shape = batch_shape<T>(c);
    // 1. forward transforms of every domain (FP64: in place, with the
    //    boundary assembly and scaling fused into the load; FP32: from the
    //    assembled FP64 field into the FP32 scratch)
    std::vector<TransformJob<T>> fw(nd), iv(nd);
    for (int d = 0; d < nd; d++) {
        const long long  n = dom_n(c, d);
        TransformJob<T>& a = fw[d];
        TransformJob<T>& b = iv[d];
        a       = {};
        a.plan  = &tplan<T>(c->nodes[d]->plan);
        a.nx    = c->nx[d];
        a.nrhs  = nrhs;
        a.dst64 = f64;
        apply_batch_shape(a, shape);
        b       = a;
        a.data  = v.fld[d];
        a.ld    = n;
        b.data  = f[d];
        b.ld    = n;
        b.dst64 = 1;
        if constexpr (f64) {
            if (!assembled) {
                a.asmb = 1;
                a.hx2  = c->hx[d] * c->hx[d];
                face_terms(c, d, ghost, a.g, a.ga);
            }
            if (rhs && rhs[d] != f[d]) {
                a.src = rhs[d];
                a.lds = n;
            }
        } else {
            a.src   = f[d];
            a.lds   = n;
            a.src64 = 1;
            b.src   = v.fld[d];
            b.lds   = n;
        }
    }
    launch_ytransform_batch<T>(fw.data(), nd, 0, s);
    const int       nxr = c->nx[root], nyr = c->ny[root];
    const long long nr  = dom_n(c, root);
    if (R.ch.empty()) {
        launch_tridiag<T>(R.plan, v.fld[root], nr, nrhs, s);
        launch_ytransform_batch<T>(iv.data(), nd, 1, s);
        return;
    }
    FastLines<T>& w  = fast_lines<T>(c);
    const size_t  nc = R.ch.size();
    FP_REQUIRE(nc <= 8, "too many children");

    // 2. children's upward solves, only the line next to the root written:
    //    root face XNEG (row 0) meets the child's row nx-1, XPOS its row 0;
    //    root face YNEG (column 0) meets the child's column ny-1 (phi1 sums),
    //    YPOS its column 0 (phi0 sums)
    std::vector<SweepJob<T>> up(nc);
    ColsumBatchT<T>          ycs;
    GemvBatchT<T>            qup;
    for (size_t k = 0; k < nc; k++) {
        const ChildLink& L   = R.ch[k];
        const Node&      C   = *c->nodes[L.dom];
        const int        nyc = c->ny[L.dom];
        SweepJob<T>&     j   = up[k];
        j.sp   = &C.plan;
        j.nrhs = nrhs;
        j.f    = v.fld[L.dom];
        j.ld   = dom_n(c, L.dom);
        j.phi  = v.phi[L.dom];
        j.fin  = 1;
        j.fout = 0;
        if (L.loc == XNEG || L.loc == XPOS) {
            T* slot = w.LBc.p + (L.loc == XNEG ? 0 : nyr);
            T* dst  = v.Q[k] ? w.Xs[k].p : slot;
            j.ldxo  = v.Q[k] ? nyc : R.ldl;
            if (L.loc == XNEG) j.xo1 = dst;
            else j.xo0 = dst;
            if (v.Q[k]) qup.j[qup.n++] = {v.Q[k], nyc, w.Xs[k].p, nyc, slot, R.ldl};
        } else {
            j.cp = w.CPc[k].p;
            ycs.j[ycs.n++] = {w.CPc[k].p, sweep_blocks(C.plan), c->nx[L.dom], L.loc == YNEG ? 1 : 0,
                              w.LBc.p + 2LL * nyr + (L.loc == YNEG ? 0 : nxr), R.ldl};
        }
    }
    launch_sweep_batch<T>(up.data(), (int)nc, s);
    launch_exchange(ycs, qup, LineCopyBatchT<T>{}, nrhs, s);

    // 3. the root's interface lines of A_r^-1 b' (the GMRES right-hand side)
    {
        SweepJob<T> j{};
        j.sp   = &R.plan;
        j.nrhs = nrhs;
        j.f    = v.fld[root];
        j.ld   = nr;
        j.lbm  = w.LBc.p;
        j.ldl  = R.ldl;
        j.phi  = v.phi[root];
        j.xo0  = v.XU;
        j.xo1  = v.XU + nyr;
        j.ldxo = 2LL * nyr;
        j.cp   = R.has_y ? v.CP : nullptr;
        j.fin  = 1;
        j.fout = 0;
        launch_sweep_batch<T>(&j, 1, s);
        launch_lines_assemble<T>(nullptr, R.C.p, R.K, R.kface.p, R.kpos.p, v.XU, nyr, v.CP, sweep_blocks(R.plan),
                                 nxr, nrhs, s, 0, 1);
    }
    // 4. the root's solution: sweep of F_r - child lines + E S y, written as
    //    coefficients, with its interface rows and column sums for the children
    //    (E S y is still in the line buffer when GMRES ended on a residual
    //    check of y by the line-sourced matvec)
    auto final_sweep = [&](bool schur, T* fo) {
        if (schur) launch_schur_lines(schur_lines_args(c, R, R.X.p), nrhs, s);
        if constexpr (!f64) launch_to_f32(R.LB.p, w.LB.p, (long long)nrhs * R.ldl, s);
        SweepJob<T> j{};
        j.sp   = &R.plan;
        j.nrhs = nrhs;
        j.f    = v.fld[root];
        j.fo   = fo;
        j.ld   = nr;
        j.lbp  = v.LB;
        j.lbm  = w.LBc.p;
        j.ldl  = R.ldl;
        j.phi  = v.phi[root];
        j.xo0  = v.XU;
        j.xo1  = v.XU + nyr;
        j.ldxo = 2LL * nyr;
        j.cp   = R.has_y ? v.CP : nullptr;
        j.fin  = 1;
        j.fout = 1;
        launch_sweep_batch<T>(&j, 1, s);
    };
    bool swept = false;
    if constexpr (f64) {
        if (R.precond) {
            // FP64 with the explicit-inverse start x0 = M C: the final sweep
            // runs on x0 at once, into the scratch field G so that F_r survives
            // a restart, and doubles as the residual check.  Its interface
            // values are R U_r = C + R A^-1 E S x0, so B x0 = x0 + C - R U_r
            // and the first check costs no separate operator application.
            bool           fused;
            const GmresDev g = gmres_begin(c, R, nrhs, s, fused);
            FP_REQUIRE(fused || !c->graph_mode, "graph capture needs the one-workgroup GMRES start");
            if (c->no_wait) {
                // a correction solve of an FP64 refinement: x0 = M C with no
                // residual check (the outer FP64 residual decides)
                final_sweep(true, nullptr);
            } else if (fused) {
                final_sweep(true, R.G.p);
                launch_lines_assemble<double>(R.X.p, R.Wv.p, R.K, R.kface.p, R.kpos.p, R.XU.p, nyr, R.CP.p,
                                              sweep_blocks(R.plan), nxr, nrhs, s, 0, 2, R.C.p);
                gm_first(g, c->tol, c->tol_rel, s);
                if (!c->graph_mode) {
                    FP_HIP(hipStreamSynchronize(s));
                    bool done = true;
                    for (int r = 0; r < nrhs; r++)
                        if (R.h_state[4 * r + 2] == 0) done = false;
                    if (done) gmres_stats(c, R, nrhs, 0);
                    else final_sweep(!gmres_loop(c, R, nrhs, s, g, fused, false), R.G.p);
                }
                iv[root].src = R.G.p;
                iv[root].lds = nr;
            } else {
                final_sweep(!gmres_loop(c, R, nrhs, s, g, fused, true), nullptr);
            }
            swept = true;
        }
    }
    if (!swept) final_sweep(!gmres_iterate(c, R, nrhs, s), nullptr);

    // 5. the root's line next to each child into the child's right-hand side:
    //    x-lines from the solved rows (through Q^T when the bases differ),
    //    y-lines from the column sums; then the children's final solves and
    //    every domain's inverse transform in one batch
    ColsumBatchT<T>   ydn;
    GemvBatchT<T>     qdn;
    LineCopyBatchT<T> lc;
    for (size_t k = 0; k < nc; k++) {
        const ChildLink& L   = R.ch[k];
        const Node&      C   = *c->nodes[L.dom];
        const int        nxc = c->nx[L.dom], nyc = c->ny[L.dom];
        if (L.loc == XNEG || L.loc == XPOS) {
            // root row 0 (XNEG) -> child's row nx-1 (slot x1); root row nx-1 (XPOS) -> child's row 0 (x0)
            const T* src = v.XU + (L.loc == XNEG ? 0 : nyr);
            T*       dst = w.LBr[k].p + (L.loc == XNEG ? nyc : 0);
            if (v.Qt[k]) qdn.j[qdn.n++] = {v.Qt[k], nyc, src, 2LL * nyr, dst, C.ldl};
            else lc.c[lc.n++] = {src, 2LL * nyr, 1, nyc, dst, C.ldl, T(1)};
        } else {
            // root column 0 (YNEG) -> child's column ny-1 (slot y1); root column ny-1 (YPOS) -> child's column 0 (y0)
            ydn.j[ydn.n++] = {v.CP, sweep_blocks(R.plan), nxr, L.loc == YNEG ? 0 : 1,
                              w.LBr[k].p + 2LL * nyc + (L.loc == YNEG ? nxc : 0), C.ldl};
        }
    }
    launch_exchange(ydn, qdn, lc, nrhs, s);
    std::vector<SweepJob<T>> dn(nc);
    for (size_t k = 0; k < nc; k++) {
        const int    cd = R.ch[k].dom;
        const Node&  C  = *c->nodes[cd];
        SweepJob<T>& j  = dn[k];
        j.sp   = &C.plan;
        j.nrhs = nrhs;
        j.f    = v.fld[cd];
        j.ld   = dom_n(c, cd);
        j.lbm  = w.LBr[k].p;
        j.ldl  = C.ldl;
        j.phi  = v.phi[cd];
        j.fin  = 1;
        j.fout = 1;
    }
    launch_sweep_batch<T>(dn.data(), (int)nc, s);
    launch_ytransform_batch<T>(iv.data(), nd, 1, s);
}
static void cat_sweep_fast(fp_cat_s* c, double* const* f, const double* const* ghost, int nrhs, hipStream_t s,
                           const double* const* rhs = nullptr)
{
    cat_sweep_fast_t<double>(c, f, ghost, nrhs, s, rhs);
}

// ConcatPoissonSolver2D::solve (pe/concat/concat_solver2d.cpp:101-157).
// rhs != null: the right-hand sides are read from there (unchanged) and the
// solution written to f; the two-level FP64 path reads them in its forward
// transforms, the other paths start from a copy.
static void cat_solve(fp_cat_s* c, double* const* f, const double* const* ghost, int nrhs, hipStream_t s,
                      const double* const* rhs = nullptr)
{
    c->last_refine_steps = 0;
    c->last_refine_est   = 0.0;
    c->last_refine_corr  = 0.0;
    if (rhs && c->precision == FP_F64 && fast_ok(c)) {
        cat_sweep_fast(c, f, ghost, nrhs, s, rhs);
        return;
    }
    if (rhs)
        for (int d = 0; d < c->nd; d++)
            if (rhs[d] != f[d])
                FP_HIP(hipMemcpyAsync(f[d], rhs[d], sizeof(double) * dom_n(c, d) * nrhs, hipMemcpyDeviceToDevice, s));
    if (c->precision != FP_F64) {
        cat_assemble(c, f, ghost, nrhs, s);
        cat_refine(c, f, nrhs, s);
    } else if (fast_ok(c))
        cat_sweep_fast(c, f, ghost, nrhs, s);
    else {
        cat_assemble(c, f, ghost, nrhs, s);
        cat_sweep(c, f, nrhs, s);
    }
}

}  // namespace fp

// ---------------------------------------------------------------------------
// C ABI

namespace {
using fp::guarded;
}  // namespace

namespace fp {
std::string& last_error_slot()
{
    thread_local std::string s;
    return s;
}
}  // namespace fp

namespace {
template <typename T>
void solve_batched(fp_ctx ctx, int nprob, const fp_sub* subs, T* const* d_f, const int* nrhs, void* stream)
{
    FP_REQUIRE(ctx && nprob >= 0 && (nprob == 0 || (subs && d_f && nrhs)), "null argument");
    FP_HIP(hipSetDevice(ctx->device));
    std::vector<fp::SubPlan*> plans(nprob);
    std::vector<long long>    ld(nprob);
    for (int k = 0; k < nprob; k++) {
        FP_REQUIRE(subs[k] && d_f[k] && nrhs[k] >= 0, "null problem");
        plans[k] = &subs[k]->plan;
        ld[k]    = (long long)subs[k]->plan.nx * subs[k]->plan.ny;
    }
    fp::sub_solve_batch<T>(plans.data(), d_f, ld.data(), nrhs, nprob, (hipStream_t)stream, nprob > 1);
}
}  // namespace

extern "C" {

int fp_version(void) { return 1; }

const char* fp_status_str(int st)
{
    switch (st) {
        case FP_OK: return "ok";
        case FP_E_INVALID: return "invalid argument";
        case FP_E_NOMEM: return "out of memory";
        case FP_E_HIP: return "HIP runtime error";
        case FP_E_NOT_CONVERGED: return "GMRES did not reach the tolerance";
        case FP_E_NUMERIC: return "non-finite value in the iteration";
        case FP_E_UNSUPPORTED: return "unsupported configuration";
        default: return "unknown status";
    }
}

const char* fp_last_error(void) { return fp::last_error_slot().c_str(); }

int fp_ctx_create(int device, fp_ctx* out)
{
    return guarded([&] {
        FP_REQUIRE(out != nullptr, "null output handle");
        int n = 0;
        FP_HIP(hipGetDeviceCount(&n));
        FP_REQUIRE(device >= 0 && device < n, "no such HIP device");
        FP_HIP(hipSetDevice(device));
        auto* c   = new fp_ctx_s();
        c->device = device;
        // the device address of this GPU's scratch array (its name in host
        // code is not a device address)
        void*            ptr = nullptr;
        const hipError_t e   = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            delete c;
            FP_HIP(e);
        }
        c->scratch = static_cast<unsigned char*>(ptr);
        *out       = c;
    });
}

void fp_ctx_destroy(fp_ctx ctx) { delete ctx; }

int fp_ctx_set_planning(fp_ctx ctx, int rigor)
{
    return guarded([&] {
        FP_REQUIRE(ctx, "null context");
        FP_REQUIRE(rigor == FP_PLAN_MEASURE || rigor == FP_PLAN_ESTIMATE, "unknown planning rigor");
        ctx->measure = rigor == FP_PLAN_MEASURE;
    });
}

int fp_sub_create(fp_ctx ctx, int nx, int ny, double hx, double hy, const unsigned char bc[4], fp_sub* out)
{
    return guarded([&] {
        FP_REQUIRE(ctx && out && bc, "null argument");
        FP_HIP(hipSetDevice(ctx->device));
        auto s = std::make_unique<fp_sub_s>();
        s->ctx = ctx;
        int b[4] = {bc[0], bc[1], bc[2], bc[3]};
        s->plan.measure = ctx->measure;
        s->plan.build(nx, ny, hx, hy, b);
        *out = s.release();
    });
}

void fp_sub_destroy(fp_sub sub) { delete sub; }

int fp_sub_solve(fp_sub sub, double* d_f, int nrhs, long long ld, void* stream)
{
    return guarded([&] {
        FP_REQUIRE(sub && d_f && nrhs >= 0, "null argument");
        FP_REQUIRE(ld >= (long long)sub->plan.nx * sub->plan.ny || nrhs <= 1, "ld smaller than nx*ny");
        FP_HIP(hipSetDevice(sub->ctx->device));
        fp::sub_solve<double>(sub->plan, d_f, ld, nrhs, (hipStream_t)stream);
    });
}

int fp_sub_transform(fp_sub sub, double* d_f, int nrhs, long long ld, int inverse, void* stream)
{
    return guarded([&] {
        FP_REQUIRE(sub && d_f && nrhs >= 0, "null argument");
        FP_REQUIRE(ld >= (long long)sub->plan.nx * sub->plan.ny || nrhs <= 1, "ld smaller than nx*ny");
        FP_HIP(hipSetDevice(sub->ctx->device));
        if (nrhs) fp::launch_ytransform<double>(sub->plan.t64, d_f, ld, sub->plan.nx, nrhs, inverse ? 1 : 0,
                                                (hipStream_t)stream);
    });
}

int fp_sub_chasing(fp_sub sub, double* d_f, int nrhs, long long ld, void* stream)
{
    return guarded([&] {
        FP_REQUIRE(sub && d_f && nrhs >= 0, "null argument");
        FP_REQUIRE(ld >= (long long)sub->plan.nx * sub->plan.ny || nrhs <= 1, "ld smaller than nx*ny");
        FP_HIP(hipSetDevice(sub->ctx->device));
        if (nrhs) fp::launch_tridiag<double>(sub->plan, d_f, ld, nrhs, (hipStream_t)stream);
    });
}

int fp_sub_solve_f32(fp_sub sub, float* d_f, int nrhs, long long ld, void* stream)
{
    return guarded([&] {
        FP_REQUIRE(sub && d_f && nrhs >= 0, "null argument");
        FP_REQUIRE(ld >= (long long)sub->plan.nx * sub->plan.ny || nrhs <= 1, "ld smaller than nx*ny");
        FP_HIP(hipSetDevice(sub->ctx->device));
        fp::sub_solve<float>(sub->plan, d_f, ld, nrhs, (hipStream_t)stream);
    });
}

int fp_sub_solve_refine(fp_sub sub, double* d_f, int nrhs, long long ld, double tol, int max_steps, void* stream,
                        fp_solve_info* info)
{
    int extra = FP_OK;
    int st    = guarded([&] {
        FP_REQUIRE(sub && d_f && nrhs >= 1, "null argument");
        FP_REQUIRE(ld >= (long long)sub->plan.nx * sub->plan.ny || nrhs <= 1, "ld smaller than nx*ny");
        FP_REQUIRE(tol >= 0.0, "negative tolerance");
        FP_HIP(hipSetDevice(sub->ctx->device));
        const long long     ldd = nrhs <= 1 ? (long long)sub->plan.nx * sub->plan.ny : ld;
        const fp::RefineStats r = fp::refine_solve(sub->plan, d_f, ldd, nrhs, tol, max_steps, (hipStream_t)stream);
        if (info) {
            *info                 = fp_solve_info{};
            info->refine_steps      = r.steps;
            info->refine_residual   = r.correction;
            info->refine_correction = r.correction;
        }
        if (r.numeric) extra = FP_E_NUMERIC;
        else if (r.correction > tol) extra = FP_E_NOT_CONVERGED;
    });
    return st != FP_OK ? st : extra;
}

int fp_sub_solve_batched(fp_ctx ctx, int nprob, const fp_sub* subs, double* const* d_f, const int* nrhs, void* stream)
{
    return guarded([&] { solve_batched<double>(ctx, nprob, subs, d_f, nrhs, stream); });
}

int fp_sub_solve_batched_f32(fp_ctx ctx, int nprob, const fp_sub* subs, float* const* d_f, const int* nrhs,
                             void* stream)
{
    return guarded([&] { solve_batched<float>(ctx, nprob, subs, d_f, nrhs, stream); });
}

int fp_sub_solve_batched_noalloc(fp_ctx ctx, int nprob, const fp_sub* subs, double* const* d_f, const int* nrhs,
                                 void* stream)
{
    return guarded([&] {
        FP_REQUIRE(ctx && ctx->scratch, "null context");
        FP_REQUIRE(nprob <= FP_NOALLOC_MAX_PROBLEMS, "more problems than the no-allocation entry point takes");
        fp::TableArena ar;
        ar.base = ctx->scratch;
        ar.cap  = kNoAllocScratchBytes;
        fp::TableArena*& cur = fp::table_arena();
        cur                  = &ar;
        try {
            solve_batched<double>(ctx, nprob, subs, d_f, nrhs, stream);
        } catch (...) {
            cur = nullptr;
            throw;
        }
        cur = nullptr;
    });
}

int fp_cat_create(fp_ctx ctx, const fp_geometry_desc* g, int m, double tol, int tol_is_relative, int max_iter,
                  int precision, fp_cat* out)
{
    return guarded([&] {
        FP_REQUIRE(ctx && g && out, "null argument");
        FP_REQUIRE(g->nd >= 1 && g->nx && g->ny && g->hx && g->hy && g->bc, "empty geometry");
        FP_REQUIRE(m >= 1 && m <= 1000, "m must be in [1, 1000]");
        FP_REQUIRE(max_iter >= 1 && tol >= 0.0, "invalid GMRES parameters");
        FP_REQUIRE(precision == FP_F64 || precision == FP_F32_REFINE || precision == FP_F64_REFINE,
                   "unknown precision");
        FP_HIP(hipSetDevice(ctx->device));
        auto c       = std::make_unique<fp_cat_s>();
        c->ctx       = ctx;
        c->nd        = g->nd;
        c->m         = m;
        c->tol       = tol;
        c->tol_rel   = tol_is_relative ? 1 : 0;
        c->maxiter   = max_iter;
        c->precision = precision;
        // FP64 refinement: one correction takes the FP64 floor to about 1e-14
        if (precision == FP_F64_REFINE) c->refine_max = 2;
        const int nd = g->nd;
        c->nx.assign(g->nx, g->nx + nd);
        c->ny.assign(g->ny, g->ny + nd);
        c->hx.assign(g->hx, g->hx + nd);
        c->hy.assign(g->hy, g->hy + nd);
        c->bc.assign(g->bc, g->bc + 4 * nd);
        c->adj.assign(4 * nd, -1);
        for (int k = 0; k < g->nconn; k++) {
            const int a = g->conn[3 * k], dir = g->conn[3 * k + 1], b = g->conn[3 * k + 2];
            FP_REQUIRE(a >= 0 && a < nd && b >= 0 && b < nd && a != b && dir >= 0 && dir < 4, "bad connection");
            c->adj[4 * a + dir]       = b;
            c->adj[4 * b + (dir ^ 1)] = a;
            c->bc[4 * a + dir]        = fp::ADJACENTED;
            c->bc[4 * b + (dir ^ 1)]  = fp::ADJACENTED;
        }
        if (const char* e = std::getenv("FP_INTERFACE_PRECOND")) c->precond_kind = std::atoi(e);
        // No Dirichlet face anywhere (Neumann and periodic faces only): the
        // assembled operator has the constants as its null space, and so has
        // the interface operator of a composite.  Its explicit inverse does not
        // exist; the interface GMRES runs unpreconditioned, as the
        // reference's does (gmres_solver2d.cpp:123-252), and converges for a
        // compatible right-hand side (zero total).  A lone domain is gauged by
        // its sweep (x_0 = 0 in the zero mode, chasing_method_base.cpp:146-154).
        bool any_dirichlet = false;
        for (int k = 0; k < 4 * nd; k++)
            if (c->bc[k] == fp::DIRICHLET) any_dirichlet = true;
        if (nd > 1 && !any_dirichlet) {
            if (precision != FP_F64)
                throw fp::Error(FP_E_UNSUPPORTED, "refinement of a composite without a Dirichlet face (singular operator)");
            c->precond_kind = 0;
        }
        fp::build_tree(c.get());
        c->nodes.resize(nd);
        // construct_solver_map: deepest level first, root last (concat_solver2d.cpp:88-99)
        hipStream_t s = nullptr;
        for (int l = c->nlev - 1; l >= 0; l--)
            for (int d : c->levels[l]) fp::build_node(c.get(), d, s);
        if (nd > 1 && fp::fast_shape_ok(c.get())) {
            if (precision == FP_F32_REFINE) fp::batch_shape<float>(c.get());
            else fp::batch_shape<double>(c.get());
        }
        FP_HIP(hipDeviceSynchronize());
        *out = c.release();
    });
}

void fp_cat_destroy(fp_cat cat) { delete cat; }

int fp_cat_root(fp_cat cat) { return cat ? cat->root : -1; }

int fp_cat_transform(fp_cat cat, double* const* d_fields, int nrhs, int inverse, void* stream)
{
    return guarded([&] {
        FP_REQUIRE(cat && d_fields && nrhs >= 0, "null argument");
        for (int d = 0; d < cat->nd; d++) FP_REQUIRE(d_fields[d], "null field");
        FP_HIP(hipSetDevice(cat->ctx->device));
        if (nrhs == 0) return;
        // the batched launch of the solve's first and last stages: every
        // domain in one launch with the handle's shared shape
        const auto                        shape = fp::batch_shape<double>(cat);
        std::vector<fp::TransformJob<double>> jobs(cat->nd);
        for (int d = 0; d < cat->nd; d++) {
            fp::TransformJob<double>& j = jobs[d];
            j       = {};
            j.plan  = &cat->nodes[d]->plan.t64;
            j.data  = d_fields[d];
            j.ld    = fp::dom_n(cat, d);
            j.nx    = cat->nx[d];
            j.nrhs  = nrhs;
            j.dst64 = 1;
            fp::apply_batch_shape(j, shape);
        }
        fp::launch_ytransform_batch<double>(jobs.data(), cat->nd, inverse ? 1 : 0, (hipStream_t)stream);
    });
}

static int cat_solve_entry(fp_cat cat, const double* const* d_rhs, double* const* d_fields,
                           const double* const* d_ghost, int nrhs, void* stream, fp_solve_info* info)
{
    int st = guarded([&] {
        FP_REQUIRE(cat && d_fields && nrhs >= 1, "null argument");
        for (int d = 0; d < cat->nd; d++) FP_REQUIRE(d_fields[d] && (!d_rhs || d_rhs[d]), "null field");
        FP_HIP(hipSetDevice(cat->ctx->device));
        fp::cat_solve(cat, d_fields, d_ghost, nrhs, (hipStream_t)stream, d_rhs);
        if (info) {
            info->iterations      = cat->last_steps;
            info->restarts        = cat->last_restarts;
            info->residual        = cat->last_res;
            info->rel_residual    = cat->last_rel;
            info->refine_steps      = cat->last_refine_steps;
            info->refine_residual   = cat->last_refine_est;
            info->refine_correction = cat->last_refine_corr;
        }
    });
    if (st == FP_OK && cat && cat->last_status) st = cat->last_status;
    return st;
}

int fp_cat_solve(fp_cat cat, double* const* d_fields, const double* const* d_ghost, int nrhs, void* stream,
                 fp_solve_info* info)
{
    return cat_solve_entry(cat, nullptr, d_fields, d_ghost, nrhs, stream, info);
}

int fp_cat_solve_into(fp_cat cat, const double* const* d_rhs, double* const* d_out, const double* const* d_ghost,
                      int nrhs, void* stream, fp_solve_info* info)
{
    if (!d_rhs) return guarded([] { FP_REQUIRE(false, "null argument"); });
    return cat_solve_entry(cat, d_rhs, d_out, d_ghost, nrhs, stream, info);
}

// Records the FP64 solve of fp_cat_solve(_into) into a HIP graph, with
// `pre` and `post` (may be empty) enqueued on the capturing stream before and
// after it, into the handle's slot gexec (proj = 0) or pexec, recorded for
// the projection whose id is proj.  Throws on error.
extern "C++" void fp::cat_capture(fp_cat_s* cat, const double* const* d_rhs, double* const* d_fields,
                     const double* const* d_ghost, int nrhs, hipStream_t s,
                     const std::function<void(hipStream_t)>& pre, const std::function<void(hipStream_t)>& post,
                     unsigned long long proj)
{
    {
        FP_REQUIRE(cat && d_fields && nrhs >= 1, "null argument");
        FP_REQUIRE(s, "graph capture needs a stream other than the null stream");
        for (int d = 0; d < cat->nd; d++) FP_REQUIRE(d_fields[d] && (!d_rhs || d_rhs[d]), "null field");
        const bool refine = cat->precision == FP_F64_REFINE;
        FP_REQUIRE((cat->precision == FP_F64 || refine) && fp::fast_ok(cat),
                   "graph capture covers the FP64 solve (plain or refined) of a root with leaf children");
        fp::Node& R = *cat->nodes[cat->root];
        FP_REQUIRE(R.ch.empty() || R.precond, "graph capture needs the explicit-inverse interface preconditioner");
        FP_HIP(hipSetDevice(cat->ctx->device));
        hipGraphExec_t& slot = proj ? cat->pexec : cat->gexec;
        // every allocation before the capture
        if (!R.ch.empty()) {
            fp::ensure_ws(cat, R, nrhs);
            fp::ensure_fast_ws<double>(cat, nrhs);
        }
        if (refine) fp::ensure_refine_ws(cat, nrhs);
        FP_HIP(hipDeviceSynchronize());
        if (slot) {
            (void)hipGraphExecDestroy(slot);
            slot = nullptr;
        }
        hipGraph_t graph = nullptr;
        FP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        cat->graph_mode = true;
        try {
            if (pre) pre(s);
            if (refine) {
                if (d_rhs)
                    for (int d = 0; d < cat->nd; d++)
                        if (d_rhs[d] != d_fields[d])
                            FP_HIP(hipMemcpyAsync(d_fields[d], d_rhs[d], sizeof(double) * fp::dom_n(cat, d) * nrhs,
                                                  hipMemcpyDeviceToDevice, s));
                fp::cat_assemble(cat, d_fields, d_ghost, nrhs, s);
                fp::cat_refine(cat, d_fields, nrhs, s, true);
            } else
                fp::cat_sweep_fast(cat, d_fields, d_ghost, nrhs, s, d_rhs);
            if (post) post(s);
        } catch (...) {
            cat->graph_mode = false;
            (void)hipStreamEndCapture(s, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        cat->graph_mode = false;
        FP_HIP(hipStreamEndCapture(s, &graph));
        hipError_t e = hipGraphInstantiate(&slot, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        FP_HIP(e);
        cat->gnrhs  = nrhs;
        cat->gsteps = refine ? cat->last_refine_steps : 0;
        if (proj) cat->pexec_id = proj;
    }
}

static int graph_capture(fp_cat cat, const double* const* d_rhs, double* const* d_fields,
                         const double* const* d_ghost, int nrhs, void* stream)
{
    return guarded([&] { fp::cat_capture(cat, d_rhs, d_fields, d_ghost, nrhs, (hipStream_t)stream, {}, {}, 0); });
}

// Replays the graph in slot `proj` ? pexec : gexec (pexec only when the
// projection `proj` recorded it); with info, waits and reports the recorded
// residual check.  Returns the status.
extern "C++" int fp::cat_replay(fp_cat_s* cat, unsigned long long proj, void* stream, fp_solve_info* info)
{
    int st = guarded([&] {
        FP_REQUIRE(cat, "null argument");
        hipGraphExec_t exec = proj ? (cat->pexec_id == proj ? cat->pexec : nullptr) : cat->gexec;
        FP_REQUIRE(exec, proj ? "no captured projection" : "no captured solve");
        FP_HIP(hipSetDevice(cat->ctx->device));
        hipStream_t s = (hipStream_t)stream;
        FP_HIP(hipGraphLaunch(exec, s));
        cat->last_status = 0;
        if (!info) return;
        FP_HIP(hipStreamSynchronize(s));
        fp::Node& R = *cat->nodes[cat->root];
        cat->last_steps = cat->last_restarts = 0;
        cat->last_res = cat->last_rel = 0.0;
        if (!R.ch.empty()) fp::gmres_stats(cat, R, cat->gnrhs, 0);
        info->iterations      = cat->last_steps;
        info->restarts        = cat->last_restarts;
        info->residual        = cat->last_res;
        info->rel_residual    = cat->last_rel;
        // a replayed refinement took its recorded number of sweeps; no
        // estimate is formed (no host reads inside the graph)
        info->refine_steps      = cat->gsteps;
        info->refine_residual   = 0.0;
        info->refine_correction = 0.0;
    });
    if (st == FP_OK && cat && cat->last_status) st = cat->last_status;
    return st;
}

int fp_cat_graph_capture(fp_cat cat, double* const* d_fields, const double* const* d_ghost, int nrhs, void* stream)
{
    return graph_capture(cat, nullptr, d_fields, d_ghost, nrhs, stream);
}

int fp_cat_graph_capture_into(fp_cat cat, const double* const* d_rhs, double* const* d_out,
                              const double* const* d_ghost, int nrhs, void* stream)
{
    if (!d_rhs) return guarded([] { FP_REQUIRE(false, "null argument"); });
    return graph_capture(cat, d_rhs, d_out, d_ghost, nrhs, stream);
}

int fp_cat_graph_launch(fp_cat cat, void* stream, fp_solve_info* info)
{
    return fp::cat_replay(cat, 0, stream, info);
}

// The geometry of a composite handle, for the pressure projection.
extern "C++" fp::CatView fp::cat_view(const fp_cat_s* c)
{
    return {c->nd, c->nx.data(), c->ny.data(), c->hx.data(), c->hy.data(), c->bc.data(), c->adj.data(),
            c->ctx->device};
}

int fp_cat_set_refinement(fp_cat cat, double tol, int max_steps)
{
    return guarded([&] {
        FP_REQUIRE(cat, "null argument");
        FP_REQUIRE(tol > 0.0 && max_steps >= 2, "refinement needs tol > 0 and max_steps >= 2");
        cat->refine_tol = tol;
        cat->refine_max = max_steps;
    });
}

// Adjoint of fp_cat_solve.  The solve is u = A^-1 b with b_d = hx_d^2 (f_d + C_d g_d):
// A is the assembled scaled 5-point operator (symmetric: the interface
// coupling is 1 in both directions), C_d maps ghost data onto the face lines
// (Dirichlet -1/h^2, Neumann +-1/h).  Hence lambda = A^-1 dL/du (one more
// composite solve), dL/df_d = hx_d^2 lambda_d and
// dL/dg_{d,face} = hx_d^2 coef(face) lambda_d on that face's line.
int fp_cat_vjp(fp_cat cat, const double* const* d_gu, double* const* d_gf, double* const* d_gghost, int nrhs,
               void* stream, fp_solve_info* info)
{
    int st = guarded([&] {
        FP_REQUIRE(cat && d_gu && d_gf && nrhs >= 1, "null argument");
        for (int d = 0; d < cat->nd; d++) FP_REQUIRE(d_gu[d] && d_gf[d], "null field");
        FP_HIP(hipSetDevice(cat->ctx->device));
        hipStream_t s  = (hipStream_t)stream;
        const int   nd = cat->nd;
        cat->last_refine_steps = 0;
        cat->last_refine_est = cat->last_refine_corr = 0.0;
        // lambda = A^-1 dL/du into d_gf: dL/du is the assembled, scaled
        // right-hand side as it stands (no ghost terms, no hx^2 factor), read
        // by the forward transforms of the two-level path directly
        if (cat->precision == FP_F64 && fp::fast_ok(cat)) {
            fp::cat_sweep_fast_t<double>(cat, d_gf, nullptr, nrhs, s, d_gu, true);
        } else {
            for (int d = 0; d < nd; d++)
                if (d_gf[d] != d_gu[d])
                    FP_HIP(hipMemcpyAsync(d_gf[d], d_gu[d], sizeof(double) * fp::dom_n(cat, d) * nrhs,
                                          hipMemcpyDeviceToDevice, s));
            if (cat->precision != FP_F64) fp::cat_refine(cat, d_gf, nrhs, s);
            else fp::cat_sweep(cat, d_gf, nrhs, s);
        }
        for (int d = 0; d < nd; d++) {
            const long long n   = fp::dom_n(cat, d);
            const double    hx2 = cat->hx[d] * cat->hx[d], hx = cat->hx[d], hy = cat->hy[d];
            for (int face = 0; face < 4; face++) {
                double* out = d_gghost ? d_gghost[4 * d + face] : nullptr;
                if (!out) continue;
                const int t = cat->bc[4 * d + face];
                double    coef;
                if (t == fp::DIRICHLET) coef = face < 2 ? -1.0 / hx / hx : -1.0 / hy / hy;
                else if (t == fp::NEUMANN) {
                    const double h = face < 2 ? hx : hy;
                    coef           = (face == fp::XNEG || face == fp::YNEG) ? 1.0 / h : -1.0 / h;
                } else
                    coef = 0.0;  // Adjacented / Null faces take no ghost data
                fp::launch_face_extract(d_gf[d], cat->nx[d], cat->ny[d], n, face, hx2 * coef, out, nrhs, s);
            }
            // dL/df = hx^2 lambda, in place
            fp::launch_axpby(d_gf[d], d_gf[d], n * nrhs, hx2, 0.0, s);
        }
        if (info) {
            info->iterations        = cat->last_steps;
            info->restarts          = cat->last_restarts;
            info->residual          = cat->last_res;
            info->rel_residual      = cat->last_rel;
            info->refine_steps      = cat->last_refine_steps;
            info->refine_residual   = cat->last_refine_est;
            info->refine_correction = cat->last_refine_corr;
        }
    });
    if (st == FP_OK && cat && cat->last_status) st = cat->last_status;
    return st;
}

}  // extern "C"
