// The 3-D composite solver: ConcatPoissonSolver3D with GMRESSolver3D and the
// SchurMat3D_* blocks (pe/concat/concat_solver3d.cpp, gmres_solver3d.cpp,
// schur_mat3d.cpp) on the device.
//
// The tree sweep of concat_solver3d.cpp:97-157 over any tree of boxes: the
// faces shared by a parent and a child are index lists of cells (a face of
// nx x ny x nz cells is one plane), so the boundary assembly, the reference's
// field3::bond_add couplings and the Schur products are gathers and
// scatter-adds over those lists.  A non-leaf's GMRES iterates on its face
// unknowns (K = sum of its children's face cells) rather than on full root
// fields: the operator B = I - R A^-1 E S of the reference's
// (I - A^-1 sum S) x = A^-1 b restricted to the faces, with the batched
// CGS2 kernels of the 2-D solver; its matrix-vector product is the Schur
// products, one 3-D subdomain solve of the root and a gather.
#include "fp_solver3d.h"
#include "fp_vec.h"

#include <cmath>
#include <cstring>
#include <memory>

namespace fp {

// dst[r ldd + idst[t]] += a src[r lds + (isrc ? isrc[t] : t)], t < cnt: the
// face couplings (bond_add) and the ghost-data terms of the assembly.
__global__ void k_index_add(double* dst, long long ldd, const int* idst, const double* src, long long lds,
                            const int* isrc, int cnt, double a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const long long r = blockIdx.y;
    dst[r * ldd + idst[t]] += a * src[r * lds + (isrc ? isrc[t] : t)];
}
static void index_add(double* dst, long long ldd, const int* idst, const double* src, long long lds, const int* isrc,
                      int cnt, double a, int nrhs, hipStream_t s)
{
    if (cnt <= 0 || nrhs <= 0) return;
    hipLaunchKernelGGL(k_index_add, dim3((cnt + 255) / 256, nrhs), dim3(256), 0, s, dst, ldd, idst, src, lds, isrc,
                       cnt, a);
    FP_HIP(hipGetLastError());
}

// The cells of face `loc` of an nx x ny x nz box, in the reference's face
// order: x faces (j, k), y faces (i, k), z faces (i, j), last index fastest
// (schur_mat3d.cpp: row index jj * bnz + kk and the like).
static std::vector<int> face_cells(int loc, int nx, int ny, int nz)
{
    std::vector<int> v;
    auto             id = [&](int i, int j, int k) { return (i * ny + j) * nz + k; };
    if (loc == 0 || loc == 1) {
        const int i = loc == 0 ? 0 : nx - 1;
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) v.push_back(id(i, j, k));
    } else if (loc == 2 || loc == 3) {
        const int j = loc == 2 ? 0 : ny - 1;
        for (int i = 0; i < nx; i++)
            for (int k = 0; k < nz; k++) v.push_back(id(i, j, k));
    } else {
        const int k = loc == 4 ? 0 : nz - 1;
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) v.push_back(id(i, j, k));
    }
    return v;
}

struct Link3 {
    int            dom = -1, loc = 0, cn = 0, off = 0;
    DevBuf<int>    pcell, ccell;  // parent's face cells at loc, the child's at loc ^ 1 (same order)
    DevBuf<double> S;             // cn x cn Schur block of the child
};

struct Node3 {
    int                dom = -1;
    SubPlan3D          plan;
    std::vector<Link3> ch;
    int                K = 0;
    DevBuf<int>        cell;  // the K face cells of this node, children in link order
    // the distinct face cells (nu) and, per cell, the up to three positions
    // in the K face values that land on it (edge and corner cells belong to
    // two or three faces), in link order, -1 past the last
    int                nu = 0;
    DevBuf<int>        ucell, uk;
    // GMRES workspace for ws_nrhs systems
    int            ws_nrhs = 0, ws_m = 0, nblk = 1;
    DevBuf<double> V, Wv, Tv, X, C, H, Hacc, cs, sn, g, Y, part, part2, tolabs, resid, rnorm, G;
    DevBuf<int>    state;
    int*           h_state = nullptr;
    double*        h_resid = nullptr;
    int*           dh_state = nullptr;
    double*        dh_resid = nullptr;
    ~Node3()
    {
        if (h_state) (void)hipHostFree(h_state);
        if (h_resid) (void)hipHostFree(h_resid);
    }
};

}  // namespace fp

struct fp_cat3d_s {
    fp_ctx_s*                               ctx = nullptr;
    int                                     nd  = 0;
    std::vector<int>                        nx, ny, nz, bc, adj, child, parent, parent_loc, level;
    std::vector<double>                     hx, hy, hz;
    int                                     root = 0, nlev = 0;
    std::vector<std::vector<int>>           levels;
    std::vector<std::unique_ptr<fp::Node3>> nodes;
    std::vector<fp::DevBuf<double>>         temp;
    int                                     temp_nrhs = 0;
    std::vector<fp::DevBuf<int>>            fcell;  // nd*6: every face's cells (for the assembly)
    int                                     m = 20, tol_rel = 0, maxiter = 100;
    double                                  tol = 1e-8;
    int    last_steps = 0, last_restarts = 0, last_status = 0;
    double last_res = 0.0, last_rel = 0.0;
};

namespace fp {

using Cat3 = fp_cat3d_s;

static long long n3(const Cat3* c, int d) { return (long long)c->nx[d] * c->ny[d] * c->nz[d]; }

static void ensure_ws3(Cat3* c, Node3& nd, int nrhs)
{
    if (nd.ws_nrhs >= nrhs && nd.ws_m == c->m) return;
    const int K = nd.K, m = c->m;
    nd.nblk     = (K + 63) / 64;
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
    nd.G.alloc((size_t)nrhs * n3(c, nd.dom));
    FP_HIP(hipDeviceSynchronize());
    if (nd.h_state) (void)hipHostFree(nd.h_state);
    if (nd.h_resid) (void)hipHostFree(nd.h_resid);
    nd.h_state = nullptr;
    nd.h_resid = nullptr;
    FP_HIP(hipHostMalloc((void**)&nd.h_state, sizeof(int) * 4 * nrhs, hipHostMallocMapped | hipHostMallocCoherent));
    FP_HIP(hipHostMalloc((void**)&nd.h_resid, sizeof(double) * 2 * nrhs, hipHostMallocMapped | hipHostMallocCoherent));
    FP_HIP(hipHostGetDevicePointer((void**)&nd.dh_state, nd.h_state, 0));
    FP_HIP(hipHostGetDevicePointer((void**)&nd.dh_resid, nd.h_resid, 0));
    nd.ws_nrhs = nrhs;
    nd.ws_m    = m;
}

static GmresDev gdev3(Cat3* c, Node3& nd, int nrhs)
{
    GmresDev g{};
    g.K = nd.K; g.nrhs = nrhs; g.m = c->m; g.nblk = nd.nblk;
    g.V = nd.V.p; g.Wv = nd.Wv.p; g.X = nd.X.p; g.C = nd.C.p; g.H = nd.H.p; g.Hacc = nd.Hacc.p;
    g.cs = nd.cs.p; g.sn = nd.sn.p; g.g = nd.g.p; g.Y = nd.Y.p; g.part = nd.part.p; g.part2 = nd.part2.p;
    g.tolabs = nd.tolabs.p; g.resid = nd.resid.p; g.rnorm = nd.rnorm.p; g.state = nd.state.p;
    g.hstate = nd.dh_state; g.hres = nd.dh_resid;
    return g;
}

static void node3_solve(Cat3* c, int d, double* f, int nrhs, hipStream_t s);

// G[ucell[t]] = the sum, in link order, of w at the up to three positions
// uk[3t..3t+2] (one thread per distinct face cell, so no two write one cell)
__global__ void k_face_scatter(double* G, long long ldG, const int* ucell, const int* uk, int nu, const double* w,
                               long long ldw)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nu) return;
    const long long r  = blockIdx.y;
    const double*   wr = w + r * ldw;
    double          v  = wr[uk[3 * t]];
    if (uk[3 * t + 1] >= 0) v += wr[uk[3 * t + 1]];
    if (uk[3 * t + 2] >= 0) v += wr[uk[3 * t + 2]];
    G[r * ldG + ucell[t]] = v;
}

// G = E S v scattered onto the root's face cells: edge and corner cells
// shared by two or three faces receive every contribution, in the order of
// the reference's accumulation of the S products (child by child).
static void schur_scatter(Cat3* c, Node3& nd, const double* v, int nrhs, hipStream_t s)
{
    const long long         n = n3(c, nd.dom);
    std::vector<SchurChild> ch;
    for (auto& L : nd.ch) ch.push_back({L.S.p, L.cn, L.off, L.off, 0});
    launch_schur_gemv_all(ch.data(), (int)ch.size(), v, nd.K, nd.Tv.p, nd.K, nrhs, s);  // every face in one launch
    launch_fill<double>(nd.G.p, n * nrhs, 0.0, s);
    if (nd.nu > 0) {
        hipLaunchKernelGGL(k_face_scatter, dim3((nd.nu + 255) / 256, nrhs), dim3(256), 0, s, nd.G.p, n, nd.ucell.p,
                           nd.uk.p, nd.nu, nd.Tv.p, (long long)nd.K);
        FP_HIP(hipGetLastError());
    }
}

// out = v - R A^-1 E S v  (GMRESSolver3D::Afun on face vectors)
static void matvec3(Cat3* c, Node3& nd, const double* v, double* out, int nrhs, hipStream_t s)
{
    const long long n = n3(c, nd.dom);
    schur_scatter(c, nd, v, nrhs, s);
    sub3d_solve(nd.plan, nd.G.p, n, nrhs, s);
    launch_gather(nd.G.p, n, nd.cell.p, nd.K, v, out, nd.K, 1.0, -1.0, nrhs, s);
}

static void read_state3(Node3& nd, int nrhs, hipStream_t s)
{
    GmresDev g{};
    g.nrhs  = nrhs;
    g.state = nd.state.p;
    g.resid = nd.resid.p;
    g.rnorm = nd.rnorm.p;
    gm_publish(g, nd.dh_state, nd.dh_resid, s);
    FP_HIP(hipStreamSynchronize(s));
}

// GMRESSolver3D::solve (gmres_solver3d.cpp): x0 = the right-hand side, the
// reference's tolerance semantics, restarts of length m.
static void gmres3(Cat3* c, Node3& nd, int nrhs, hipStream_t s)
{
    const int K = nd.K, m = c->m;
    FP_HIP(hipMemcpyAsync(nd.X.p, nd.C.p, sizeof(double) * K * nrhs, hipMemcpyDeviceToDevice, s));
    const GmresDev g = gdev3(c, nd, nrhs);
    gm_init(g, c->tol, c->tol_rel, s);
    int restarts = 0;
    for (int outer = 0; outer < c->maxiter; outer++) {
        matvec3(c, nd, nd.X.p, nd.Wv.p, nrhs, s);
        if (gm_resid(g, s)) FP_HIP(hipStreamSynchronize(s));
        else read_state3(nd, nrhs, s);
        bool all_done = true;
        for (int r = 0; r < nrhs; r++)
            if (nd.h_state[4 * r + 2] == 0) all_done = false;
        if (all_done) break;
        const long long ldv = (long long)nrhs * K;
        for (int j = 0; j < m; j++) {
            matvec3(c, nd, nd.V.p + j * ldv, nd.Wv.p, nrhs, s);
            gm_arnoldi(g, j, s);
            if ((j + 1) % 4 == 0 || j == m - 1) {
                read_state3(nd, nrhs, s);
                bool any = false;
                for (int r = 0; r < nrhs; r++)
                    if (nd.h_state[4 * r + 0]) any = true;
                if (!any) break;
            }
        }
        gm_finish_cycle(g, s);
        restarts++;
    }
    read_state3(nd, nrhs, s);
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

// A non-leaf: b <- A^-1 b; C = R b; GMRES for the face values y; then
// u = A^-1 b + A^-1 E S y.
static void gmres_node3(Cat3* c, Node3& nd, double* b, int nrhs, hipStream_t s)
{
    ensure_ws3(c, nd, nrhs);
    const long long n = n3(c, nd.dom);
    sub3d_solve(nd.plan, b, n, nrhs, s);
    launch_gather(b, n, nd.cell.p, nd.K, nullptr, nd.C.p, nd.K, 0.0, 1.0, nrhs, s);
    gmres3(c, nd, nrhs, s);
    schur_scatter(c, nd, nd.X.p, nrhs, s);
    sub3d_solve(nd.plan, nd.G.p, n, nrhs, s);
    launch_axpby(nd.G.p, b, n * nrhs, 1.0, 1.0, s);
}

static void node3_solve(Cat3* c, int d, double* f, int nrhs, hipStream_t s)
{
    Node3& nd = *c->nodes[d];
    if (nd.ch.empty()) sub3d_solve(nd.plan, f, n3(c, d), nrhs, s);
    else gmres_node3(c, nd, f, nrhs, s);
}

// GMRESSolver3D::schur_mat_construct with SchurMat3D_*::construct
// (schur_mat3d.cpp): unit sources on the child's face next to the parent,
// solved in batches, the response on that face collected column by column.
static void build_schur3(Cat3* c, Link3& L, hipStream_t s)
{
    const long long nc = n3(c, L.dom);
    const int       cn = L.cn;
    L.S.alloc((size_t)cn * cn);
    const long long budget = 1LL << 27;  // doubles of scratch (1 GiB)
    const int       chunk  = (int)std::max<long long>(1, std::min<long long>(cn, budget / nc));
    DevBuf<double>  buf((size_t)chunk * nc);
    for (int i0 = 0; i0 < cn; i0 += chunk) {
        const int cnt = std::min(chunk, cn - i0);
        launch_fill<double>(buf.p, nc * cnt, 0.0, s);
        launch_set_units(buf.p, nc, L.ccell.p, i0, cnt, s);
        node3_solve(c, L.dom, buf.p, cnt, s);
        launch_schur_collect(buf.p, nc, L.ccell.p, cn, i0, cnt, L.S.p, s);
    }
    FP_HIP(hipStreamSynchronize(s));
}

static void build_node3(Cat3* c, int d, hipStream_t s)
{
    auto nd = std::make_unique<Node3>();
    nd->dom = d;
    int bc6[6];
    for (int f = 0; f < 6; f++) bc6[f] = c->bc[6 * d + f];
    nd->plan.measure = c->ctx->measure;
    nd->plan.build(c->nx[d], c->ny[d], c->nz[d], c->hx[d], c->hy[d], c->hz[d], bc6);
    std::vector<int> cells;
    for (int loc = 0; loc < 6; loc++) {
        const int chd = c->child[6 * d + loc];
        if (chd < 0) continue;
        Link3 L;
        L.dom = chd;
        L.loc = loc;
        std::vector<int> pc = face_cells(loc, c->nx[d], c->ny[d], c->nz[d]);
        std::vector<int> cc = face_cells(loc ^ 1, c->nx[chd], c->ny[chd], c->nz[chd]);
        FP_REQUIRE(pc.size() == cc.size(), "connected faces must have the same cells");
        L.cn  = (int)pc.size();
        L.off = (int)cells.size();
        cells.insert(cells.end(), pc.begin(), pc.end());
        L.pcell.alloc(pc.size());
        L.pcell.upload(pc.data(), pc.size());
        L.ccell.alloc(cc.size());
        L.ccell.upload(cc.data(), cc.size());
        nd->ch.push_back(std::move(L));
    }
    nd->K = (int)cells.size();
    if (nd->K) {
        nd->cell.alloc(cells.size());
        nd->cell.upload(cells.data(), cells.size());
        std::vector<int> first(n3(c, d), -1), uc, uk;
        for (int k = 0; k < nd->K; k++) {
            int& u = first[cells[k]];
            if (u < 0) {
                u = (int)uc.size();
                uc.push_back(cells[k]);
                uk.insert(uk.end(), {k, -1, -1});
            } else {
                int q = 1;
                while (q < 3 && uk[3 * u + q] >= 0) q++;
                FP_REQUIRE(q < 3, "a face cell on more than three faces");
                uk[3 * u + q] = k;
            }
        }
        nd->nu = (int)uc.size();
        nd->ucell.alloc(uc.size());
        nd->ucell.upload(uc.data(), uc.size());
        nd->uk.alloc(uk.size());
        nd->uk.upload(uk.data(), uk.size());
    }
    c->nodes[d] = std::move(nd);
    for (auto& L : c->nodes[d]->ch) build_schur3(c, L, s);
}

// TreeUtils root choice and BFS levels (base/domain/geometry_tree.hpp:22-200)
// over the six faces; a two-centre tie goes to the lower domain index.
static void build_tree3(Cat3* c)
{
    const int nd = c->nd;
    c->root      = 0;
    if (nd > 1) {
        std::vector<int> deg(nd, 0), q;
        for (int a = 0; a < nd; a++)
            for (int l = 0; l < 6; l++)
                if (c->adj[6 * a + l] >= 0) deg[a]++;
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
                for (int l = 0; l < 6; l++) {
                    const int nb = c->adj[6 * leaf + l];
                    if (nb >= 0 && --deg[nb] == 1) q.push_back(nb);
                }
            }
        }
        if (qh < q.size()) {
            c->root = q[qh];
            if (qh + 1 < q.size() && q[qh + 1] < c->root) c->root = q[qh + 1];
        }
    }
    c->child.assign(6 * nd, -1);
    c->parent.assign(nd, -1);
    c->parent_loc.assign(nd, -1);
    c->level.assign(nd, -1);
    std::vector<int> order{c->root};
    c->level[c->root] = 0;
    for (size_t h = 0; h < order.size(); h++) {
        const int a = order[h];
        for (int l = 0; l < 6; l++) {
            const int b = c->adj[6 * a + l];
            if (b < 0 || b == c->parent[a] || c->level[b] >= 0) continue;
            c->child[6 * a + l] = b;
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

// ConcatPoissonSolver3D::boundary_assembly and the hx^2 scaling
// (concat_solver3d.cpp:105-111, 159-247): Dirichlet faces f -= g / h^2,
// Neumann faces f += g / h (negative faces) or f -= g / h (positive faces).
static void assemble3(Cat3* c, double* const* f, const double* const* ghost, int nrhs, hipStream_t s)
{
    for (int d = 0; d < c->nd; d++) {
        const long long n = n3(c, d);
        const double    h[3] = {c->hx[d], c->hy[d], c->hz[d]};
        for (int face = 0; face < 6; face++) {
            const double* g = ghost ? ghost[6 * d + face] : nullptr;
            const int     t = c->bc[6 * d + face];
            if (!g || (t != DIRICHLET && t != NEUMANN)) continue;
            const double hh = h[face / 2];
            const double a  = t == DIRICHLET ? -1.0 / (hh * hh) : ((face & 1) ? -1.0 / hh : 1.0 / hh);
            const int    cnt = (int)c->fcell[6 * d + face].n;
            index_add(f[d], n, c->fcell[6 * d + face].p, g, cnt, nullptr, cnt, a, nrhs, s);
        }
        launch_axpby(f[d], f[d], n * nrhs, c->hx[d] * c->hx[d], 0.0, s);
    }
}

// field3::bond_add(loc, a, nb) between a parent and a child: the face cells of
// `dst` at its side `loc` receive a times the neighbour's adjacent face.
static void bond3(Cat3* c, int dst, double* fd, int src, const double* fs, int loc, double a, int nrhs, hipStream_t s)
{
    (void)loc;  // the parent-child link holds both index lists
    const Link3* L = nullptr;
    if (c->parent[src] == dst) {
        for (const Link3& l : c->nodes[dst]->ch)
            if (l.dom == src) L = &l;
        index_add(fd, n3(c, dst), L->pcell.p, fs, n3(c, src), L->ccell.p, L->cn, a, nrhs, s);
    } else {
        for (const Link3& l : c->nodes[src]->ch)
            if (l.dom == dst) L = &l;
        index_add(fd, n3(c, dst), L->ccell.p, fs, n3(c, src), L->pcell.p, L->cn, a, nrhs, s);
    }
}

// ConcatPoissonSolver3D::solve (concat_solver3d.cpp:97-157) on assembled,
// scaled right-hand sides.
static void sweep3(Cat3* c, double* const* f, int nrhs, hipStream_t s)
{
    const int nd = c->nd;
    if (c->temp_nrhs < nrhs) {
        c->temp.clear();
        c->temp.resize(nd);
        for (int d = 0; d < nd; d++)
            if (d != c->root) c->temp[d].alloc((size_t)n3(c, d) * nrhs);
        c->temp_nrhs = nrhs;
    }
    for (int l = c->nlev - 1; l >= 1; l--) {
        for (int d : c->levels[l]) {
            FP_HIP(hipMemcpyAsync(c->temp[d].p, f[d], sizeof(double) * n3(c, d) * nrhs, hipMemcpyDeviceToDevice, s));
            for (int loc = 0; loc < 6; loc++) {
                const int ch = c->child[6 * d + loc];
                if (ch < 0) continue;
                bond3(c, d, c->temp[d].p, ch, c->temp[ch].p, loc, -1.0, nrhs, s);
                bond3(c, d, f[d], ch, c->temp[ch].p, loc, -1.0, nrhs, s);
            }
        }
        for (int d : c->levels[l]) node3_solve(c, d, c->temp[d].p, nrhs, s);
    }
    const int r = c->root;
    for (int loc = 0; loc < 6; loc++) {
        const int ch = c->child[6 * r + loc];
        if (ch >= 0) bond3(c, r, f[r], ch, c->temp[ch].p, loc, -1.0, nrhs, s);
    }
    c->last_steps = c->last_restarts = c->last_status = 0;
    c->last_res = c->last_rel = 0.0;
    node3_solve(c, r, f[r], nrhs, s);
    for (int l = 1; l < c->nlev; l++) {
        for (int d : c->levels[l]) {
            const int p = c->parent[d];
            bond3(c, d, f[d], p, f[p], c->parent_loc[d], -1.0, nrhs, s);
        }
        for (int d : c->levels[l]) node3_solve(c, d, f[d], nrhs, s);
    }
}

}  // namespace fp

extern "C" {

int fp_cat3d_create(fp_ctx ctx, const fp_geometry3d_desc* g, int m, double tol, int tol_is_relative, int max_iter,
                    fp_cat3d* out)
{
    return fp::guarded([&] {
        FP_REQUIRE(ctx && g && out, "null argument");
        FP_REQUIRE(g->nd >= 1 && g->nx && g->ny && g->nz && g->hx && g->hy && g->hz && g->bc, "empty geometry");
        FP_REQUIRE(m >= 1 && m <= 1000, "m must be in [1, 1000]");
        FP_REQUIRE(max_iter >= 1 && tol >= 0.0, "invalid GMRES parameters");
        FP_HIP(hipSetDevice(ctx->device));
        auto c     = std::make_unique<fp_cat3d_s>();
        c->ctx     = ctx;
        c->nd      = g->nd;
        c->m       = m;
        c->tol     = tol;
        c->tol_rel = tol_is_relative ? 1 : 0;
        c->maxiter = max_iter;
        const int nd = g->nd;
        c->nx.assign(g->nx, g->nx + nd);
        c->ny.assign(g->ny, g->ny + nd);
        c->nz.assign(g->nz, g->nz + nd);
        c->hx.assign(g->hx, g->hx + nd);
        c->hy.assign(g->hy, g->hy + nd);
        c->hz.assign(g->hz, g->hz + nd);
        c->bc.assign(g->bc, g->bc + 6 * nd);
        c->adj.assign(6 * nd, -1);
        for (int k = 0; k < g->nconn; k++) {
            const int a = g->conn[3 * k], dir = g->conn[3 * k + 1], b = g->conn[3 * k + 2];
            FP_REQUIRE(a >= 0 && a < nd && b >= 0 && b < nd && a != b && dir >= 0 && dir < 6, "bad connection");
            c->adj[6 * a + dir]       = b;
            c->adj[6 * b + (dir ^ 1)] = a;
            // Variable3D::set_geometry: connected faces become Adjacented
            c->bc[6 * a + dir]       = fp::ADJACENTED;
            c->bc[6 * b + (dir ^ 1)] = fp::ADJACENTED;
        }
        // no Dirichlet face anywhere: the assembled operator is singular
        bool any_dirichlet = false;
        for (int k = 0; k < 6 * nd; k++)
            if (c->bc[k] == fp::DIRICHLET) any_dirichlet = true;
        if (nd > 1 && !any_dirichlet)
            throw fp::Error(FP_E_UNSUPPORTED, "a 3-D composite without a Dirichlet face (singular operator)");
        c->fcell.resize(6 * nd);
        for (int d = 0; d < nd; d++)
            for (int f = 0; f < 6; f++) {
                const std::vector<int> v = fp::face_cells(f, c->nx[d], c->ny[d], c->nz[d]);
                c->fcell[6 * d + f].alloc(v.size());
                c->fcell[6 * d + f].upload(v.data(), v.size());
            }
        fp::build_tree3(c.get());
        c->nodes.resize(nd);
        hipStream_t s = nullptr;
        for (int l = c->nlev - 1; l >= 0; l--)
            for (int d : c->levels[l]) fp::build_node3(c.get(), d, s);
        FP_HIP(hipDeviceSynchronize());
        *out = c.release();
    });
}

void fp_cat3d_destroy(fp_cat3d cat) { delete cat; }

int fp_cat3d_root(fp_cat3d cat) { return cat ? cat->root : -1; }

int fp_cat3d_solve(fp_cat3d cat, double* const* d_fields, const double* const* d_ghost, int nrhs, void* stream,
                   fp_solve_info* info)
{
    int st = fp::guarded([&] {
        FP_REQUIRE(cat && d_fields && nrhs >= 1, "null argument");
        for (int d = 0; d < cat->nd; d++) FP_REQUIRE(d_fields[d], "null field");
        FP_HIP(hipSetDevice(cat->ctx->device));
        hipStream_t s = (hipStream_t)stream;
        fp::assemble3(cat, d_fields, d_ghost, nrhs, s);
        fp::sweep3(cat, d_fields, nrhs, s);
        if (info) {
            *info              = fp_solve_info{};
            info->iterations   = cat->last_steps;
            info->restarts     = cat->last_restarts;
            info->residual     = cat->last_res;
            info->rel_residual = cat->last_rel;
        }
    });
    if (st == FP_OK && cat && cat->last_status) st = cat->last_status;
    return st;
}

}  // extern "C"
