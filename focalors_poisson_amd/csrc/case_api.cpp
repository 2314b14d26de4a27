// C++ case API over the C ABI (see include/focalors/case_api.h for the
// reference classes each one mirrors).
#include "focalors/case_api.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <mutex>
#include <queue>
#include <unordered_set>

namespace {

void check_status(int st)
{
    if (st != FP_OK) throw FocalorsError(st, std::string(fp_status_str(st)) + ": " + fp_last_error());
}

void check_hip(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw FocalorsError(FP_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int    g_device = 0;
fp_ctx g_ctx    = nullptr;
std::mutex g_ctx_mutex;

int bc_code(PDEBoundaryType t) { return (int)t; }

}  // namespace

LocationType opposite(LocationType location)
{
    switch (location) {
        case LocationType::XNegative: return LocationType::XPositive;
        case LocationType::XPositive: return LocationType::XNegative;
        case LocationType::YNegative: return LocationType::YPositive;
        case LocationType::YPositive: return LocationType::YNegative;
        case LocationType::ZNegative: return LocationType::ZPositive;
        case LocationType::ZPositive: return LocationType::ZNegative;
    }
    throw std::invalid_argument("Invalid location type");
}

bool isDirLike(PDEBoundaryType t) { return t == PDEBoundaryType::Dirichlet || t == PDEBoundaryType::Adjacented; }

void focalors_set_device(int device)
{
    std::lock_guard<std::mutex> lk(g_ctx_mutex);
    if (g_ctx) throw std::runtime_error("focalors_set_device: context already created");
    g_device = device;
}

fp_ctx focalors_context()
{
    std::lock_guard<std::mutex> lk(g_ctx_mutex);
    if (!g_ctx) check_status(fp_ctx_create(g_device, &g_ctx));
    return g_ctx;
}

// ---- field2 (base/field/field2.cpp) ----------------------------------------

void field2::init(int in_nx, int in_ny, const std::string& in_name)
{
    release_device();
    nx      = in_nx;
    ny      = in_ny;
    name    = in_name;
    host_ok = true;
    value.assign((size_t)nx * ny, 0.0);
}

void field2::release_device()
{
    if (dev) (void)hipFree(dev);
    dev    = nullptr;
    dev_ok = false;
}

field2::~field2() { release_device(); }

field2::field2(const field2& o) { *this = o; }

field2& field2::operator=(const field2& o)
{
    if (this == &o) return *this;
    if (nx != o.nx || ny != o.ny) {
        release_device();
        nx = o.nx;
        ny = o.ny;
        value.assign((size_t)nx * ny, 0.0);
    }
    name = o.name;
    if (o.dev_ok) {
        // device to device: the copy stays where the values are; queued on the
        // null stream, ordered before the solves and host copies that follow
        if (!dev) {
            (void)focalors_context();
            check_hip(hipMalloc(&dev, sizeof(double) * std::max<size_t>(1, value.size())), "hipMalloc");
        }
        check_hip(hipMemcpyAsync(dev, o.dev, sizeof(double) * value.size(), hipMemcpyDeviceToDevice, nullptr),
                  "hipMemcpyAsync");
        dev_ok  = true;
        host_ok = false;
    } else {
        value   = o.value;
        host_ok = true;
        dev_ok  = false;
    }
    return *this;
}

field2::field2(field2&& o) noexcept { *this = std::move(o); }

field2& field2::operator=(field2&& o) noexcept
{
    if (this == &o) return *this;
    release_device();
    nx      = o.nx;
    ny      = o.ny;
    value   = std::move(o.value);
    name    = std::move(o.name);
    dev     = o.dev;
    host_ok = o.host_ok;
    dev_ok  = o.dev_ok;
    o.dev   = nullptr;
    o.dev_ok = false;
    o.nx = o.ny = 0;
    o.value.clear();
    o.host_ok = true;
    return *this;
}

const double* field2::device_data() const
{
    if (!dev) {
        (void)focalors_context();  // the device of the case API
        check_hip(hipMalloc(&dev, sizeof(double) * std::max<size_t>(1, value.size())), "hipMalloc");
        dev_ok = false;
    }
    if (!dev_ok) {
        check_hip(hipMemcpy(dev, value.data(), sizeof(double) * value.size(), hipMemcpyHostToDevice), "hipMemcpy");
        dev_ok = true;
    }
    return dev;
}

double* field2::device_data()
{
    const double* d = static_cast<const field2*>(this)->device_data();
    host_ok         = false;
    return const_cast<double*>(d);
}

void field2::sync_host() const
{
    if (host_ok) return;
    check_hip(hipMemcpy(value.data(), dev, sizeof(double) * value.size(), hipMemcpyDeviceToHost), "hipMemcpy");
    host_ok = true;
}

void field2::clear(double v)
{
    host_write();
    std::fill(value.begin(), value.end(), v);
}

field2& field2::operator*=(double a)
{
    host_write();
    for (auto& x : value) x *= a;
    return *this;
}

void field2::add_affine_transform(double a, const field2& x, double b)
{
    host_write();
    x.host_read();
    for (size_t q = 0; q < value.size(); q++) value[q] += a * x.value[q] + b;
}

double field2::dot(const field2& o) const
{
    host_read();
    o.host_read();
    double s = 0.0;
    for (size_t q = 0; q < value.size(); q++) s += value[q] * o.value[q];
    return s;
}

double field2::norm() const { return std::sqrt(dot(*this)); }

void field2::write_binary(const std::string& path) const
{
    host_read();
    FILE* fh = std::fopen(path.c_str(), "wb");
    if (!fh) throw std::runtime_error("field2::write_binary: cannot open " + path);
    const int32_t dims[2] = {nx, ny};
    const bool    ok      = std::fwrite("FPF2", 1, 4, fh) == 4 && std::fwrite(dims, sizeof(int32_t), 2, fh) == 2 &&
                   std::fwrite(value.data(), sizeof(double), value.size(), fh) == value.size();
    if (std::fclose(fh) != 0 || !ok) throw std::runtime_error("field2::write_binary: write failed for " + path);
}

void field2::read_binary(const std::string& path)
{
    FILE* fh = std::fopen(path.c_str(), "rb");
    if (!fh) throw std::runtime_error("field2::read_binary: cannot open " + path);
    char    magic[4];
    int32_t dims[2];
    bool    ok = std::fread(magic, 1, 4, fh) == 4 && std::memcmp(magic, "FPF2", 4) == 0 &&
              std::fread(dims, sizeof(int32_t), 2, fh) == 2 && dims[0] > 0 && dims[1] > 0;
    if (ok) {
        if (dims[0] != nx || dims[1] != ny) init(dims[0], dims[1], name);
        host_write();
        ok = std::fread(value.data(), sizeof(double), value.size(), fh) == value.size();
    }
    std::fclose(fh);
    if (!ok) throw std::runtime_error("field2::read_binary: not a field file or truncated: " + path);
}

double field2::sum() const
{
    host_read();
    double s = 0.0;
    for (double v : value) s += v;
    return s;
}

// field2::print (field2.cpp:382-398): comma separated rows
void field2::print() const
{
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) std::cout << (*this)(i, j) << (j == ny - 1 ? '\n' : ',');
}

// ---- Domain2DUniform / Geometry2D -------------------------------------------

Domain2DUniform::Domain2DUniform(int in_nx, int in_ny, double in_lx, double in_ly, const std::string& in_name)
    : nx(in_nx), ny(in_ny), lx(in_lx), ly(in_ly), name(in_name)
{
    hx = in_lx / in_nx;
    hy = in_ly / in_ny;
}

void Geometry2D::add_domain(Domain2DUniform* s)
{
    if (!s) throw std::invalid_argument("add_domain: nullptr");
    if (std::find(domains.begin(), domains.end(), s) == domains.end()) domains.push_back(s);
    s->parent = this;
}

void Geometry2D::add_domain(std::initializer_list<Domain2DUniform*> list)
{
    for (auto* s : list) add_domain(s);
}

void Geometry2D::connect(Domain2DUniform* a, LocationType dir, Domain2DUniform* b)
{
    add_domain({a, b});
    if (dir == LocationType::ZNegative || dir == LocationType::ZPositive)
        throw std::invalid_argument("Geometry2D does not support ZNegative/ZPositive");
    adjacency[a][dir]           = b;
    adjacency[b][opposite(dir)] = a;
    connections.push_back({{a, dir}, b});
    if (dir == LocationType::XNegative || dir == LocationType::XPositive) {
        b->set_ly(a->ly);
        b->set_ny(a->ny);
    } else {
        b->set_lx(a->lx);
        b->set_nx(a->nx);
    }
}

bool Geometry2D::is_single_connected() const
{
    if (domains.size() <= 1) return true;
    std::unordered_set<Domain2DUniform*> seen{domains[0]};
    std::queue<Domain2DUniform*>         q;
    q.push(domains[0]);
    while (!q.empty()) {
        auto* u = q.front();
        q.pop();
        auto it = adjacency.find(u);
        if (it == adjacency.end()) continue;
        for (auto& kv : it->second)
            if (seen.insert(kv.second).second) q.push(kv.second);
    }
    return seen.size() == domains.size();
}

void Geometry2D::check()
{
    for (auto* s : domains)
        if (!s->check_profile()) throw std::runtime_error("Domain " + s->name + " has invalid profile");
    if (!is_single_connected()) throw std::runtime_error("Geometry is not single connected");
    is_checked = true;
}

void Geometry2D::solve_prepare()
{
    if (!is_checked) throw std::runtime_error("Geometry2D is not checked");
    is_prepared = true;  // the tree is built by the solver (fp_cat_create)
}

void Geometry2D::set_global_spatial_step(double hx, double hy)
{
    for (auto* s : domains) s->set_spatial_step(hx, hy);
}

// Geometry2D::axis / update_offset_x / update_offset_y (geometry2d.cpp:186-270)
void Geometry2D::axis(Domain2DUniform* d, LocationType loc)
{
    if (std::find(domains.begin(), domains.end(), d) == domains.end())
        throw std::runtime_error("axis: Domain is not in geometry");
    if (loc == LocationType::XNegative || loc == LocationType::XPositive) {
        d->set_offset_x(loc == LocationType::XNegative ? 0.0 : -d->get_lx());
        update_offset(d, true);
    } else if (loc == LocationType::YNegative || loc == LocationType::YPositive) {
        d->set_offset_y(loc == LocationType::YNegative ? 0.0 : -d->get_ly());
        update_offset(d, false);
    } else
        throw std::runtime_error("axis: Invalid LocationType for 2D");
}

void Geometry2D::update_offset(Domain2DUniform* d, bool x)
{
    std::unordered_set<Domain2DUniform*> visited{d};
    std::queue<Domain2DUniform*>         q;
    q.push(d);
    while (!q.empty()) {
        auto* u = q.front();
        q.pop();
        auto it = adjacency.find(u);
        if (it == adjacency.end()) continue;
        for (auto& kv : it->second) {
            auto* v = kv.second;
            if (visited.count(v)) continue;
            if (x) {
                double nxo = u->get_offset_x();
                if (kv.first == LocationType::XPositive) nxo += u->get_lx();
                else if (kv.first == LocationType::XNegative) nxo -= v->get_lx();
                v->set_offset_x(nxo);
            } else {
                double nyo = u->get_offset_y();
                if (kv.first == LocationType::YPositive) nyo += u->get_ly();
                else if (kv.first == LocationType::YNegative) nyo -= v->get_ly();
                v->set_offset_y(nyo);
            }
            visited.insert(v);
            q.push(v);
        }
    }
}

// ---- Variable2D (base/domain/variable2d.cpp) ----------------------------------

void Variable2D::check_geometry(Domain2DUniform* s) const
{
    if (!geometry) throw std::runtime_error("Variable2D has no geometry set");
    if (s->parent != geometry) throw std::runtime_error("Domain not found in geometry");
    if (s->nx <= 0 || s->ny <= 0) throw std::runtime_error("Domain mesh size invalid");
}

void Variable2D::set_geometry(Geometry2D& g)
{
    geometry = &g;
    for (auto& kv : g.adjacency)
        for (auto& lv : kv.second) boundary_type_map[kv.first][lv.first] = PDEBoundaryType::Adjacented;
}

void Variable2D::set_center_field(Domain2DUniform* s, field2& f)
{
    check_geometry(s);
    f.init(s->nx, s->ny, name + "_" + s->name);
    field_map[s] = &f;
    auto& b      = buffer_map[s];
    b[LocationType::XNegative].assign(s->ny, 0.0);
    b[LocationType::XPositive].assign(s->ny, 0.0);
    b[LocationType::YNegative].assign(s->nx, 0.0);
    b[LocationType::YPositive].assign(s->nx, 0.0);
    position_type = VariablePositionType::Center;
}

void Variable2D::set_boundary_type(Domain2DUniform* s, LocationType loc, PDEBoundaryType type)
{
    check_geometry(s);
    auto it = geometry->adjacency.find(s);
    if (it != geometry->adjacency.end() && it->second.count(loc)) {
        if (type != PDEBoundaryType::Adjacented)
            throw std::runtime_error("Attempt to override an adjacented face with non-Adjacented boundary on domain " +
                                     s->name);
        boundary_type_map[s][loc] = PDEBoundaryType::Adjacented;
        return;
    }
    auto& m = boundary_type_map[s];
    if (m.count(loc) && m[loc] == PDEBoundaryType::Adjacented && type != PDEBoundaryType::Adjacented)
        throw std::runtime_error("Attempt to change previously Adjacented boundary to another type on domain " +
                                 s->name);
    m[loc] = type;
}

void Variable2D::set_boundary_type(Domain2DUniform* s,
                                   std::initializer_list<std::pair<LocationType, PDEBoundaryType>> list)
{
    for (auto& p : list) set_boundary_type(s, p.first, p.second);
}

void Variable2D::set_boundary_type(PDEBoundaryType type)
{
    if (!geometry) throw std::runtime_error("Variable2D has no geometry set");
    for (auto* d : geometry->domains) {
        auto& m = boundary_type_map[d];
        for (LocationType loc : kBoundaryLocations2D)
            if (!m.count(loc)) set_boundary_type(d, loc, type);
    }
}

// cell centres: offset + (i + 1/2) h  (variable2d.cpp:467-511)
void Variable2D::set_value(const std::function<double(double, double)>& func)
{
    for (auto& kv : field_map) {
        Domain2DUniform* s = kv.first;
        field2&          f = *kv.second;
        for (int i = 0; i < f.get_nx(); i++)
            for (int j = 0; j < f.get_ny(); j++)
                f(i, j) = func(s->get_offset_x() + (0.5 + i) * s->get_hx(), s->get_offset_y() + (0.5 + j) * s->get_hy());
    }
}

// ghost centres: i = -1 / nx or j = -1 / ny  (variable2d.cpp:345-416)
void Variable2D::set_buffer_value(Domain2DUniform* s, LocationType loc, const std::function<double(double, double)>& f)
{
    check_geometry(s);
    auto& b = buffer_map[s][loc];
    if (loc == LocationType::XNegative || loc == LocationType::XPositive) {
        const int i = loc == LocationType::XNegative ? -1 : s->nx;
        b.resize(s->ny);
        for (int j = 0; j < s->ny; j++)
            b[j] = f(s->get_offset_x() + (0.5 + i) * s->get_hx(), s->get_offset_y() + (0.5 + j) * s->get_hy());
    } else {
        const int j = loc == LocationType::YNegative ? -1 : s->ny;
        b.resize(s->nx);
        for (int i = 0; i < s->nx; i++)
            b[i] = f(s->get_offset_x() + (0.5 + i) * s->get_hx(), s->get_offset_y() + (0.5 + j) * s->get_hy());
    }
}

void Variable2D::set_buffer(const std::function<double(double, double)>& f)
{
    if (!geometry) throw std::runtime_error("Variable2D has no geometry set");
    for (auto& kv : field_map) {
        auto& tm = boundary_type_map[kv.first];
        for (LocationType loc : kBoundaryLocations2D) {
            auto it = tm.find(loc);
            if (it == tm.end() || it->second == PDEBoundaryType::Adjacented || it->second == PDEBoundaryType::Null)
                continue;
            set_buffer_value(kv.first, loc, f);
        }
    }
}

// ---- PoissonSolver2D -----------------------------------------------------------

PoissonSolver2D::PoissonSolver2D(int in_nx, int in_ny, double hx, double hy, PDEBoundaryType xneg,
                                 PDEBoundaryType xpos, PDEBoundaryType yneg, PDEBoundaryType ypos)
    : nx(in_nx), ny(in_ny)
{
    init(hx, hy, {xneg, xpos, yneg, ypos});
}

// an unset face reads as Dirichlet, as the reference's unordered_map::operator[]
// default (poisson_solver2d.cpp:28-31)
PoissonSolver2D::PoissonSolver2D(Domain2DUniform* domain, Variable2D* variable) : nx(domain->nx), ny(domain->ny)
{
    std::array<PDEBoundaryType, 4> bc{PDEBoundaryType::Dirichlet, PDEBoundaryType::Dirichlet,
                                      PDEBoundaryType::Dirichlet, PDEBoundaryType::Dirichlet};
    if (variable) {
        auto& m = variable->boundary_type_map[domain];
        for (int f = 0; f < 4; f++) {
            auto it = m.find(kBoundaryLocations2D[f]);
            if (it != m.end()) bc[f] = it->second;
        }
    }
    init(domain->hx, domain->hy, bc);
}

void PoissonSolver2D::init(double hx, double hy, const std::array<PDEBoundaryType, 4>& bc)
{
    const unsigned char b[4] = {(unsigned char)bc_code(bc[0]), (unsigned char)bc_code(bc[1]),
                                (unsigned char)bc_code(bc[2]), (unsigned char)bc_code(bc[3])};
    check_status(fp_sub_create(focalors_context(), nx, ny, hx, hy, b, &sub));
}

PoissonSolver2D::~PoissonSolver2D() { fp_sub_destroy(sub); }

// in place on the field's device copy (uploaded only if the host copy is newer)
void PoissonSolver2D::solve(field2& f)
{
    if (f.get_nx() != nx || f.get_ny() != ny) throw std::invalid_argument("PoissonSolver2D::solve: field size");
    check_status(fp_sub_solve(sub, f.device_data(), 1, (long long)nx * ny, nullptr));
}

// ---- ConcatPoissonSolver2D --------------------------------------------------------

ConcatPoissonSolver2D::ConcatPoissonSolver2D(Variable2D* in_variable) : variable(in_variable)
{
    if (!variable || !variable->geometry) throw std::runtime_error("ConcatPoissonSolver2D: _variable->geometry is null");
    if (!variable->geometry->is_checked) variable->geometry->check();
    if (!variable->geometry->is_prepared) variable->geometry->solve_prepare();
    rebuild();
}

ConcatPoissonSolver2D::~ConcatPoissonSolver2D() { release(); }

void ConcatPoissonSolver2D::release()
{
    for (auto* p : d_ghost)
        if (p) (void)hipFree(p);
    d_fields.clear();
    d_ghost.clear();
    d_ghost_c.clear();
    ghost_sent.clear();
    if (cat) fp_cat_destroy(cat);
    cat = nullptr;
}

void ConcatPoissonSolver2D::set_parameter(int in_m, double in_tol, int in_maxIter)
{
    const bool changed = m != in_m || tol != in_tol || maxIter != in_maxIter;
    m                  = in_m;
    tol                = in_tol;
    maxIter            = in_maxIter;
    if (changed) rebuild();
}

void ConcatPoissonSolver2D::set_relative_tolerance(bool relative)
{
    if (rel != relative) {
        rel = relative;
        rebuild();
    }
}

void ConcatPoissonSolver2D::rebuild()
{
    release();
    Geometry2D& g = *variable->geometry;
    order         = g.domains;
    std::unordered_map<Domain2DUniform*, int> idx;
    for (size_t k = 0; k < order.size(); k++) idx[order[k]] = (int)k;
    const int              nd = (int)order.size();
    std::vector<int>       nx(nd), ny(nd), conn;
    std::vector<double>    hx(nd), hy(nd);
    std::vector<unsigned char> bc(4 * nd, (unsigned char)PDEBoundaryType::Dirichlet);
    for (int d = 0; d < nd; d++) {
        auto* s = order[d];
        nx[d]   = s->nx;
        ny[d]   = s->ny;
        hx[d]   = s->hx;
        hy[d]   = s->hy;
        auto& m = variable->boundary_type_map[s];
        for (int f = 0; f < 4; f++) {
            auto it = m.find(kBoundaryLocations2D[f]);
            if (it != m.end()) bc[4 * d + f] = (unsigned char)bc_code(it->second);
        }
    }
    for (auto& c : g.connections) {
        conn.push_back(idx.at(c.first.first));
        conn.push_back((int)c.first.second);
        conn.push_back(idx.at(c.second));
    }
    fp_geometry_desc desc{nd, nx.data(), ny.data(), hx.data(), hy.data(), bc.data(), (int)(conn.size() / 3),
                          conn.data()};
    check_status(fp_cat_create(focalors_context(), &desc, m, tol, rel ? 1 : 0, maxIter, FP_F64, &cat));
    d_fields.assign(nd, nullptr);
    d_ghost.assign(4 * nd, nullptr);
    d_ghost_c.assign(4 * nd, nullptr);
    ghost_sent.assign(4 * nd, {});
    for (int d = 0; d < nd; d++) {
        for (int f = 0; f < 4; f++) {
            const PDEBoundaryType t = (PDEBoundaryType)bc[4 * d + f];
            if (t != PDEBoundaryType::Dirichlet && t != PDEBoundaryType::Neumann) continue;
            const int len = f < 2 ? ny[d] : nx[d];
            check_hip(hipMalloc(&d_ghost[4 * d + f], sizeof(double) * len), "hipMalloc");
            d_ghost_c[4 * d + f] = d_ghost[4 * d + f];
        }
    }
}

Domain2DUniform* ConcatPoissonSolver2D::tree_root() const
{
    const int r = fp_cat_root(cat);
    return r >= 0 && r < (int)order.size() ? order[r] : nullptr;
}

// The fields are solved in place on their device copies (uploaded only when
// the host copy is newer) and stay there: the host copy is brought back when
// the caller next reads it.  Ghost data are uploaded when they changed.
void ConcatPoissonSolver2D::solve()
{
    prepare();
    const int st = fp_cat_solve(cat, d_fields.data(), d_ghost_c.data(), 1, nullptr, &info);
    // the reference returns silently after maxIter; so does this layer
    if (st != FP_OK && st != FP_E_NOT_CONVERGED) check_status(st);
}

void ConcatPoissonSolver2D::prepare()
{
    const int nd = (int)order.size();
    for (int d = 0; d < nd; d++) {
        auto* s  = order[d];
        auto  it = variable->field_map.find(s);
        if (it == variable->field_map.end()) throw std::runtime_error("ConcatPoissonSolver2D: domain without a field");
        field2& f = *it->second;
        if (f.get_nx() != s->nx || f.get_ny() != s->ny)
            throw std::runtime_error("ConcatPoissonSolver2D: field size differs from its domain");
        d_fields[d] = f.device_data();
        for (int k = 0; k < 4; k++) {
            if (!d_ghost[4 * d + k]) continue;
            auto&        b   = variable->buffer_map[s][kBoundaryLocations2D[k]];
            const size_t len = k < 2 ? s->ny : s->nx;
            std::vector<double> host(len, 0.0);
            std::copy(b.begin(), b.begin() + std::min(len, b.size()), host.begin());
            if (host == ghost_sent[4 * d + k]) continue;
            check_hip(hipMemcpy(d_ghost[4 * d + k], host.data(), sizeof(double) * len, hipMemcpyHostToDevice),
                      "hipMemcpy");
            ghost_sent[4 * d + k] = std::move(host);
        }
    }
}

// ---- pressure projection --------------------------------------------------

PressureProjection2D::PressureProjection2D(ConcatPoissonSolver2D& s) : solver(&s)
{
    check_hip(hipStreamCreate(reinterpret_cast<hipStream_t*>(&stream)), "hipStreamCreate");
}

PressureProjection2D::~PressureProjection2D()
{
    if (proj) fp_proj_destroy(proj);
    if (stream) (void)hipStreamDestroy(reinterpret_cast<hipStream_t>(stream));
}

void PressureProjection2D::set_velocity(Domain2DUniform* s, field2& u, field2& v)
{
    if (u.get_nx() != s->nx + 1 || u.get_ny() != s->ny || v.get_nx() != s->nx || v.get_ny() != s->ny + 1)
        throw std::runtime_error("PressureProjection2D: u needs (nx+1) x ny values and v nx x (ny+1)");
    vel[s]   = {&u, &v};
    captured = false;
}

// device pointers of every field, the pressure's ghost data uploaded; a
// rebuilt solver (new parameters) gets a new projection handle
void PressureProjection2D::bind()
{
    if (proj && cat != solver->cat) {
        fp_proj_destroy(proj);
        proj     = nullptr;
        captured = false;
    }
    if (!proj) {
        check_status(fp_proj_create(solver->cat, &proj));
        cat = solver->cat;
    }
    solver->prepare();
    const int nd = (int)solver->order.size();
    du.assign(nd, nullptr);
    dv.assign(nd, nullptr);
    for (int d = 0; d < nd; d++) {
        auto it = vel.find(solver->order[d]);
        if (it == vel.end()) throw std::runtime_error("PressureProjection2D: a domain without velocity fields");
        du[d] = it->second.first->device_data();
        dv[d] = it->second.second->device_data();
    }
}

void PressureProjection2D::project(double dt)
{
    bind();
    const int st = fp_proj_step(proj, du.data(), dv.data(), solver->d_fields.data(), solver->d_ghost_c.data(), dt,
                                stream, &info);
    check_hip(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)), "hipStreamSynchronize");
    if (st != FP_OK && st != FP_E_NOT_CONVERGED) check_status(st);
}

void PressureProjection2D::capture(double dt)
{
    bind();
    check_status(fp_proj_capture(proj, du.data(), dv.data(), solver->d_fields.data(), solver->d_ghost_c.data(), dt,
                                 stream));
    cu       = du;
    cv       = dv;
    cp       = solver->d_fields;
    captured = true;
}

void PressureProjection2D::replay()
{
    if (!captured) throw std::runtime_error("PressureProjection2D: nothing captured");
    bind();
    if (!captured || du != cu || dv != cv || solver->d_fields != cp)
        throw std::runtime_error("PressureProjection2D: the fields moved since capture(): capture again");
    const int st = fp_proj_launch(proj, stream, &info);
    if (st != FP_OK && st != FP_E_NOT_CONVERGED) check_status(st);
}
