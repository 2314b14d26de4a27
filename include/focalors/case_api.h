// C++ case API of the solve path, with the reference's class names and call
// semantics, running every solve through the C ABI (focalors_poisson.h) on a
// HIP device.  Mirrors, per class:
//   field2             base/field/field2.h:6-82 (row-major [i][j], j contiguous)
//   Domain2DUniform    base/domain/domain2d.cpp:1-116 (h = l / n, offsets)
//   Geometry2D         base/domain/geometry2d.cpp:12-310 (add_domain, connect,
//                      check, solve_prepare, axis, set_global_spatial_step)
//   Variable2D         base/domain/variable2d.cpp:25-511 (boundary types,
//                      cell-centre values, ghost-centre buffers)
//   DomainSolver2D     pe/concat/domain_solver.h:9-15
//   PoissonSolver2D    pe/poisson/poisson_solver2d.h:10-34
//   ConcatPoissonSolver2D  pe/concat/concat_solver2d.h:18-56
// field2 keeps a device mirror: the solvers work on it in place, so a field
// crosses PCIe only when the host reads or writes it between solves.
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <initializer_list>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "focalors_poisson.h"

enum class LocationType : std::uint8_t { XNegative, XPositive, YNegative, YPositive, ZNegative, ZPositive };
enum class PDEBoundaryType : std::uint8_t { Dirichlet, Neumann, Periodic, Adjacented, Null };
enum class VariablePositionType : std::uint8_t { Center, XFace, YFace, ZFace, Corner, Null };

constexpr std::array<LocationType, 4> kBoundaryLocations2D = {LocationType::XNegative, LocationType::XPositive,
                                                              LocationType::YNegative, LocationType::YPositive};

LocationType opposite(LocationType location);
bool         isDirLike(PDEBoundaryType t);

// Status codes of the C ABI surface as this exception in the C++ layer.
struct FocalorsError : std::runtime_error {
    int status;
    FocalorsError(int st, const std::string& m) : std::runtime_error(m), status(st) {}
};

class field2
{
public:
    field2() = default;
    field2(int in_nx, int in_ny, const std::string& in_name = "Default") { init(in_nx, in_ny, in_name); }
    // copies take the source's current values and stay on the side where they
    // are: a device-current source is copied device to device
    field2(const field2& o);
    field2& operator=(const field2& o);
    field2(field2&& o) noexcept;
    field2& operator=(field2&& o) noexcept;
    ~field2();
    void init(int in_nx, int in_ny, const std::string& in_name = "Default");

    // host access brings the host copy up to date first; a writable access
    // makes the device copy stale
    double&       operator()(int i, int j) { host_write(); return value[(size_t)i * ny + j]; }
    double        operator()(int i, int j) const { host_read(); return value[(size_t)i * ny + j]; }
    int           get_nx() const { return nx; }
    int           get_ny() const { return ny; }
    int           get_size_n() const { return nx * ny; }
    double*       data() { host_write(); return value.data(); }
    const double* data() const { host_read(); return value.data(); }
    std::string   get_name() const { return name; }
    void          clear(double v = 0.0);
    field2&       operator*=(double a);
    void          add_affine_transform(double a, const field2& x, double b);
    double        dot(const field2& o) const;
    double        norm() const;
    double        sum() const;
    void          print() const;

    // Device mirror (no reference counterpart; it mirrors the host storage of
    // base/field/field2.cpp:55-72): the values are current on the host, on the
    // device of focalors_context(), or on both, and each side is brought up to
    // date when it is used, so that fields stay on the device across solves
    // and cross PCIe only when the host touches them.
    double*       device_data();        // current device copy, for writing: the host copy becomes stale
    const double* device_data() const;  // current device copy, read only
    void          sync_host() const;    // the host copy brought up to date
    bool          host_current() const { return host_ok; }
    bool          device_current() const { return dev_ok; }

    // Field I/O (no reference counterpart in this tree; a time-stepping
    // caller writes and restores its fields): a little-endian binary file,
    // the 4 bytes "FPF2", nx and ny as 32-bit integers, then the nx*ny values
    // row-major as IEEE doubles.  read_binary takes the stored size.
    void write_binary(const std::string& path) const;
    void read_binary(const std::string& path);

private:
    void host_read() const
    {
        if (!host_ok) sync_host();
    }
    void host_write()
    {
        host_read();
        dev_ok = false;
    }
    void release_device();

    int                         nx = 0, ny = 0;
    mutable std::vector<double> value;
    std::string                 name = "Default";
    mutable double*             dev     = nullptr;
    mutable bool                host_ok = true, dev_ok = false;
};

class Geometry2D;

class Domain2DUniform
{
public:
    Domain2DUniform() = default;
    explicit Domain2DUniform(const std::string& in_name) : name(in_name) {}
    Domain2DUniform(int in_nx, int in_ny, double in_lx, double in_ly, const std::string& in_name);
    Domain2DUniform(int in_nx, int in_ny, const std::string& in_name) : nx(in_nx), ny(in_ny), name(in_name) {}

    void set_nx(int v) { nx = v; hx = lx / v; }
    void set_ny(int v) { ny = v; hy = ly / v; }
    void set_lx(double v) { lx = v; hx = v / nx; }
    void set_ly(double v) { ly = v; hy = v / ny; }
    void set_spatial_step(double in_hx, double in_hy) { hx = in_hx; hy = in_hy; lx = nx * hx; ly = ny * hy; }
    void set_offset_x(double v) { offset_x = v; }
    void set_offset_y(double v) { offset_y = v; }
    bool check_profile() const { return nx > 0 && ny > 0 && lx > 0.0 && ly > 0.0; }

    int    get_nx() const { return nx; }
    int    get_ny() const { return ny; }
    double get_hx() const { return hx; }
    double get_hy() const { return hy; }
    double get_lx() const { return lx; }
    double get_ly() const { return ly; }
    double get_offset_x() const { return offset_x; }
    double get_offset_y() const { return offset_y; }

    int         nx = 0, ny = 0;
    double      lx = 0.0, ly = 0.0, hx = 0.0, hy = 0.0;
    double      offset_x = 0.0, offset_y = 0.0;
    std::string name;
    Geometry2D* parent = nullptr;
};

class Geometry2D
{
public:
    void add_domain(Domain2DUniform* s);
    void add_domain(std::initializer_list<Domain2DUniform*> list);
    // connect(a, dir, b): b sits at a's side `dir`; b inherits the shared
    // extent (geometry2d.cpp:47-72)
    void connect(Domain2DUniform* a, LocationType dir, Domain2DUniform* b);
    void check();
    void solve_prepare();
    void set_global_spatial_step(double hx, double hy);
    void axis(Domain2DUniform* d, LocationType loc);

    std::vector<Domain2DUniform*>                                                            domains;
    std::unordered_map<Domain2DUniform*, std::map<LocationType, Domain2DUniform*>>          adjacency;
    std::vector<std::pair<std::pair<Domain2DUniform*, LocationType>, Domain2DUniform*>>      connections;
    bool is_checked  = false;
    bool is_prepared = false;

private:
    bool is_single_connected() const;
    void update_offset(Domain2DUniform* d, bool x);
};

class Variable2D
{
public:
    explicit Variable2D(const std::string& in_name = "Default") : name(in_name) {}
    void set_geometry(Geometry2D& g);
    void set_center_field(Domain2DUniform* s, field2& f);
    void set_boundary_type(Domain2DUniform* s, LocationType loc, PDEBoundaryType type);
    void set_boundary_type(Domain2DUniform* s, std::initializer_list<std::pair<LocationType, PDEBoundaryType>> list);
    void set_boundary_type(PDEBoundaryType type);
    void set_value(const std::function<double(double, double)>& func);
    void set_buffer_value(Domain2DUniform* s, LocationType loc, const std::function<double(double, double)>& f);
    void set_buffer(const std::function<double(double, double)>& f);

    std::string                                                                         name;
    Geometry2D*                                                                         geometry = nullptr;
    std::unordered_map<Domain2DUniform*, field2*>                                       field_map;
    std::unordered_map<Domain2DUniform*, std::map<LocationType, PDEBoundaryType>>       boundary_type_map;
    std::unordered_map<Domain2DUniform*, std::map<LocationType, std::vector<double>>>   buffer_map;
    VariablePositionType position_type = VariablePositionType::Center;

private:
    void check_geometry(Domain2DUniform* s) const;
};

class DomainSolver2D
{
public:
    virtual ~DomainSolver2D() = default;
    virtual void solve(field2& f) = 0;
};

// The device (fp_ctx) the case API solves on; device 0 unless set first.
void   focalors_set_device(int device);
fp_ctx focalors_context();

class PoissonSolver2D : public DomainSolver2D
{
public:
    PoissonSolver2D(int nx, int ny, double hx, double hy, PDEBoundaryType xneg, PDEBoundaryType xpos,
                    PDEBoundaryType yneg, PDEBoundaryType ypos);
    PoissonSolver2D(Domain2DUniform* domain, Variable2D* variable = nullptr);
    ~PoissonSolver2D() override;
    PoissonSolver2D(const PoissonSolver2D&)            = delete;
    PoissonSolver2D& operator=(const PoissonSolver2D&) = delete;
    // f: hx^2-scaled right-hand side in, solution out (poisson_solver2d.cpp:62-90)
    void solve(field2& f) override;

private:
    void   init(double hx, double hy, const std::array<PDEBoundaryType, 4>& bc);
    int    nx = 0, ny = 0;
    fp_sub sub = nullptr;
};

class ConcatPoissonSolver2D
{
public:
    explicit ConcatPoissonSolver2D(Variable2D* in_variable);
    ~ConcatPoissonSolver2D();
    ConcatPoissonSolver2D(const ConcatPoissonSolver2D&)            = delete;
    ConcatPoissonSolver2D& operator=(const ConcatPoissonSolver2D&) = delete;
    // GMRES restart length, tolerance and restarts (concat_solver2d.cpp:7-17);
    // tol is absolute unless set_relative_tolerance(true).
    void set_parameter(int in_m, double in_tol, int in_maxIter);
    void set_relative_tolerance(bool relative);
    // boundary assembly, h^2 scaling, tree sweep; fields are overwritten with
    // the solution (concat_solver2d.cpp:101-157)
    void          solve();
    fp_solve_info last_info() const { return info; }
    Domain2DUniform* tree_root() const;

    Variable2D* variable = nullptr;

private:
    friend class PressureProjection2D;
    void rebuild();
    void release();
    // device copies of the fields and the ghost data as the solve needs them
    void prepare();

    int                              m = 20, maxIter = 100;
    double                           tol = 1e-8;
    bool                             rel = false;
    fp_cat                           cat = nullptr;
    std::vector<Domain2DUniform*>    order;
    std::vector<double*>             d_fields;
    std::vector<const double*>       d_ghost_c;
    std::vector<double*>             d_ghost;
    std::vector<std::vector<double>> ghost_sent;  // the ghost data last uploaded, per face
    fp_solve_info                    info{};
};

// The pressure projection of an incompressible flow step on a staggered grid
// (fp_proj_*; no reference counterpart in this tree: it is the Navier–Stokes
// caller of ConcatPoissonSolver2D): D G p = D u* / dt, then u = u* - dt G p,
// so that the discrete divergence of u vanishes to the solve's accuracy.
// The pressure is the variable of `solver` (its centre fields receive p; its
// Dirichlet faces are outlets whose buffer_map holds the pressure there, its
// Neumann faces walls or inlets whose normal velocity stays as given).  Each
// domain's face velocities are field2s of (nx+1) x ny (u, x-faces) and
// nx x (ny+1) (v, y-faces); a face shared by two domains is held by both.
// Needs hx == hy in every domain.  The fields stay on the device between
// steps (field2's mirror): capture() records one step into a HIP graph and
// replay() runs it each time step on the same field objects, refreshed in
// place; project() is the eager step.
class PressureProjection2D
{
public:
    explicit PressureProjection2D(ConcatPoissonSolver2D& solver);
    ~PressureProjection2D();
    PressureProjection2D(const PressureProjection2D&)            = delete;
    PressureProjection2D& operator=(const PressureProjection2D&) = delete;
    void          set_velocity(Domain2DUniform* s, field2& u, field2& v);
    void          project(double dt);
    void          capture(double dt);
    void          replay();
    fp_solve_info last_info() const { return info; }

private:
    void bind();

    ConcatPoissonSolver2D*                                          solver = nullptr;
    fp_cat                                                          cat    = nullptr;  // the handle proj was made for
    fp_proj                                                         proj   = nullptr;
    void*                                                           stream = nullptr;
    std::unordered_map<Domain2DUniform*, std::pair<field2*, field2*>> vel;
    std::vector<double*>                                            du, dv, cu, cv, cp;  // bound / captured pointers
    bool                                                            captured = false;
    fp_solve_info                                                   info{};
};
