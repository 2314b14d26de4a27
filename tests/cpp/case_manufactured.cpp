// Manufactured solution on an L-shaped three-domain geometry through the case
// API's coordinate helpers (axis offsets, Variable2D::set_value,
// set_buffer_value / set_buffer) and the composite solver.
//   u(x, y) = sin(1.3 x + 0.4) cos(0.7 y) + x y,  lap u = -2.18 sin(1.3 x + 0.4) cos(0.7 y)
// Domains: A = [0,1]x[0,1], B = [1,2]x[0,1] at A's x+, C = [0,1]x[1,1.5] at
// A's y+.  Outer faces Dirichlet with u at the ghost centres, except C's y+
// face: Neumann with du/dy on the wall.  Prints the max-norm error for each
// n given on the command line (cells per unit length).
#include <cmath>
#include <cstdlib>
#include <iostream>

#include "focalors/case_api.h"

static double exact(double x, double y) { return std::sin(1.3 * x + 0.4) * std::cos(0.7 * y) + x * y; }
static double lap(double x, double y) { return -2.18 * std::sin(1.3 * x + 0.4) * std::cos(0.7 * y); }
static double dudy(double x, double y) { return -0.7 * std::sin(1.3 * x + 0.4) * std::sin(0.7 * y) + x; }

static double run(int n)
{
    using L = LocationType;
    using P = PDEBoundaryType;
    Domain2DUniform A(n, n, 1.0, 1.0, "A"), B("B"), C("C");
    B.set_nx(n);
    B.set_lx(1.0);
    C.set_ny(n / 2);
    C.set_ly(0.5);
    Geometry2D geo;
    geo.connect(&A, L::XPositive, &B);
    geo.connect(&A, L::YPositive, &C);
    geo.check();
    geo.axis(&A, L::XNegative);
    geo.axis(&A, L::YNegative);
    geo.solve_prepare();

    Variable2D u("u");
    u.set_geometry(geo);
    field2 fA, fB, fC;
    u.set_center_field(&A, fA);
    u.set_center_field(&B, fB);
    u.set_center_field(&C, fC);
    u.set_boundary_type(&C, L::YPositive, P::Neumann);
    u.set_boundary_type(P::Dirichlet);
    u.set_value(lap);
    u.set_buffer(exact);  // Dirichlet faces: u at the ghost centres
    auto& top = u.buffer_map[&C][L::YPositive];
    for (int i = 0; i < C.get_nx(); i++)
        top[i] = dudy(C.get_offset_x() + (i + 0.5) * C.get_hx(), C.get_offset_y() + C.get_ly());

    ConcatPoissonSolver2D solver(&u);
    solver.set_parameter(30, 1e-13, 50);
    solver.set_relative_tolerance(true);
    solver.solve();

    double err = 0.0;
    for (auto* d : {&A, &B, &C}) {
        field2& f = *u.field_map[d];
        for (int i = 0; i < f.get_nx(); i++)
            for (int j = 0; j < f.get_ny(); j++) {
                const double x = d->get_offset_x() + (i + 0.5) * d->get_hx();
                const double y = d->get_offset_y() + (j + 0.5) * d->get_hy();
                err            = std::max(err, std::fabs(f(i, j) - exact(x, y)));
            }
    }
    return err;
}

int main(int argc, char** argv)
{
    try {
        for (int k = 1; k < argc; k++) {
            const int n = std::atoi(argv[k]);
            std::cout << "n " << n << " maxerr " << run(n) << "\n";
        }
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
