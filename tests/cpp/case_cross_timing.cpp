// Times the case-API solve path on the cross (centre n x n, arms n/2 deep,
// the boundary table of case_cross_io, h = 1/n): ConcatPoissonSolver2D::solve
// with the fields device-resident between solves (refilled from resident
// copies by field2 assignment, device to device) and with a host round trip
// per solve (right-hand side written on the host, solution read back on the
// host: PCIe included).  Prints one line: n, steps, ms per solve of each.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <iostream>

#include "focalors/case_api.h"

int main(int argc, char** argv)
{
    try {
        const int    n     = argc > 1 ? std::atoi(argv[1]) : 1024;
        const int    steps = argc > 2 ? std::atoi(argv[2]) : 20;
        const double h     = 1.0 / n;
        using L            = LocationType;
        using P            = PDEBoundaryType;
        Domain2DUniform C(n, n, n * h, n * h, "C");
        Domain2DUniform Ld("L"), Rd("R"), Bd("B"), Td("T");
        Ld.set_nx(n / 2);
        Ld.set_lx(n / 2 * h);
        Rd.set_nx(n / 2);
        Rd.set_lx(n / 2 * h);
        Bd.set_ny(n / 2);
        Bd.set_ly(n / 2 * h);
        Td.set_ny(n / 2);
        Td.set_ly(n / 2 * h);
        Geometry2D geo;
        geo.connect(&C, L::XNegative, &Ld);
        geo.connect(&C, L::XPositive, &Rd);
        geo.connect(&C, L::YNegative, &Bd);
        geo.connect(&C, L::YPositive, &Td);
        geo.check();
        geo.solve_prepare();
        Variable2D v("p");
        v.set_geometry(geo);
        field2           fC, fL, fR, fB, fT;
        Domain2DUniform* doms[5] = {&C, &Ld, &Rd, &Bd, &Td};
        field2*          fs[5]   = {&fC, &fL, &fR, &fB, &fT};
        for (int d = 0; d < 5; d++) v.set_center_field(doms[d], *fs[d]);
        v.set_boundary_type(&Ld, {{L::XNegative, P::Dirichlet}, {L::YNegative, P::Neumann}, {L::YPositive, P::Neumann}});
        v.set_boundary_type(&Rd, {{L::XPositive, P::Neumann}, {L::YNegative, P::Dirichlet}, {L::YPositive, P::Dirichlet}});
        v.set_boundary_type(&Bd, {{L::XNegative, P::Dirichlet}, {L::XPositive, P::Dirichlet}, {L::YNegative, P::Neumann}});
        v.set_boundary_type(&Td, {{L::XNegative, P::Neumann}, {L::XPositive, P::Dirichlet}, {L::YPositive, P::Neumann}});
        v.set_value([](double x, double y) { return std::sin(2 * M_PI * x) * std::cos(M_PI * y); });
        field2 rhs[5];
        for (int d = 0; d < 5; d++) rhs[d] = *fs[d];  // host copies of the right-hand sides

        ConcatPoissonSolver2D solver(&v);
        solver.set_parameter(200, 1e-12, 100);
        solver.set_relative_tolerance(true);
        // resident right-hand sides: uploaded once
        field2 res[5];
        for (int d = 0; d < 5; d++) {
            res[d] = rhs[d];
            (void)static_cast<const field2&>(res[d]).device_data();
        }
        auto timed = [&](bool roundtrip) {
            auto step = [&] {
                for (int d = 0; d < 5; d++) *fs[d] = roundtrip ? rhs[d] : res[d];
                if (roundtrip)
                    for (int d = 0; d < 5; d++) (void)fs[d]->data();  // the host writes the right-hand side
                solver.solve();
                if (roundtrip)
                    for (int d = 0; d < 5; d++) (void)static_cast<const field2*>(fs[d])->data();  // and reads u
            };
            step();
            step();
            const auto t0 = std::chrono::steady_clock::now();
            for (int s = 0; s < steps; s++) step();
            for (int d = 0; d < 5; d++) fs[d]->sync_host();
            const auto t1 = std::chrono::steady_clock::now();
            return std::chrono::duration<double, std::milli>(t1 - t0).count() / steps;
        };
        const double dev_ms  = timed(false);
        double       norm_dev = 0.0;
        for (int d = 0; d < 5; d++) norm_dev += fs[d]->dot(*fs[d]);
        const double host_ms = timed(true);
        double       norm_host = 0.0;
        for (int d = 0; d < 5; d++) norm_host += fs[d]->dot(*fs[d]);
        std::cout << "case_cross_timing n " << n << " steps " << steps << " device_resident_ms " << dev_ms
                  << " host_roundtrip_ms " << host_ms << " solution_norm " << std::sqrt(norm_dev) << " "
                  << std::sqrt(norm_host) << "\n";
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
