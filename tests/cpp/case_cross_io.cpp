// A five-domain cross through the case API with mixed outer faces and ghost
// data read from a file.  Centre C (n x n) with arms L (x-, n/2 x n), R (x+),
// B (y-, n x n/2) and T (y+).  Outer faces (xneg, xpos, yneg, ypos):
//   L: D . N N    R: . N D D    B: D D N .    T: N D . N
// Input argv[2]: the five fields (C, L, R, B, T), then for each domain in
// that order and each face in (xneg, xpos, yneg, ypos) order that is
// Dirichlet or Neumann, its ghost data (ny values on x faces, nx on y faces).
// Output argv[1]: the five solved fields.  argv[3] = n.
#include <cstdlib>
#include <iostream>

#include "case_io.h"

int main(int argc, char** argv)
{
    if (argc < 4) {
        std::cerr << "usage: " << argv[0] << " out.bin in.bin n\n";
        return 2;
    }
    try {
        const int    n = std::atoi(argv[3]);
        const double h = 1.0 / n;
        using L        = LocationType;
        using P        = PDEBoundaryType;
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
        field2 fC, fL, fR, fB, fT;
        Domain2DUniform* doms[5] = {&C, &Ld, &Rd, &Bd, &Td};
        field2*          fs[5]   = {&fC, &fL, &fR, &fB, &fT};
        for (int d = 0; d < 5; d++) v.set_center_field(doms[d], *fs[d]);
        v.set_boundary_type(&Ld, {{L::XNegative, P::Dirichlet}, {L::YNegative, P::Neumann}, {L::YPositive, P::Neumann}});
        v.set_boundary_type(&Rd, {{L::XPositive, P::Neumann}, {L::YNegative, P::Dirichlet}, {L::YPositive, P::Dirichlet}});
        v.set_boundary_type(&Bd, {{L::XNegative, P::Dirichlet}, {L::XPositive, P::Dirichlet}, {L::YNegative, P::Neumann}});
        v.set_boundary_type(&Td, {{L::XNegative, P::Neumann}, {L::XPositive, P::Dirichlet}, {L::YPositive, P::Neumann}});

        RawIn in(argv[2]);
        for (int d = 0; d < 5; d++) in.read(fs[d]->data(), (size_t)fs[d]->get_size_n());
        for (int d = 0; d < 5; d++)
            for (LocationType loc : kBoundaryLocations2D) {
                const P t = v.boundary_type_map[doms[d]][loc];
                if (t != P::Dirichlet && t != P::Neumann) continue;
                auto& b = v.buffer_map[doms[d]][loc];
                in.read(b.data(), b.size());
            }

        ConcatPoissonSolver2D solver(&v);
        solver.set_parameter(20, 1e-13, 100);
        solver.set_relative_tolerance(true);
        solver.solve();
        const fp_solve_info info = solver.last_info();
        std::cout << "root " << solver.tree_root()->name << " iterations " << info.iterations << " rel_residual "
                  << info.rel_residual << "\n";
        RawOut out(argv[1]);
        for (int d = 0; d < 5; d++) out.write(*fs[d]);
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
