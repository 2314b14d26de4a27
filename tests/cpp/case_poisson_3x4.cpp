// The single-domain validation case of case/validation/2d/
// poisson_solver_reference.cpp:3-95: nx=3, ny=4, hx=1, hy=2,
// f(i,j) = i*ny + j + 1, every Dirichlet/Neumann combination of the four
// faces that keeps a Dirichlet face (the all-Neumann case has no solution for
// this right-hand side).  Combinations run in lexicographic order of
// (xneg, xpos, yneg, ypos) with Dirichlet < Neumann; the solutions are
// appended to argv[1].
#include <iostream>

#include "case_io.h"

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " out.bin\n";
        return 2;
    }
    try {
        RawOut out(argv[1]);
        const int nx = 3, ny = 4;
        const double hx = 1.0, hy = 2.0;
        for (int mask = 0; mask < 15; mask++) {  // 15 = all Neumann, excluded
            PDEBoundaryType b[4];
            for (int k = 0; k < 4; k++)
                b[k] = ((mask >> (3 - k)) & 1) ? PDEBoundaryType::Neumann : PDEBoundaryType::Dirichlet;
            PoissonSolver2D solver(nx, ny, hx, hy, b[0], b[1], b[2], b[3]);
            field2 f(nx, ny);
            for (int i = 0; i < nx; i++)
                for (int j = 0; j < ny; j++) f(i, j) = i * ny + j + 1;
            solver.solve(f);
            out.write(f);
        }
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
