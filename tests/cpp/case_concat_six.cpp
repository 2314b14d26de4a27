// The six-domain composite case of case/validation/2d/
// concat_solver_reference.cpp:42-108, written against the same classes and
// calls: T1 8x7 in the middle, T2 (7 wide) at its x-, T3 (8 wide) at x+, T4
// (8 high) at y-, T5 (7 high) at y+, T6 (8 wide) at T3's x+; h = 1, every
// outer face Dirichlet with zero data, f(i,j) = i*ny + j + 1.  Fields T1..T6
// are appended to argv[1]; the tree root's name is printed.
#include <iostream>

#include "case_io.h"

static void fill(field2& f)
{
    for (int i = 0; i < f.get_nx(); i++)
        for (int j = 0; j < f.get_ny(); j++) f(i, j) = i * f.get_ny() + j + 1;
}

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " out.bin\n";
        return 2;
    }
    try {
        Geometry2D geo;
        const int nx_1 = 8, ny_1 = 7, nx_2 = 7, nx_3 = 8, ny_4 = 8, ny_5 = 7, nx_6 = 8;
        Domain2DUniform T1(nx_1, ny_1, nx_1, ny_1, "T1");
        Domain2DUniform T2("T2");
        T2.set_nx(nx_2);
        T2.set_lx(nx_2);
        Domain2DUniform T3("T3");
        T3.set_nx(nx_3);
        T3.set_lx(nx_3);
        Domain2DUniform T4("T4");
        T4.set_ny(ny_4);
        T4.set_ly(ny_4);
        Domain2DUniform T5("T5");
        T5.set_ny(ny_5);
        T5.set_ly(ny_5);
        Domain2DUniform T6("T6");
        T6.set_nx(nx_6);
        T6.set_lx(nx_6);

        geo.add_domain({&T1, &T2, &T3, &T4, &T5, &T6});
        geo.connect(&T1, LocationType::XNegative, &T2);
        geo.connect(&T1, LocationType::XPositive, &T3);
        geo.connect(&T1, LocationType::YNegative, &T4);
        geo.connect(&T1, LocationType::YPositive, &T5);
        geo.connect(&T3, LocationType::XPositive, &T6);
        geo.check();
        geo.solve_prepare();

        Variable2D v("v");
        v.set_geometry(geo);
        field2 v_T1, v_T2, v_T3, v_T4, v_T5, v_T6;
        v.set_center_field(&T1, v_T1);
        v.set_center_field(&T2, v_T2);
        v.set_center_field(&T3, v_T3);
        v.set_center_field(&T4, v_T4);
        v.set_center_field(&T5, v_T5);
        v.set_center_field(&T6, v_T6);
        v.set_boundary_type(PDEBoundaryType::Dirichlet);
        for (auto kv : v.field_map) fill(*kv.second);

        ConcatPoissonSolver2D solver(&v);
        solver.set_parameter(20, 1e-13, 100);
        solver.solve();
        std::cout << "root " << solver.tree_root()->name << "\n";

        RawOut out(argv[1]);
        for (field2* f : {&v_T1, &v_T2, &v_T3, &v_T4, &v_T5, &v_T6}) out.write(*f);
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
