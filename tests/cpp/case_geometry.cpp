// Host-only checks of the case API's geometry and variable bookkeeping (no
// device is touched): connect() inheritance of the shared extent, axis()
// offsets, boundary-type rules for adjacented faces, cell-centre set_value
// and ghost-centre set_buffer_value coordinates.  Prints "ok" and exits 0 when
// every check holds; prints the failed check and exits 1 otherwise.
#include <cmath>
#include <iostream>

#include "focalors/case_api.h"

static int failures = 0;
#define CHECK(c)                                                         \
    do {                                                                 \
        if (!(c)) {                                                      \
            std::cout << "FAILED line " << __LINE__ << ": " #c "\n";     \
            failures++;                                                  \
        }                                                                \
    } while (0)

static bool near(double a, double b) { return std::fabs(a - b) < 1e-12; }

int main()
{
    using L = LocationType;
    using P = PDEBoundaryType;
    // centre 4x6 of size 2x3, arms at all four sides and one more beyond x+
    Domain2DUniform c(4, 6, 2.0, 3.0, "c"), w("w"), e("e"), s("s"), n("n"), ee("ee");
    w.set_nx(2);
    w.set_lx(1.0);
    e.set_nx(3);
    e.set_lx(1.5);
    s.set_ny(2);
    s.set_ly(1.0);
    n.set_ny(4);
    n.set_ly(2.0);
    ee.set_nx(2);
    ee.set_lx(1.0);
    Geometry2D g;
    g.connect(&c, L::XNegative, &w);
    g.connect(&c, L::XPositive, &e);
    g.connect(&c, L::YNegative, &s);
    g.connect(&c, L::YPositive, &n);
    g.connect(&e, L::XPositive, &ee);
    CHECK(g.domains.size() == 6);
    CHECK(w.ny == 6 && near(w.ly, 3.0) && near(w.hy, 0.5) && near(w.hx, 0.5));
    CHECK(s.nx == 4 && near(s.lx, 2.0) && near(s.hx, 0.5));
    CHECK(ee.ny == 6 && near(ee.hy, 0.5));
    CHECK(g.adjacency[&w][L::XPositive] == &c && g.adjacency[&c][L::XNegative] == &w);
    CHECK(g.connections.size() == 5);
    g.check();
    CHECK(g.is_checked);

    // axis at the west arm's x- face and the south arm's y- face
    g.axis(&w, L::XNegative);
    g.axis(&s, L::YNegative);
    CHECK(near(w.offset_x, 0.0) && near(c.offset_x, 1.0) && near(e.offset_x, 3.0) && near(ee.offset_x, 4.5));
    CHECK(near(s.offset_y, 0.0) && near(c.offset_y, 1.0) && near(n.offset_y, 4.0));
    // axis at the x+ face puts that face at 0
    g.axis(&ee, L::XPositive);
    CHECK(near(ee.offset_x, -1.0) && near(e.offset_x, -2.5) && near(c.offset_x, -4.5) && near(w.offset_x, -5.5));
    g.axis(&w, L::XNegative);

    // a geometry that is not connected fails check()
    {
        Domain2DUniform a(2, 2, 1.0, 1.0, "a"), b(2, 2, 1.0, 1.0, "b");
        Geometry2D      g2;
        g2.add_domain({&a, &b});
        bool threw = false;
        try {
            g2.check();
        } catch (const std::exception&) {
            threw = true;
        }
        CHECK(threw);
    }

    Variable2D v("v");
    v.set_geometry(g);
    CHECK(v.boundary_type_map[&c][L::XNegative] == P::Adjacented);
    CHECK(v.boundary_type_map[&e][L::XPositive] == P::Adjacented);
    field2 fc, fw, fe, fs, fn, fee;
    v.set_center_field(&c, fc);
    v.set_center_field(&w, fw);
    v.set_center_field(&e, fe);
    v.set_center_field(&s, fs);
    v.set_center_field(&n, fn);
    v.set_center_field(&ee, fee);
    CHECK(fc.get_nx() == 4 && fc.get_ny() == 6 && fc.get_name() == "v_c");
    CHECK(v.buffer_map[&w][L::XNegative].size() == 6 && v.buffer_map[&w][L::YNegative].size() == 2);

    bool threw = false;
    try {
        v.set_boundary_type(&c, L::XNegative, P::Dirichlet);
    } catch (const std::exception&) {
        threw = true;
    }
    CHECK(threw);
    v.set_boundary_type(&w, L::XNegative, P::Neumann);
    v.set_boundary_type(P::Dirichlet);
    CHECK(v.boundary_type_map[&w][L::XNegative] == P::Neumann);
    CHECK(v.boundary_type_map[&w][L::YPositive] == P::Dirichlet);
    CHECK(v.boundary_type_map[&c][L::YPositive] == P::Adjacented);

    // cell centres: offset + (i + 1/2) h
    v.set_value([](double x, double y) { return 10.0 * x + y; });
    CHECK(near(fc(0, 0), 10.0 * (1.0 + 0.25) + (1.0 + 0.25)));
    CHECK(near(fee(1, 5), 10.0 * (4.5 + 0.75) + (1.0 + 2.75)));
    // ghost centres: i = -1 or nx, j = -1 or ny
    v.set_buffer_value(&w, L::XNegative, [](double x, double y) { return 10.0 * x + y; });
    CHECK(near(v.buffer_map[&w][L::XNegative][0], 10.0 * (-0.25) + 1.25));
    v.set_buffer_value(&n, L::YPositive, [](double x, double y) { return 10.0 * x + y; });
    CHECK(near(v.buffer_map[&n][L::YPositive][3], 10.0 * (1.0 + 1.75) + (4.0 + 2.25)));
    // set_buffer fills Dirichlet/Neumann faces only
    v.set_buffer([](double, double) { return 7.0; });
    CHECK(near(v.buffer_map[&e][L::YNegative][2], 7.0));
    CHECK(near(v.buffer_map[&c][L::XNegative][0], 0.0));

    // field2 arithmetic
    field2 a(2, 3), b(2, 3);
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 3; j++) {
            a(i, j) = i * 3 + j;
            b(i, j) = 1.0;
        }
    a.add_affine_transform(2.0, b, 1.0);
    CHECK(near(a(1, 2), 5.0 + 3.0));
    CHECK(near(a.sum(), 15.0 + 18.0));
    CHECK(near(b.dot(b), 6.0) && near(b.norm(), std::sqrt(6.0)));
    a *= 0.5;
    CHECK(near(a(0, 0), 1.5));
    CHECK(opposite(L::YNegative) == L::YPositive && isDirLike(P::Adjacented) && !isDirLike(P::Neumann));

    if (failures) return 1;
    std::cout << "ok\n";
    return 0;
}
