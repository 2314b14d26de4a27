// The caller side of the solve path through the case API: the pressure
// projection of an incompressible flow step, reused every time step.  A
// channel of two domains (inlet with a parabolic profile on the left,
// walls top and bottom, outlet at pressure 0 on the right); each step adds a
// forcing to the velocity and projects it, the step replayed from one HIP
// graph with the fields resident on the device.  Checks: the first replay
// equals the eager step, the projected velocity is divergence-free to the
// solve's accuracy after every checked step, the inlet and wall velocities
// stay as given, the two copies of the interface faces agree, and the
// pressure survives a write_binary / read_binary round trip.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <iostream>

#include "focalors/case_api.h"

namespace {

const double kPi = 3.14159265358979323846;

double max_divergence(const field2& u, const field2& v, double h)
{
    double m = 0.0;
    for (int i = 0; i < v.get_nx(); i++)
        for (int j = 0; j < u.get_ny(); j++)
            m = std::max(m, std::fabs((u(i + 1, j) - u(i, j)) / h + (v(i, j + 1) - v(i, j)) / h));
    return m;
}

double max_diff(const field2& a, const field2& b)
{
    double m = 0.0;
    for (int i = 0; i < a.get_nx(); i++)
        for (int j = 0; j < a.get_ny(); j++) m = std::max(m, std::fabs(a(i, j) - b(i, j)));
    return m;
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        const int    n = 64, steps = 20;
        const double h = 1.0 / n, dt = 0.01;
        Domain2DUniform left(n, n, 1.0, 1.0, "left"), right(n / 2, n, 0.5, 1.0, "right");
        Geometry2D      geo;
        geo.add_domain({&left, &right});
        geo.connect(&left, LocationType::XPositive, &right);
        Variable2D pvar("p");
        pvar.set_geometry(geo);
        field2 pl(n, n, "p_left"), pr(n / 2, n, "p_right");
        pvar.set_center_field(&left, pl);
        pvar.set_center_field(&right, pr);
        pvar.set_boundary_type(&left, {{LocationType::XNegative, PDEBoundaryType::Neumann},
                                       {LocationType::YNegative, PDEBoundaryType::Neumann},
                                       {LocationType::YPositive, PDEBoundaryType::Neumann}});
        pvar.set_boundary_type(&right, {{LocationType::XPositive, PDEBoundaryType::Dirichlet},
                                        {LocationType::YNegative, PDEBoundaryType::Neumann},
                                        {LocationType::YPositive, PDEBoundaryType::Neumann}});
        pvar.set_buffer_value(&right, LocationType::XPositive, [](double, double) { return 0.0; });
        ConcatPoissonSolver2D solver(&pvar);
        solver.set_parameter(200, 1e-13, 100);
        solver.set_relative_tolerance(true);

        // face velocities: parabolic inflow, no flow through the walls, a
        // smooth interior field with divergence
        auto init = [&](field2& u, field2& v, int nx, double x0) {
            for (int i = 0; i <= nx; i++)
                for (int j = 0; j < n; j++) {
                    const double x = x0 + i * h, y = (j + 0.5) * h;
                    u(i, j) = 4.0 * y * (1.0 - y) + (i > 0 || x0 > 0 ? 0.1 * std::sin(2 * kPi * x) * std::cos(kPi * y) : 0.0);
                }
            for (int i = 0; i < nx; i++)
                for (int j = 0; j <= n; j++) {
                    const double x = x0 + (i + 0.5) * h, y = j * h;
                    v(i, j) = (j == 0 || j == n) ? 0.0 : 0.2 * std::sin(kPi * y) * std::cos(3 * kPi * x);
                }
        };
        field2 ul(n + 1, n), vl(n, n + 1), ur(n / 2 + 1, n), vr(n / 2, n + 1);
        init(ul, vl, n, 0.0);
        init(ur, vr, n / 2, 1.0);
        for (int j = 0; j < n; j++) ur(0, j) = ul(n, j);  // one value on the shared face
        // a forcing field with divergence, added every step (zero on the
        // inlet and wall faces)
        field2 fl(n + 1, n), fr(n / 2 + 1, n);
        for (int i = 0; i <= n; i++)
            for (int j = 0; j < n; j++) fl(i, j) = i == 0 ? 0.0 : std::cos(kPi * i * h) * std::sin(2 * kPi * (j + 0.5) * h);
        for (int i = 0; i <= n / 2; i++)
            for (int j = 0; j < n; j++) fr(i, j) = std::cos(kPi * (1.0 + i * h)) * std::sin(2 * kPi * (j + 0.5) * h);
        for (int j = 0; j < n; j++) fr(0, j) = fl(n, j);

        PressureProjection2D proj(solver);
        proj.set_velocity(&left, ul, vl);
        proj.set_velocity(&right, ur, vr);

        // the eager step on copies, for the first replay to match
        field2 el = ul, evl = vl, er = ur, evr = vr;
        {
            PressureProjection2D eager(solver);
            eager.set_velocity(&left, el, evl);
            eager.set_velocity(&right, er, evr);
            eager.project(dt);
        }
        const double d_star = std::max(max_divergence(ul, vl, h), max_divergence(ur, vr, h));

        proj.capture(dt);
        double worst_rel = 0.0, first_diff = -1.0, replay_ms = 0.0;
        for (int s = 0; s < steps; s++) {
            if (s > 0)  // u* = u + dt f on the host copy (uploaded by the next replay)
                for (int i = 0; i <= n; i++)
                    for (int j = 0; j < n; j++) {
                        ul(i, j) += dt * fl(i, j);
                        if (i <= n / 2) ur(i, j) += dt * fr(i, j);
                    }
            const double d0 = (s % 5 == 0) ? std::max(max_divergence(ul, vl, h), max_divergence(ur, vr, h)) : 0.0;
            const auto t0 = std::chrono::steady_clock::now();
            proj.replay();
            replay_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (s == 0) first_diff = std::max({max_diff(ul, el), max_diff(vl, evl), max_diff(ur, er), max_diff(vr, evr)});
            if (s % 5 == 0) {
                const double d1 = std::max(max_divergence(ul, vl, h), max_divergence(ur, vr, h));
                worst_rel = std::max(worst_rel, d1 / d0);
            }
        }
        double inlet = 0.0, walls = 0.0, shared = 0.0;
        for (int j = 0; j < n; j++) {
            const double y = (j + 0.5) * h;
            inlet  = std::max(inlet, std::fabs(ul(0, j) - 4.0 * y * (1.0 - y)));
            shared = std::max(shared, std::fabs(ul(n, j) - ur(0, j)));
        }
        for (int i = 0; i < n; i++) walls = std::max({walls, std::fabs(vl(i, 0)), std::fabs(vl(i, n))});
        // field I/O round trip of the pressure
        const std::string path = argc > 1 ? argv[1] : "case_projection_p.fpf2";
        pl.write_binary(path);
        field2 back;
        back.read_binary(path);
        const double io = max_diff(pl, back);
        std::remove(path.c_str());
        bool bad_rejected = false;
        {
            FILE* fh = std::fopen(path.c_str(), "wb");
            std::fputs("not a field", fh);
            std::fclose(fh);
            try {
                back.read_binary(path);
            } catch (const std::exception&) {
                bad_rejected = true;
            }
            std::remove(path.c_str());
        }
        const fp_solve_info info = proj.last_info();
        std::printf("projection: %d steps, %.3f ms per replayed step (incl. uploads of the forced velocity), "
                    "divergence %.2e of %.2e at worst, first replay vs eager %.2e, inlet %.1e walls %.1e "
                    "shared faces %.1e, pressure file round trip %.1e, interface residual %.2e\n",
                    steps, replay_ms / steps, worst_rel, d_star, first_diff, inlet, walls, shared, io,
                    info.rel_residual);
        const bool ok = worst_rel < 1e-11 && first_diff < 1e-12 && inlet == 0.0 && walls == 0.0 && shared < 1e-13 &&
                        io == 0.0 && bad_rejected;
        std::cout << (ok ? "projection ok" : "projection FAILED") << "\n";
        return ok ? 0 : 1;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}
