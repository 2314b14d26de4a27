// field2's device mirror through the case API: the same two solves of one
// domain (PoissonSolver2D, then ConcatPoissonSolver2D on a two-domain
// geometry) with the host copy read between them and with no host access
// until the end give identical results; a device-current field copies device
// to device; host writes after a solve reach the next solve.
#include <cmath>
#include <iostream>

#include "focalors/case_api.h"

static double maxdiff(const field2& a, const field2& b)
{
    double m = 0.0;
    for (int i = 0; i < a.get_nx(); i++)
        for (int j = 0; j < a.get_ny(); j++) m = std::max(m, std::fabs(a(i, j) - b(i, j)));
    return m;
}

int main()
{
    try {
        const int nx = 40, ny = 28;
        field2    f0(nx, ny);
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) f0(i, j) = std::sin(0.3 * i + 0.7 * j) * 1e-3;
        PoissonSolver2D ps(nx, ny, 1.0 / nx, 1.0 / nx, PDEBoundaryType::Dirichlet, PDEBoundaryType::Neumann,
                           PDEBoundaryType::Dirichlet, PDEBoundaryType::Dirichlet);
        // with host reads between the solves
        field2 a = f0;
        ps.solve(a);
        const double probe = static_cast<const field2&>(a)(3, 4);  // a read: download, both copies current
        ps.solve(a);                                               // no upload needed
        // with no host access until the end
        field2 b = f0;
        ps.solve(b);
        ps.solve(b);
        const bool dev_between = !b.host_current() && b.device_current();
        const double d1 = maxdiff(a, b);
        // a device-current field copies device to device
        field2 c = b;
        const bool copied_on_device = c.device_current() && !c.host_current();
        const double d2 = maxdiff(c, b);
        // a host write after a solve reaches the next solve
        field2 e = f0;
        ps.solve(e);
        e(5, 5) += 1.0;
        ps.solve(e);
        field2 g = f0;
        ps.solve(g);
        g.sync_host();
        g(5, 5) += 1.0;
        field2 gh(nx, ny);
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) gh(i, j) = g(i, j);  // a host-only copy
        ps.solve(gh);
        const double d3 = maxdiff(e, gh);
        std::cout << "probe " << probe << " d1 " << d1 << " d2 " << d2 << " d3 " << d3 << " dev_between "
                  << dev_between << " copied_on_device " << copied_on_device << "\n";
        const bool ok = d1 == 0.0 && d2 == 0.0 && d3 <= 1e-15 && dev_between && copied_on_device;
        std::cout << (ok ? "mirror ok" : "mirror FAILED") << "\n";
        return ok ? 0 : 1;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}
