// Isentropic-vortex driver written against the reference-named C++ API (dgfem_amd.hpp), the
// counterpart of the reference's examples/euler_vortex_test.cpp:24-103: build a periodic mesh,
// construct the solver, project the vortex, run SSP-RK3 and report diagnostics.
//
// usage: euler_vortex [n] [order] [steps] [fp32|fp64] [monomial|dubiner]
// prints one line of key=value pairs: elements, dt, steps, the four conservation sums before
// and after, the density range and the relative L2 density error against the exact vortex.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dgfem_amd.hpp"

using namespace dgfem_amd;

static std::array<double, 4> vortex(double x, double y, double t, double L) {
    const double g = 1.4, beta = 5.0, xc = 5.0, yc = 5.0;
    x = std::fmod(std::fmod(x - t, L) + L, L);
    double dx = x - xc, dy = y - yc, r2 = dx * dx + dy * dy;
    double f = std::exp(0.5 * (1.0 - r2));
    double u = 1.0 - beta / (2 * M_PI) * f * dy, v = beta / (2 * M_PI) * f * dx;
    double T = 1.0 - (g - 1.0) * beta * beta / (8 * g * M_PI * M_PI) * std::exp(1.0 - r2);
    double rho = std::pow(T, 1.0 / (g - 1.0)), p = std::pow(rho, g);
    return {rho, rho * u, rho * v, p / (g - 1.0) + 0.5 * rho * (u * u + v * v)};
}

int main(int argc, char** argv) {
    int n = argc > 1 ? std::atoi(argv[1]) : 32;
    int order = argc > 2 ? std::atoi(argv[2]) : 2;
    int steps = argc > 3 ? std::atoi(argv[3]) : 100;
    dgx_precision prec = (argc > 4 && std::strcmp(argv[4], "fp32") == 0) ? DGX_FP32 : DGX_FP64;
    dgx_basis basis = (argc > 5 && std::strcmp(argv[5], "monomial") == 0) ? DGX_BASIS_MONOMIAL : DGX_BASIS_DUBINER;
    const double L = 10.0;
    try {
        DGMesh mesh = DGMesh::split_quads(n, L);
        mesh.set_periodic_boundaries("Left", "Right");
        mesh.set_periodic_boundaries("Bottom", "Top");
        EulerWeakFormulation wf(1.4);
        DGContext ctx(mesh, order, wf, prec, basis);
        EulerDGSolver solver(ctx);

        auto ic = [&](double x, double y) { return vortex(x, y, 0.0, L); };
        StateVector u0 = solver.project_initial_condition(ic);
        // dt from the CFL measure of the reference scaled by 1/(2p+1), CFL 0.3
        double lam = solver.compute_max_cfl(u0, 1.0);
        double dt = 0.3 / ((2 * order + 1) * lam);
        double s0[4], s1[4];
        check(dgx_conservation_sums(ctx.get(), s0));
        // T_final = steps dt and save_every = steps, as the reference's driver loop counts them
        // (its FP64 accumulation of t may add one step; solver.steps_taken() reports the count)
        auto frames = solver.solve(ic, steps * dt, dt, steps);
        const int taken = solver.steps_taken();
        check(dgx_conservation_sums(ctx.get(), s1));
        // the final state is the context's (with an aligned step count it is not a frame)
        auto [rmin, rmax] = solver.compute_density_range(solver.get_state());

        // density error against the exact vortex at the volume points
        const auto& I = ctx.info();
        std::vector<double> xy(size_t(I.n_elements) * I.n_volume_points * 2), Uq(size_t(I.n_elements) * I.n_volume_points * 4);
        check(dgx_volume_points(ctx.get(), xy.data()));
        check(dgx_eval_volume(ctx.get(), Uq.data()));
        double num = 0, den = 0;
        for (size_t k = 0; k < xy.size() / 2; ++k) {
            double r = vortex(xy[2 * k], xy[2 * k + 1], taken * dt, L)[0];
            num += (Uq[4 * k] - r) * (Uq[4 * k] - r);
            den += r * r;
        }
        std::printf("elements=%d order=%d dt=%.17g steps=%d frames=%zu", I.n_elements, order, dt, taken, frames.size());
        for (int v = 0; v < 4; ++v) std::printf(" sum0_%d=%.17g sum1_%d=%.17g", v, s0[v], v, s1[v]);
        std::printf(" rho_min=%.17g rho_max=%.17g l2_rho=%.6e\n", rmin, rmax, std::sqrt(num / den));
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
