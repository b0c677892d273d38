// Navier-Stokes shear-flow driver written against the reference-named C++ API (dgfem_amd.hpp),
// the counterpart of the reference's tests/solver/shear_flow_accuracy_test.cpp:89-137: the
// manufactured shear W = (1, y, 0, 1) on [0,1]^2 (dx = 0.05), far-field
// condition W(x) on every boundary tag as a function of position, mu = 5e-6, Pr = 0.72,
// penalty 1e4, order 2, dt = 1e-5 to T = 5e-4 with frames every 50 steps, then the relative L2
// error of each primitive field (compute_variable_error, :32-79) on the device.
//
// usage: ns_shear_flow [n] [fp32|fp64] [reference|correct] [tri|quad]
//   n: squares per side (20 = dx 0.05); the third argument selects where the Euler boundary
//   flux evaluates the condition; the last one the element type (the reference test's
//   transfinite quads, or each square split into two triangles).
// prints one line of key=value pairs: elements, steps, frames, err_0..err_3.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dgfem_amd.hpp"

using namespace dgfem_amd;

static std::array<double, 4> shear_primitive(double, double y) { return {1.0, y, 0.0, 1.0}; }

static std::array<double, 4> shear_conserved(double x, double y) {
    const double g = 1.4;
    const auto W = shear_primitive(x, y);
    return {W[0], W[0] * W[1], W[0] * W[2], W[3] / (g - 1.0) + 0.5 * W[0] * (W[1] * W[1] + W[2] * W[2])};
}

int main(int argc, char** argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 20;
    const dgx_precision prec = (argc > 2 && std::strcmp(argv[2], "fp32") == 0) ? DGX_FP32 : DGX_FP64;
    const bool reference_points = !(argc > 3 && std::strcmp(argv[3], "correct") == 0);
    const bool quads = argc > 4 && std::strcmp(argv[4], "quad") == 0;
    try {
        DGMesh mesh = quads ? DGMesh::rectangular_quads(1.0 / n, 0.0, 1.0, 0.0, 1.0) : DGMesh::split_quads(n, 1.0);
        BoundaryConditionEuler shear_bc(DGX_BC_FAR_FIELD, shear_conserved);
        for (const char* tag : {"Bottom", "Right", "Top", "Left"}) mesh.set_boundary_condition_euler(tag, shear_bc);
        NavierStokesWeakFormulation wf(1.4, 5.0e-6, 0.72, 1.0e4);
        wf.set_reference_bc_points(reference_points);
        DGContext ctx(mesh, 2, wf, prec, DGX_BASIS_DUBINER);
        NavierStokesDGSolver solver(ctx);
        const double dt = 1.0e-5, T_final = 5.0e-4;
        auto frames = solver.solve(shear_conserved, T_final, dt, 50);

        // exact primitive values at the volume points, error of the final state
        const auto& I = ctx.info();
        std::vector<double> xy(size_t(I.n_elements) * I.n_volume_points * 2), W(size_t(I.n_elements) * I.n_volume_points * 4);
        check(dgx_volume_points(ctx.get(), xy.data()));
        for (size_t k = 0; k < xy.size() / 2; ++k) {
            const auto w = shear_primitive(xy[2 * k], xy[2 * k + 1]);
            for (int v = 0; v < 4; ++v) W[4 * k + v] = w[v];
        }
        double err[4];
        check(dgx_l2_error(ctx.get(), W.data(), 1, err));
        std::printf("elements=%d steps=%d frames=%zu", I.n_elements, solver.steps_taken(), frames.size());
        for (int v = 0; v < 4; ++v) std::printf(" err_%d=%.17g", v, err[v]);
        std::printf("\n");
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
