// dgfem_amd.hpp — C++20 mirror of the reference's compressible-solver interface on top of the
// dgx C ABI (dgx.h).  Class and method names follow the reference so a caller of
//   EulerWeakFormulation / NavierStokesWeakFormulation   (I/weak_forms/*.hpp)
//   DGAssembler::assemble_euler_residual                  (I/solver/assembler.hpp:58-59)
//   EulerDGSolver / NavierStokesDGSolver::solve           (I/solver/dg_solver.hpp:223-296)
//   apply_mass_inv / time_step_ssp_rk3 / compute_max_cfl / compute_density_range
//                                                         (S/solver/compressible_solver_base.cpp)
// can switch to the GPU path.  Eigen is not required: a state is std::vector<double> in the
// reference's DGSolution layout [K][Np][4] (S/core/solution.cpp:143-145), and a frame is the
// flattened (K, 4 Np) row-major matrix with column i*4+v (flatten_frames, :340-368).
// Errors are re-thrown as std::runtime_error / std::invalid_argument like the reference.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "dgx.h"

namespace dgfem_amd {

using StateVector = std::vector<double>;  // [K][Np][4]

inline void check(dgx_status st) {
    if (st == DGX_OK) return;
    std::string msg = dgx_last_error();
    if (st == DGX_E_INVALID) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// --------------------------------------------------------------------------------------------
// Mesh: vertices + triangles + face connectivity (periodic pairing or boundary tags).
// --------------------------------------------------------------------------------------------
// BoundaryConditionEuler (I/boundary/conditions.hpp:75-79): a constant state or a function of
// position.  FAR_FIELD: conserved state; NO_SLIP: primitive (rho, u, v, p).
struct BoundaryConditionEuler {
    using Function = std::function<std::array<double, 4>(double x, double y)>;
    BoundaryConditionEuler() = default;
    BoundaryConditionEuler(dgx_bc_type t, const std::array<double, 4>& v) : type(t), value(v) {}
    BoundaryConditionEuler(dgx_bc_type t, Function f) : type(t), function(std::move(f)) {}
    dgx_bc_type type = DGX_BC_FAR_FIELD;
    std::array<double, 4> value{};
    Function function;  // when set, replaces value
};

class DGMesh {
public:
    // elements: [K][nv] vertex indices, counter-clockwise; nv = 3 (triangles) or 4 (quads)
    DGMesh(std::vector<double> vertices, std::vector<int32_t> elements, int vertices_per_element = 3)
        : vertices_(std::move(vertices)), elements_(std::move(elements)), nv_(vertices_per_element) {
        if (nv_ != 3 && nv_ != 4) throw std::invalid_argument("elements must be triangles or quads");
        neighbors_.assign(elements_.size() * 2, -1);
        check(dgx_mesh_connect_elements(n_vertices(), vertices_.data(), n_elements(), elements_.data(), nv_, 0,
                                        neighbors_.data()));
    }
    // MeshCreator::create_rectangular_mesh(dx, use_triangles = false, xmin, xmax, ymin, ymax)
    // (S/utils/mesh_creation.cpp:49-74,212-243): int((xmax-xmin)/dx) x int((ymax-ymin)/dx)
    // structured quads with the tags Bottom/Right/Top/Left.
    static DGMesh rectangular_quads(double dx, double xmin, double xmax, double ymin, double ymax) {
        const int nx = int((xmax - xmin) / dx), ny = int((ymax - ymin) / dx);
        if (nx < 1 || ny < 1) throw std::invalid_argument("dx larger than the domain");
        std::vector<double> v(2 * size_t(nx + 1) * (ny + 1));
        std::vector<int32_t> q(4 * size_t(nx) * ny);
        check(dgx_mesh_rect_quads(nx, ny, xmin, xmax, ymin, ymax, v.data(), q.data()));
        return DGMesh(std::move(v), std::move(q), 4);
    }
    // n x n squares on [x0, x0+L]^2 split into 2 n^2 CCW triangles.
    static DGMesh split_quads(int n, double L, double x0 = 0.0, double y0 = 0.0) {
        std::vector<double> v(2 * (n + 1) * (n + 1));
        std::vector<int32_t> t(6 * n * n);
        check(dgx_mesh_split_quads(n, L, x0, y0, v.data(), t.data()));
        return DGMesh(std::move(v), std::move(t));
    }
    // DGMesh::set_periodic_boundaries("Left","Right") / ("Bottom","Top")
    void set_periodic_boundaries(const std::string& a, const std::string& b) {
        bool lr = (a == "Left" || a == "Right") && (b == "Left" || b == "Right");
        bool bt = (a == "Bottom" || a == "Top") && (b == "Bottom" || b == "Top");
        if (!lr && !bt) throw std::invalid_argument("Unknown boundary tag pair: " + a + "/" + b);
        periodic_ |= lr ? 1 : 2;
        check(dgx_mesh_connect_elements(n_vertices(), vertices_.data(), n_elements(), elements_.data(), nv_,
                                        periodic_, neighbors_.data()));
    }
    // A GMSH ASCII mesh (MSH 2.2 / 4.1) with the names of its 1-D physical groups as the
    // boundary tags (MeshCreator::create_dg_mesh_from_gmsh, S/utils/mesh_creation.cpp:195-330).
    static DGMesh read_gmsh(const std::string& path) {
        int32_t nv = 0, K = 0, nvpe = 0, nt = 0;
        check(dgx_mesh_read_gmsh(path.c_str(), &nv, &K, &nvpe, &nt, nullptr, nullptr, nullptr, nullptr));
        std::vector<double> v(2 * size_t(nv));
        std::vector<int32_t> e(size_t(K) * nvpe), nb(size_t(K) * nvpe * 2);
        std::vector<char> names(64 * size_t(std::max(nt, 1)));
        check(dgx_mesh_read_gmsh(path.c_str(), &nv, &K, &nvpe, &nt, v.data(), e.data(), nb.data(), names.data()));
        DGMesh m(std::move(v), std::move(e), nvpe);
        m.neighbors_ = std::move(nb);
        m.tags_.clear();
        for (int k = 0; k < nt; ++k) m.tags_.push_back(std::string(&names[64 * size_t(k)]));
        m.resize_tags();
        return m;
    }
    // DGMesh::set_boundary_condition_euler(tag, bc)
    void set_boundary_condition_euler(const std::string& tag, const BoundaryConditionEuler& bc) {
        for (size_t k = 0; k < tags_.size(); ++k)
            if (tag == tags_[k]) {
                if (bc.type == DGX_BC_PERIODIC)
                    throw std::runtime_error("PERIODIC BC: use set_periodic_boundaries");
                bc_types_[k] = bc.type;
                for (int v = 0; v < 4; ++v) bc_values_[4 * k + v] = bc.value[v];
                bc_functions_[k] = bc.function;
                return;
            }
        throw std::invalid_argument("Unknown boundary tag: " + tag);
    }
    const std::vector<std::string>& boundary_tags() const { return tags_; }
    int n_vertices() const { return int(vertices_.size() / 2); }
    int n_elements() const { return int(elements_.size() / nv_); }
    int vertices_per_element() const { return nv_; }
    const std::vector<double>& vertices() const { return vertices_; }
    // the function-valued condition of boundary tag k (Bottom, Right, Top, Left), or empty
    const BoundaryConditionEuler::Function& bc_function(int k) const { return bc_functions_[k]; }

    dgx_mesh_desc desc() const {
        dgx_mesh_desc d{};
        d.n_vertices = n_vertices();
        d.vertices = vertices_.data();
        d.n_elements = n_elements();
        d.elements = elements_.data();
        d.face_neighbors = neighbors_.data();
        d.n_bcs = int32_t(tags_.size());
        d.bc_types = bc_types_.data();
        d.bc_values = bc_values_.data();
        d.vertices_per_element = nv_;
        return d;
    }

private:
    std::vector<double> vertices_;
    std::vector<int32_t> elements_;
    std::vector<int32_t> neighbors_;
    int nv_ = 3;
    int periodic_ = 0;
    void resize_tags() {
        bc_types_.assign(tags_.size(), -1);
        bc_values_.assign(4 * tags_.size(), 0.0);
        bc_functions_.assign(tags_.size(), BoundaryConditionEuler::Function());
    }
    // boundary tags: the positional Bottom/Right/Top/Left of dgx_mesh_connect_elements, or a
    // GMSH file's physical group names
    std::vector<std::string> tags_{"Bottom", "Right", "Top", "Left"};
    std::vector<int32_t> bc_types_ = std::vector<int32_t>(4, -1);
    std::vector<double> bc_values_ = std::vector<double>(16, 0.0);
    std::vector<BoundaryConditionEuler::Function> bc_functions_ = std::vector<BoundaryConditionEuler::Function>(4);
};

// --------------------------------------------------------------------------------------------
// Weak forms: physics parameters and constructor checks.
// --------------------------------------------------------------------------------------------
class EulerWeakFormulation {
public:
    explicit EulerWeakFormulation(double gamma = 1.4) : gamma_(gamma) {}
    virtual ~EulerWeakFormulation() = default;
    double get_gamma() const { return gamma_; }
    int get_n_vars() const { return 4; }
    virtual bool has_viscous_terms() const { return false; }
    virtual dgx_physics physics() const { return dgx_physics{DGX_EULER, gamma_, 0.0, 0.72, 5.0, bc_point_quirk()}; }
    // The Legendre derivative at +1 as the reference evaluates it (default; enters the
    // Navier-Stokes face gradients on quads) or P'_n(1) = n(n+1)/2.
    void set_reference_legendre_endpoint(bool reference) { legendre_endpoint_reference_ = reference; }
    // Where the Euler boundary flux evaluates a function-valued condition: the reference's
    // row-major read of the face points (default, euler_weak_formulation.cpp:188-194) or the
    // face points themselves.
    void set_reference_bc_points(bool reference) { bc_points_reference_ = reference; }

protected:
    uint32_t bc_point_quirk() const {
        return (bc_points_reference_ ? uint32_t(DGX_QUIRK_EULER_BC_POINTS_REFERENCE) : 0u) |
               (legendre_endpoint_reference_ ? uint32_t(DGX_QUIRK_LEGENDRE_ENDPOINT_REFERENCE) : 0u);
    }
    double gamma_;
    bool bc_points_reference_ = true;
    bool legendre_endpoint_reference_ = true;
};

class NavierStokesWeakFormulation : public EulerWeakFormulation {
public:
    NavierStokesWeakFormulation(double gamma = 1.4, double dynamic_viscosity = 1.0e-3,
                                double prandtl = 0.72, double penalty_prefactor = 5.0,
                                bool reference_face_gradients = true, bool reference_viscous_average = true,
                                bool reference_heat_flux = true)
        : EulerWeakFormulation(gamma), mu_(dynamic_viscosity), pr_(prandtl), sigma0_(penalty_prefactor),
          quirk_(reference_face_gradients), avg_quirk_(reference_viscous_average), heat_quirk_(reference_heat_flux) {
        if (mu_ <= 0.0) throw std::invalid_argument("Dynamic viscosity must be positive for Navier-Stokes weak formulation");
        if (pr_ <= 0.0) throw std::invalid_argument("Prandtl number must be positive for Navier-Stokes weak formulation");
        if (sigma0_ <= 0.0) throw std::invalid_argument("Penalty prefactor must be positive for Navier-Stokes weak formulation");
    }
    bool has_viscous_terms() const override { return true; }
    double compute_penalty_parameter(int p, double h) const {
        return sigma0_ * mu_ * (p + 1) * (p + 1) / std::max(h, 1e-12);
    }
    dgx_physics physics() const override {
        return dgx_physics{DGX_NAVIER_STOKES, gamma_, mu_, pr_, sigma0_,
                           (quirk_ ? uint32_t(DGX_QUIRK_NS_FACE_GRAD_REFERENCE) : 0u) |
                               (avg_quirk_ ? uint32_t(DGX_QUIRK_NS_VISCOUS_AVERAGE_REFERENCE) : 0u) |
                               (heat_quirk_ ? uint32_t(DGX_QUIRK_NS_HEAT_FLUX_REFERENCE) : 0u) | bc_point_quirk()};
    }

private:
    double mu_, pr_, sigma0_;
    bool quirk_, avg_quirk_, heat_quirk_;
};

// --------------------------------------------------------------------------------------------
// Context (owns the device state) and the reference-named operations.
// --------------------------------------------------------------------------------------------
class DGContext {
public:
    // basis: Dubiner by default on triangles; quads always use the Legendre basis
    DGContext(const DGMesh& mesh, int order, const EulerWeakFormulation& wf, dgx_precision precision = DGX_FP32,
              dgx_basis basis = DGX_BASIS_DUBINER, dgx_rule rule = DGX_RULE_2P)
        : mesh_(mesh) {
        dgx_mesh_desc m = mesh_.desc();
        if (mesh_.vertices_per_element() == 4) basis = DGX_BASIS_LEGENDRE;
        dgx_space_desc s{order, basis, rule};
        dgx_physics ph = wf.physics();
        dgx_ctx* c = nullptr;
        check(dgx_create(&m, &s, &ph, precision, &c));
        ctx_.reset(c);
        check(dgx_info(c, &info_));
        // function-valued conditions: the library calls back while it tabulates the boundary
        // face points; the functions live on the heap so a moved context keeps their addresses
        fns_->resize(mesh_.boundary_tags().size());
        for (int k = 0; k < int(mesh_.boundary_tags().size()); ++k)
            if (mesh_.bc_function(k)) {
                (*fns_)[k] = mesh_.bc_function(k);
                check(dgx_set_bc_function(c, k, &call_bc, &(*fns_)[k]));
            }
    }
    dgx_ctx* get() const { return ctx_.get(); }
    const dgx_info_t& info() const { return info_; }
    const DGMesh& mesh() const { return mesh_; }

private:
    struct Del { void operator()(dgx_ctx* c) const { dgx_destroy(c); } };
    static void call_bc(double x, double y, double out[4], void* user) {
        const auto U = (*static_cast<BoundaryConditionEuler::Function*>(user))(x, y);
        for (int v = 0; v < 4; ++v) out[v] = U[v];
    }
    DGMesh mesh_;
    // heap storage: a moved context keeps the functions' addresses (the library holds them)
    std::unique_ptr<std::vector<BoundaryConditionEuler::Function>> fns_ =
        std::make_unique<std::vector<BoundaryConditionEuler::Function>>();
    std::unique_ptr<dgx_ctx, Del> ctx_;
    dgx_info_t info_{};
};

class DGAssembler {
public:
    explicit DGAssembler(DGContext& ctx) : ctx_(ctx) {}
    // residuals_out = M dU/dt, layout [K][Np][4]
    void assemble_euler_residual(const StateVector& u_coeffs, StateVector& residuals_out) {
        size_t n = size_t(ctx_.info().n_elements) * ctx_.info().n_modes * 4;
        if (u_coeffs.size() != n) throw std::invalid_argument("state has the wrong size");
        residuals_out.assign(n, 0.0);
        check(dgx_residual(ctx_.get(), u_coeffs.data(), residuals_out.data()));
    }

private:
    DGContext& ctx_;
};

// VTKWriter::write_euler_solution (S/utils/vtk_writer.cpp:72-176): `solution` is a frame in
// the (K, 4 Np) layout that solve() returns; the file is `filename`.vtk.
struct VTKWriter {
    static void write_euler_solution(DGContext& ctx, const std::vector<double>& solution, const std::string& filename,
                                     double gamma = 1.4, int resolution_factor = 3) {
        if (solution.size() != size_t(ctx.info().n_elements) * ctx.info().n_modes * 4)
            throw std::invalid_argument("solution has the wrong size");
        check(dgx_write_vtk(ctx.get(), solution.data(), filename.c_str(), gamma, resolution_factor));
    }
};

// CompressibleDGSolverBase / EulerDGSolver / NavierStokesDGSolver.
class CompressibleDGSolver {
public:
    using InitialCondition = std::function<std::array<double, 4>(double x, double y)>;

    explicit CompressibleDGSolver(DGContext& ctx) : ctx_(ctx), assembler_(ctx) {}

    DGAssembler& assembler() { return assembler_; }
    size_t state_size() const { return size_t(ctx_.info().n_elements) * ctx_.info().n_modes * 4; }

    StateVector project_initial_condition(const InitialCondition& ic) {
        const auto& I = ctx_.info();
        std::vector<double> xy(size_t(I.n_elements) * I.n_volume_points * 2);
        check(dgx_volume_points(ctx_.get(), xy.data()));
        std::vector<double> Uq(size_t(I.n_elements) * I.n_volume_points * 4);
        for (size_t k = 0; k < xy.size() / 2; ++k) {
            auto U = ic(xy[2 * k], xy[2 * k + 1]);
            for (int v = 0; v < 4; ++v) Uq[4 * k + v] = U[v];
        }
        check(dgx_project(ctx_.get(), Uq.data()));
        return get_state();
    }
    StateVector apply_mass_inv(const StateVector& vec) {
        StateVector r = vec;
        check(dgx_apply_minv(ctx_.get(), r.data()));
        return r;
    }
    StateVector time_step_ssp_rk3(const StateVector& u_n, double dt) {
        check(dgx_set_state(ctx_.get(), u_n.data()));
        check(dgx_rk3_steps(ctx_.get(), dt, 1));
        return get_state();
    }
    double compute_max_cfl(const StateVector& u, double dt) {
        check(dgx_set_state(ctx_.get(), u.data()));
        double c = 0;
        check(dgx_max_cfl(ctx_.get(), dt, &c));
        return c;
    }
    std::pair<double, double> compute_density_range(const StateVector& u) {
        check(dgx_set_state(ctx_.get(), u.data()));
        double r[2];
        check(dgx_density_range(ctx_.get(), r));
        return {r[0], r[1]};
    }
    StateVector get_state() {
        StateVector u(state_size());
        check(dgx_get_state(ctx_.get(), u.data()));
        return u;
    }

    // solve(ic, T_final, dt, save_every) as run_time_integration does it
    // (S/solver/compressible_solver_base.cpp:201-338): the step count of its
    // `while (t < T_final) { ...; t += dt; }` loop, a frame at step 0 and at the top of every
    // iteration whose step is a positive multiple of save_every (only when another step
    // follows), the NaN abort after each step, and a last frame when the final step count is
    // not a multiple of save_every.  Each frame is the (K, 4 Np) row-major matrix with column
    // i*4+v, i.e. the [K][Np][4] state itself.  steps_taken() reports the steps run.
    std::vector<std::vector<double>> solve(const InitialCondition& ic, double T_final, double dt, int save_every) {
        if (save_every <= 0) throw std::invalid_argument("save_every must be positive");
        const int n_steps = reference_step_count(T_final, dt);
        std::vector<std::vector<double>> frames;
        frames.push_back(project_initial_condition(ic));
        int step = 0;
        while (step < n_steps) {
            if (step % save_every == 0 && step > 0) frames.push_back(get_state());
            const int chunk = std::min(save_every - step % save_every, n_steps - step);
            int32_t done = 0;
            const dgx_status st = dgx_rk3_run(ctx_.get(), dt, chunk, &done);
            step += done;
            if (st == DGX_E_NONFINITE) break;
            check(st);
        }
        if (step % save_every != 0) frames.push_back(get_state());
        steps_taken_ = step;
        return frames;
    }
    int steps_taken() const { return steps_taken_; }

    // Iterations of the reference's `while (t < T_final) { ...; t += dt; }` from t = 0
    // (compressible_solver_base.cpp:242,258,284): the FP64 accumulation can add a step beyond
    // ceil(T_final / dt).
    static int reference_step_count(double T_final, double dt) {
        if (!(dt > 0.0) || !std::isfinite(T_final)) throw std::invalid_argument("dt must be positive and T_final finite");
        double t = 0.0;
        int n = 0;
        while (t < T_final) {
            t += dt;
            ++n;
        }
        return n;
    }

private:
    DGContext& ctx_;
    DGAssembler assembler_;
    int steps_taken_ = 0;
};

using EulerDGSolver = CompressibleDGSolver;
using NavierStokesDGSolver = CompressibleDGSolver;

}  // namespace dgfem_amd
