/*
 * dgx.h — C ABI of the MI355X-native explicit DG residual for the 2-D compressible Euler and
 * Navier-Stokes equations (modal DG on triangles, Rusanov flux, SSP-RK3).
 *
 * The reference exposes this path as C++ virtuals and member functions; each entry point below
 * names the reference interface it replaces (paths relative to the reference root:
 * S/ = dgfem/src, I/ = dgfem/include/dgfem).
 *
 * Conventions
 *   - Plain C types only.  Every call returns dgx_status; dgx_last_error() describes the most
 *     recent failure on the calling thread (the reference throws std::runtime_error /
 *     std::invalid_argument instead; the C++ mirror in dgfem_amd.hpp re-throws).
 *   - Host state layout is [K][Np][4] (element, mode, variable), FP64: the reference's
 *     DGSolution flat index e*Np*4 + i*4 + v (S/core/solution.cpp:143-145) and its frame
 *     column order i*4+v (S/solver/compressible_solver_base.cpp:355-361).  The device layout
 *     is private to the library.
 *   - One context = one mesh + space + physics + precision, one HIP stream, device buffers
 *     owned by the context.  Contexts are not thread-safe; use one per host thread.
 */
#ifndef DGX_H
#define DGX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DGX_ABI_VERSION 1

typedef int32_t dgx_status;
enum {
    DGX_OK = 0,
    DGX_E_INVALID = 1,     /* bad argument (the reference's std::invalid_argument)        */
    DGX_E_UNSUPPORTED = 2, /* valid request outside what this build implements            */
    DGX_E_DEVICE = 3,      /* HIP runtime error, or the HIP library/device is unavailable  */
    DGX_E_NONFINITE = 4,   /* NaN in the state after a step (the reference aborts its loop) */
    DGX_E_NOMEM = 5
};

/* Triangles: monomial or Dubiner; quads: the Legendre tensor basis (LegendreBasis,
 * S/basis/legendre.cpp:69-112), the only quad basis of the reference (S/core/space.cpp:309-318). */
enum dgx_basis { DGX_BASIS_MONOMIAL = 0, DGX_BASIS_DUBINER = 1, DGX_BASIS_LEGENDRE = 2 };
/* Volume rule on triangles: DGX_RULE_2P = Dunavant exact to degree 2p, as the reference
 * (S/core/space.cpp:37-46); DGX_RULE_2P1 = exact to 2p+1 (variant).  Quads: tensor
 * Gauss-Legendre with p+1 points per direction (:47-55; DGX_RULE_2P only).  Face rule: GL, p+1
 * points. */
enum dgx_rule { DGX_RULE_2P = 0, DGX_RULE_2P1 = 1 };
/* Precision of the state and the bulk arithmetic. Geometry, inverse-mass scaling and the
 * conservation sums are FP64 in every mode. DGX_FP32_BF16VOL rounds the operands of the
 * volume contraction to bfloat16 (FP32 accumulation). */
enum dgx_precision { DGX_FP64 = 0, DGX_FP32 = 1, DGX_FP32_BF16VOL = 2 };
enum dgx_physics_kind { DGX_EULER = 0, DGX_NAVIER_STOKES = 1 };
/* BCTypeEuler (I/boundary/conditions.hpp:27-32). */
enum dgx_bc_type { DGX_BC_FAR_FIELD = 0, DGX_BC_SLIP_WALL = 1, DGX_BC_NO_SLIP_WALL = 2, DGX_BC_PERIODIC = 3 };
/* Compatibility switches for reference behaviours. */
enum dgx_quirk {
    /* NS face gradients read through the reference's column-major Map of the interleaved
     * dphi_dx_face storage (S/core/space.cpp:143,153-158 vs S/core/mesh.cpp:187-195). */
    DGX_QUIRK_NS_FACE_GRAD_REFERENCE = 1u << 0,
    /* Euler boundary faces evaluate the condition at the reference's row-major read of the
     * column-major face-point storage (S/core/space.cpp:141,148,164 vs
     * S/weak_forms/euler_weak_formulation.cpp:188-194); without it, at the face points.  Only
     * position-dependent conditions (dgx_set_bc_function) see a difference; the viscous
     * boundary flux reads the points correctly either way (navier_stokes_...cpp:244-247). */
    DGX_QUIRK_EULER_BC_POINTS_REFERENCE = 1u << 1,
    /* Legendre derivative at x = +1 from the reference's endpoint formula, which returns
     * -P'_n(1) for odd n (S/basis/legendre.cpp:31-42); it enters the Navier-Stokes face
     * gradients on quads.  Without it, P'_n(1) = n(n+1)/2. */
    DGX_QUIRK_LEGENDRE_ENDPOINT_REFERENCE = 1u << 2,
    /* NS viscous face fluxes as the reference forms them: on interior faces each side's flux
     * along its own outward normal, 1/2 Fv_L.n_L + 1/2 Fv_R.n_R = the half difference
     * 1/2 (Fv_L - Fv_R).n_L, and on every face subtracted like the inviscid flux
     * (S/weak_forms/navier_stokes_weak_formulation.cpp:192-203,262-271).  Without it, the
     * consistent flux +{Fv}.n: M dU/dt gains 1/2 (Fv_L + Fv_R).n (interior) and Fv.n
     * (boundary); the penalty is the same either way. */
    DGX_QUIRK_NS_VISCOUS_AVERAGE_REFERENCE = 1u << 3,
    /* NS energy viscous flux u.tau + q with q = -kappa grad T as the reference forms it
     * (S/weak_forms/navier_stokes_weak_formulation.cpp:76-77,91,95): heat flows up the
     * temperature gradient, so the temperature anti-diffuses at a rate ~ kappa/h^2.  Without
     * it, u.tau - q (conduction). */
    DGX_QUIRK_NS_HEAT_FLUX_REFERENCE = 1u << 4
};

/* A boundary condition as a function of position (BoundaryConditionEuler's function,
 * S/boundary/conditions.cpp:47-60): out = the FAR_FIELD conserved state or the NO_SLIP
 * primitive state at (x, y). */
typedef void (*dgx_bc_fn)(double x, double y, double out[4], void* user);

typedef struct dgx_ctx dgx_ctx;

/* Mesh: straight-sided triangles or quadrilaterals (nv = 3 or 4 vertices per element,
 * counter-clockwise).  face_neighbors[(e*nv+f)*2 + {0,1}] = (neighbour element, neighbour
 * local face) for local face f = edge (v_f, v_{(f+1)%nv}), or (-1, bc index) on a boundary
 * face.  Build it with dgx_mesh_connect / dgx_mesh_connect_elements
 * (DGMesh::build_face_connectivity + set_periodic_boundaries, S/core/mesh.cpp:312-359,482-588). */
typedef struct {
    int32_t n_vertices;
    const double* vertices;       /* [n_vertices][2] */
    int32_t n_elements;
    const int32_t* elements;      /* [K][nv] */
    const int32_t* face_neighbors;/* [K][nv][2] */
    int32_t n_bcs;                /* boundary-condition table referenced by boundary faces */
    const int32_t* bc_types;      /* [n_bcs] dgx_bc_type */
    const double* bc_values;      /* [n_bcs][4]: FAR_FIELD conserved state, NO_SLIP primitive */
    int32_t vertices_per_element; /* nv: 3 (or 0) triangles, 4 quadrilaterals */
} dgx_mesh_desc;

typedef struct {
    int32_t order; /* p = 1..4 */
    int32_t basis; /* dgx_basis */
    int32_t rule;  /* dgx_rule */
} dgx_space_desc;

/* EulerWeakFormulation(gamma) / NavierStokesWeakFormulation(gamma, mu, Pr, sigma0)
 * (I/solver/dg_solver.hpp:264,283-285; S/weak_forms/navier_stokes_weak_formulation.cpp:18-34). */
typedef struct {
    int32_t kind;   /* dgx_physics_kind */
    double gamma;
    double mu;
    double prandtl;
    double sigma0;
    uint32_t quirks; /* dgx_quirk bits */
} dgx_physics;

typedef struct {
    int32_t n_elements, n_modes, n_volume_points, n_face_points;
    int32_t order, basis, precision, physics;
    int32_t n_boundary_faces;
    int32_t device;
    double bytes_per_stage_alg; /* algorithmic HBM bytes of one fused stage launch */
    int32_t stage_kernel;       /* 0: one element per lane; 1: 64-element matrix-core tiles
                                   (FP32 triangles: Euler p = 1..4, Navier-Stokes p = 3, 4); 2: four
                                   lanes per element (Euler quadrilaterals, p <= 3);
                                   DGX_STAGE_KERNEL=lane at creation selects 0 */
} dgx_info_t;

/* ---- mesh helpers (host) ------------------------------------------------------------------ */

/* Structured periodic mesh: n x n squares on [x0,x0+L]^2, each split along its (i,j)->(i+1,j+1)
 * diagonal into two CCW triangles; vertices [(n+1)^2][2], elements [2 n^2][3]. Replaces the
 * GMSH rectangle generator (S/utils/mesh_creation.cpp:49-74) for the benchmark meshes. */
dgx_status dgx_mesh_split_quads(int32_t n, double length, double x0, double y0, double* vertices,
                                int32_t* elements);

/* Face connectivity by sorted vertex pair, boundary faces tagged Bottom/Right/Top/Left by
 * position (bc index 0..3), periodic pairing by transverse face-midpoint coordinate.
 * periodic_mask bit0 pairs Left/Right, bit1 pairs Bottom/Top.
 * Replaces DGMesh::build_face_connectivity, identify_boundary_faces and
 * set_periodic_boundaries (S/core/mesh.cpp:312-402,482-588). */
dgx_status dgx_mesh_connect(int32_t n_vertices, const double* vertices, int32_t n_elements,
                            const int32_t* elements, int32_t periodic_mask,
                            int32_t* face_neighbors);
/* The same for elements of nv = 3 or 4 vertices: face_neighbors [K][nv][2]. */
dgx_status dgx_mesh_connect_elements(int32_t n_vertices, const double* vertices, int32_t n_elements,
                                     const int32_t* elements, int32_t vertices_per_element,
                                     int32_t periodic_mask, int32_t* face_neighbors);
/* Structured quadrilaterals on [xmin, xmax] x [ymin, ymax], nx x ny cells: vertex (i, j) at
 * (xmin + i (xmax - xmin)/nx, ymin + j (ymax - ymin)/ny), quad (i, j) = (v_ij, v_i+1j, v_i+1j+1,
 * v_ij+1) counter-clockwise, row-major; vertices [(nx+1)(ny+1)][2], elements [nx ny][4].  The
 * layout of MeshCreator::create_rectangular_mesh(dx, use_triangles = false), whose transfinite
 * lines take nx = int((xmax - xmin)/dx), ny = int((ymax - ymin)/dx) divisions
 * (S/utils/mesh_creation.cpp:49-74,212-243). */
dgx_status dgx_mesh_rect_quads(int32_t nx, int32_t ny, double xmin, double xmax, double ymin, double ymax,
                               double* vertices, int32_t* elements);

/* GMSH ASCII mesh (MSH 2.2 or 4.1) as MeshCreator::extract_mesh_data reads a GMSH model
 * (S/utils/mesh_creation.cpp:212-330): the nodes in file order, the 2-D elements of the first
 * 2-D type (3-node triangles or 4-node quadrilaterals), and the boundary faces named by the 1-D
 * physical groups.  face_neighbors [K][nv][2]: interior (neighbour element, face), boundary
 * (-1, index of the face's physical group in tag_names, or -1 without one); no periodic
 * pairing.  Call first with null arrays for the four sizes, then with arrays of those sizes;
 * tag_names: n_tags names of the 1-D physical groups in tag order, 64 bytes each
 * (NUL-terminated, at most 63 characters). */
dgx_status dgx_mesh_read_gmsh(const char* path, int32_t* n_vertices, int32_t* n_elements,
                              int32_t* vertices_per_element, int32_t* n_tags, double* vertices,
                              int32_t* elements, int32_t* face_neighbors, char* tag_names);

/* ---- context ------------------------------------------------------------------------------- */

/* Builds the space (DGSpace ctor, S/core/space.cpp:20-67), per-element geometry
 * (compute_element_data / compute_face_data, :69-171), face permutations (:173-228), mass
 * inverse (S/solver/compressible_solver_base.cpp:41-52) and uploads everything to the device. */
dgx_status dgx_create(const dgx_mesh_desc* mesh, const dgx_space_desc* space,
                      const dgx_physics* physics, int32_t precision, dgx_ctx** out);
void dgx_destroy(dgx_ctx* ctx);
const char* dgx_last_error(void);
dgx_status dgx_info(const dgx_ctx* ctx, dgx_info_t* out);
/* The largest element count dgx_create accepts for elements of this order and vertex count
 * (3 or 0: triangles, 4: quadrilaterals): the kernels index the state in 32 bits, so
 * 4 Np Kp <= 2^31 with Kp the element count rounded up to 256.  dgx_create rejects larger meshes
 * with DGX_E_UNSUPPORTED before allocating anything.  No counterpart in the reference (Eigen
 * indexes with 64-bit Index). */
dgx_status dgx_max_elements(int32_t order, int32_t vertices_per_element, int32_t* max_elements);

/* ---- state ---------------------------------------------------------------------------------- */

dgx_status dgx_set_state(dgx_ctx* ctx, const double* u);  /* [K][Np][4] -> device */
dgx_status dgx_get_state(dgx_ctx* ctx, double* u);        /* device -> [K][Np][4] */

/* Physical coordinates of the volume quadrature points, [K][Nq][2]. */
dgx_status dgx_volume_points(dgx_ctx* ctx, double* xy);
/* The per-element geometry the kernels read, built on the device (any pointer may be NULL):
 * geo [K][14] FP64 = dxi/dx row-major (4), |det J|, outward unit normals of faces 0..2 (6;
 * on an interior face the canonical side's normal, negated on the other side), face lengths
 * (3); hmin [K] = smallest edge length; codes [K] = 4 bits per face f at bit 4f: neighbour's
 * local face (3 = boundary), reversed point order (bit 2), canonical side (bit 3).  The data
 * of DGMesh::get_element_data / get_face_data (I/core/mesh.hpp:56-62; S/core/space.cpp:104-137,
 * 173-228; S/reference/mapping.cpp:22-50). */
dgx_status dgx_get_geometry(dgx_ctx* ctx, double* geo, double* hmin, uint32_t* codes);
/* L2 projection of point values Uq [K][Nq][4] into the state (project_initial_condition,
 * S/solver/compressible_solver_base.cpp:63-104). */
dgx_status dgx_project(dgx_ctx* ctx, const double* Uq);
/* Current state evaluated at the volume points, [K][Nq][4]. */
dgx_status dgx_eval_volume(dgx_ctx* ctx, double* Uq);
/* Relative L2 error of the primitive fields (rho, u, v, p) of the current state at the volume
 * points, weighted by w_q |J|, against exact values there ([K][Nq][4], conserved, or primitive
 * when exact_is_primitive != 0); err[f] = sqrt(sum w|J| (W - W_ex)^2 / sum w|J| W_ex^2), the
 * absolute root when the exact field's norm is below 1e-14.  Computed on the device in FP64.
 * Replaces compute_variable_error of T/solver/shear_flow_accuracy_test.cpp:32-79 and
 * X/taylor_green_test.cpp:44-86. */
dgx_status dgx_l2_error(dgx_ctx* ctx, const double* exact, int32_t exact_is_primitive, double err[4]);
/* Make boundary condition `bc_index` (a row of the mesh description's bc_types / bc_values) a
 * function of position: fn is evaluated now, on the host, at every face quadrature point of
 * the faces carrying it (and at the reference's read of those points for the Euler boundary
 * flux when DGX_QUIRK_EULER_BC_POINTS_REFERENCE is set) and replaces the constant value.
 * Replaces DGMesh::set_boundary_condition_euler with a function-valued condition
 * (S/core/mesh.cpp:366-401, S/boundary/conditions.cpp:47-60). */
dgx_status dgx_set_bc_function(dgx_ctx* ctx, int32_t bc_index, dgx_bc_fn fn, void* user);
/* Legacy-VTK file `filename`.vtk of an Euler state (host layout [K][Np][4], or the context's
 * current state when u is NULL): each element subdivided at `resolution_factor` (triangles
 * into VTK_TRIANGLE, quads into VTK_QUAD cells), fields
 * density, u_velocity, v_velocity, velocity_magnitude, pressure, temperature, mach, rho,
 * rho_u, rho_v, E at the subdivision points.  Replaces VTKWriter::write_euler_solution
 * (S/utils/vtk_writer.cpp:72-176; defaults gamma 1.4, resolution 3 in
 * I/utils/vtk_writer.hpp:78-81). */
dgx_status dgx_write_vtk(dgx_ctx* ctx, const double* u, const char* filename, double gamma,
                         int32_t resolution_factor);

/* ---- the hot path -------------------------------------------------------------------------- */

/* R = M dU/dt of u (or of the current state when u is NULL), [K][Np][4]:
 * DGAssembler::assemble_euler_residual (S/solver/assembler.cpp:137-279; I/solver/assembler.hpp:58-59). */
dgx_status dgx_residual(dgx_ctx* ctx, const double* u, double* r);
/* r <- M^{-1} r in place, [K][Np][4]: CompressibleDGSolverBase::apply_mass_inv
 * (S/solver/compressible_solver_base.cpp:54-61). */
dgx_status dgx_apply_minv(dgx_ctx* ctx, double* r);
/* Up to nsteps SSP-RK3 steps of the device state with the reference's NaN abort: the steps of
 * run_time_integration's loop (time_step_ssp_rk3, S/solver/compressible_solver_base.cpp:
 * 106-112,279-301; I/solver/time_stepping.hpp:39-55), which checks every element's coefficients
 * for NaN (Eigen hasNaN; infinite values do not count) after each step and stops at the first
 * step that has one.  Returns DGX_E_NONFINITE in that case, with the context holding the state
 * of that step and *steps_done its number (1-based); otherwise DGX_OK and *steps_done = nsteps.
 * steps_done may be NULL.  Costs one device copy of the state per call plus one reduction
 * (and, only when a NaN appears, a replay of the call's steps with a check after each). */
dgx_status dgx_rk3_run(dgx_ctx* ctx, double dt, int32_t nsteps, int32_t* steps_done);
/* dgx_rk3_run without the step count. */
dgx_status dgx_rk3_steps(dgx_ctx* ctx, double dt, int32_t nsteps);
/* Same, without the host round trip of the non-finite check; for timing loops and capture. */
dgx_status dgx_rk3_steps_async(dgx_ctx* ctx, double dt, int32_t nsteps);
dgx_status dgx_synchronize(dgx_ctx* ctx);

/* Runs five warm-up steps and nsteps timed steps and reports the mean duration of one fused
 * stage launch (HIP events around the 3 nsteps back-to-back launches on the context stream,
 * divided by their number; ms) and the event time of the whole timed loop (ms). */
dgx_status dgx_rk3_profile(dgx_ctx* ctx, double dt, int32_t nsteps, double* stage_ms,
                           double* total_ms);

/* ---- diagnostics (FP64) -------------------------------------------------------------------- */

/* Integral of each conserved variable over the mesh, FP64 deterministic reduction. */
dgx_status dgx_conservation_sums(dgx_ctx* ctx, double out[4]);
/* The conservation sums of dgx_conservation_sums written to device memory: out_dev = 4 FP64
 * values on the context's device, equal bit for bit to dgx_conservation_sums.  No counterpart in
 * the reference.  For callers that cannot allocate (code under HIP graph capture, tight serving
 * loops): it allocates nothing, sets no memory and does not synchronise; it is ordered on the
 * context's stream, so read out_dev after dgx_synchronize (or stream-ordered work).  Its block
 * partial sums live in one fixed 1 MiB device array per GPU shared by every call on that GPU:
 * calls that overlap on one GPU (several contexts on concurrent streams) race.  Meshes of more
 * than 8388608 elements are rejected with DGX_E_UNSUPPORTED. */
dgx_status dgx_conservation_sums_device(dgx_ctx* ctx, double* out_dev);
/* compute_max_cfl / compute_density_range (S/solver/compressible_solver_base.cpp:114-199):
 * CFL = dt * max_q(|u|+c) / h_min(e); returns +inf for rho<=0 or p<=0 like the reference. */
dgx_status dgx_max_cfl(dgx_ctx* ctx, double dt, double* cfl);
dgx_status dgx_density_range(dgx_ctx* ctx, double out[2]);
/* Kinetic energy functional J = sum_e sum_q w_q |J_e| (m_x^2+m_y^2)/(2 rho), FP64. */
dgx_status dgx_kinetic_energy(dgx_ctx* ctx, double* J);

/* ---- discrete adjoint (Euler) ---------------------------------------------------------------
 * No counterpart in the reference, which has no adjoint; the backward calls of the residual
 * (assemble_euler_residual), of apply_mass_inv o assemble_euler_residual and of
 * time_step_ssp_rk3.  The Rusanov max() takes the L-side branch at ties.  DGX_E_UNSUPPORTED for
 * Navier-Stokes contexts and for meshes with slip-wall faces. */

/* out = (dR/du)^T lam at u (or at the current state when u is NULL), host layout [K][Np][4]. */
dgx_status dgx_residual_vjp(dgx_ctx* ctx, const double* u, const double* lam, double* out);
/* out = (dL/du)^T lam with L = M^{-1} R (the RK right-hand side). */
dgx_status dgx_rhs_vjp(dgx_ctx* ctx, const double* u, const double* lam, double* out);
/* J of the current state and dJ/du (host layout; grad may be NULL). */
dgx_status dgx_kinetic_energy_grad(dgx_ctx* ctx, double* J, double* grad);
/* Gradient of a functional of the state after nsteps SSP-RK3 steps from the current state with
 * respect to that state: lambda_0 = (du_T/du_0)^T lambda_T.  The forward trajectory is stored
 * at every ckpt_every-th step and recomputed segment by segment in the reverse sweep.  On
 * return the context holds u_T.  lambda_T NULL seeds with dJ/du_T of the kinetic energy. */
dgx_status dgx_rk3_vjp(dgx_ctx* ctx, double dt, int32_t nsteps, int32_t ckpt_every,
                       const double* lambda_T, double* lambda_0);
/* Times, from the current state on the context stream, a plain forward run of nsteps (fwd_ms)
 * and the two phases of dgx_rk3_vjp of the same run seeded with the kinetic-energy gradient:
 * the forward sweep that stores the checkpoints (ckpt_fwd_ms) and the reverse sweep, i.e. the
 * per-segment recompute and the stage adjoints (rev_ms).  The state is restored afterwards. */
dgx_status dgx_rk3_vjp_profile(dgx_ctx* ctx, double dt, int32_t nsteps, int32_t ckpt_every,
                               double* fwd_ms, double* ckpt_fwd_ms, double* rev_ms);

/* ---- batched instances ---------------------------------------------------------------------
 * No counterpart in the reference (one solver per process there).  Many independent Euler
 * instances of mixed mesh size and order (periodic or with boundary conditions), one state and
 * dt per instance, stepped together: each RK stage is ONE persistent launch over all instances
 * (or one per order, as the tuner decides), and a two-step window of launches is captured in a
 * hipGraph and replayed. */
typedef struct dgx_batch dgx_batch;

typedef struct {
    int32_t layout;      /* 0 = one launch per stage for all orders, 1 = one launch per order,
                            2 = one launch per order, each order's stages on a stream of its own
                            (the orders' launches overlap) */
    int32_t chunk[4];    /* elements per work item = 64 * chunk[p-1] */
    int32_t blocks_per_cu;
} dgx_batch_tuning;

dgx_status dgx_batch_create(int32_t n_instances, const dgx_mesh_desc* meshes, const dgx_space_desc* spaces,
                            const dgx_physics* physics, int32_t precision, dgx_batch** out);
void dgx_batch_destroy(dgx_batch* batch);
/* The context of one instance (owned by the batch): its state, projection and diagnostics go
 * through the single-context calls. */
dgx_status dgx_batch_instance(dgx_batch* batch, int32_t instance, dgx_ctx** ctx);
dgx_status dgx_batch_set_dt(dgx_batch* batch, const double* dt);  /* [n_instances] */
dgx_status dgx_batch_rk3_steps(dgx_batch* batch, int32_t nsteps);
/* Mean wall time per step (ms) of nsteps batched steps, or of the same steps run instance after
 * instance with the single-context launches (sequential != 0); the states are restored. */
dgx_status dgx_batch_profile(dgx_batch* batch, int32_t nsteps, int32_t sequential, double* ms_per_step);
dgx_status dgx_batch_get_tuning(const dgx_batch* batch, dgx_batch_tuning* t);
dgx_status dgx_batch_set_tuning(dgx_batch* batch, const dgx_batch_tuning* t);
/* Autotuner: times trial_steps for each candidate (layout, per-order chunk, blocks per CU) by
 * coordinate descent, keeps the fastest and restores the states. */
dgx_status dgx_batch_autotune(dgx_batch* batch, int32_t trial_steps, dgx_batch_tuning* chosen);

#ifdef __cplusplus
}
#endif

#endif /* DGX_H */
