// refcpu — FP64 CPU restatement of the reference's explicit DG residual path.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP product is checked against.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
// and only as the checker or as the timed CPU baseline; the product never links it.
//
// Every routine follows the reference's loop structure and cites the reference file:line
// it restates (paths relative to the reference root; S/ = dgfem/src, I/ = dgfem/include/dgfem).
//
// Parity pinning (tests/test_oracle_*.py):
//   quadrature      : GL/Dunavant exactness, weight sums, GL 2-point KAT
//   bases           : monomial ordering/values/gradients KAT, Dubiner Gram = 2 I,
//                     FD gradients, projection of quadratics
//   geometry        : mapping det / dxi_dx KAT, J^T dxi_dx = I, face integrals (sqrt2/30 ...)
//   Euler residual  : free-stream preservation, global conservation, Rusanov consistency and
//                     antisymmetry, brute-force face/volume integrals on a 2-element mesh,
//                     basis independence (monomial vs Dubiner fields)
//   NS residual     : the reference's NS weak-form KATs (penalty, uniform flow, shear-flow
//                     volume/face/boundary formulas, no-slip residual)
//   mass / RK3      : exact monomial moments a!b!/(a+b+2)!, RK3 amplification polynomial
//   driver          : isentropic-vortex convergence, CFL by hand
#pragma once
#include <cstdint>

extern "C" {

typedef void (*ref_bc_fn)(double x, double y, double* out4, void* user);

// Mesh + space + physics.  periodic_mask bit0 = pair Left/Right, bit1 = pair Bottom/Top.
// basis: 0 = monomial (reference default), 1 = Dubiner (triangles), 2 = Legendre (quads).
// vol_rule: 0 = degree 2p (reference), 1 = degree 2p+1 (triangles).  physics: 0 = Euler,
// 1 = Navier-Stokes.  ns_face_grad: 0 = reference (scrambled Map read of dphi_dx_face),
// 1 = correct gradients.  nvpe: 3 = triangles, 4 = quads (elems [K][nvpe]).
// legendre_endpoint_ref: 1 = the reference's Legendre derivative at x = +1 (quirk Q16).
void* ref_create(int nv, const double* verts, int K, const int32_t* elems, int order, int basis,
                 int vol_rule, int periodic_mask, int physics, double gamma, double mu,
                 double prandtl, double sigma0, int ns_face_grad, int nvpe, int legendre_endpoint_ref);
void ref_destroy(void* h);
const char* ref_last_error(void);

// info: [K, Np, Nq, Nfq, n_interior_faces, n_boundary_faces]
void ref_info(void* h, int32_t* out6);

// Boundary conditions by tag name ("Bottom","Right","Top","Left").  type: 0 FAR_FIELD,
// 1 SLIP_WALL, 2 NO_SLIP_WALL.  fn may be null -> constant value4.
int ref_set_euler_bc_points(void* h, int reference);
int ref_set_ns_average(void* h, int reference);  // Q4: 1 reference viscous face fluxes, 0 consistent
int ref_set_heat_flux(void* h, int reference);   // Q17: 1 energy flux u.tau + q (reference), 0 u.tau - q
int ref_set_bc(void* h, const char* tag, int type, const double* value4, ref_bc_fn fn,
               void* user);

// Residual R = M dU/dt (assemble_euler_residual), layout [K][Np][4].
int ref_residual(void* h, const double* u, double* r);
int ref_apply_minv(void* h, double* r);
int ref_rhs(void* h, const double* u, double* k);
// SSP-RK3, nsteps steps in place; returns seconds spent in the loop (steady clock).
double ref_rk3(void* h, double* u, double dt, int nsteps);
double ref_rk3_threads(void* h, double* u, double dt, int nsteps, int nthreads);
int ref_project(void* h, const double* Uq, double* u);  // Uq: [K][Nq][4] at volume points
int ref_volume_points(void* h, double* xy);              // [K][Nq][2]
int ref_eval_volume(void* h, const double* u, double* Uq);  // [K][Nq][4]
double ref_max_cfl(void* h, const double* u, double dt);
int ref_density_range(void* h, const double* u, double* out2);
int ref_conservation_sums(void* h, const double* u, double* out4);
double ref_kinetic_energy(void* h, const double* u);

// Tables for known-answer tests.
int ref_quad_tri(int degree, double* xi, double* eta, double* w);  // returns n points
int ref_gauss_legendre(int n, double* x, double* w);
// kind 0 monomial, 1 Dubiner, 2 Legendre (6: Legendre with P'_n(+1) = n(n+1)/2); returns Np
int ref_basis_eval(int kind, int order, double xi, double eta, double* phi, double* grad);
int ref_mapping(int nvpe, const double* verts, double xi, double eta, double* det, double* dxi_dx4,
                double* x2);
int ref_face_point(int nvpe, int face, double s, double* xi2);

// Element / face level pieces (the weak-form virtuals), for KATs.
int ref_elem_data(void* h, int e, double* jdet, double* dphi_dx);  // jdet[Nq], dphi_dx[Nq*Np*2]
int ref_face_geom(void* h, int e, int f, double* normal2, double* length);
int ref_interior_faces(void* h, int32_t* out5);  // [n][eL,fL,eR,fR,reversed]
int ref_boundary_faces(void* h, int32_t* out3);  // [n][e,f,has_bc]
int ref_neighbors(void* h, int32_t* out);        // [K][3][2] (elem, face) or -1
int ref_volume_residual(void* h, int e, const double* ue, double* R);
int ref_viscous_volume_residual(void* h, int e, const double* ue, double* R);
int ref_interior_face_residual(void* h, int iface, const double* uL, const double* uR, double* RL,
                               double* RR, int viscous);
int ref_boundary_face_residual(void* h, int ibface, const double* ue, double* R, int viscous);
int ref_rusanov(double gamma, const double* UL, const double* UR, const double* n, double* H);
int ref_viscous_flux(void* h, const double* U, const double* gradU8, double* Fv, double* Gv);
double ref_penalty(void* h, double length);
}
