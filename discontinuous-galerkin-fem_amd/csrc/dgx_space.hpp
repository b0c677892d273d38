// Host-side reference-element tables for the DG space: quadrature rules and triangle bases,
// evaluated in FP64 once per context.
#pragma once
#include <vector>

namespace dgx {

struct Rule1D {
    std::vector<double> x, w;  // Gauss-Legendre on [-1,1], ascending
};
struct RuleTri {
    std::vector<double> xi, eta, w;  // reference triangle (0,0),(1,0),(0,1); weights sum to 1/2
};

// Gauss-Legendre with n points (the reference's QuadratureFactory::gauss_legendre_1d,
// S/quadrature/factory.cpp:16-31,75-236).
Rule1D gauss_legendre(int n);
// Symmetric Dunavant rule exact to `degree` (QuadratureFactory::dunavant_triangle,
// S/quadrature/factory.cpp:61-73,238-499), degree 1..9.
RuleTri dunavant(int degree);

// Triangle bases of order p: kind 0 = monomial x^i y^(d-i) by total degree
// (S/basis/monomial.cpp:30-72), kind 1 = Dubiner (S/basis/dubiner.cpp:24-168).
struct TriBasis {
    int kind, p, np;
    TriBasis(int kind, int p);
    void values(double xi, double eta, double* phi) const;       // phi[np]
    void gradients(double xi, double eta, double* g) const;      // g[np*2], (d/dxi, d/deta)
};

// Reference-triangle point of face f at GL abscissa s (DGSpace::map_face_quad_point,
// S/core/space.cpp:259-271): face 0 (t,0), face 1 (1-t,t), face 2 (0,1-t), t = (s+1)/2.
void face_point(int f, double s, double& xi, double& eta);

// Legendre tensor basis on [-1,1]^2 (LegendreBasis, S/basis/legendre.cpp:69-112):
// phi_{i + (p+1) j} = P_i(xi) P_j(eta).  `ref_endpoint` reproduces the reference's derivative
// at xi = +1 for odd degree (-P'_n(1), S/basis/legendre.cpp:31-42).
struct QuadBasis {
    int p, np;
    bool ref_endpoint;
    QuadBasis(int p, bool ref_endpoint);
    void values(double xi, double eta, double* phi) const;
    void gradients(double xi, double eta, double* g) const;  // g[np*2]
};

// Reference-square point of face f at GL abscissa s (map_face_quad_point, S/core/space.cpp:
// 273-287): (s,-1), (1,s), (-s,1), (-1,-s) along the edges (v_f, v_{f+1}).
void face_point_quad(int f, double s, double& xi, double& eta);

}  // namespace dgx
