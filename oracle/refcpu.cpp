// refcpu — FP64 CPU restatement of the reference's explicit DG residual path.
// TEST INFRASTRUCTURE ONLY (the oracle): see refcpu.hpp for the rules and what pins it.
//
// Reference paths are relative to the reference repository root:
//   S/ = dgfem/src, I/ = dgfem/include/dgfem, T/ = dgfem/tests.
#include "refcpu.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <omp.h>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_err;

// ---------------------------------------------------------------------------------------------
// Quadrature.
// Gauss-Legendre on [-1,1], ascending nodes (S/quadrature/factory.cpp:16-31,75-236 tabulates
// the same rule for n = 1..8).  Computed here by Newton on P_n from Chebyshev guesses and
// symmetrised so that x[n-1-i] = -x[i] and w[n-1-i] = w[i] hold bit for bit.
// ---------------------------------------------------------------------------------------------
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    if (n == 1) { w[0] = 2.0; return; }
    const double pi = 3.14159265358979323846;
    // P_n(z) and P_{n-1}(z) by the three-term recurrence
    auto legendre = [n](double z, double& pn, double& pnm1) {
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= n; ++k) {
            double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        pn = p1;
        pnm1 = p0;
    };
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = std::cos(pi * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double pn, pnm1;
            legendre(z, pn, pnm1);
            double dp = n * (z * pn - pnm1) / (z * z - 1.0);
            double dz = pn / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-17) break;
        }
        double pn, pnm1;
        legendre(z, pn, pnm1);
        double dp = n * (z * pn - pnm1) / (z * z - 1.0);
        double wi = 2.0 / ((1.0 - z * z) * dp * dp);
        // z is the i-th largest node; ascending order puts it at n-1-i, mirror at i
        x[n - 1 - i] = z;
        x[i] = -z;
        w[n - 1 - i] = wi;
        w[i] = wi;
    }
    if (n % 2 == 1) x[n / 2] = 0.0;
}

// Symmetric triangle rules exact to a given degree (Dunavant 1985, Int. J. Numer. Meth. Eng.
// 21:1129-1148), the rules the reference tabulates at S/quadrature/factory.cpp:238-499.
// Orbits carry barycentric parameters and weights on the reference triangle (area 1/2):
// degrees 4-6 with the full-precision published values (Intrepid2 cubature tables),
// degrees 7-9 with Dunavant's 15-digit values, degrees 1-3 exact.
struct Orbit {
    int type;  // 0 = centroid, 1 = (1-2b, b, b), 2 = (a, b, 1-a-b)
    double a, b, w;
};

std::vector<Orbit> dunavant_orbits(int degree) {
    switch (degree) {
    case 0:
    case 1: return {{0, 0, 0, 0.5}};
    case 2: return {{1, 0, 1.0 / 6.0, 1.0 / 6.0}};
    case 3: return {{0, 0, 0, -9.0 / 32.0}, {1, 0, 0.2, 25.0 / 96.0}};
    case 4: return {{1, 0, 4.4594849091596487577332043252695176084800e-1, 1.1169079483900573053953942803038657817028e-1},
                    {1, 0, 9.1576213509770745141706146463816673424873e-2, 5.4975871827660936395761590801109078322738e-2}};
    case 5: return {{0, 0, 0, 1.125e-1},
                    {1, 0, 1.0128650732345633880098736191512382805558e-1, 6.2969590272413576297841972750090666828820e-2},
                    {1, 0, 4.7014206410511508977044120951344760051585e-1, 6.6197076394253090368824693916575999837847e-2}};
    case 6: return {{1, 0, 6.3089014491502228340331602870819157341003e-2, 2.5422453185103408460468404553434492023395e-2},
                    {1, 0, 2.4928674517091042129163855310701907608796e-1, 5.8393137863189683012644805692789720663043e-2},
                    {2, 5.3145049844816939902261738355299128796183e-2, 3.1035245103378441286759566723265117140082e-1,
                     4.1425537809186787596776728210221226990114e-2}};
    case 7: return {{0, 0, 0, -0.5 * 0.149570044467682},
                    {1, 0, 0.260345966079040, 0.5 * 0.175615257433208},
                    {1, 0, 0.065130102902216, 0.5 * 0.053347235608838},
                    {2, 0.048690315425316, 0.312865496004874, 0.5 * 0.077113760890257}};
    case 8: return {{0, 0, 0, 0.5 * 0.144315607677787},
                    {1, 0, 0.459292588292723, 0.5 * 0.095091634267285},
                    {1, 0, 0.170569307751760, 0.5 * 0.103217370534718},
                    {1, 0, 0.050547228317031, 0.5 * 0.032458497623198},
                    {2, 0.008394777409958, 0.263112829634638, 0.5 * 0.027230314174435}};
    case 9: return {{0, 0, 0, 0.5 * 0.097135796282799},
                    {1, 0, 0.489682519198738, 0.5 * 0.031334700227139},
                    {1, 0, 0.437089591492937, 0.5 * 0.077827541004774},
                    {1, 0, 0.188203535619033, 0.5 * 0.079647738927210},
                    {1, 0, 0.044729513394453, 0.5 * 0.025577675658698},
                    {2, 0.036838412054736, 0.221962989160766, 0.5 * 0.043283539377289}};
    default: throw std::invalid_argument("unsupported triangle rule degree");
    }
}

void triangle_rule(int degree, std::vector<double>& xi, std::vector<double>& eta, std::vector<double>& w) {
    xi.clear(); eta.clear(); w.clear();
    for (const auto& o : dunavant_orbits(degree)) {
        if (o.type == 0) {
            xi.push_back(1.0 / 3.0); eta.push_back(1.0 / 3.0); w.push_back(o.w);
        } else if (o.type == 1) {
            double b = o.b, a = 1.0 - 2.0 * b;
            // barycentric (a,b,b),(b,a,b),(b,b,a) -> (xi,eta) = (l2,l3)
            double l[3][3] = {{a, b, b}, {b, a, b}, {b, b, a}};
            for (auto& p : l) { xi.push_back(p[1]); eta.push_back(p[2]); w.push_back(o.w); }
        } else {
            double a = o.a, b = o.b, c = 1.0 - o.a - o.b;
            double l[6][3] = {{a, b, c}, {a, c, b}, {b, a, c}, {b, c, a}, {c, a, b}, {c, b, a}};
            for (auto& p : l) { xi.push_back(p[1]); eta.push_back(p[2]); w.push_back(o.w); }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Bases.
// Monomial x^i y^(d-i), ordered by total degree d, i = 0..d (S/basis/monomial.cpp:30-72).
// Dubiner with j outer / i inner index order, collapsed coordinates r = 2 xi/(1-eta) - 1,
// s = 2 eta - 1, normalisation 2 sqrt((2i+1)(i+j+1)) (S/basis/dubiner.cpp:24-168).
// ---------------------------------------------------------------------------------------------
double jacobi(int n, double al, double be, double x) {
    if (n == 0) return 1.0;
    double pm1 = 1.0, p = 0.5 * (al - be + (al + be + 2.0) * x);
    if (n == 1) return p;
    for (int k = 2; k <= n; ++k) {
        double a1 = 2.0 * k * (k + al + be) * (2.0 * k + al + be - 2.0);
        double a2 = (2.0 * k + al + be - 1.0) * (al * al - be * be);
        double a3 = (2.0 * k + al + be - 2.0) * (2.0 * k + al + be - 1.0) * (2.0 * k + al + be);
        double a4 = 2.0 * (k + al - 1.0) * (k + be - 1.0) * (2.0 * k + al + be);
        double pn = ((a2 + a3 * x) * p - a4 * pm1) / a1;
        pm1 = p;
        p = pn;
    }
    return p;
}
double jacobi_d(int n, double al, double be, double x) {
    if (n == 0) return 0.0;
    return 0.5 * (n + al + be + 1.0) * jacobi(n - 1, al + 1.0, be + 1.0, x);
}

// Legendre polynomials by the three-term recurrence (I/basis/legendre.hpp:55-72) and their
// derivative (S/basis/legendre.cpp:31-49): n (x P_n - P_{n-1}) / (x^2 - 1) inside, and at
// |x| = 1 the endpoint value 0.5 n (n+1) times parity(n) times sign(x), which is -P'_n(1) for
// odd n at x = +1 (quirk Q16); `ref_endpoint` = false gives P'_n(+1) = n(n+1)/2 instead.
double legendre(int n, double x) {
    if (n == 0) return 1.0;
    if (n == 1) return x;
    double pm1 = 1.0, pn = x;
    for (int i = 2; i <= n; ++i) {
        double pp = ((2 * i - 1) * x * pn - (i - 1) * pm1) / i;
        pm1 = pn;
        pn = pp;
    }
    return pn;
}
double legendre_d(int n, double x, bool ref_endpoint) {
    if (n == 0) return 0.0;
    if (std::fabs(std::fabs(x) - 1.0) < 1e-12) {
        double sign = x > 0 ? 1.0 : -1.0, parity = (n % 2 == 0) ? 1.0 : -1.0;
        if (!ref_endpoint && x > 0) parity = 1.0;
        return 0.5 * n * (n + 1) * parity * sign;
    }
    return n * (x * legendre(n, x) - legendre(n - 1, x)) / (x * x - 1.0);
}

// kind 0 monomial, 1 Dubiner (triangles), 2 Legendre tensor product (quads).
struct Basis {
    int kind = 0, p = 1, np = 3;
    bool ref_endpoint = true;  // Q16 for the Legendre derivative at x = +1
    Basis() = default;
    Basis(int kind_, int p_) : kind(kind_), p(p_), np(kind_ == 2 ? (p_ + 1) * (p_ + 1) : (p_ + 1) * (p_ + 2) / 2) {}

    void ij(int idx, int& i, int& j) const {  // Dubiner index order (S/basis/dubiner.cpp:77-90)
        int k = 0;
        for (int jj = 0; jj <= p; ++jj)
            for (int ii = 0; ii <= p - jj; ++ii) {
                if (k == idx) { i = ii; j = jj; return; }
                ++k;
            }
        i = j = 0;
    }
    static void clamp_to_triangle(double& x, double& y) {  // S/reference/elements.cpp:23-25,58-61
        if (!(x >= 0 && y >= 0 && (x + y) <= 1)) {
            x = std::max(0.0, std::min(1.0, x));
            y = std::max(0.0, std::min(1.0 - x, y));
        }
    }
    void eval(double x, double y, double* phi) const {
        if (kind == 2) {  // phi_{i,j} = P_i(x) P_j(y), j outer (S/basis/legendre.cpp:69-86)
            int k = 0;
            for (int j = 0; j <= p; ++j)
                for (int i = 0; i <= p; ++i) phi[k++] = legendre(i, x) * legendre(j, y);
            return;
        }
        if (kind == 0) {
            int k = 0;
            for (int d = 0; d <= p; ++d)
                for (int i = 0; i <= d; ++i) {
                    int j = d - i;
                    phi[k++] = (i == 0 ? 1.0 : std::pow(x, i)) * (j == 0 ? 1.0 : std::pow(y, j));
                }
            return;
        }
        clamp_to_triangle(x, y);
        double ome = 1.0 - y;
        double r = (std::fabs(ome) < 1e-14) ? -1.0 : 2.0 * x / ome - 1.0;
        double s = 2.0 * y - 1.0;
        for (int k = 0; k < np; ++k) {
            int i, j;
            ij(k, i, j);
            double Pi = jacobi(i, 0, 0, r), Pj = jacobi(j, 2.0 * i + 1.0, 0, s);
            double nrm = 2.0 * std::sqrt((2.0 * i + 1.0) * (i + j + 1.0));
            phi[k] = nrm * Pi * Pj * std::pow(0.5 * (1.0 - s), i);
        }
    }
    // grad[k*2 + d]
    void grad(double x, double y, double* g) const {
        if (kind == 2) {  // (S/basis/legendre.cpp:88-112)
            int k = 0;
            for (int j = 0; j <= p; ++j)
                for (int i = 0; i <= p; ++i) {
                    g[2 * k] = legendre_d(i, x, ref_endpoint) * legendre(j, y);
                    g[2 * k + 1] = legendre(i, x) * legendre_d(j, y, ref_endpoint);
                    ++k;
                }
            return;
        }
        if (kind == 0) {
            int k = 0;
            for (int d = 0; d <= p; ++d)
                for (int i = 0; i <= d; ++i) {
                    int j = d - i;
                    double px = (i == 0) ? 1.0 : std::pow(x, i), py = (j == 0) ? 1.0 : std::pow(y, j);
                    g[2 * k] = (i == 0) ? 0.0 : i * (i - 1 == 0 ? 1.0 : std::pow(x, i - 1)) * py;
                    g[2 * k + 1] = (j == 0) ? 0.0 : px * j * (j - 1 == 0 ? 1.0 : std::pow(y, j - 1));
                    ++k;
                }
            return;
        }
        clamp_to_triangle(x, y);
        double ome = 1.0 - y;
        double r = (std::fabs(ome) < 1e-14) ? -1.0 : 2.0 * x / ome - 1.0;
        double s = 2.0 * y - 1.0;
        double dr_dx = 0, dr_dy = 0, ds_dx = 0, ds_dy = 0;
        if (std::fabs(ome) >= 1e-14) { dr_dx = 2.0 / ome; dr_dy = 2.0 * x / (ome * ome); ds_dy = 2.0; }
        for (int k = 0; k < np; ++k) {
            int i, j;
            ij(k, i, j);
            double Pi = jacobi(i, 0, 0, r), Pj = jacobi(j, 2.0 * i + 1.0, 0, s);
            double dPi = i > 0 ? jacobi_d(i, 0, 0, r) : 0.0;
            double dPj = j > 0 ? jacobi_d(j, 2.0 * i + 1.0, 0, s) : 0.0;
            double nrm = 2.0 * std::sqrt((2.0 * i + 1.0) * (i + j + 1.0));
            double sc = std::pow(0.5 * (1.0 - s), i);
            double dsc = i > 0 ? -0.5 * i * std::pow(0.5 * (1.0 - s), i - 1) : 0.0;
            double dphi_dr = nrm * dPi * Pj * sc;
            double dphi_ds = nrm * Pi * (dPj * sc + Pj * dsc);
            g[2 * k] = dphi_dr * dr_dx + dphi_ds * ds_dx;
            g[2 * k + 1] = dphi_dr * dr_dy + dphi_ds * ds_dy;
        }
    }
};

// Reference triangle (0,0),(1,0),(0,1); edges (0,1),(1,2),(2,0) (S/reference/elements.cpp:18-49).
// Reference quad [-1,1]^2, vertices (-1,-1),(1,-1),(1,1),(-1,1), edges (f, f+1) (:61-87).
// Face point for s in [-1,1] (map_face_quad_point, S/core/space.cpp:259-287).
void face_point(int nvpe, int f, double s, double& x, double& y) {
    if (nvpe == 4) {
        switch (f) {
        case 0: x = s; y = -1.0; break;
        case 1: x = 1.0; y = s; break;
        case 2: x = -s; y = 1.0; break;
        case 3: x = -1.0; y = -s; break;
        default: throw std::out_of_range("Invalid face for quad");
        }
        return;
    }
    double t = 0.5 * (s + 1.0);
    switch (f) {
    case 0: x = t; y = 0.0; break;
    case 1: x = 1.0 - t; y = t; break;
    case 2: x = 0.0; y = 1.0 - t; break;
    default: throw std::out_of_range("Invalid face for triangle");
    }
}

// Mapping (S/reference/mapping.cpp:22-50): x = X^T N, J_T = dN^T X, det, dxi_dx = inv(J_T);
// linear shape functions on the triangle, bilinear on the quad (S/reference/elements.cpp:33-47,
// 89-108).
struct Mapping {
    double det, A[4], x, y;  // A row-major: A[r*2+c]
};
Mapping compute_mapping(int nvpe, const double* V, double xi, double eta) {
    Mapping m;
    double N[4], dN[4][2];
    if (nvpe == 4) {
        N[0] = 0.25 * (1.0 - xi) * (1.0 - eta);
        N[1] = 0.25 * (1.0 + xi) * (1.0 - eta);
        N[2] = 0.25 * (1.0 + xi) * (1.0 + eta);
        N[3] = 0.25 * (1.0 - xi) * (1.0 + eta);
        dN[0][0] = -0.25 * (1.0 - eta); dN[0][1] = -0.25 * (1.0 - xi);
        dN[1][0] = 0.25 * (1.0 - eta);  dN[1][1] = -0.25 * (1.0 + xi);
        dN[2][0] = 0.25 * (1.0 + eta);  dN[2][1] = 0.25 * (1.0 + xi);
        dN[3][0] = -0.25 * (1.0 + eta); dN[3][1] = 0.25 * (1.0 - xi);
    } else {
        N[0] = 1.0 - xi - eta; N[1] = xi; N[2] = eta;
        dN[0][0] = -1.0; dN[0][1] = -1.0; dN[1][0] = 1.0; dN[1][1] = 0.0; dN[2][0] = 0.0; dN[2][1] = 1.0;
    }
    m.x = 0.0; m.y = 0.0;
    // J_T[a][b] = sum_n dN_n/dxi_a * X_n,b
    double JT00 = 0.0, JT01 = 0.0, JT10 = 0.0, JT11 = 0.0;
    for (int n = 0; n < nvpe; ++n) {
        m.x += N[n] * V[2 * n];
        m.y += N[n] * V[2 * n + 1];
        JT00 += dN[n][0] * V[2 * n];
        JT01 += dN[n][0] * V[2 * n + 1];
        JT10 += dN[n][1] * V[2 * n];
        JT11 += dN[n][1] * V[2 * n + 1];
    }
    m.det = JT00 * JT11 - JT01 * JT10;
    if (std::fabs(m.det) < 1e-12) throw std::runtime_error("Singular Jacobian matrix");
    double id = 1.0 / m.det;
    m.A[0] = JT11 * id;
    m.A[1] = -JT01 * id;
    m.A[2] = -JT10 * id;
    m.A[3] = JT00 * id;
    return m;
}

// ---------------------------------------------------------------------------------------------
// Euler / NS pointwise physics.
// ---------------------------------------------------------------------------------------------
using V4 = std::array<double, 4>;

// conserved_to_primitive (S/weak_forms/euler_weak_formulation.cpp:279-296)
V4 cons2prim(const V4& U, double g) {
    double rho = U[0], ri = 1.0 / rho, u = U[1] * ri, v = U[2] * ri;
    double p = (g - 1.0) * (U[3] - 0.5 * rho * (u * u + v * v));
    return {rho, u, v, p};
}
// primitive_to_conserved (:261-277)
V4 prim2cons(const V4& W, double g) {
    double rho = W[0], u = W[1], v = W[2], p = W[3];
    return {rho, rho * u, rho * v, p / (g - 1.0) + 0.5 * rho * (u * u + v * v)};
}
// get_fluxes_and_primitive (:27-43)
void fluxes(const V4& U, double g, V4& F, V4& G, V4& W) {
    W = cons2prim(U, g);
    double rho = W[0], u = W[1], v = W[2], p = W[3];
    F = {U[1], rho * u * u + p, rho * u * v, u * (U[3] + p)};
    G = {U[2], rho * u * v, rho * v * v + p, v * (U[3] + p)};
}
// rusanov_flux (:51-70)
V4 rusanov(const V4& UL, const V4& UR, const double* n, double g) {
    V4 FL, GL, WL, FR, GR, WR;
    fluxes(UL, g, FL, GL, WL);
    fluxes(UR, g, FR, GR, WR);
    double cL = std::sqrt(g * WL[3] / WL[0]), cR = std::sqrt(g * WR[3] / WR[0]);
    double unL = WL[1] * n[0] + WL[2] * n[1], unR = WR[1] * n[0] + WR[2] * n[1];
    double smax = std::max(std::fabs(unL) + cL, std::fabs(unR) + cR);
    V4 H;
    for (int v = 0; v < 4; ++v) {
        double FnL = FL[v] * n[0] + GL[v] * n[1], FnR = FR[v] * n[0] + GR[v] * n[1];
        H[v] = 0.5 * (FnL + FnR) - 0.5 * smax * (UR[v] - UL[v]);
    }
    return H;
}

// heat_sign: +1 the reference's energy flux u.tau + q with q = -kappa grad T
// (navier_stokes_weak_formulation.cpp:76-77,91,95), -1 the conducting u.tau - q.
struct NSParams { double gamma, mu, pr, sigma0, R, heat_sign = 1.0; };
// compute_primitive_gradients + compute_viscous_fluxes
// (S/weak_forms/navier_stokes_weak_formulation.cpp:36-98); gradU[v*2+d].
void viscous_flux(const V4& U, const double* gU, const NSParams& P, V4& Fv, V4& Gv) {
    double rho_safe = std::max(U[0], 1e-12);
    V4 W = cons2prim(U, P.gamma);
    double rho = W[0], u = W[1], v = W[2], p = W[3];
    double grho[2] = {gU[0], gU[1]}, gru[2] = {gU[2], gU[3]}, grv[2] = {gU[4], gU[5]}, gE[2] = {gU[6], gU[7]};
    double ri = 1.0 / rho_safe;
    double gu[2], gv[2], gp[2], gT[2];
    for (int d = 0; d < 2; ++d) { gu[d] = (gru[d] - u * grho[d]) * ri; gv[d] = (grv[d] - v * grho[d]) * ri; }
    double ksq = u * u + v * v;
    for (int d = 0; d < 2; ++d) {
        double mom = rho * (u * gu[d] + v * gv[d]);
        gp[d] = (P.gamma - 1.0) * (gE[d] - 0.5 * ksq * grho[d] - mom);
    }
    double rho_sq = rho * rho;
    for (int d = 0; d < 2; ++d) gT[d] = (rho * gp[d] - p * grho[d]) / (rho_sq * P.R);
    double div = gu[0] + gv[1];
    double lam = -2.0 * P.mu / 3.0;
    double txx = 2.0 * P.mu * gu[0] + lam * div;
    double tyy = 2.0 * P.mu * gv[1] + lam * div;
    double txy = P.mu * (gu[1] + gv[0]);
    double kappa = P.mu * P.gamma / (P.pr * (P.gamma - 1.0));
    double qx = -kappa * gT[0], qy = -kappa * gT[1];
    Fv = {0.0, txx, txy, u * txx + v * txy + P.heat_sign * qx};
    Gv = {0.0, txy, tyy, u * txy + v * tyy + P.heat_sign * qy};
}

// ---------------------------------------------------------------------------------------------
// Mesh, space and geometry.
// ---------------------------------------------------------------------------------------------
struct BC {
    int type = -1;  // 0 FAR_FIELD, 1 SLIP, 2 NO_SLIP
    double value[4] = {0, 0, 0, 0};
    ref_bc_fn fn = nullptr;
    void* user = nullptr;
    V4 eval(double x, double y) const {
        V4 r;
        if (fn) fn(x, y, r.data(), user);
        else for (int v = 0; v < 4; ++v) r[v] = value[v];
        return r;
    }
};

struct FaceData {  // S/core/space.cpp:104-171
    double normal[2];
    double length;
    std::vector<double> qp_flat;       // [x_0..x_{n-1}, y_0..y_{n-1}] (Eigen column-major)
    std::vector<double> dphi_dx_face;  // [(q*Np+i)*2 + d]
};
struct ElemData {  // S/core/space.cpp:69-102
    double V[8];
    std::vector<double> jdet;     // Nq
    std::vector<double> dphi_dx;  // [(q*Np+i)*2 + d]
};
struct IFace { int eL, fL, eR, fR; std::vector<int> perm; };
struct BFace { int e, f; std::string tag; };

struct Ctx {
    int K = 0, p = 1, np = 3, nq = 0, nfq = 0, physics = 0, ns_face_grad = 0;
    int nv = 3;  // vertices = faces per element: 3 triangles, 4 quads
    int euler_bc_points_reference = 1;  // Q5: Euler boundary points read through the transposed map
    int ns_average_reference = 1;       // Q4: viscous face fluxes as the reference signs them
    Basis basis;
    std::vector<double> vx, vy;
    std::vector<int> tri;             // K*nv
    std::vector<int> nbr;             // K*nv*2  (elem, face) or -1
    std::vector<std::string> btag;    // K*nv boundary tag name ("" if none / interior)
    std::vector<double> qxi, qeta, qw, gs, gw;
    std::vector<double> phi_vol;      // nq*np
    std::vector<double> dphi_ref;     // nq*np*2
    std::vector<double> phi_face;     // nv*nfq*np
    std::vector<ElemData> ed;
    std::vector<FaceData> fd;         // K*nv
    std::vector<IFace> ifaces;
    std::vector<BFace> bfaces;
    std::map<std::string, BC> bcs;
    std::vector<double> minv;         // K*np*np
    NSParams ns{1.4, 1e-3, 0.72, 5.0, 1.0};
    std::vector<double> rbuf, kbuf;
    // threaded timing path: per-face contributions and, per element, its faces in face order
    std::vector<double> fl, fr, fvl, fvr, fb, fvb;
    std::vector<std::vector<std::pair<int, int>>> eif;  // (interior face, side 0 = L / 1 = R)
    std::vector<std::vector<int>> ebf;                   // boundary faces
};

// Connectivity through a map keyed by the sorted vertex pair (S/core/mesh.cpp:312-359 builds
// the same neighbour table with an O(K^2) rescan; this one is O(K log K)).
void build_connectivity(Ctx& c) {
    const int nv = c.nv;
    c.nbr.assign(c.K * nv * 2, -1);
    std::map<std::pair<int, int>, std::vector<std::pair<int, int>>> fm;
    for (int e = 0; e < c.K; ++e)
        for (int f = 0; f < nv; ++f) {
            int v1 = c.tri[e * nv + f], v2 = c.tri[e * nv + (f + 1) % nv];
            fm[{std::min(v1, v2), std::max(v1, v2)}].push_back({e, f});
        }
    for (auto& kv : fm) {
        if (kv.second.size() == 2) {
            auto a = kv.second[0], b = kv.second[1];
            c.nbr[(a.first * nv + a.second) * 2] = b.first;
            c.nbr[(a.first * nv + a.second) * 2 + 1] = b.second;
            c.nbr[(b.first * nv + b.second) * 2] = a.first;
            c.nbr[(b.first * nv + b.second) * 2 + 1] = a.second;
        }
    }
}

// Boundary tags by position on the bounding box, with the GMSH physical names Bottom/Right/
// Top/Left (S/utils/mesh_creation.cpp:256-260).
void tag_boundaries(Ctx& c) {
    double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
    for (size_t i = 0; i < c.vx.size(); ++i) {
        xmin = std::min(xmin, c.vx[i]); xmax = std::max(xmax, c.vx[i]);
        ymin = std::min(ymin, c.vy[i]); ymax = std::max(ymax, c.vy[i]);
    }
    double tol = 1e-10 * std::max(xmax - xmin, ymax - ymin);
    const int nv = c.nv;
    c.btag.assign(c.K * nv, "");
    for (int e = 0; e < c.K; ++e)
        for (int f = 0; f < nv; ++f) {
            if (c.nbr[(e * nv + f) * 2] >= 0) continue;
            int v1 = c.tri[e * nv + f], v2 = c.tri[e * nv + (f + 1) % nv];
            double x1 = c.vx[v1], x2 = c.vx[v2], y1 = c.vy[v1], y2 = c.vy[v2];
            std::string t;
            if (std::fabs(y1 - ymin) < tol && std::fabs(y2 - ymin) < tol) t = "Bottom";
            else if (std::fabs(x1 - xmax) < tol && std::fabs(x2 - xmax) < tol) t = "Right";
            else if (std::fabs(y1 - ymax) < tol && std::fabs(y2 - ymax) < tol) t = "Top";
            else if (std::fabs(x1 - xmin) < tol && std::fabs(x2 - xmin) < tol) t = "Left";
            c.btag[e * nv + f] = t;
        }
}

// Periodic pairing by the transverse coordinate of the face midpoint (S/core/mesh.cpp:482-588:
// Left/Right match y, Bottom/Top match x).  Sorted instead of the reference's O(n^2) scan.
void pair_periodic(Ctx& c, const char* t1, const char* t2, bool match_y) {
    const int nv = c.nv;
    std::vector<std::pair<double, int>> a, b;
    for (int k = 0; k < c.K * nv; ++k) {
        if (c.nbr[k * 2] >= 0) continue;
        int e = k / nv, f = k % nv;
        int v1 = c.tri[e * nv + f], v2 = c.tri[e * nv + (f + 1) % nv];
        double m = match_y ? 0.5 * (c.vy[v1] + c.vy[v2]) : 0.5 * (c.vx[v1] + c.vx[v2]);
        if (c.btag[k] == t1) a.push_back({m, k});
        else if (c.btag[k] == t2) b.push_back({m, k});
    }
    if (a.empty() || b.empty() || a.size() != b.size())
        throw std::runtime_error("Cannot set periodic boundaries: unmatched boundary faces");
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    for (size_t i = 0; i < a.size(); ++i) {
        int ka = a[i].second, kb = b[i].second;
        c.nbr[ka * 2] = kb / nv; c.nbr[ka * 2 + 1] = kb % nv;
        c.nbr[kb * 2] = ka / nv; c.nbr[kb * 2 + 1] = ka % nv;
        c.btag[ka] = ""; c.btag[kb] = "";
    }
}

void build_space(Ctx& c, int vol_rule) {
    // Triangles: volume rule exact to degree 2p; quads: tensor Gauss-Legendre with p+1 points
    // per direction, j outer (S/quadrature/factory.cpp:33-59); faces: GL with p+1 points
    // (S/core/space.cpp:37-55).
    gauss_legendre(c.p + 1, c.gs, c.gw);
    if (c.nv == 4) {
        c.qxi.clear(); c.qeta.clear(); c.qw.clear();
        for (int j = 0; j <= c.p; ++j)
            for (int i = 0; i <= c.p; ++i) {
                c.qxi.push_back(c.gs[i]);
                c.qeta.push_back(c.gs[j]);
                c.qw.push_back(c.gw[i] * c.gw[j]);
            }
    } else {
        int deg = 2 * c.p + (vol_rule ? 1 : 0);
        triangle_rule(deg, c.qxi, c.qeta, c.qw);
    }
    c.nq = (int)c.qw.size();
    c.nfq = (int)c.gw.size();
    c.phi_vol.assign(c.nq * c.np, 0);
    c.dphi_ref.assign(c.nq * c.np * 2, 0);
    for (int q = 0; q < c.nq; ++q) {
        c.basis.eval(c.qxi[q], c.qeta[q], &c.phi_vol[q * c.np]);
        c.basis.grad(c.qxi[q], c.qeta[q], &c.dphi_ref[q * c.np * 2]);
    }
    c.phi_face.assign(c.nv * c.nfq * c.np, 0);
    for (int f = 0; f < c.nv; ++f)
        for (int q = 0; q < c.nfq; ++q) {
            double x, y;
            face_point(c.nv, f, c.gs[q], x, y);
            c.basis.eval(x, y, &c.phi_face[(f * c.nfq + q) * c.np]);
        }
}

void build_geometry(Ctx& c) {
    const int nv = c.nv;
    c.ed.resize(c.K);
    c.fd.resize(c.K * nv);
    std::vector<double> g(c.np * 2);
    for (int e = 0; e < c.K; ++e) {
        ElemData& E = c.ed[e];
        for (int k = 0; k < nv; ++k) { E.V[2 * k] = c.vx[c.tri[e * nv + k]]; E.V[2 * k + 1] = c.vy[c.tri[e * nv + k]]; }
        E.jdet.resize(c.nq);
        E.dphi_dx.resize(c.nq * c.np * 2);
        for (int q = 0; q < c.nq; ++q) {
            Mapping m = compute_mapping(nv, E.V, c.qxi[q], c.qeta[q]);
            E.jdet[q] = m.det;
            for (int i = 0; i < c.np; ++i) {
                const double* gr = &c.dphi_ref[(q * c.np + i) * 2];
                E.dphi_dx[(q * c.np + i) * 2 + 0] = m.A[0] * gr[0] + m.A[1] * gr[1];
                E.dphi_dx[(q * c.np + i) * 2 + 1] = m.A[2] * gr[0] + m.A[3] * gr[1];
            }
        }
        // centroid = the vertex mean (colwise().mean())
        double cx = 0.0, cy = 0.0;
        for (int k = 0; k < nv; ++k) { cx += E.V[2 * k]; cy += E.V[2 * k + 1]; }
        cx /= nv;
        cy /= nv;
        for (int f = 0; f < nv; ++f) {
            FaceData& F = c.fd[e * nv + f];
            int a = f, b = (f + 1) % nv;
            double v1x = E.V[2 * a], v1y = E.V[2 * a + 1], v2x = E.V[2 * b], v2y = E.V[2 * b + 1];
            double tx = v2x - v1x, ty = v2y - v1y;
            double len = std::sqrt(tx * tx + ty * ty);
            if (len < 1e-14) throw std::runtime_error("Degenerate edge detected");
            double nx = ty / len, ny = -tx / len;
            double mx = 0.5 * (v1x + v2x) - cx, my = 0.5 * (v1y + v2y) - cy;
            if (nx * mx + ny * my < 0) { nx = -nx; ny = -ny; }
            F.normal[0] = nx; F.normal[1] = ny; F.length = len;
            F.qp_flat.assign(2 * c.nfq, 0);
            F.dphi_dx_face.assign(c.nfq * c.np * 2, 0);
            for (int q = 0; q < c.nfq; ++q) {
                double t = 0.5 * (c.gs[q] + 1.0);
                F.qp_flat[q] = (1.0 - t) * v1x + t * v2x;
                F.qp_flat[q + c.nfq] = (1.0 - t) * v1y + t * v2y;
                double xi, eta;
                face_point(nv, f, c.gs[q], xi, eta);
                Mapping m = compute_mapping(nv, E.V, xi, eta);
                c.basis.grad(xi, eta, g.data());
                for (int i = 0; i < c.np; ++i) {
                    int row = q * c.np + i;
                    F.dphi_dx_face[row * 2] = m.A[0] * g[2 * i] + m.A[1] * g[2 * i + 1];
                    F.dphi_dx_face[row * 2 + 1] = m.A[2] * g[2 * i] + m.A[3] * g[2 * i + 1];
                }
            }
        }
    }
}

// compute_face_permutation (S/core/space.cpp:173-228)
std::vector<int> face_permutation(const Ctx& c, int eL, int fL, int eR, int fR) {
    const double* VL = c.ed[eL].V;
    const double* VR = c.ed[eR].V;
    int a1 = fL, a2 = (fL + 1) % c.nv, b1 = fR, b2 = (fR + 1) % c.nv;
    double e1v1[2] = {VL[2 * a1], VL[2 * a1 + 1]}, e1v2[2] = {VL[2 * a2], VL[2 * a2 + 1]};
    double e2v1[2] = {VR[2 * b1], VR[2 * b1 + 1]}, e2v2[2] = {VR[2 * b2], VR[2 * b2 + 1]};
    auto dist = [](const double* a, const double* b) { return std::hypot(a[0] - b[0], a[1] - b[1]); };
    const double tol = 1e-10;
    bool rev_match = dist(e1v1, e2v2) < tol && dist(e1v2, e2v1) < tol;
    bool same_match = dist(e1v1, e2v1) < tol && dist(e1v2, e2v2) < tol;
    bool reverse;
    if (rev_match || same_match) reverse = rev_match;
    else {
        double t1[2] = {e1v2[0] - e1v1[0], e1v2[1] - e1v1[1]}, t2[2] = {e2v2[0] - e2v1[0], e2v2[1] - e2v1[1]};
        double n1 = std::hypot(t1[0], t1[1]), n2 = std::hypot(t2[0], t2[1]);
        reverse = (t1[0] / n1 * t2[0] / n2 + t1[1] / n1 * t2[1] / n2) < 0.0;
    }
    std::vector<int> perm(c.nfq);
    for (int i = 0; i < c.nfq; ++i) perm[i] = reverse ? c.nfq - 1 - i : i;
    return perm;
}

// build_precomputed_faces (S/core/mesh.cpp:404-480): interior faces deduplicated by element
// pair in element/face order; boundary faces in the same sweep.
void build_face_lists(Ctx& c) {
    c.ifaces.clear();
    c.bfaces.clear();
    std::set<std::pair<int, int>> done;
    for (int eL = 0; eL < c.K; ++eL)
        for (int fL = 0; fL < c.nv; ++fL) {
            int eR = c.nbr[(eL * c.nv + fL) * 2], fR = c.nbr[(eL * c.nv + fL) * 2 + 1];
            if (eR >= 0) {
                auto sig = std::make_pair(std::min(eL, eR), std::max(eL, eR));
                if (done.count(sig)) continue;
                done.insert(sig);
                c.ifaces.push_back({eL, fL, eR, fR, face_permutation(c, eL, fL, eR, fR)});
            } else {
                c.bfaces.push_back({eL, fL, c.btag[eL * c.nv + fL]});
            }
        }
}

// compute_mass_integral + inverse (S/weak_forms/weak_formulation_base.cpp:198-217,
// S/solver/compressible_solver_base.cpp:41-52); Gauss-Jordan with partial pivoting.
void build_minv(Ctx& c) {
    int n = c.np;
    c.minv.assign(c.K * n * n, 0);
    std::vector<double> M(n * n), I(n * n);
    for (int e = 0; e < c.K; ++e) {
        std::fill(M.begin(), M.end(), 0.0);
        for (int q = 0; q < c.nq; ++q) {
            double w = c.qw[q] * std::fabs(c.ed[e].jdet[q]);
            const double* ph = &c.phi_vol[q * n];
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) M[i * n + j] += w * (ph[i] * ph[j]);
        }
        for (int i = 0; i < n * n; ++i) I[i] = (i / n == i % n) ? 1.0 : 0.0;
        for (int col = 0; col < n; ++col) {
            int piv = col;
            for (int r = col + 1; r < n; ++r) if (std::fabs(M[r * n + col]) > std::fabs(M[piv * n + col])) piv = r;
            if (piv != col) for (int k = 0; k < n; ++k) { std::swap(M[col * n + k], M[piv * n + k]); std::swap(I[col * n + k], I[piv * n + k]); }
            double d = 1.0 / M[col * n + col];
            for (int k = 0; k < n; ++k) { M[col * n + k] *= d; I[col * n + k] *= d; }
            for (int r = 0; r < n; ++r) {
                if (r == col) continue;
                double f = M[r * n + col];
                if (f == 0.0) continue;
                for (int k = 0; k < n; ++k) { M[r * n + k] -= f * M[col * n + k]; I[r * n + k] -= f * I[col * n + k]; }
            }
        }
        std::copy(I.begin(), I.end(), c.minv.begin() + (size_t)e * n * n);
    }
}

// ---------------------------------------------------------------------------------------------
// Weak-form pieces.  Element coefficient blocks are [Np][4] row-major (u[i*4+v]), the
// reference's Np x 4 MatrixXd indexed (i, v).
// ---------------------------------------------------------------------------------------------

// EulerWeakFormulation::volume_residual (S/weak_forms/euler_weak_formulation.cpp:72-115)
void volume_residual(const Ctx& c, int e, const double* u, double* R) {
    const int np = c.np;
    std::fill(R, R + np * 4, 0.0);
    const ElemData& E = c.ed[e];
    for (int q = 0; q < c.nq; ++q) {
        V4 U = {0, 0, 0, 0};
        for (int i = 0; i < np; ++i)
            for (int v = 0; v < 4; ++v) U[v] += c.phi_vol[q * np + i] * u[i * 4 + v];
        V4 F, G, W;
        fluxes(U, c.ns.gamma, F, G, W);
        double w = c.qw[q] * std::fabs(E.jdet[q]);
        for (int i = 0; i < np; ++i) {
            double gx = E.dphi_dx[(q * np + i) * 2], gy = E.dphi_dx[(q * np + i) * 2 + 1];
            for (int v = 0; v < 4; ++v) R[i * 4 + v] += w * (gx * F[v] + gy * G[v]);
        }
    }
}

// Face-gradient read used by the NS face terms.  The reference stores dphi_dx_face interleaved
// (S/core/space.cpp:143,153-158) and reads it back through a column-major (Nfq*Np) x 2
// Eigen::Map (S/core/mesh.cpp:187-195), so row r reads (vec[r], vec[r + Nfq*Np]).
void face_grad(const Ctx& c, const FaceData& F, int q, int i, double* g) {
    int r = q * c.np + i;
    if (c.ns_face_grad == 0) { g[0] = F.dphi_dx_face[r]; g[1] = F.dphi_dx_face[r + c.nfq * c.np]; }
    else { g[0] = F.dphi_dx_face[2 * r]; g[1] = F.dphi_dx_face[2 * r + 1]; }
}

// NavierStokesWeakFormulation::viscous_volume_residual (navier_stokes_weak_formulation.cpp:100-140)
void viscous_volume_residual(const Ctx& c, int e, const double* u, double* R) {
    const int np = c.np;
    std::fill(R, R + np * 4, 0.0);
    const ElemData& E = c.ed[e];
    for (int q = 0; q < c.nq; ++q) {
        V4 U = {0, 0, 0, 0};
        for (int i = 0; i < np; ++i)
            for (int v = 0; v < 4; ++v) U[v] += c.phi_vol[q * np + i] * u[i * 4 + v];
        double gU[8] = {0};
        for (int i = 0; i < np; ++i) {
            double gx = E.dphi_dx[(q * np + i) * 2], gy = E.dphi_dx[(q * np + i) * 2 + 1];
            for (int v = 0; v < 4; ++v) { gU[v * 2] += u[i * 4 + v] * gx; gU[v * 2 + 1] += u[i * 4 + v] * gy; }
        }
        V4 Fv, Gv;
        viscous_flux(U, gU, c.ns, Fv, Gv);
        double w = c.qw[q] * std::fabs(E.jdet[q]);
        for (int i = 0; i < np; ++i) {
            double gx = E.dphi_dx[(q * np + i) * 2], gy = E.dphi_dx[(q * np + i) * 2 + 1];
            for (int v = 0; v < 4; ++v) R[i * 4 + v] -= w * (Fv[v] * gx + Gv[v] * gy);
        }
    }
}

// interior_face_residual (euler_weak_formulation.cpp:117-175)
void interior_face_residual(const Ctx& c, const IFace& fc, const double* uL, const double* uR,
                            double* RL, double* RR) {
    const int np = c.np;
    std::fill(RL, RL + np * 4, 0.0);
    std::fill(RR, RR + np * 4, 0.0);
    const FaceData& FL = c.fd[fc.eL * c.nv + fc.fL];
    const double* phL = &c.phi_face[fc.fL * c.nfq * np];
    const double* phR = &c.phi_face[fc.fR * c.nfq * np];
    for (int q = 0; q < c.nfq; ++q) {
        int qR = fc.perm[q];
        V4 UL = {0, 0, 0, 0}, UR = {0, 0, 0, 0};
        for (int i = 0; i < np; ++i)
            for (int v = 0; v < 4; ++v) {
                UL[v] += phL[q * np + i] * uL[i * 4 + v];
                UR[v] += phR[qR * np + i] * uR[i * 4 + v];
            }
        V4 H = rusanov(UL, UR, FL.normal, c.ns.gamma);
        double w = c.gw[q] * FL.length * 0.5;
        for (int i = 0; i < np; ++i) {
            double sL = w * phL[q * np + i], sR = w * phR[qR * np + i];
            for (int v = 0; v < 4; ++v) { RL[i * 4 + v] -= sL * H[v]; RR[i * 4 + v] += sR * H[v]; }
        }
    }
}

double penalty(const Ctx& c, double h) {  // compute_penalty_parameter (:275-278)
    double hs = std::max(h, 1e-12);
    return c.ns.sigma0 * c.ns.mu * (c.p + 1) * (c.p + 1) / hs;
}

// viscous_interior_face_residual (navier_stokes_weak_formulation.cpp:142-207)
void viscous_interior_face_residual(const Ctx& c, const IFace& fc, const double* uL,
                                    const double* uR, double* RL, double* RR) {
    const int np = c.np;
    std::fill(RL, RL + np * 4, 0.0);
    std::fill(RR, RR + np * 4, 0.0);
    const FaceData& FL = c.fd[fc.eL * c.nv + fc.fL];
    const FaceData& FR = c.fd[fc.eR * c.nv + fc.fR];
    const double* phL = &c.phi_face[fc.fL * c.nfq * np];
    const double* phR = &c.phi_face[fc.fR * c.nfq * np];
    double sigma = penalty(c, FL.length);
    for (int q = 0; q < c.nfq; ++q) {
        int qR = fc.perm[q];
        V4 UL = {0, 0, 0, 0}, UR = {0, 0, 0, 0};
        for (int i = 0; i < np; ++i)
            for (int v = 0; v < 4; ++v) {
                UL[v] += phL[q * np + i] * uL[i * 4 + v];
                UR[v] += phR[qR * np + i] * uR[i * 4 + v];
            }
        double gL[8] = {0}, gR[8] = {0};
        for (int i = 0; i < np; ++i) {
            double a[2], b[2];
            face_grad(c, FL, q, i, a);
            face_grad(c, FR, qR, i, b);
            for (int v = 0; v < 4; ++v) {
                gL[v * 2] += uL[i * 4 + v] * a[0]; gL[v * 2 + 1] += uL[i * 4 + v] * a[1];
                gR[v * 2] += uR[i * 4 + v] * b[0]; gR[v * 2 + 1] += uR[i * 4 + v] * b[1];
            }
        }
        V4 FvL, GvL, FvR, GvR;
        viscous_flux(UL, gL, c.ns, FvL, GvL);
        viscous_flux(UR, gR, c.ns, FvR, GvR);
        const double* n = FL.normal;
        V4 Fn;
        // Q4: the reference takes each side's flux along its own outward normal (:192-196), a
        // half difference, and subtracts it like the inviscid flux; the consistent flux enters
        // M dU/dt as +{Fv}.n, i.e. Fn = -1/2 (Fv_L + Fv_R).n_L here
        const double sL = c.ns_average_reference ? 1.0 : -1.0, sR = -1.0;
        for (int v = 0; v < 4; ++v) {
            double avg = 0.5 * sL * (FvL[v] * n[0] + GvL[v] * n[1]);
            avg += 0.5 * (FvR[v] * (sR * n[0]) + GvR[v] * (sR * n[1]));
            Fn[v] = avg - sigma * (UR[v] - UL[v]);
        }
        double w = c.gw[q] * FL.length * 0.5;
        for (int i = 0; i < np; ++i)
            for (int v = 0; v < 4; ++v) {
                RL[i * 4 + v] -= w * phL[q * np + i] * Fn[v];
                RR[i * 4 + v] += w * phR[qR * np + i] * Fn[v];
            }
    }
}

// boundary_face_residual (euler_weak_formulation.cpp:177-250).  The face points are read
// through a row-major n x 2 Map of the column-major storage (:188-194).
void boundary_face_residual(const Ctx& c, const BFace& bf, const BC& bc, const double* u, double* R) {
    const int np = c.np;
    std::fill(R, R + np * 4, 0.0);
    const FaceData& F = c.fd[bf.e * c.nv + bf.f];
    const double* ph = &c.phi_face[bf.f * c.nfq * np];
    const double* n = F.normal;
    for (int q = 0; q < c.nfq; ++q) {
        V4 UL = {0, 0, 0, 0};
        for (int i = 0; i < np; ++i)
            for (int v = 0; v < 4; ++v) UL[v] += ph[q * np + i] * u[i * 4 + v];
        V4 WL = cons2prim(UL, c.ns.gamma);
        // Q5 (the reference's read, default) or the point itself
        double xq = F.qp_flat[2 * q], yq = F.qp_flat[2 * q + 1];
        if (!c.euler_bc_points_reference) { xq = F.qp_flat[q]; yq = F.qp_flat[q + c.nfq]; }
        V4 UR;
        if (bc.type == 0) UR = bc.eval(xq, yq);
        else if (bc.type == 1) {
            double un = WL[1] * n[0] + WL[2] * n[1];
            double unx = un * n[0], uny = un * n[1];
            double utx = WL[1] - unx, uty = WL[2] - uny;
            UR = prim2cons({WL[0], utx - unx, uty - uny, WL[3]}, c.ns.gamma);
        } else if (bc.type == 2) UR = prim2cons(bc.eval(xq, yq), c.ns.gamma);
        else throw std::runtime_error("Unsupported BC type for Euler equations");
        V4 H = rusanov(UL, UR, n, c.ns.gamma);
        double w = c.gw[q] * F.length * 0.5;
        for (int i = 0; i < np; ++i) {
            double s = w * ph[q * np + i];
            for (int v = 0; v < 4; ++v) R[i * 4 + v] -= s * H[v];
        }
    }
}

// viscous_boundary_face_residual (navier_stokes_weak_formulation.cpp:209-273)
void viscous_boundary_face_residual(const Ctx& c, const BFace& bf, const BC* bc, const double* u, double* R) {
    const int np = c.np;
    std::fill(R, R + np * 4, 0.0);
    const FaceData& F = c.fd[bf.e * c.nv + bf.f];
    const double* ph = &c.phi_face[bf.f * c.nfq * np];
    const double* n = F.normal;
    double sigma = penalty(c, F.length);
    for (int q = 0; q < c.nfq; ++q) {
        V4 UL = {0, 0, 0, 0};
        for (int i = 0; i < np; ++i)
            for (int v = 0; v < 4; ++v) UL[v] += ph[q * np + i] * u[i * 4 + v];
        double gL[8] = {0};
        for (int i = 0; i < np; ++i) {
            double a[2];
            face_grad(c, F, q, i, a);
            for (int v = 0; v < 4; ++v) { gL[v * 2] += u[i * 4 + v] * a[0]; gL[v * 2 + 1] += u[i * 4 + v] * a[1]; }
        }
        V4 FvL, GvL;
        viscous_flux(UL, gL, c.ns, FvL, GvL);
        V4 Ubc = UL;
        if (bc) {
            V4 d = bc->eval(F.qp_flat[q], F.qp_flat[q + c.nfq]);
            Ubc = (bc->type == 2) ? prim2cons(d, c.ns.gamma) : d;
        }
        double w = c.gw[q] * F.length * 0.5;
        for (int v = 0; v < 4; ++v) {
            const double sb = c.ns_average_reference ? 1.0 : -1.0;  // Q4 (the consistent sign)
            double Fn = sb * (FvL[v] * n[0] + GvL[v] * n[1]) - sigma * (Ubc[v] - UL[v]);
            for (int i = 0; i < np; ++i) R[i * 4 + v] -= w * ph[q * np + i] * Fn;
        }
    }
}

const BC* face_bc(const Ctx& c, const BFace& bf) {
    auto it = c.bcs.find(bf.tag);
    if (it == c.bcs.end()) return nullptr;
    return &it->second;
}

// DGAssembler::assemble_euler_residual (S/solver/assembler.cpp:137-279)
void assemble(Ctx& c, const double* u, double* R) {
    const int nb = c.np * 4;
    std::vector<double> a(nb), b(nb);
    for (int e = 0; e < c.K; ++e) volume_residual(c, e, u + (size_t)e * nb, R + (size_t)e * nb);
    const bool visc = c.physics == 1;
    if (visc)
        for (int e = 0; e < c.K; ++e) {
            viscous_volume_residual(c, e, u + (size_t)e * nb, a.data());
            for (int k = 0; k < nb; ++k) R[(size_t)e * nb + k] += a[k];
        }
    for (const auto& fc : c.ifaces) {
        interior_face_residual(c, fc, u + (size_t)fc.eL * nb, u + (size_t)fc.eR * nb, a.data(), b.data());
        for (int k = 0; k < nb; ++k) { R[(size_t)fc.eL * nb + k] += a[k]; R[(size_t)fc.eR * nb + k] += b[k]; }
    }
    if (visc)
        for (const auto& fc : c.ifaces) {
            viscous_interior_face_residual(c, fc, u + (size_t)fc.eL * nb, u + (size_t)fc.eR * nb, a.data(), b.data());
            for (int k = 0; k < nb; ++k) { R[(size_t)fc.eL * nb + k] += a[k]; R[(size_t)fc.eR * nb + k] += b[k]; }
        }
    for (const auto& bf : c.bfaces) {
        const BC* bc = face_bc(c, bf);
        if (!bc) continue;
        boundary_face_residual(c, bf, *bc, u + (size_t)bf.e * nb, a.data());
        for (int k = 0; k < nb; ++k) R[(size_t)bf.e * nb + k] += a[k];
    }
    if (visc)
        for (const auto& bf : c.bfaces) {
            const BC* bc = face_bc(c, bf);
            if (!bc) continue;
            viscous_boundary_face_residual(c, bf, bc, u + (size_t)bf.e * nb, a.data());
            for (int k = 0; k < nb; ++k) R[(size_t)bf.e * nb + k] += a[k];
        }
}

// apply_mass_inv (S/solver/compressible_solver_base.cpp:54-61)
void apply_minv(const Ctx& c, const double* R, double* out) {
    const int n = c.np;
    for (int e = 0; e < c.K; ++e) {
        const double* M = &c.minv[(size_t)e * n * n];
        const double* r = R + (size_t)e * n * 4;
        double* o = out + (size_t)e * n * 4;
        for (int i = 0; i < n; ++i)
            for (int v = 0; v < 4; ++v) {
                double s = 0.0;
                for (int j = 0; j < n; ++j) s += M[i * n + j] * r[j * 4 + v];
                o[i * 4 + v] = s;
            }
    }
}

void rhs(Ctx& c, const double* u, double* k) {
    c.rbuf.resize((size_t)c.K * c.np * 4);
    assemble(c, u, c.rbuf.data());
    apply_minv(c, c.rbuf.data(), k);
}

// The same residual with threads, for the multi-core CPU baseline: elements and faces are
// processed in parallel and every face contribution is written to its own buffer, then each
// element adds its contributions in the order of the sequential assembly (volume, viscous
// volume, interior faces by index, viscous interior faces, boundary faces, viscous boundary
// faces), so the result is bitwise that of assemble().
void assemble_threaded(Ctx& c, const double* u, double* R) {
    const int nb = c.np * 4;
    const bool visc = c.physics == 1;
    const int nif = (int)c.ifaces.size(), nbf = (int)c.bfaces.size();
    if (c.eif.empty()) {
        c.eif.assign(c.K, {});
        c.ebf.assign(c.K, {});
        for (int k = 0; k < nif; ++k) {
            c.eif[c.ifaces[k].eL].push_back({k, 0});
            c.eif[c.ifaces[k].eR].push_back({k, 1});
        }
        for (int k = 0; k < nbf; ++k) c.ebf[c.bfaces[k].e].push_back(k);
        c.fl.assign((size_t)nif * nb, 0.0);
        c.fr.assign((size_t)nif * nb, 0.0);
        c.fb.assign((size_t)nbf * nb, 0.0);
        if (visc) {
            c.fvl.assign((size_t)nif * nb, 0.0);
            c.fvr.assign((size_t)nif * nb, 0.0);
            c.fvb.assign((size_t)nbf * nb, 0.0);
        }
    }
#pragma omp parallel
    {
#pragma omp for schedule(static)
        for (int k = 0; k < nif; ++k) {
            const auto& fc = c.ifaces[k];
            interior_face_residual(c, fc, u + (size_t)fc.eL * nb, u + (size_t)fc.eR * nb, &c.fl[(size_t)k * nb],
                                   &c.fr[(size_t)k * nb]);
            if (visc)
                viscous_interior_face_residual(c, fc, u + (size_t)fc.eL * nb, u + (size_t)fc.eR * nb,
                                               &c.fvl[(size_t)k * nb], &c.fvr[(size_t)k * nb]);
        }
#pragma omp for schedule(static)
        for (int k = 0; k < nbf; ++k) {
            const BC* bc = face_bc(c, c.bfaces[k]);
            if (!bc) continue;
            boundary_face_residual(c, c.bfaces[k], *bc, u + (size_t)c.bfaces[k].e * nb, &c.fb[(size_t)k * nb]);
            if (visc) viscous_boundary_face_residual(c, c.bfaces[k], bc, u + (size_t)c.bfaces[k].e * nb,
                                                     &c.fvb[(size_t)k * nb]);
        }
#pragma omp for schedule(static)
        for (int e = 0; e < c.K; ++e) {
            double* r = R + (size_t)e * nb;
            std::vector<double> a(nb);
            volume_residual(c, e, u + (size_t)e * nb, r);
            if (visc) {
                viscous_volume_residual(c, e, u + (size_t)e * nb, a.data());
                for (int k = 0; k < nb; ++k) r[k] += a[k];
            }
            for (const auto& [k, side] : c.eif[e]) {
                const double* x = &(side ? c.fr : c.fl)[(size_t)k * nb];
                for (int i = 0; i < nb; ++i) r[i] += x[i];
            }
            if (visc)
                for (const auto& [k, side] : c.eif[e]) {
                    const double* x = &(side ? c.fvr : c.fvl)[(size_t)k * nb];
                    for (int i = 0; i < nb; ++i) r[i] += x[i];
                }
            for (int k : c.ebf[e]) {
                if (!face_bc(c, c.bfaces[k])) continue;
                for (int i = 0; i < nb; ++i) r[i] += c.fb[(size_t)k * nb + i];
            }
            if (visc)
                for (int k : c.ebf[e]) {
                    if (!face_bc(c, c.bfaces[k])) continue;
                    for (int i = 0; i < nb; ++i) r[i] += c.fvb[(size_t)k * nb + i];
                }
        }
    }
}

void rk3_step_threaded(Ctx& c, double* u, double dt) {
    const size_t n = (size_t)c.K * c.np * 4;
    const int nb = c.np * 4;
    std::vector<double> k(n), u1(n), u2(n), r(n);
    auto L = [&](const double* x) {
        assemble_threaded(c, x, r.data());
#pragma omp parallel for schedule(static)
        for (int e = 0; e < c.K; ++e) {
            const int np = c.np;
            const double* M = &c.minv[(size_t)e * np * np];
            for (int i = 0; i < np; ++i)
                for (int v = 0; v < 4; ++v) {
                    double s = 0.0;
                    for (int j = 0; j < np; ++j) s += M[i * np + j] * r[(size_t)e * nb + j * 4 + v];
                    k[(size_t)e * nb + i * 4 + v] = s;
                }
        }
    };
    L(u);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) u1[i] = u[i] + dt * k[i];
    L(u1.data());
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) u2[i] = 0.75 * u[i] + 0.25 * (u1[i] + dt * k[i]);
    L(u2.data());
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) u[i] = (1.0 / 3.0) * u[i] + (2.0 / 3.0) * (u2[i] + dt * k[i]);
}

// SSP_RK::step_rk3 (I/solver/time_stepping.hpp:39-55) with the reference's operation order:
// u1 = u + dt k1; u2 = 0.75 u + 0.25 (u1 + dt k2); u' = (1/3) u + (2/3) (u2 + dt k3).
void rk3_step(Ctx& c, double* u, double dt) {
    const size_t n = (size_t)c.K * c.np * 4;
    std::vector<double> k(n), u1(n), u2(n);
    rhs(c, u, k.data());
    for (size_t i = 0; i < n; ++i) u1[i] = u[i] + dt * k[i];
    rhs(c, u1.data(), k.data());
    for (size_t i = 0; i < n; ++i) u2[i] = 0.75 * u[i] + 0.25 * (u1[i] + dt * k[i]);
    rhs(c, u2.data(), k.data());
    for (size_t i = 0; i < n; ++i) u[i] = (1.0 / 3.0) * u[i] + (2.0 / 3.0) * (u2[i] + dt * k[i]);
}

}  // namespace

// =============================================================================================
// C API
// =============================================================================================
extern "C" {

const char* ref_last_error(void) { return g_err.c_str(); }

void* ref_create(int nv, const double* verts, int K, const int32_t* elems, int order, int basis,
                 int vol_rule, int periodic_mask, int physics, double gamma, double mu,
                 double prandtl, double sigma0, int ns_face_grad, int nvpe, int legendre_endpoint_ref) {
    try {
        if (order < 1 || order > 4) throw std::invalid_argument("order must be 1..4");
        if (nvpe != 3 && nvpe != 4) throw std::invalid_argument("elements must be triangles or quads");
        // the reference's quad space always uses the Legendre basis (S/core/space.cpp:309-318)
        if ((nvpe == 4) != (basis == 2)) throw std::invalid_argument("Legendre basis on quads only");
        if (physics == 1) {  // NavierStokesWeakFormulation ctor checks (:18-34)
            if (mu <= 0) throw std::invalid_argument("Dynamic viscosity must be positive");
            if (prandtl <= 0) throw std::invalid_argument("Prandtl number must be positive");
            if (sigma0 <= 0) throw std::invalid_argument("Penalty prefactor must be positive");
        }
        auto* c = new Ctx();
        c->K = K;
        c->p = order;
        c->nv = nvpe;
        c->basis = Basis(basis, order);
        c->basis.ref_endpoint = legendre_endpoint_ref != 0;
        c->np = c->basis.np;
        c->physics = physics;
        c->ns_face_grad = ns_face_grad;
        c->ns = {gamma, mu, prandtl, sigma0, 1.0};
        c->vx.resize(nv); c->vy.resize(nv);
        for (int i = 0; i < nv; ++i) { c->vx[i] = verts[2 * i]; c->vy[i] = verts[2 * i + 1]; }
        c->tri.assign(elems, elems + (size_t)nvpe * K);
        build_connectivity(*c);
        tag_boundaries(*c);
        if (periodic_mask & 1) pair_periodic(*c, "Left", "Right", true);
        if (periodic_mask & 2) pair_periodic(*c, "Bottom", "Top", false);
        build_space(*c, vol_rule);
        build_geometry(*c);
        build_face_lists(*c);
        build_minv(*c);
        return c;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return nullptr;
    }
}

void ref_destroy(void* h) { delete static_cast<Ctx*>(h); }

void ref_info(void* h, int32_t* o) {
    Ctx& c = *static_cast<Ctx*>(h);
    o[0] = c.K; o[1] = c.np; o[2] = c.nq; o[3] = c.nfq;
    o[4] = (int32_t)c.ifaces.size(); o[5] = (int32_t)c.bfaces.size();
}

// Q5 switch: 1 = the reference's transposed read of the Euler boundary points, 0 = the points.
int ref_set_euler_bc_points(void* h, int reference) {
    if (!h) return -1;
    static_cast<Ctx*>(h)->euler_bc_points_reference = reference ? 1 : 0;
    return 0;
}

// Q17 switch: 1 = the reference's heat-flux sign in the energy viscous flux, 0 = conduction.
int ref_set_heat_flux(void* h, int reference) {
    if (!h) return -1;
    static_cast<Ctx*>(h)->ns.heat_sign = reference ? 1.0 : -1.0;
    return 0;
}

// Q4 switch: 1 = the reference's viscous face fluxes (interior half difference, subtracted like
// the inviscid flux), 0 = the consistent +{Fv}.n (interior average, boundary Fv.n).
int ref_set_ns_average(void* h, int reference) {
    if (!h) return -1;
    static_cast<Ctx*>(h)->ns_average_reference = reference ? 1 : 0;
    return 0;
}

int ref_set_bc(void* h, const char* tag, int type, const double* value4, ref_bc_fn fn, void* user) {
    Ctx& c = *static_cast<Ctx*>(h);
    BC bc;
    bc.type = type;
    if (value4) for (int v = 0; v < 4; ++v) bc.value[v] = value4[v];
    bc.fn = fn;
    bc.user = user;
    c.bcs[tag] = bc;
    return 0;
}

#define REF_TRY(...) try { __VA_ARGS__; return 0; } catch (const std::exception& ex) { g_err = ex.what(); return 1; }

int ref_residual(void* h, const double* u, double* r) { REF_TRY(assemble(*static_cast<Ctx*>(h), u, r)) }
int ref_apply_minv(void* h, double* r) {
    REF_TRY({
        Ctx& c = *static_cast<Ctx*>(h);
        std::vector<double> t(r, r + (size_t)c.K * c.np * 4);
        apply_minv(c, t.data(), r);
    })
}
int ref_rhs(void* h, const double* u, double* k) { REF_TRY(rhs(*static_cast<Ctx*>(h), u, k)) }

double ref_rk3(void* h, double* u, double dt, int nsteps) {
    Ctx& c = *static_cast<Ctx*>(h);
    auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < nsteps; ++s) rk3_step(c, u, dt);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// nsteps with the threaded residual on `nthreads` OpenMP threads (bitwise equal to ref_rk3).
double ref_rk3_threads(void* h, double* u, double dt, int nsteps, int nthreads) {
    Ctx& c = *static_cast<Ctx*>(h);
    omp_set_num_threads(nthreads > 0 ? nthreads : 1);
    auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < nsteps; ++s) rk3_step_threaded(c, u, dt);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// project_initial_condition (S/solver/compressible_solver_base.cpp:63-104), IC values given at
// the volume points.
int ref_project(void* h, const double* Uq, double* u) {
    REF_TRY({
        Ctx& c = *static_cast<Ctx*>(h);
        const int np = c.np;
        std::vector<double> F((size_t)c.K * np * 4, 0.0);
        for (int e = 0; e < c.K; ++e)
            for (int q = 0; q < c.nq; ++q) {
                double w = c.qw[q] * std::fabs(c.ed[e].jdet[q]);
                const double* U = Uq + ((size_t)e * c.nq + q) * 4;
                for (int i = 0; i < np; ++i)
                    for (int v = 0; v < 4; ++v) F[((size_t)e * np + i) * 4 + v] += w * c.phi_vol[q * np + i] * U[v];
            }
        apply_minv(c, F.data(), u);
    })
}

int ref_volume_points(void* h, double* xy) {
    REF_TRY({
        Ctx& c = *static_cast<Ctx*>(h);
        for (int e = 0; e < c.K; ++e)
            for (int q = 0; q < c.nq; ++q) {
                Mapping m = compute_mapping(c.nv, c.ed[e].V, c.qxi[q], c.qeta[q]);
                xy[((size_t)e * c.nq + q) * 2] = m.x;
                xy[((size_t)e * c.nq + q) * 2 + 1] = m.y;
            }
    })
}

int ref_eval_volume(void* h, const double* u, double* Uq) {
    REF_TRY({
        Ctx& c = *static_cast<Ctx*>(h);
        for (int e = 0; e < c.K; ++e)
            for (int q = 0; q < c.nq; ++q)
                for (int v = 0; v < 4; ++v) {
                    double s = 0.0;
                    for (int i = 0; i < c.np; ++i) s += c.phi_vol[q * c.np + i] * u[((size_t)e * c.np + i) * 4 + v];
                    Uq[((size_t)e * c.nq + q) * 4 + v] = s;
                }
    })
}

// compute_max_cfl (S/solver/compressible_solver_base.cpp:114-170)
double ref_max_cfl(void* h, const double* u, double dt) {
    Ctx& c = *static_cast<Ctx*>(h);
    double g = c.ns.gamma, mx = 0.0;
    for (int e = 0; e < c.K; ++e) {
        const double* V = c.ed[e].V;
        double he = 1e300;
        for (int i = 0; i < c.nv; ++i)
            for (int j = i + 1; j < c.nv; ++j) he = std::min(he, std::hypot(V[2 * i] - V[2 * j], V[2 * i + 1] - V[2 * j + 1]));
        double ws = 0.0;
        for (int q = 0; q < c.nq; ++q) {
            double U[4] = {0, 0, 0, 0};
            for (int i = 0; i < c.np; ++i)
                for (int v = 0; v < 4; ++v) U[v] += c.phi_vol[q * c.np + i] * u[((size_t)e * c.np + i) * 4 + v];
            double rho = U[0];
            if (rho <= 0.0 || !std::isfinite(rho)) return INFINITY;
            double uu = U[1] / rho, vv = U[2] / rho;
            double p = (g - 1.0) * (U[3] - rho * 0.5 * (uu * uu + vv * vv));
            if (p <= 0.0 || !std::isfinite(p)) return INFINITY;
            double cs = std::sqrt(std::max(g * p / rho, 0.0));
            ws = std::max(ws, std::sqrt(uu * uu + vv * vv) + cs);
        }
        mx = std::max(mx, dt * ws / he);
    }
    return mx;
}

// compute_density_range (:172-199)
int ref_density_range(void* h, const double* u, double* out2) {
    Ctx& c = *static_cast<Ctx*>(h);
    double lo = INFINITY, hi = -INFINITY;
    for (int e = 0; e < c.K; ++e)
        for (int q = 0; q < c.nq; ++q) {
            double r = 0.0;
            for (int i = 0; i < c.np; ++i) r += c.phi_vol[q * c.np + i] * u[((size_t)e * c.np + i) * 4];
            if (!std::isfinite(r)) continue;
            lo = std::min(lo, r); hi = std::max(hi, r);
        }
    if (!std::isfinite(lo) || !std::isfinite(hi)) { out2[0] = out2[1] = NAN; return 0; }
    out2[0] = lo; out2[1] = hi;
    return 0;
}

// Integral of each conserved variable over the domain: sum_e sum_q w_q |J| U_q.
int ref_conservation_sums(void* h, const double* u, double* out4) {
    Ctx& c = *static_cast<Ctx*>(h);
    for (int v = 0; v < 4; ++v) out4[v] = 0.0;
    for (int e = 0; e < c.K; ++e)
        for (int q = 0; q < c.nq; ++q) {
            double w = c.qw[q] * std::fabs(c.ed[e].jdet[q]);
            for (int v = 0; v < 4; ++v) {
                double s = 0.0;
                for (int i = 0; i < c.np; ++i) s += c.phi_vol[q * c.np + i] * u[((size_t)e * c.np + i) * 4 + v];
                out4[v] += w * s;
            }
        }
    return 0;
}

// J = sum_e sum_q w_q |J_e| (m_x^2 + m_y^2) / (2 rho)
double ref_kinetic_energy(void* h, const double* u) {
    Ctx& c = *static_cast<Ctx*>(h);
    double J = 0.0;
    for (int e = 0; e < c.K; ++e)
        for (int q = 0; q < c.nq; ++q) {
            double U[4] = {0, 0, 0, 0};
            for (int i = 0; i < c.np; ++i)
                for (int v = 0; v < 4; ++v) U[v] += c.phi_vol[q * c.np + i] * u[((size_t)e * c.np + i) * 4 + v];
            J += c.qw[q] * std::fabs(c.ed[e].jdet[q]) * 0.5 * (U[1] * U[1] + U[2] * U[2]) / U[0];
        }
    return J;
}

int ref_quad_tri(int degree, double* xi, double* eta, double* w) {
    try {
        std::vector<double> a, b, c;
        triangle_rule(degree, a, b, c);
        if (xi) { std::copy(a.begin(), a.end(), xi); std::copy(b.begin(), b.end(), eta); std::copy(c.begin(), c.end(), w); }
        return (int)c.size();
    } catch (const std::exception& ex) { g_err = ex.what(); return -1; }
}

int ref_gauss_legendre(int n, double* x, double* w) {
    std::vector<double> a, b;
    gauss_legendre(n, a, b);
    std::copy(a.begin(), a.end(), x);
    std::copy(b.begin(), b.end(), w);
    return n;
}

int ref_basis_eval(int kind, int order, double xi, double eta, double* phi, double* grad) {
    Basis b(kind % 4, order);
    b.ref_endpoint = kind < 4;  // kind + 4: Legendre derivative with P'_n(+1) = n(n+1)/2
    if (phi) b.eval(xi, eta, phi);
    if (grad) b.grad(xi, eta, grad);
    return b.np;
}

int ref_mapping(int nvpe, const double* V, double xi, double eta, double* det, double* A, double* x2) {
    REF_TRY({
        Mapping m = compute_mapping(nvpe, V, xi, eta);
        *det = m.det;
        for (int k = 0; k < 4; ++k) A[k] = m.A[k];
        x2[0] = m.x; x2[1] = m.y;
    })
}

int ref_face_point(int nvpe, int f, double s, double* xi2) { REF_TRY(face_point(nvpe, f, s, xi2[0], xi2[1])) }

int ref_elem_data(void* h, int e, double* jdet, double* dphi) {
    Ctx& c = *static_cast<Ctx*>(h);
    std::copy(c.ed[e].jdet.begin(), c.ed[e].jdet.end(), jdet);
    std::copy(c.ed[e].dphi_dx.begin(), c.ed[e].dphi_dx.end(), dphi);
    return 0;
}

int ref_face_geom(void* h, int e, int f, double* n2, double* len) {
    Ctx& c = *static_cast<Ctx*>(h);
    const FaceData& F = c.fd[e * c.nv + f];
    n2[0] = F.normal[0]; n2[1] = F.normal[1]; *len = F.length;
    return 0;
}

int ref_interior_faces(void* h, int32_t* out5) {
    Ctx& c = *static_cast<Ctx*>(h);
    for (size_t k = 0; k < c.ifaces.size(); ++k) {
        const auto& f = c.ifaces[k];
        out5[k * 5] = f.eL; out5[k * 5 + 1] = f.fL; out5[k * 5 + 2] = f.eR; out5[k * 5 + 3] = f.fR;
        out5[k * 5 + 4] = (c.nfq > 1 && f.perm[0] != 0) ? 1 : 0;
    }
    return (int)c.ifaces.size();
}

int ref_neighbors(void* h, int32_t* out) {
    Ctx& c = *static_cast<Ctx*>(h);
    std::copy(c.nbr.begin(), c.nbr.end(), out);
    return 0;
}

int ref_boundary_faces(void* h, int32_t* out3) {
    Ctx& c = *static_cast<Ctx*>(h);
    for (size_t k = 0; k < c.bfaces.size(); ++k) {
        out3[k * 3] = c.bfaces[k].e; out3[k * 3 + 1] = c.bfaces[k].f;
        out3[k * 3 + 2] = face_bc(c, c.bfaces[k]) ? 1 : 0;
    }
    return (int)c.bfaces.size();
}

int ref_volume_residual(void* h, int e, const double* ue, double* R) {
    REF_TRY(volume_residual(*static_cast<Ctx*>(h), e, ue, R))
}
int ref_viscous_volume_residual(void* h, int e, const double* ue, double* R) {
    REF_TRY(viscous_volume_residual(*static_cast<Ctx*>(h), e, ue, R))
}
int ref_interior_face_residual(void* h, int k, const double* uL, const double* uR, double* RL,
                               double* RR, int viscous) {
    REF_TRY({
        Ctx& c = *static_cast<Ctx*>(h);
        if (viscous) viscous_interior_face_residual(c, c.ifaces.at(k), uL, uR, RL, RR);
        else interior_face_residual(c, c.ifaces.at(k), uL, uR, RL, RR);
    })
}
int ref_boundary_face_residual(void* h, int k, const double* ue, double* R, int viscous) {
    REF_TRY({
        Ctx& c = *static_cast<Ctx*>(h);
        const BFace& bf = c.bfaces.at(k);
        const BC* bc = face_bc(c, bf);
        if (viscous) viscous_boundary_face_residual(c, bf, bc, ue, R);
        else {
            if (!bc) throw std::runtime_error("boundary face has no boundary condition");
            boundary_face_residual(c, bf, *bc, ue, R);
        }
    })
}
int ref_rusanov(double gamma, const double* UL, const double* UR, const double* n, double* H) {
    V4 a = {UL[0], UL[1], UL[2], UL[3]}, b = {UR[0], UR[1], UR[2], UR[3]};
    V4 r = rusanov(a, b, n, gamma);
    for (int v = 0; v < 4; ++v) H[v] = r[v];
    return 0;
}
int ref_viscous_flux(void* h, const double* U, const double* g, double* Fv, double* Gv) {
    Ctx& c = *static_cast<Ctx*>(h);
    V4 a = {U[0], U[1], U[2], U[3]}, F, G;
    viscous_flux(a, g, c.ns, F, G);
    for (int v = 0; v < 4; ++v) { Fv[v] = F[v]; Gv[v] = G[v]; }
    return 0;
}
double ref_penalty(void* h, double length) { return penalty(*static_cast<Ctx*>(h), length); }

}  // extern "C"
