// Quadrature rules and triangle bases for the device tables (FP64, host, once per context).
#include "dgx_space.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace dgx {

Rule1D gauss_legendre(int n) {
    if (n < 1 || n > 16) throw std::invalid_argument("Gauss-Legendre order must be 1..16");
    Rule1D r;
    r.x.assign(n, 0.0);
    r.w.assign(n, 0.0);
    if (n == 1) { r.w[0] = 2.0; return r; }
    // Newton on P_n from the Tricomi-style guess cos(pi (k + 3/4)/(n + 1/2)); derivative from
    // (z^2 - 1) P_n' = n (z P_n - P_{n-1}).
    for (int k = 0; k < (n + 1) / 2; ++k) {
        double z = std::cos(M_PI * (k + 0.75) / (n + 0.5));
        double pn = 0, pm = 0;
        for (int it = 0; it < 64; ++it) {
            pm = 1.0;
            pn = z;
            for (int m = 2; m <= n; ++m) {
                double t = ((2 * m - 1) * z * pn - (m - 1) * pm) / m;
                pm = pn;
                pn = t;
            }
            double d = n * (z * pn - pm) / (z * z - 1.0);
            double step = pn / d;
            z -= step;
            if (std::fabs(step) <= 1e-16) break;
        }
        pm = 1.0;
        pn = z;
        for (int m = 2; m <= n; ++m) {
            double t = ((2 * m - 1) * z * pn - (m - 1) * pm) / m;
            pm = pn;
            pn = t;
        }
        double d = n * (z * pn - pm) / (z * z - 1.0);
        double w = 2.0 / ((1.0 - z * z) * d * d);
        r.x[k] = -z;          // mirrored pair written together keeps the rule exactly symmetric
        r.x[n - 1 - k] = z;
        r.w[k] = w;
        r.w[n - 1 - k] = w;
    }
    if (n & 1) r.x[n / 2] = 0.0;
    return r;
}

namespace {

// Dunavant (1985) symmetric orbits: kind 0 centroid, kind 1 (a,b,b) with b = (1-a)/2,
// kind 2 (a,b,c) with c = 1-a-b; weights normalised to a unit sum.  Published to 15 digits;
// polished below to double precision on the moment equations.
struct Orb { int kind; double a, b, w; };

std::vector<Orb> orbits(int degree) {
    switch (degree) {
    case 1: return {{0, 0, 0, 1.0}};
    case 2: return {{1, 2.0 / 3.0, 0, 1.0 / 3.0}};
    case 3: return {{0, 0, 0, -0.5625}, {1, 0.6, 0, 25.0 / 48.0}};
    case 4: return {{1, 0.108103018168070, 0, 0.223381589678011}, {1, 0.816847572980459, 0, 0.109951743655322}};
    case 5: return {{0, 0, 0, 0.225}, {1, 0.059715871789770, 0, 0.132394152788506},
                    {1, 0.797426985353087, 0, 0.125939180544827}};
    case 6: return {{1, 0.501426509658179, 0, 0.116786275726379}, {1, 0.873821971016996, 0, 0.050844906370207},
                    {2, 0.053145049844817, 0.310352451033784, 0.082851075618374}};
    case 7: return {{0, 0, 0, -0.149570044467682}, {1, 0.479308067841920, 0, 0.175615257433208},
                    {1, 0.869739794195568, 0, 0.053347235608838},
                    {2, 0.048690315425316, 0.312865496004874, 0.077113760890257}};
    case 8: return {{0, 0, 0, 0.144315607677787}, {1, 0.081414823414554, 0, 0.095091634267285},
                    {1, 0.658861384496480, 0, 0.103217370534718}, {1, 0.898905543365938, 0, 0.032458497623198},
                    {2, 0.008394777409958, 0.263112829634638, 0.027230314174435}};
    case 9: return {{0, 0, 0, 0.097135796282799}, {1, 0.020634961602525, 0, 0.031334700227139},
                    {1, 0.125820817014127, 0, 0.077827541004774}, {1, 0.623592928761935, 0, 0.079647738927210},
                    {1, 0.910540973211095, 0, 0.025577675658698},
                    {2, 0.036838412054736, 0.221962989160766, 0.043283539377289}};
    default: throw std::invalid_argument("triangle rule degree must be 1..9");
    }
}

RuleTri expand(const std::vector<Orb>& o) {
    RuleTri r;
    auto add = [&](double l1, double l2, double l3, double w) {
        (void)l1;
        r.xi.push_back(l2);
        r.eta.push_back(l3);
        r.w.push_back(0.5 * w);
    };
    for (const Orb& s : o) {
        if (s.kind == 0) add(1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, s.w);
        else if (s.kind == 1) {
            double b = 0.5 * (1.0 - s.a);
            add(s.a, b, b, s.w); add(b, s.a, b, s.w); add(b, b, s.a, s.w);
        } else {
            double c = 1.0 - s.a - s.b;
            add(s.a, s.b, c, s.w); add(s.a, c, s.b, s.w); add(s.b, s.a, c, s.w);
            add(s.b, c, s.a, s.w); add(c, s.a, s.b, s.w); add(c, s.b, s.a, s.w);
        }
    }
    return r;
}

// exact moments of the reference triangle: int xi^a eta^b = a! b! / (a+b+2)!
double tri_moment(int a, int b) {
    double num = 1.0;
    for (int k = 2; k <= a; ++k) num *= k;
    for (int k = 2; k <= b; ++k) num *= k;
    double den = 1.0;
    for (int k = 2; k <= a + b + 2; ++k) den *= k;
    return num / den;
}

std::vector<double> residuals(const std::vector<Orb>& o, int degree) {
    RuleTri r = expand(o);
    std::vector<double> res;
    for (int d = 0; d <= degree; ++d)
        for (int a = 0; a <= d; ++a) {
            double s = 0.0;
            for (size_t q = 0; q < r.w.size(); ++q) s += r.w[q] * std::pow(r.xi[q], a) * std::pow(r.eta[q], d - a);
            res.push_back(s - tri_moment(a, d - a));
        }
    return res;
}

// Gauss-Newton polish of the orbit parameters (finite-difference Jacobian, normal equations
// solved by Cholesky).
void polish(std::vector<Orb>& o, int degree) {
    std::vector<double*> par;
    for (Orb& s : o) {
        if (s.kind >= 1) par.push_back(&s.a);
        if (s.kind == 2) par.push_back(&s.b);
        par.push_back(&s.w);
    }
    const int n = (int)par.size();
    for (int it = 0; it < 6; ++it) {
        std::vector<double> r0 = residuals(o, degree);
        const int m = (int)r0.size();
        std::vector<double> J((size_t)m * n);
        for (int j = 0; j < n; ++j) {
            double save = *par[j], h = 1e-6;
            *par[j] = save + h;
            std::vector<double> rp = residuals(o, degree);
            *par[j] = save - h;
            std::vector<double> rm = residuals(o, degree);
            *par[j] = save;
            for (int i = 0; i < m; ++i) J[(size_t)i * n + j] = (rp[i] - rm[i]) / (2 * h);
        }
        std::vector<double> N((size_t)n * n, 0.0), g(n, 0.0);
        for (int i = 0; i < m; ++i)
            for (int a = 0; a < n; ++a) {
                g[a] -= J[(size_t)i * n + a] * r0[i];
                for (int b = 0; b < n; ++b) N[(size_t)a * n + b] += J[(size_t)i * n + a] * J[(size_t)i * n + b];
            }
        // Cholesky N = L L^T
        bool ok = true;
        for (int j = 0; j < n && ok; ++j) {
            double d = N[(size_t)j * n + j];
            for (int k = 0; k < j; ++k) d -= N[(size_t)j * n + k] * N[(size_t)j * n + k];
            if (d <= 0) { ok = false; break; }
            d = std::sqrt(d);
            N[(size_t)j * n + j] = d;
            for (int i = j + 1; i < n; ++i) {
                double s = N[(size_t)i * n + j];
                for (int k = 0; k < j; ++k) s -= N[(size_t)i * n + k] * N[(size_t)j * n + k];
                N[(size_t)i * n + j] = s / d;
            }
        }
        if (!ok) break;
        for (int i = 0; i < n; ++i) {
            double s = g[i];
            for (int k = 0; k < i; ++k) s -= N[(size_t)i * n + k] * g[k];
            g[i] = s / N[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = g[i];
            for (int k = i + 1; k < n; ++k) s -= N[(size_t)k * n + i] * g[k];
            g[i] = s / N[(size_t)i * n + i];
        }
        for (int j = 0; j < n; ++j) *par[j] += g[j];
    }
}

}  // namespace

RuleTri dunavant(int degree) {
    std::vector<Orb> o = orbits(degree);
    polish(o, degree);
    RuleTri r = expand(o);
    for (double e : residuals(o, degree))
        if (std::fabs(e) > 1e-14) throw std::runtime_error("triangle rule failed its moment check");
    return r;
}

TriBasis::TriBasis(int kind_, int p_) : kind(kind_), p(p_), np((p_ + 1) * (p_ + 2) / 2) {
    if (kind != 0 && kind != 1) throw std::invalid_argument("unknown basis kind");
}

namespace {
// Jacobi P_n^{(al,0)}(x) and its derivative by the standard three-term recurrence
// (Szego; as S/basis/dubiner.cpp:24-53 with beta = 0).
void jacobi_al0(int n, double al, double x, double& P, double& dP) {
    auto eval = [](int nn, double a, double b, double xx) {
        if (nn == 0) return 1.0;
        double p0 = 1.0, p1 = 0.5 * (a - b + (a + b + 2.0) * xx);
        for (int k = 2; k <= nn; ++k) {
            double c = 2.0 * k + a + b;
            double a1 = 2.0 * k * (k + a + b) * (c - 2.0);
            double a2 = (c - 1.0) * (a * a - b * b);
            double a3 = (c - 2.0) * (c - 1.0) * c;
            double a4 = 2.0 * (k + a - 1.0) * (k + b - 1.0) * c;
            double p2 = ((a2 + a3 * xx) * p1 - a4 * p0) / a1;
            p0 = p1;
            p1 = p2;
        }
        return p1;
    };
    P = eval(n, al, 0.0, x);
    dP = (n == 0) ? 0.0 : 0.5 * (n + al + 1.0) * eval(n - 1, al + 1.0, 1.0, x);
}
}  // namespace

void TriBasis::values(double xi, double eta, double* phi) const {
    if (kind == 0) {
        int k = 0;
        for (int d = 0; d <= p; ++d)
            for (int i = 0; i <= d; ++i) phi[k++] = std::pow(xi, i) * std::pow(eta, d - i);
        return;
    }
    // Dubiner: clamp into the triangle (ReferenceTriangle::project_to_bounds), collapse, evaluate
    if (!(xi >= 0 && eta >= 0 && xi + eta <= 1)) {
        xi = std::max(0.0, std::min(1.0, xi));
        eta = std::max(0.0, std::min(1.0 - xi, eta));
    }
    double om = 1.0 - eta;
    double r = std::fabs(om) < 1e-14 ? -1.0 : 2.0 * xi / om - 1.0;
    double s = 2.0 * eta - 1.0;
    int k = 0;
    for (int j = 0; j <= p; ++j)
        for (int i = 0; i <= p - j; ++i) {
            double Pi, dPi, Pj, dPj;
            jacobi_al0(i, 0.0, r, Pi, dPi);
            jacobi_al0(j, 2.0 * i + 1.0, s, Pj, dPj);
            phi[k++] = 2.0 * std::sqrt((2.0 * i + 1.0) * (i + j + 1.0)) * Pi * Pj * std::pow(0.5 * (1.0 - s), i);
        }
}

void TriBasis::gradients(double xi, double eta, double* g) const {
    if (kind == 0) {
        int k = 0;
        for (int d = 0; d <= p; ++d)
            for (int i = 0; i <= d; ++i, ++k) {
                int j = d - i;
                g[2 * k] = i ? i * std::pow(xi, i - 1) * std::pow(eta, j) : 0.0;
                g[2 * k + 1] = j ? j * std::pow(xi, i) * std::pow(eta, j - 1) : 0.0;
            }
        return;
    }
    if (!(xi >= 0 && eta >= 0 && xi + eta <= 1)) {
        xi = std::max(0.0, std::min(1.0, xi));
        eta = std::max(0.0, std::min(1.0 - xi, eta));
    }
    double om = 1.0 - eta;
    bool apex = std::fabs(om) < 1e-14;
    double r = apex ? -1.0 : 2.0 * xi / om - 1.0;
    double s = 2.0 * eta - 1.0;
    double rx = apex ? 0.0 : 2.0 / om, ry = apex ? 0.0 : 2.0 * xi / (om * om), sy = apex ? 0.0 : 2.0;
    int k = 0;
    for (int j = 0; j <= p; ++j)
        for (int i = 0; i <= p - j; ++i, ++k) {
            double Pi, dPi, Pj, dPj;
            jacobi_al0(i, 0.0, r, Pi, dPi);
            jacobi_al0(j, 2.0 * i + 1.0, s, Pj, dPj);
            double c = 2.0 * std::sqrt((2.0 * i + 1.0) * (i + j + 1.0));
            double h = std::pow(0.5 * (1.0 - s), i);
            double dh = i ? -0.5 * i * std::pow(0.5 * (1.0 - s), i - 1) : 0.0;
            double dr = c * dPi * Pj * h;
            double ds = c * Pi * (dPj * h + Pj * dh);
            g[2 * k] = dr * rx;
            g[2 * k + 1] = dr * ry + ds * sy;
        }
}

void face_point(int f, double s, double& xi, double& eta) {
    double t = 0.5 * (s + 1.0);
    if (f == 0) { xi = t; eta = 0.0; }
    else if (f == 1) { xi = 1.0 - t; eta = t; }
    else if (f == 2) { xi = 0.0; eta = 1.0 - t; }
    else throw std::out_of_range("Invalid face for triangle");
}

// P_n and P'_n together by the three-term recurrence (P'_{k} = k P_{k-1} + x P'_{k-1}), which
// is exact at x = +-1; the reference's endpoint rule is applied on top when asked for.
static void legendre_pd(int n, double x, bool ref_endpoint, double* P, double* dP) {
    P[0] = 1.0;
    dP[0] = 0.0;
    if (n >= 1) { P[1] = x; dP[1] = 1.0; }
    for (int k = 2; k <= n; ++k) {
        P[k] = ((2 * k - 1) * x * P[k - 1] - (k - 1) * P[k - 2]) / k;
        dP[k] = k * P[k - 1] + x * dP[k - 1];
    }
    if (ref_endpoint && std::fabs(std::fabs(x) - 1.0) < 1e-12 && x > 0)
        for (int k = 1; k <= n; k += 2) dP[k] = -dP[k];
}

QuadBasis::QuadBasis(int p_, bool ref) : p(p_), np((p_ + 1) * (p_ + 1)), ref_endpoint(ref) {
    if (p < 1 || p > 4) throw std::invalid_argument("order out of range");
}

void QuadBasis::values(double xi, double eta, double* phi) const {
    double Px[5], dPx[5], Py[5], dPy[5];
    legendre_pd(p, xi, ref_endpoint, Px, dPx);
    legendre_pd(p, eta, ref_endpoint, Py, dPy);
    for (int j = 0; j <= p; ++j)
        for (int i = 0; i <= p; ++i) phi[j * (p + 1) + i] = Px[i] * Py[j];
}

void QuadBasis::gradients(double xi, double eta, double* g) const {
    double Px[5], dPx[5], Py[5], dPy[5];
    legendre_pd(p, xi, ref_endpoint, Px, dPx);
    legendre_pd(p, eta, ref_endpoint, Py, dPy);
    for (int j = 0; j <= p; ++j)
        for (int i = 0; i <= p; ++i) {
            g[(j * (p + 1) + i) * 2] = dPx[i] * Py[j];
            g[(j * (p + 1) + i) * 2 + 1] = Px[i] * dPy[j];
        }
}

void face_point_quad(int f, double s, double& xi, double& eta) {
    switch (f) {
    case 0: xi = s; eta = -1.0; break;
    case 1: xi = 1.0; eta = s; break;
    case 2: xi = -s; eta = 1.0; break;
    case 3: xi = -1.0; eta = -s; break;
    default: throw std::out_of_range("Invalid face for quad");
    }
}

}  // namespace dgx
