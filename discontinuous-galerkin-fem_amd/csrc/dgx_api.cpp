// C ABI of the MI355X DG residual path: context construction (space tables, FP64 geometry,
// connectivity codes), state transfer, the fused RK3 stepping loop and FP64 diagnostics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dgx.h"
#include "dgx_internal.hpp"
#include "dgx_space.hpp"
#include "dgx_ctx.hpp"
#include "dgx_mfma.hpp"

namespace dgx_detail {
thread_local std::string g_error;
}  // namespace dgx_detail

namespace {

// The kernels' block renumbering must visit every block once.
constexpr bool xcd_tile_is_bijection(int n) {
    uint64_t seen[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // n <= 512
    for (int b = 0; b < n; ++b) {
        const int t = dgx::xcd_tile(b, n);
        if (t < 0 || t >= n || ((seen[t / 64] >> (t % 64)) & 1u)) return false;
        seen[t / 64] |= uint64_t(1) << (t % 64);
    }
    return true;
}
static_assert(xcd_tile_is_bijection(1) && xcd_tile_is_bijection(5) && xcd_tile_is_bijection(8) &&
                  xcd_tile_is_bijection(13) && xcd_tile_is_bijection(256) && xcd_tile_is_bijection(509),
              "xcd_tile is not a permutation");

// Inverse of a small dense SPD/general matrix by Gauss-Jordan with partial pivoting.
std::vector<double> invert(std::vector<double> M, int n) {
    std::vector<double> I(n * n, 0.0);
    for (int i = 0; i < n; ++i) I[i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(M[r * n + c]) > std::fabs(M[piv * n + c])) piv = r;
        if (std::fabs(M[piv * n + c]) < 1e-300) throw std::runtime_error("singular reference mass matrix");
        if (piv != c)
            for (int k = 0; k < n; ++k) { std::swap(M[c * n + k], M[piv * n + k]); std::swap(I[c * n + k], I[piv * n + k]); }
        double d = 1.0 / M[c * n + c];
        for (int k = 0; k < n; ++k) { M[c * n + k] *= d; I[c * n + k] *= d; }
        for (int r = 0; r < n; ++r) {
            if (r == c || M[r * n + c] == 0.0) continue;
            double f = M[r * n + c];
            for (int k = 0; k < n; ++k) { M[r * n + k] -= f * M[c * n + k]; I[r * n + k] -= f * I[c * n + k]; }
        }
    }
    return I;
}

}  // namespace

namespace {

// Boundary data rows (bc_row) of every boundary face point: the condition's type, its value
// at the point the Euler boundary face reads (with DGX_QUIRK_EULER_BC_POINTS_REFERENCE the
// reference's row-major read of the column-major storage [x_0..x_n-1, y_0..y_n-1],
// euler_weak_formulation.cpp:188-194, else the point) and at the point itself (the viscous
// boundary flux, navier_stokes_weak_formulation.cpp:244-247).  Face points run along the face
// from vertex f to vertex f+1 at t = (s+1)/2 (S/core/space.cpp:141-164).
void upload_bc_table(dgx_ctx& c) {
    const int nfq = c.nfq;
    const Rule1D fr = gauss_legendre(nfq);
    const bool q5 = (c.quirks & DGX_QUIRK_EULER_BC_POINTS_REFERENCE) != 0;
    std::vector<double> tab((size_t)std::max(c.n_bfaces, 1) * nfq * kBcRow, 0.0);
    std::vector<double> flat(2 * nfq);
    for (int bf = 0; bf < c.n_bfaces; ++bf) {
        const auto& B = c.h_bfaces[bf];
        const double* hv = &c.h_verts[(size_t)B.e * 2 * c.nf];
        const int a = B.f, b = (B.f + 1) % c.nf;
        for (int q = 0; q < nfq; ++q) {
            const double t = 0.5 * (fr.x[q] + 1.0);
            flat[q] = (1.0 - t) * hv[2 * a] + t * hv[2 * b];
            flat[q + nfq] = (1.0 - t) * hv[2 * a + 1] + t * hv[2 * b + 1];
        }
        const int type = c.h_bctype[B.bci];
        const auto fn = c.h_bcfn[B.bci];
        for (int q = 0; q < nfq; ++q) {
            double* row = &tab[((size_t)bf * nfq + q) * kBcRow];
            row[0] = type;
            double ve[4], vv[4];
            if (fn.first) {
                const double xe = q5 ? flat[2 * q] : flat[q], ye = q5 ? flat[2 * q + 1] : flat[q + nfq];
                fn.first(xe, ye, ve, fn.second);
                fn.first(flat[q], flat[q + nfq], vv, fn.second);
            } else {
                for (int v = 0; v < 4; ++v) ve[v] = vv[v] = c.h_bcval[(size_t)B.bci * 4 + v];
            }
            for (int v = 0; v < 4; ++v) {
                row[1 + v] = ve[v];
                row[5 + v] = vv[v];
            }
        }
    }
    c.upload(c.bc, tab.data(), tab.size() * sizeof(double));
}

// FP64 geometry, element sizes and face codes on the device from the mesh arrays (validated
// on the host); a singular mapping or a degenerate edge is reported like the reference.
void build_geometry(dgx_ctx& c, const dgx_mesh_desc& m) {
    DevBuf dV, dT, dN, dP, derr;
    if (!c.perm.empty()) c.upload(dP, c.perm.data(), c.perm.size() * sizeof(int32_t));
    dV.alloc((size_t)m.n_vertices * 2 * sizeof(double));
    dT.alloc((size_t)c.K * 3 * sizeof(int32_t));
    dN.alloc((size_t)c.K * 6 * sizeof(int32_t));
    derr.alloc(sizeof(int));
    c.geo.alloc((size_t)kGeoRows * c.Kp * sizeof(double));
    c.hmin.alloc((size_t)c.Kp * sizeof(double));
    c.code.alloc((size_t)c.Kp * sizeof(uint32_t));
    hip_check(hipMemcpyAsync(dV.p, m.vertices, (size_t)m.n_vertices * 2 * sizeof(double), hipMemcpyHostToDevice, c.stream), "copy");
    hip_check(hipMemcpyAsync(dT.p, m.elements, (size_t)c.K * 3 * sizeof(int32_t), hipMemcpyHostToDevice, c.stream), "copy");
    hip_check(hipMemcpyAsync(dN.p, m.face_neighbors, (size_t)c.K * 6 * sizeof(int32_t), hipMemcpyHostToDevice, c.stream), "copy");
    hip_check(hipMemsetAsync(derr.p, 0, sizeof(int), c.stream), "memset");
    hip_check(launch_geometry(dV.as<double>(), static_cast<const int32_t*>(dT.p), static_cast<const int32_t*>(dN.p),
                              static_cast<const int32_t*>(dP.p), c.K, c.Kp, c.geo.as<double>(), c.hmin.as<double>(), static_cast<uint32_t*>(c.code.p),
                              static_cast<int*>(derr.p), c.stream),
              "geometry");
    int err = 0;
    hip_check(hipMemcpyAsync(&err, derr.p, sizeof(int), hipMemcpyDeviceToHost, c.stream), "copy");
    hip_check(hipStreamSynchronize(c.stream), "sync");
    if (err & 1) throw std::invalid_argument("Singular Jacobian matrix");
    if (err & 2) throw std::invalid_argument("Degenerate edge detected");
}

// The boundary-condition table of the mesh descriptor (one extra "no condition" row): types,
// constant values, and no functions yet.
void init_bc_types(dgx_ctx& c, const dgx_mesh_desc& m) {
    c.h_bctype.assign(m.n_bcs + 1, -1);
    c.h_bcval.assign((size_t)(m.n_bcs + 1) * 4, 0.0);
    c.h_bcfn.assign(m.n_bcs + 1, {nullptr, nullptr});
    for (int b = 0; b < m.n_bcs; ++b) {
        const int ty = m.bc_types ? m.bc_types[b] : -1;
        if (ty == DGX_BC_PERIODIC)
            throw std::invalid_argument("PERIODIC BC encountered on a boundary face - periodic boundaries should be paired");
        if (ty < -1 || ty > DGX_BC_NO_SLIP_WALL) throw std::invalid_argument("unknown boundary condition type");
        c.h_bctype[b] = ty;
        for (int v = 0; v < 4; ++v) c.h_bcval[b * 4 + v] = m.bc_values ? m.bc_values[b * 4 + v] : 0.0;
    }
}

// The context's device, its non-blocking stream and timing events.
void open_device(dgx_ctx& c) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw Fail(DGX_E_DEVICE, "no HIP device available");
    hip_check(hipGetDevice(&c.device), "hipGetDevice");
    hip_check(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreate(&c.ev0), "hipEventCreate");
    hip_check(hipEventCreate(&c.ev1), "hipEventCreate");
}

// Inverse of the n x n matrix M (row-major) by Gauss-Jordan elimination with partial pivoting.
std::vector<double> invert_gj(std::vector<double> M, int n) {
    std::vector<double> I(n * n, 0.0);
    for (int i = 0; i < n; ++i) I[i * n + i] = 1.0;
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (std::fabs(M[r * n + col]) > std::fabs(M[piv * n + col])) piv = r;
        if (M[piv * n + col] == 0.0) throw std::runtime_error("singular mass matrix");
        if (piv != col)
            for (int k = 0; k < n; ++k) { std::swap(M[col * n + k], M[piv * n + k]); std::swap(I[col * n + k], I[piv * n + k]); }
        const double d = 1.0 / M[col * n + col];
        for (int k = 0; k < n; ++k) { M[col * n + k] *= d; I[col * n + k] *= d; }
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            const double f = M[r * n + col];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) { M[r * n + k] -= f * M[col * n + k]; I[r * n + k] -= f * I[col * n + k]; }
        }
    }
    return I;
}

// Bilinear map of the reference square (S/reference/elements.cpp:89-108, mapping.cpp:22-50):
// x = sum N_n X_n, J_T = dN^T X, dxi/dx = inv(J_T).
struct QuadMap {
    double x, y, det, A[4];
};
QuadMap quad_map(const double* V, double xi, double eta) {
    const double N[4] = {0.25 * (1.0 - xi) * (1.0 - eta), 0.25 * (1.0 + xi) * (1.0 - eta),
                         0.25 * (1.0 + xi) * (1.0 + eta), 0.25 * (1.0 - xi) * (1.0 + eta)};
    const double dN[4][2] = {{-0.25 * (1.0 - eta), -0.25 * (1.0 - xi)}, {0.25 * (1.0 - eta), -0.25 * (1.0 + xi)},
                             {0.25 * (1.0 + eta), 0.25 * (1.0 + xi)}, {-0.25 * (1.0 + eta), 0.25 * (1.0 - xi)}};
    QuadMap m{0, 0, 0, {0, 0, 0, 0}};
    double j00 = 0, j01 = 0, j10 = 0, j11 = 0;
    for (int n = 0; n < 4; ++n) {
        m.x += N[n] * V[2 * n];
        m.y += N[n] * V[2 * n + 1];
        j00 += dN[n][0] * V[2 * n];
        j01 += dN[n][0] * V[2 * n + 1];
        j10 += dN[n][1] * V[2 * n];
        j11 += dN[n][1] * V[2 * n + 1];
    }
    m.det = j00 * j11 - j01 * j10;
    if (std::fabs(m.det) < 1e-12) throw std::invalid_argument("Singular Jacobian matrix");
    const double id = 1.0 / m.det;
    m.A[0] = j11 * id;
    m.A[1] = -j01 * id;
    m.A[2] = -j10 * id;
    m.A[3] = j00 * id;
    return m;
}

// The quad space (DGSpace for element type "quad", S/core/space.cpp:47-55): tables, per-point
// geometry and per-element inverse mass matrices on the host, then uploaded.
// Largest element count whose state indices (i Kp + e) 4 + v <= 4 Np Kp - 1, with Kp = K rounded up to 256,
// stay below 2^31 (sidx, the int index of the per-lane, quad, pack and reduction kernels).
int max_elements(int order, int nvpe) {
    const long long np = nvpe == 4 ? (long long)(order + 1) * (order + 1) : (long long)(order + 1) * (order + 2) / 2;
    const long long kp = (1LL << 31) / (4 * np) / 256 * 256;  // largest multiple of 256 with 4 np kp <= 2^31
    return (int)std::min<long long>(kp, INT32_MAX);
}

void build_quad_context(dgx_ctx& c, const dgx_mesh_desc& m, const dgx_space_desc& s, const dgx_physics& ph, int precision) {
    if (s.basis != DGX_BASIS_LEGENDRE) throw std::invalid_argument("quadrilateral elements use the Legendre basis");
    if (s.rule != DGX_RULE_2P) throw Fail(DGX_E_UNSUPPORTED, "quadrilateral elements use the tensor Gauss-Legendre rule");
    if (precision == DGX_FP32_BF16VOL) throw Fail(DGX_E_UNSUPPORTED, "the BF16-input volume variant covers triangles");
    const int K = c.K, Kp = c.Kp, p = c.p;
    const QuadDims D(p);
    c.nf = 4;
    c.np = D.np;
    c.nq = D.nq;
    c.nfq = D.nfq;
    c.phi0 = 1.0;
    const int np = D.np, nq = D.nq, nfq = D.nfq;
    const Rule1D fr = gauss_legendre(nfq);
    const QuadBasis B(p, (ph.quirks & DGX_QUIRK_LEGENDRE_ENDPOINT_REFERENCE) != 0);

    // ---------------- tables (tensor points j outer, i inner) ----------------
    std::vector<double> t(D.total, 0.0), wq(nq), g(np * 2);
    std::vector<double> qxi(nq), qeta(nq);
    for (int j = 0; j < nfq; ++j)
        for (int i = 0; i < nfq; ++i) {
            const int q = j * nfq + i;
            qxi[q] = fr.x[i];
            qeta[q] = fr.x[j];
            wq[q] = fr.w[i] * fr.w[j];
            B.values(qxi[q], qeta[q], &t[D.phiV + q * np]);
            B.gradients(qxi[q], qeta[q], g.data());
            for (int k = 0; k < np; ++k) {
                t[D.GXV + q * np + k] = g[2 * k];
                t[D.GYV + q * np + k] = g[2 * k + 1];
            }
        }
    std::vector<double> phf(np);
    for (int f = 0; f < 4; ++f)
        for (int q = 0; q < nfq; ++q) {
            double xi, eta;
            face_point_quad(f, fr.x[q], xi, eta);
            B.values(xi, eta, phf.data());
            B.gradients(xi, eta, g.data());
            for (int k = 0; k < np; ++k) {
                t[D.phiF + (f * nfq + q) * np + k] = phf[k];
                t[D.LF + (f * nfq + q) * np + k] = -0.5 * fr.w[q] * phf[k];
                t[D.GF + ((f * nfq + q) * np + k) * 2] = g[2 * k];
                t[D.GF + ((f * nfq + q) * np + k) * 2 + 1] = g[2 * k + 1];
            }
        }
    // one-dimensional factors of the four-lane kernel: P_a(x) and P_a'(x) are the (a, 0) entries
    // of the tensor basis and its xi-derivative (P_0 = 1)
    {
        std::vector<double> ph(np), gr(np * 2);
        auto P1 = [&](double x, int a) { B.values(x, 0.0, ph.data()); return ph[a]; };
        auto dP1 = [&](double x, int a) { B.gradients(x, 0.0, gr.data()); return gr[2 * a]; };
        for (int a = 0; a < nfq; ++a) {
            double m = 0.0;
            for (int q = 0; q < nfq; ++q) {
                t[D.T1 + a * nfq + q] = P1(fr.x[q], a);
                t[D.D1 + a * nfq + q] = dP1(fr.x[q], a);
                m += fr.w[q] * (t[D.T1 + a * nfq + q] * t[D.T1 + a * nfq + q]);
            }
            t[D.IM1 + a] = 1.0 / m;
            t[D.X1 + a] = fr.x[a];
            t[D.W1 + a] = fr.w[a];
        }
        for (int f = 0; f < 4; ++f)
            for (int q = 0; q < nfq; ++q) {
                double xi, eta;
                face_point_quad(f, fr.x[q], xi, eta);
                for (int a = 0; a < nfq; ++a) {
                    const int k = (f * nfq + q) * nfq + a;
                    t[D.FX + k] = P1(xi, a);
                    t[D.FY + k] = P1(eta, a);
                    t[D.LX + k] = -0.5 * fr.w[q] * t[D.FX + k];
                }
            }
    }
    c.h_qw = wq;

    // ---------------- connectivity (host) ----------------
    const double* V = m.vertices;
    const int32_t* T = m.elements;
    const int32_t* N = m.face_neighbors;
    std::vector<int32_t> nbr((size_t)4 * Kp, 0);
    c.h_verts.assign((size_t)K * 8, 0.0);
    for (int e = 0; e < K; ++e)
        for (int k = 0; k < 4; ++k) {
            const int id = T[e * 4 + k];
            if (id < 0 || id >= m.n_vertices) throw std::invalid_argument("element references a missing vertex");
            c.h_verts[(size_t)e * 8 + 2 * k] = V[2 * id];
            c.h_verts[(size_t)e * 8 + 2 * k + 1] = V[2 * id + 1];
        }
    // outward unit normal and length of face f = edge (v_f, v_f+1), centroid = vertex mean
    // (DGSpace::compute_face_data, S/core/space.cpp:104-137)
    auto face_geom = [&](int e, int f, double& nx, double& ny, double& len) {
        const double* hv = &c.h_verts[(size_t)e * 8];
        const int a = f, b = (f + 1) % 4;
        const double tx = hv[2 * b] - hv[2 * a], ty = hv[2 * b + 1] - hv[2 * a + 1];
        len = std::sqrt(tx * tx + ty * ty);
        if (len < 1e-14) throw std::invalid_argument("Degenerate edge detected");
        nx = ty / len;
        ny = -tx / len;
        const double cx = (hv[0] + hv[2] + hv[4] + hv[6]) / 4.0, cy = (hv[1] + hv[3] + hv[5] + hv[7]) / 4.0;
        if (nx * (0.5 * (hv[2 * a] + hv[2 * b]) - cx) + ny * (0.5 * (hv[2 * a + 1] + hv[2 * b + 1]) - cy) < 0) {
            nx = -nx;
            ny = -ny;
        }
    };
    std::vector<double> geo((size_t)D.rows * Kp, 0.0), hmin(Kp, 1.0), minve((size_t)np * np * Kp, 0.0);
    std::vector<uint32_t> code(Kp, 0);
    for (int e = K; e < Kp; ++e)  // padding: unit weights and identity mass
        for (int q = 0; q < nq; ++q) geo[(size_t)(D.gV + 5 * q + 4) * Kp + e] = 1.0;
    int nbf = 0;
    for (int e = 0; e < K; ++e) {
        int en4[4];
        for (int f = 0; f < 4; ++f) {
            const int en = N[(e * 4 + f) * 2], fn = N[(e * 4 + f) * 2 + 1];
            en4[f] = en;
            double nx, ny, len;
            face_geom(e, f, nx, ny, len);
            uint32_t cf;
            if (en < 0) {
                int bci = fn < 0 ? m.n_bcs : fn;
                if (bci > m.n_bcs) throw std::invalid_argument("boundary face references a missing boundary condition");
                if (bci < m.n_bcs && m.bc_types && m.bc_types[bci] == DGX_BC_SLIP_WALL) c.has_slip = true;
                nbr[f * Kp + e] = nbf++;
                c.h_bfaces.push_back({e, f, bci});
                cf = (1u << 3) | (1u << 4);
            } else {
                if (en >= K || fn < 0 || fn > 3) throw std::invalid_argument("face neighbour out of range");
                if (N[(en * 4 + fn) * 2] != e || N[(en * 4 + fn) * 2 + 1] != f)
                    throw std::invalid_argument("face_neighbors is not symmetric");
                const bool isL = (e < en) || (e == en && f < fn);
                const int eL = isL ? e : en, fL = isL ? f : fn, eR = isL ? en : e, fR = isL ? fn : f;
                double nLx, nLy, lenL;
                face_geom(eL, fL, nLx, nLy, lenL);
                // compute_face_permutation (S/core/space.cpp:173-228)
                const double* VL = &c.h_verts[(size_t)eL * 8];
                const double* VR = &c.h_verts[(size_t)eR * 8];
                const int a1 = fL, a2 = (fL + 1) % 4, b1 = fR, b2 = (fR + 1) % 4;
                auto d = [](const double* P, int i, const double* Q, int j) {
                    const double dx = P[2 * i] - Q[2 * j], dy = P[2 * i + 1] - Q[2 * j + 1];
                    return std::sqrt(dx * dx + dy * dy);
                };
                const double tol = 1e-10;
                const bool rev = d(VL, a1, VR, b2) < tol && d(VL, a2, VR, b1) < tol;
                const bool same = d(VL, a1, VR, b1) < tol && d(VL, a2, VR, b2) < tol;
                bool reverse;
                if (rev || same) {
                    reverse = rev;
                } else {
                    const double t1x = VL[2 * a2] - VL[2 * a1], t1y = VL[2 * a2 + 1] - VL[2 * a1 + 1];
                    const double t2x = VR[2 * b2] - VR[2 * b1], t2y = VR[2 * b2 + 1] - VR[2 * b1 + 1];
                    const double n1 = std::sqrt(t1x * t1x + t1y * t1y), n2 = std::sqrt(t2x * t2x + t2y * t2y);
                    reverse = (t1x / n1 * t2x / n2 + t1y / n1 * t2y / n2) < 0.0;
                }
                nx = isL ? nLx : -nLx;
                ny = isL ? nLy : -nLy;
                len = lenL;
                nbr[f * Kp + e] = en;
                cf = uint32_t(fn) | (reverse ? 4u : 0u) | (isL ? 8u : 0u);
            }
            geo[(size_t)(D.gF + 3 * f) * Kp + e] = nx;
            geo[(size_t)(D.gF + 3 * f + 1) * Kp + e] = ny;
            geo[(size_t)(D.gF + 3 * f + 2) * Kp + e] = len;
            code[e] |= cf << (5 * f);
        }
        for (int f = 0; f < 4; ++f) {
            if (en4[f] < 0) continue;
            int shared = 0;
            for (int h = 0; h < 4; ++h) shared += en4[h] == en4[f];
            if (shared > (en4[f] == e ? 2 : 1))
                throw Fail(DGX_E_UNSUPPORTED, "two elements share more than one face (periodic mesh with n < 2)");
        }
        // volume points: dxi/dx and w |J|; face points: dxi/dx; mass matrix
        const double* hv = &c.h_verts[(size_t)e * 8];
        for (int k = 0; k < 8; ++k) geo[(size_t)(D.gX + k) * Kp + e] = hv[k];
        if (hv[0] + hv[4] == hv[2] + hv[6] && hv[1] + hv[5] == hv[3] + hv[7]) code[e] |= kQuadAffineBit;
        std::vector<double> M(np * np, 0.0);
        for (int q = 0; q < nq; ++q) {
            const QuadMap qm = quad_map(hv, qxi[q], qeta[q]);
            const double wd = wq[q] * std::fabs(qm.det);
            for (int k = 0; k < 4; ++k) geo[(size_t)(D.gV + 5 * q + k) * Kp + e] = qm.A[k];
            geo[(size_t)(D.gV + 5 * q + 4) * Kp + e] = wd;
            const double* ph = &t[D.phiV + q * np];
            for (int i = 0; i < np; ++i)
                for (int j = 0; j < np; ++j) M[i * np + j] += wd * (ph[i] * ph[j]);
        }
        for (int f = 0; f < 4; ++f)
            for (int q = 0; q < nfq; ++q) {
                double xi, eta;
                face_point_quad(f, fr.x[q], xi, eta);
                const QuadMap qm = quad_map(hv, xi, eta);
                for (int k = 0; k < 4; ++k) geo[(size_t)(D.gFA + (f * nfq + q) * 4 + k) * Kp + e] = qm.A[k];
            }
        const std::vector<double> Mi = invert_gj(M, np);
        for (int k = 0; k < np * np; ++k) minve[(size_t)k * Kp + e] = Mi[k];
        // element size: smallest distance between two vertices (compute_max_cfl,
        // S/solver/compressible_solver_base.cpp:125-135)
        double h = 1e300;
        for (int a = 0; a < 4; ++a)
            for (int b = a + 1; b < 4; ++b) {
                const double dx = hv[2 * a] - hv[2 * b], dy = hv[2 * a + 1] - hv[2 * b + 1];
                h = std::min(h, std::sqrt(dx * dx + dy * dy));
            }
        hmin[e] = h;
    }
    for (int e = K; e < Kp; ++e)
        for (int i = 0; i < np; ++i) minve[(size_t)(i * np + i) * Kp + e] = 1.0;
    c.n_bfaces = nbf;
    init_bc_types(c, m);

    // ---------------- device ----------------
    open_device(c);
    c.upload(c.geo, geo.data(), geo.size() * sizeof(double));
    c.upload(c.hmin, hmin.data(), hmin.size() * sizeof(double));
    c.upload(c.code, code.data(), code.size() * sizeof(uint32_t));
    c.upload(c.nbr, nbr.data(), nbr.size() * sizeof(int32_t));
    c.upload(c.minve, minve.data(), minve.size() * sizeof(double));
    upload_bc_table(c);
    c.upload(c.tab64, t.data(), t.size() * sizeof(double));
    c.upload(c.qw, wq.data(), wq.size() * sizeof(double));
    if (precision == DGX_FP64) {
        c.upload_real<double>(c.tabL, t);
        c.upload_real<double>(c.tabR, t);
    } else {
        c.upload_real<float>(c.tabL, t);
        c.upload_real<float>(c.tabR, t);
    }
    c.u.alloc(c.state_bytes());
    c.u1.alloc(c.state_bytes());
    c.u2.alloc(c.state_bytes());
    hip_check(hipMemset(c.u.p, 0, c.state_bytes()), "hipMemset");
    hip_check(hipMemset(c.u1.p, 0, c.state_bytes()), "hipMemset");
    hip_check(hipMemset(c.u2.p, 0, c.state_bytes()), "hipMemset");
    c.stage.alloc((size_t)K * std::max(np, nq) * 4 * sizeof(double));
    c.partial.alloc((size_t)c.red_blocks() * 8 * sizeof(double));
    // Euler, p <= 3: four lanes per element (DGX_STAGE_KERNEL=lane keeps one element per lane)
    const char* kv = std::getenv("DGX_STAGE_KERNEL");
    c.qtensor = c.physics == DGX_EULER && p <= 3 && !(kv && std::string(kv) == "lane");
}

// Arrays of the matrix-core tile kernels: the FP32 geometry (both modes, formed on the device
// from the FP64 geometry), and per 64-element tile the face connections and the list of
// external faces (neighbour outside the tile first, boundary faces after).  nbr [3][Kp] holds
// the neighbour element of each face or a boundary face's running index; the face codes
// (neighbour's local face, reversed order, canonical side) come from the device geometry.
void build_tile_arrays(dgx_ctx& c, const std::vector<int32_t>& nbr) {
    const int K = c.K, Kp = c.Kp, tiles = (K + 63) / 64, nfq = c.nfq;
    const int nrec = 3 * nfq * 64;  // records of a tile's own face points
    c.geo32L.alloc((size_t)kGeo32NsRows * Kp * sizeof(float));
    c.geo32R.alloc((size_t)kGeo32NsRows * Kp * sizeof(float));
    hip_check(launch_geometry32(c.geo.as<double>(), K, Kp, 0, c.geo32L.as<float>(), c.stream), "geometry32");
    hip_check(launch_geometry32(c.geo.as<double>(), K, Kp, 1, c.geo32R.as<float>(), c.stream), "geometry32");
    std::vector<uint32_t> code(Kp);
    hip_check(hipMemcpyAsync(code.data(), c.code.p, Kp * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream), "copy");
    hip_check(hipStreamSynchronize(c.stream), "sync");
    std::vector<uint32_t> conn((size_t)2 * Kp, 0u);
    std::vector<int32_t> ext, off(tiles + 1), nint(tiles);
    int cap = 0;
    for (int t = 0; t < tiles; ++t) {
        const int lo = 64 * t, hi = std::min(K, lo + 64);
        off[t] = int(ext.size() / 2);
        int slots = 0;
        for (int pass = 0; pass < 2; ++pass) {  // faces with a neighbour outside the tile, then boundary faces
            for (int e = lo; e < hi; ++e)
                for (int f = 0; f < 3; ++f) {
                    const uint32_t cf = (code[e] >> (4 * f)) & 15u;
                    const uint32_t fn = cf & 3u, flip = (cf >> 2) & 1u, isL = (cf >> 3) & 1u;
                    const int en = nbr[(size_t)f * Kp + e];
                    const bool bnd = fn == kFaceBoundary;
                    const bool inside = !bnd && en >= lo && en < hi;
                    uint32_t base, step;
                    if (inside) {
                        if (pass) continue;
                        // the neighbour's own record on its face fn, its points reversed when flagged
                        base = (fn * nfq + (flip ? nfq - 1 : 0)) * 64 + uint32_t(en - lo);
                        step = flip ? 1u : 0u;
                    } else {
                        if (bnd != (pass == 1)) continue;
                        const int slot = slots++;
                        base = bnd ? uint32_t(slot) : uint32_t(nrec + slot * nfq);
                        step = 2u;
                        ext.push_back(en);
                        ext.push_back(
                            int32_t(fn | (flip << 2) | (uint32_t(f) << 3) | (uint32_t(e - lo) << 5) | (isL << 11)));
                    }
                    if (base > 2047u) throw std::runtime_error("tile record index exceeds 11 bits");
                    const uint32_t field = base | (step << 11) | (isL << 13) | (bnd ? kConnBoundary : 0u);
                    conn[2 * (size_t)e + (f >> 1)] |= field << (16 * (f & 1));
                }
            if (pass == 0) nint[t] = slots;
        }
        cap = std::max(cap, nint[t]);
    }
    off[tiles] = int(ext.size() / 2);
    // fixed-stride copy of each tile's first 64 external faces with a neighbour, so a kernel can
    // address them without first reading the tile's list offset
    std::vector<int32_t> halo((size_t)tiles * 128, 0);
    for (int t = 0; t < tiles; ++t)
        for (int k = 0; k < 64; ++k) {
            const bool has = k < nint[t];
            halo[((size_t)t * 64 + k) * 2] = has ? ext[(size_t)(off[t] + k) * 2] : -1;
            halo[((size_t)t * 64 + k) * 2 + 1] = has ? ext[(size_t)(off[t] + k) * 2 + 1] : 0;
        }
    if (ext.empty()) ext.assign(2, 0);
    c.upload(c.halo, halo.data(), halo.size() * sizeof(int32_t));
    c.upload(c.conn, conn.data(), conn.size() * sizeof(uint32_t));
    c.upload(c.ext, ext.data(), ext.size() * sizeof(int32_t));
    c.upload(c.ext_off, off.data(), off.size() * sizeof(int32_t));
    c.upload(c.ext_nint, nint.data(), nint.size() * sizeof(int32_t));
    c.halo_cap = cap;
}

// Morton (Z-order) numbering of triangles by centroid: 16-bit coordinates on one common scale
// interleaved, ties in caller order.  Sets c.perm / c.iperm and returns true, or returns false
// (caller order kept) for a mesh whose vertex references are invalid (reported later).
uint32_t spread16(uint32_t x) {
    x &= 0xffffu;
    x = (x | (x << 8)) & 0x00ff00ffu;
    x = (x | (x << 4)) & 0x0f0f0f0fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

bool morton_order(dgx_ctx& c, const dgx_mesh_desc& m) {
    const int K = m.n_elements;
    std::vector<double> cx(K), cy(K);
    double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    for (int e = 0; e < K; ++e) {
        double sx = 0.0, sy = 0.0;
        for (int k = 0; k < 3; ++k) {
            const int id = m.elements[e * 3 + k];
            if (id < 0 || id >= m.n_vertices) return false;
            sx += m.vertices[2 * id];
            sy += m.vertices[2 * id + 1];
        }
        cx[e] = sx / 3.0;
        cy[e] = sy / 3.0;
        x0 = std::min(x0, cx[e]); x1 = std::max(x1, cx[e]);
        y0 = std::min(y0, cy[e]); y1 = std::max(y1, cy[e]);
    }
    const double ext = std::max(x1 - x0, y1 - y0);
    if (!(ext > 0.0) || !std::isfinite(ext)) return false;
    const double scale = 65535.0 / ext;
    std::vector<uint64_t> key(K);
    for (int e = 0; e < K; ++e) {
        const uint32_t qx = uint32_t(std::lround((cx[e] - x0) * scale)), qy = uint32_t(std::lround((cy[e] - y0) * scale));
        key[e] = (uint64_t(spread16(qx) | (spread16(qy) << 1)) << 32) | uint32_t(e);
    }
    std::sort(key.begin(), key.end());
    c.perm.resize(K);
    c.iperm.resize(K);
    bool identity = true;
    for (int e = 0; e < K; ++e) {
        c.perm[e] = int32_t(key[e] & 0xffffffffu);
        c.iperm[c.perm[e]] = e;
        identity &= c.perm[e] == e;
    }
    if (identity) {
        c.perm.clear();
        c.iperm.clear();
    }
    return !identity;
}

void build_tri_context(dgx_ctx& c, const dgx_mesh_desc& m, const dgx_space_desc& s, const dgx_physics& ph,
                       int precision);

void build_context(dgx_ctx& c, const dgx_mesh_desc& m, const dgx_space_desc& s, const dgx_physics& ph, int precision) {
    if (!m.vertices || !m.elements || !m.face_neighbors || m.n_elements <= 0 || m.n_vertices <= 0)
        throw std::invalid_argument("mesh descriptor is incomplete");
    if (s.order < 1 || s.order > 4) throw std::invalid_argument("Element order must be between 1 and 4");
    const int nvpe = m.vertices_per_element == 0 ? 3 : m.vertices_per_element;
    if (nvpe != 3 && nvpe != 4) throw std::invalid_argument("elements must be triangles or quadrilaterals");
    // every stage, pack and reduction kernel forms state indices in 32 bits: reject larger
    // meshes before any host or device allocation, for both element types
    if (m.n_elements > max_elements(s.order, nvpe))
        throw Fail(DGX_E_UNSUPPORTED, "mesh too large: state indices exceed 32 bits (at most " +
                                          std::to_string(max_elements(s.order, nvpe)) + " elements of order " +
                                          std::to_string(s.order) + ")");
    if (nvpe == 3 && s.basis == DGX_BASIS_LEGENDRE) throw std::invalid_argument("the Legendre basis is for quadrilaterals");
    if (nvpe == 3 && s.basis != DGX_BASIS_MONOMIAL && s.basis != DGX_BASIS_DUBINER) throw std::invalid_argument("unknown basis");
    if (s.rule != DGX_RULE_2P && s.rule != DGX_RULE_2P1) throw std::invalid_argument("unknown volume rule");
    if (precision != DGX_FP64 && precision != DGX_FP32 && precision != DGX_FP32_BF16VOL)
        throw std::invalid_argument("unknown precision");
    // the bfloat16 volume products exist in the matrix-core Euler kernel only
    if (precision == DGX_FP32_BF16VOL &&
        (ph.kind != DGX_EULER || s.rule != DGX_RULE_2P || !mfma_order(s.order)))
        throw Fail(DGX_E_UNSUPPORTED, "the BF16-input volume variant covers Euler, p = 3 or 4, volume rule 2p");
    if (ph.kind != DGX_EULER && ph.kind != DGX_NAVIER_STOKES) throw std::invalid_argument("unknown physics kind");
    if (!(ph.gamma > 1.0)) throw std::invalid_argument("gamma must exceed 1");
    if (ph.kind == DGX_NAVIER_STOKES) {  // NavierStokesWeakFormulation ctor (:18-34)
        if (!(ph.mu > 0)) throw std::invalid_argument("Dynamic viscosity must be positive for Navier-Stokes weak formulation");
        if (!(ph.prandtl > 0)) throw std::invalid_argument("Prandtl number must be positive for Navier-Stokes weak formulation");
        if (!(ph.sigma0 > 0)) throw std::invalid_argument("Penalty prefactor must be positive for Navier-Stokes weak formulation");
    }

    c.K = m.n_elements;
    c.Kp = (c.K + 255) / 256 * 256;
    c.p = s.order;
    c.rule = s.rule;
    c.basis = s.basis;
    c.precision = precision;
    c.physics = ph.kind;
    c.gamma = ph.gamma;
    c.mu = ph.mu;
    c.prandtl = ph.prandtl;
    c.sigma0 = ph.sigma0;
    c.quirks = ph.quirks;
    c.rsz = precision == DGX_FP64 ? 8 : 4;
    if (nvpe == 4) {
        build_quad_context(c, m, s, ph, precision);
        return;
    }
    // triangles run in Morton order of their centroids (DGX_ELEMENT_ORDER=given keeps the
    // caller's): the mesh is renumbered here and every per-element array crossing the ABI is
    // permuted (dgx_ctx::rows_in / rows_out)
    const char* eo = std::getenv("DGX_ELEMENT_ORDER");
    std::vector<int32_t> Tp, Np_;
    dgx_mesh_desc mp = m;
    if (!(eo && std::string(eo) == "given") && morton_order(c, m)) {
        const int K = m.n_elements;
        Tp.resize((size_t)K * 3);
        Np_.resize((size_t)K * 6);
        for (int e = 0; e < K; ++e) {
            const int o = c.perm[e];
            for (int k = 0; k < 3; ++k) {
                Tp[e * 3 + k] = m.elements[o * 3 + k];
                const int en = m.face_neighbors[(o * 3 + k) * 2];
                Np_[(e * 3 + k) * 2] = (en >= 0 && en < K) ? c.iperm[en] : en;
                Np_[(e * 3 + k) * 2 + 1] = m.face_neighbors[(o * 3 + k) * 2 + 1];
            }
        }
        mp.elements = Tp.data();
        mp.face_neighbors = Np_.data();
    }
    build_tri_context(c, mp, s, ph, precision);
}

void build_tri_context(dgx_ctx& c, const dgx_mesh_desc& m, const dgx_space_desc& s, const dgx_physics& ph,
                       int precision) {
    // ---------------- reference-element tables (FP64) ----------------
    RuleTri vr = dunavant(2 * c.p + c.rule);
    Rule1D fr = gauss_legendre(c.p + 1);
    TriBasis B(c.basis, c.p);
    const int np = B.np, nq = (int)vr.w.size(), nfq = (int)fr.w.size();
    if (nq != nq_of(c.p, c.rule)) throw std::runtime_error("volume rule size mismatch");
    c.np = np;
    c.nq = nq;
    c.nfq = nfq;
    TabLayout L(c.p, c.rule);

    std::vector<double> phiV(nq * np), gV(nq * np * 2);
    for (int q = 0; q < nq; ++q) {
        B.values(vr.xi[q], vr.eta[q], &phiV[q * np]);
        B.gradients(vr.xi[q], vr.eta[q], &gV[q * np * 2]);
    }
    c.phi0 = phiV[0];
    int ndummy = 0;
    if (std::frexp(c.phi0, &ndummy) != 0.5) throw std::runtime_error("constant basis function is not a power of two");
    std::vector<double> Mref(np * np, 0.0);
    for (int q = 0; q < nq; ++q)
        for (int i = 0; i < np; ++i)
            for (int j = 0; j < np; ++j) Mref[i * np + j] += vr.w[q] * phiV[q * np + i] * phiV[q * np + j];
    std::vector<double> Mi = invert(Mref, np);
    c.h_minv = Mi;
    c.h_qw = vr.w;

    std::vector<double> phiF(3 * nfq * np), gF(3 * nfq * np * 2);
    for (int f = 0; f < 3; ++f)
        for (int q = 0; q < nfq; ++q) {
            double xi, eta;
            face_point(f, fr.x[q], xi, eta);
            B.values(xi, eta, &phiF[(f * nfq + q) * np]);
            B.gradients(xi, eta, &gF[(f * nfq + q) * np * 2]);
        }

    auto fold = [&](const std::vector<double>& raw, int rows, std::vector<double>& outv, int off) {
        // out[r][i] = sum_j Mi[i][j] raw[r][j]
        for (int r = 0; r < rows; ++r)
            for (int i = 0; i < np; ++i) {
                double sacc = 0.0;
                for (int j = 0; j < np; ++j) sacc += Mi[i * np + j] * raw[r * np + j];
                outv[off + r * np + i] = sacc;
            }
    };
    std::vector<double> tR(L.total, 0.0), tL(L.total, 0.0);
    std::vector<double> px(nq * np), py(nq * np), lf(3 * nfq * np);
    for (int q = 0; q < nq; ++q)
        for (int i = 0; i < np; ++i) {
            px[q * np + i] = vr.w[q] * gV[(q * np + i) * 2];
            py[q * np + i] = vr.w[q] * gV[(q * np + i) * 2 + 1];
        }
    for (int f = 0; f < 3; ++f)
        for (int q = 0; q < nfq; ++q)
            for (int i = 0; i < np; ++i) lf[(f * nfq + q) * np + i] = -0.5 * fr.w[q] * phiF[(f * nfq + q) * np + i];
    for (auto* t : {&tR, &tL}) {
        std::copy(phiV.begin(), phiV.end(), t->begin() + L.phiV);
        std::copy(phiF.begin(), phiF.end(), t->begin() + L.phiF);
        for (int q = 0; q < nq; ++q)
            for (int i = 0; i < np; ++i) {
                (*t)[L.DXV + q * np + i] = gV[(q * np + i) * 2];
                (*t)[L.DYV + q * np + i] = gV[(q * np + i) * 2 + 1];
            }
        // face-gradient slots (c, d, k): reference read = column-major Map over the interleaved
        // storage, so component c of row r = q*np+i is element r + c*nfq*np of the vector whose
        // entry ((q'*np+i')*2 + d) is d-th physical derivative of phi_i' at point q'.
        const bool quirk = (c.quirks & DGX_QUIRK_NS_FACE_GRAD_REFERENCE) != 0;
        for (int f = 0; f < 3; ++f)
            for (int q = 0; q < nfq; ++q)
                for (int i = 0; i < np; ++i)
                    for (int cc = 0; cc < 2; ++cc) {
                        int qs = q, is = i, ds = cc;
                        if (quirk) {
                            int k = q * np + i + cc * nfq * np;
                            ds = k % 2;
                            qs = (k / 2) / np;
                            is = (k / 2) % np;
                        }
                        for (int kk = 0; kk < 2; ++kk) {
                            int slot = (cc * 2 + ds) * 2 + kk;
                            (*t)[L.FG + ((f * nfq + q) * 8 + slot) * np + i] = gF[((f * nfq + qs) * np + is) * 2 + kk];
                        }
                    }
        for (int i = 0; i < np; ++i) {
            double sm = 0.0;
            for (int q = 0; q < nq; ++q) sm += vr.w[q] * phiV[q * np + i];
            (*t)[L.mass + i] = sm;
        }
    }
    std::copy(px.begin(), px.end(), tR.begin() + L.PX);
    std::copy(py.begin(), py.end(), tR.begin() + L.PY);
    std::copy(lf.begin(), lf.end(), tR.begin() + L.LF);
    fold(px, nq, tL, L.PX);
    fold(py, nq, tL, L.PY);
    fold(lf, 3 * nfq, tL, L.LF);
    std::vector<double> Pm(nq * np);  // projection: Mi phi w
    for (int q = 0; q < nq; ++q)
        for (int i = 0; i < np; ++i) {
            double sacc = 0.0;
            for (int j = 0; j < np; ++j) sacc += Mi[i * np + j] * vr.w[q] * phiV[q * np + j];
            Pm[q * np + i] = sacc;
        }

    // ---------------- connectivity (host: integer work and validation) ----------------
    // The FP64 geometry and the face codes are built on the device from the vertices (below).
    const int K = c.K, Kp = c.Kp;
    const double* V = m.vertices;
    const int32_t* T = m.elements;
    const int32_t* N = m.face_neighbors;
    std::vector<int32_t> nbr((size_t)3 * Kp, 0);
    c.h_verts.assign((size_t)K * 6, 0.0);
    for (int e = 0; e < K; ++e)
        for (int k = 0; k < 3; ++k) {
            const int id = T[e * 3 + k];
            if (id < 0 || id >= m.n_vertices) throw std::invalid_argument("element references a missing vertex");
            c.h_verts[(size_t)e * 6 + 2 * k] = V[2 * id];
            c.h_verts[(size_t)e * 6 + 2 * k + 1] = V[2 * id + 1];
        }
    int nbf = 0;
    for (int e = 0; e < K; ++e) {
        int en3[3];
        for (int f = 0; f < 3; ++f) {
            const int en = N[(e * 3 + f) * 2], fn = N[(e * 3 + f) * 2 + 1];
            en3[f] = en;
            if (en < 0) {
                int bci = fn;
                if (bci < 0) bci = m.n_bcs;  // "no boundary condition" row
                if (bci > m.n_bcs) throw std::invalid_argument("boundary face references a missing boundary condition");
                if (bci < m.n_bcs && m.bc_types && m.bc_types[bci] == DGX_BC_SLIP_WALL) c.has_slip = true;
                nbr[f * Kp + e] = nbf++;  // the boundary face's running index
                c.h_bfaces.push_back({e, f, bci});
            } else {
                if (en >= K || fn < 0 || fn > 2) throw std::invalid_argument("face neighbour out of range");
                if (N[(en * 3 + fn) * 2] != e || N[(en * 3 + fn) * 2 + 1] != f)
                    throw std::invalid_argument("face_neighbors is not symmetric");
                nbr[f * Kp + e] = en;
            }
        }
        // a pair of elements shares at most one face (an element paired with itself: one face
        // pair, i.e. two of its faces)
        for (int f = 0; f < 3; ++f) {
            if (en3[f] < 0) continue;
            int shared = 0;
            for (int g = 0; g < 3; ++g) shared += en3[g] == en3[f];
            if (shared > (en3[f] == e ? 2 : 1))
                throw Fail(DGX_E_UNSUPPORTED, "two elements share more than one face (periodic mesh with n < 2)");
        }
    }
    c.n_bfaces = nbf;
    init_bc_types(c, m);

    // ---------------- device ----------------
    open_device(c);
    c.upload(c.nbr, nbr.data(), nbr.size() * sizeof(int32_t));
    build_geometry(c, m);
    upload_bc_table(c);
    c.upload(c.tab64, tL.data(), tL.size() * sizeof(double));
    c.upload(c.minv, Mi.data(), Mi.size() * sizeof(double));
    c.upload(c.pm, Pm.data(), Pm.size() * sizeof(double));
    c.upload(c.qw, vr.w.data(), vr.w.size() * sizeof(double));
    if (precision == DGX_FP64) {
        c.upload_real<double>(c.tabL, tL);
        c.upload_real<double>(c.tabR, tR);
    } else {
        c.upload_real<float>(c.tabL, tL);
        c.upload_real<float>(c.tabR, tR);
    }
    // matrix-core stage kernels: FP32 Euler (p = 1..4) and Navier-Stokes (p = 3, 4), degree-2p
    // volume rule (DGX_STAGE_KERNEL=lane keeps the one-element-per-lane kernel for FP32).  At
    // p = 1 the per-lane kernel is the faster one on large meshes (stage 10.2 against 12.0 us at
    // K = 131072, 18.1 against 22.9 us at K = 294912; the tiles win below K = 75000), so single
    // FP32 contexts of more than 98304 elements keep it.
    const char* kv = std::getenv("DGX_STAGE_KERNEL");
    const bool lane_only = precision == DGX_FP32 &&
                           ((kv && std::string(kv) == "lane") || (c.p == 1 && c.K > 98304));
    // the matrix-core kernels address every buffer through 32-bit byte offsets: the largest one
    // (state or geometry) must stay below 2 GiB
    const double big = std::max({16.0 * c.np, 8.0 * kGeoRows, 4.0 * kGeo32NsRows}) * c.Kp;
    const bool fits32 = big < 2147483648.0;
    if (precision == DGX_FP32_BF16VOL && !fits32)
        throw Fail(DGX_E_UNSUPPORTED, "mesh too large for the BF16-input volume variant (2 GiB buffers)");
    if (precision != DGX_FP64 && c.physics == DGX_EULER && c.rule == 0 && mfma_euler_order(c.p) && !lane_only &&
        fits32) {
        const bool bf16 = precision == DGX_FP32_BF16VOL;
        std::vector<float> m;
        build_mfma_tables(c.p, tL, m, bf16);
        c.upload(c.tabML, m.data(), m.size() * sizeof(float));
        build_mfma_tables(c.p, tR, m, bf16);
        c.upload(c.tabMR, m.data(), m.size() * sizeof(float));
        if (mfma_order(c.p)) {  // the matrix-core stage adjoint (p = 3, 4)
            build_mfma_vjp_tables(c.p, tL, m);
            c.upload(c.tabVL, m.data(), m.size() * sizeof(float));
            build_mfma_vjp_tables(c.p, tR, m);
            c.upload(c.tabVR, m.data(), m.size() * sizeof(float));
        }
        build_tile_arrays(c, nbr);
    }
    if (precision == DGX_FP32 && c.physics == DGX_NAVIER_STOKES && c.rule == 0 && mfma_order(c.p) && !lane_only &&
        fits32) {
        std::vector<float> m;
        const bool quirk = (c.quirks & DGX_QUIRK_NS_FACE_GRAD_REFERENCE) != 0;
        build_mfma_ns_tables(c.p, tL, false, quirk, m);
        c.upload(c.tabML, m.data(), m.size() * sizeof(float));
        build_mfma_ns_tables(c.p, tR, false, quirk, m);
        c.upload(c.tabMR, m.data(), m.size() * sizeof(float));
        build_mfma_ns_tables(c.p, tL, true, quirk, m);
        c.upload(c.tabNT, m.data(), m.size() * sizeof(float));
        build_tile_arrays(c, nbr);
    }
    c.u.alloc(c.state_bytes());
    c.u1.alloc(c.state_bytes());
    c.u2.alloc(c.state_bytes());
    hip_check(hipMemset(c.u.p, 0, c.state_bytes()), "hipMemset");
    hip_check(hipMemset(c.u1.p, 0, c.state_bytes()), "hipMemset");
    hip_check(hipMemset(c.u2.p, 0, c.state_bytes()), "hipMemset");
    c.stage.alloc((size_t)K * std::max(np, nq) * 4 * sizeof(double));
    c.partial.alloc((size_t)c.red_blocks() * 4 * sizeof(double));
}

}  // namespace

extern "C" {

const char* dgx_last_error(void) { return g_error.c_str(); }

dgx_status dgx_mesh_split_quads(int32_t n, double length, double x0, double y0, double* vertices, int32_t* elements) {
    return guarded([&] {
        if (n < 1 || !(length > 0) || !vertices || !elements) throw std::invalid_argument("bad mesh size");
        const double h = length / n;
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i) {
                vertices[2 * (j * (n + 1) + i)] = x0 + i * h;
                vertices[2 * (j * (n + 1) + i) + 1] = y0 + j * h;
            }
        auto vid = [n](int i, int j) { return j * (n + 1) + i; };
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                int32_t* t = elements + 6 * (j * n + i);
                t[0] = vid(i, j); t[1] = vid(i + 1, j); t[2] = vid(i + 1, j + 1);
                t[3] = vid(i, j); t[4] = vid(i + 1, j + 1); t[5] = vid(i, j + 1);
            }
    });
}

dgx_status dgx_mesh_connect(int32_t nv, const double* V, int32_t K, const int32_t* T, int32_t periodic_mask,
                            int32_t* N) {
    return dgx_mesh_connect_elements(nv, V, K, T, 3, periodic_mask, N);
}

dgx_status dgx_mesh_connect_elements(int32_t nv, const double* V, int32_t K, const int32_t* T, int32_t nf,
                                     int32_t periodic_mask, int32_t* N) {
    return guarded([&] {
        if (nv <= 0 || K <= 0 || !V || !T || !N) throw std::invalid_argument("bad mesh");
        if (nf != 3 && nf != 4) throw std::invalid_argument("elements must have 3 or 4 vertices");
        for (size_t k = 0; k < (size_t)K * nf * 2; ++k) N[k] = -1;
        // sorted-vertex-pair signature -> face (DGMesh::build_face_connectivity, S/core/mesh.cpp:312-359)
        std::vector<std::pair<uint64_t, int>> sig;
        sig.reserve((size_t)K * nf);
        for (int e = 0; e < K; ++e)
            for (int f = 0; f < nf; ++f) {
                uint32_t a = (uint32_t)T[e * nf + f], b = (uint32_t)T[e * nf + (f + 1) % nf];
                if ((int)a >= nv || (int)b >= nv) throw std::invalid_argument("element references a missing vertex");
                uint64_t s = ((uint64_t)std::min(a, b) << 32) | std::max(a, b);
                sig.push_back({s, e * nf + f});
            }
        std::sort(sig.begin(), sig.end());
        for (size_t k = 0; k < sig.size();) {
            size_t j = k;
            while (j < sig.size() && sig[j].first == sig[k].first) ++j;
            if (j - k == 2) {
                int A = sig[k].second, B = sig[k + 1].second;
                N[A * 2] = B / nf; N[A * 2 + 1] = B % nf;
                N[B * 2] = A / nf; N[B * 2 + 1] = A % nf;
            } else if (j - k > 2) {
                throw std::invalid_argument("non-manifold mesh: an edge is shared by more than two elements");
            }
            k = j;
        }
        // boundary tags by position: 0 Bottom, 1 Right, 2 Top, 3 Left (GMSH physical names,
        // S/utils/mesh_creation.cpp:256-260); bc index = tag
        double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
        for (int i = 0; i < nv; ++i) {
            xmin = std::min(xmin, V[2 * i]); xmax = std::max(xmax, V[2 * i]);
            ymin = std::min(ymin, V[2 * i + 1]); ymax = std::max(ymax, V[2 * i + 1]);
        }
        const double tol = 1e-10 * std::max(xmax - xmin, ymax - ymin);
        std::vector<std::vector<std::pair<double, int>>> side(4);
        for (int k = 0; k < K * nf; ++k) {
            if (N[k * 2] >= 0) continue;
            int e = k / nf, f = k % nf;
            int a = T[e * nf + f], b = T[e * nf + (f + 1) % nf];
            double xa = V[2 * a], ya = V[2 * a + 1], xb = V[2 * b], yb = V[2 * b + 1];
            int tag = -1;
            if (std::fabs(ya - ymin) < tol && std::fabs(yb - ymin) < tol) tag = 0;
            else if (std::fabs(xa - xmax) < tol && std::fabs(xb - xmax) < tol) tag = 1;
            else if (std::fabs(ya - ymax) < tol && std::fabs(yb - ymax) < tol) tag = 2;
            else if (std::fabs(xa - xmin) < tol && std::fabs(xb - xmin) < tol) tag = 3;
            N[k * 2 + 1] = tag;
            if (tag >= 0) side[tag].push_back({(tag == 1 || tag == 3) ? 0.5 * (ya + yb) : 0.5 * (xa + xb), k});
        }
        // periodic pairing by transverse midpoint (DGMesh::set_periodic_boundaries, :482-588)
        auto pair_sides = [&](int s1, int s2) {
            auto& A = side[s1];
            auto& B = side[s2];
            if (A.empty() || A.size() != B.size())
                throw std::invalid_argument("Cannot set periodic boundaries: unmatched boundary faces");
            std::sort(A.begin(), A.end());
            std::sort(B.begin(), B.end());
            for (size_t i = 0; i < A.size(); ++i) {
                int a = A[i].second, b = B[i].second;
                N[a * 2] = b / nf; N[a * 2 + 1] = b % nf;
                N[b * 2] = a / nf; N[b * 2 + 1] = a % nf;
            }
        };
        if (periodic_mask & 1) pair_sides(3, 1);
        if (periodic_mask & 2) pair_sides(0, 2);
    });
}

dgx_status dgx_mesh_rect_quads(int32_t nx, int32_t ny, double xmin, double xmax, double ymin, double ymax,
                               double* vertices, int32_t* elements) {
    return guarded([&] {
        if (nx < 1 || ny < 1 || !(xmax > xmin) || !(ymax > ymin) || !vertices || !elements)
            throw std::invalid_argument("bad mesh size");
        const double hx = (xmax - xmin) / nx, hy = (ymax - ymin) / ny;
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                vertices[2 * (j * (nx + 1) + i)] = xmin + i * hx;
                vertices[2 * (j * (nx + 1) + i) + 1] = ymin + j * hy;
            }
        auto vid = [nx](int i, int j) { return j * (nx + 1) + i; };
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int32_t* q = elements + 4 * (j * nx + i);
                q[0] = vid(i, j); q[1] = vid(i + 1, j); q[2] = vid(i + 1, j + 1); q[3] = vid(i, j + 1);
            }
    });
}

dgx_status dgx_create(const dgx_mesh_desc* mesh, const dgx_space_desc* space, const dgx_physics* physics,
                      int32_t precision, dgx_ctx** out) {
    if (!out) { g_error = "out is null"; return DGX_E_INVALID; }
    *out = nullptr;
    auto* c = new (std::nothrow) dgx_ctx();
    if (!c) { g_error = "out of host memory"; return DGX_E_NOMEM; }
    dgx_status st = guarded([&] {
        if (!mesh || !space || !physics) throw std::invalid_argument("null descriptor");
        build_context(*c, *mesh, *space, *physics, precision);
    });
    if (st != DGX_OK) { delete c; return st; }
    *out = c;
    return DGX_OK;
}

void dgx_destroy(dgx_ctx* ctx) { delete ctx; }

dgx_status dgx_max_elements(int32_t order, int32_t vertices_per_element, int32_t* out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("null argument");
        if (order < 1 || order > 4) throw std::invalid_argument("Element order must be between 1 and 4");
        const int nvpe = vertices_per_element == 0 ? 3 : vertices_per_element;
        if (nvpe != 3 && nvpe != 4) throw std::invalid_argument("elements must be triangles or quadrilaterals");
        *out = max_elements(order, nvpe);
    });
}

dgx_status dgx_info(const dgx_ctx* c, dgx_info_t* o) {
    if (!c || !o) { g_error = "null argument"; return DGX_E_INVALID; }
    o->n_elements = c->K;
    o->n_modes = c->np;
    o->n_volume_points = c->nq;
    o->n_face_points = c->nfq;
    o->order = c->p;
    o->basis = c->basis;
    o->precision = c->precision;
    o->physics = c->physics;
    o->n_boundary_faces = c->n_bfaces;
    o->device = c->device;
    // per stage: read u_k, (stages 2,3) u^n, write u_{k+1}: 8 state touches per 3 stages,
    // plus 128 B of FP64 geometry + connectivity per element
    o->bytes_per_stage_alg = (double)c->K * (8.0 / 3.0 * 4.0 * c->np * (double)c->rsz + 128.0);
    o->stage_kernel = c->tabML.p ? 1 : c->qtensor ? 2 : 0;
    return DGX_OK;
}

dgx_status dgx_set_state(dgx_ctx* c, const double* u) {
    return guarded([&] {
        if (!c || !u) throw std::invalid_argument("null argument");
        size_t bytes = (size_t)c->K * c->np * 4 * sizeof(double);
        // storage offset: mean of the element averages' constant-mode values
        double U[4] = {0, 0, 0, 0};
        for (int e = 0; e < c->K; ++e)
            for (int v = 0; v < 4; ++v) U[v] += u[(size_t)e * c->np * 4 + v];
        for (int v = 0; v < 4; ++v) U[v] *= c->phi0 / c->K;
        c->set_offset(U);
        std::vector<double> tmp;
        hip_check(hipMemcpyAsync(c->stage.p, c->rows_in(u, (size_t)c->np * 4, tmp), bytes, hipMemcpyHostToDevice, c->stream),
                  "copy");
        hip_check(launch_pack(c->precision, c->np, c->stage.as<double>(), c->u.p, c->K, c->Kp, c->off, c->stream), "pack");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

dgx_status dgx_get_state(dgx_ctx* c, double* u) {
    return guarded([&] {
        if (!c || !u) throw std::invalid_argument("null argument");
        size_t bytes = (size_t)c->K * c->np * 4 * sizeof(double);
        hip_check(launch_unpack(c->precision, c->np, c->u.p, c->stage.as<double>(), c->K, c->Kp, c->off, c->stream), "unpack");
        hip_check(hipMemcpyAsync(u, c->stage.p, bytes, hipMemcpyDeviceToHost, c->stream), "copy");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        c->rows_out(u, (size_t)c->np * 4);
    });
}

dgx_status dgx_volume_points(dgx_ctx* c, double* xy) {
    return guarded([&] {
        if (!c || !xy) throw std::invalid_argument("null argument");
        // GeometricMapping::map_to_physical (S/reference/mapping.cpp:52-58)
        if (c->nf == 4) {
            const Rule1D fr = gauss_legendre(c->nfq);
            for (int e = 0; e < c->K; ++e)
                for (int j = 0; j < c->nfq; ++j)
                    for (int i = 0; i < c->nfq; ++i) {
                        const QuadMap qm = quad_map(&c->h_verts[(size_t)e * 8], fr.x[i], fr.x[j]);
                        const size_t o = ((size_t)e * c->nq + j * c->nfq + i) * 2;
                        xy[o] = qm.x;
                        xy[o + 1] = qm.y;
                    }
            return;
        }
        RuleTri vr = dunavant(2 * c->p + c->rule);
        for (int e = 0; e < c->K; ++e) {
            const double* V = &c->h_verts[(size_t)e * 6];
            for (int q = 0; q < c->nq; ++q) {
                double n0 = 1.0 - vr.xi[q] - vr.eta[q], n1 = vr.xi[q], n2 = vr.eta[q];
                xy[((size_t)e * c->nq + q) * 2] = n0 * V[0] + n1 * V[2] + n2 * V[4];
                xy[((size_t)e * c->nq + q) * 2 + 1] = n0 * V[1] + n1 * V[3] + n2 * V[5];
            }
        }
        c->rows_out(xy, (size_t)c->nq * 2);
    });
}

dgx_status dgx_get_geometry(dgx_ctx* c, double* geo, double* hmin, uint32_t* codes) {
    return guarded([&] {
        if (!c) throw std::invalid_argument("null argument");
        if (c->nf == 4) throw Fail(DGX_E_UNSUPPORTED, "the per-element geometry layout is the triangles'");
        const size_t K = c->K, Kp = c->Kp;
        if (geo) {
            std::vector<double> g((size_t)kGeoRows * Kp);
            hip_check(hipMemcpyAsync(g.data(), c->geo.p, g.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy");
            hip_check(hipStreamSynchronize(c->stream), "sync");
            for (size_t e = 0; e < K; ++e)
                for (int r = 0; r < kGeoRows; ++r) geo[e * kGeoRows + r] = g[r * Kp + e];
            c->rows_out(geo, (size_t)kGeoRows);
        }
        if (hmin)
            hip_check(hipMemcpyAsync(hmin, c->hmin.p, K * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy");
        if (codes)
            hip_check(hipMemcpyAsync(codes, c->code.p, K * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream), "copy");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        c->rows_out(hmin, 1);
        c->rows_out(codes, 1);
    });
}

dgx_status dgx_set_bc_function(dgx_ctx* c, int32_t bc_index, dgx_bc_fn fn, void* user) {
    return guarded([&] {
        if (!c || !fn) throw std::invalid_argument("null argument");
        if (bc_index < 0 || bc_index >= (int)c->h_bctype.size() - 1)
            throw std::invalid_argument("boundary condition index out of range");
        c->h_bcfn[bc_index] = {fn, user};
        upload_bc_table(*c);
    });
}

// Relative L2 error per primitive field at the volume points against caller-supplied exact
// values (compute_variable_error, T/solver/shear_flow_accuracy_test.cpp:32-79): sqrt(sum w |J|
// (W - W_exact)^2 / sum w |J| W_exact^2), or the absolute sqrt(sum w |J| (W - W_exact)^2) when
// the exact field's norm is below 1e-14 (as there).  FP64 block partials summed in order.
dgx_status dgx_l2_error(dgx_ctx* c, const double* exact, int32_t exact_is_primitive, double err[4]) {
    return guarded([&] {
        if (!c || !exact || !err) throw std::invalid_argument("null argument");
        const size_t n = (size_t)c->K * c->nq * 4;
        DevBuf ex, part;
        ex.alloc(n * sizeof(double));
        const int nb = c->red_blocks();
        part.alloc((size_t)nb * 8 * sizeof(double));
        std::vector<double> tmp;
        hip_check(hipMemcpyAsync(ex.p, c->rows_in(exact, (size_t)c->nq * 4, tmp), n * sizeof(double), hipMemcpyHostToDevice,
                                 c->stream),
                  "copy");
        hip_check(launch_eval_n(c->precision, c->np, c->nq, c->u.p, c->tab64.as<double>(), c->K, c->Kp,
                                c->stage.as<double>(), c->uref, c->stream),
                  "eval");
        hip_check(launch_l2_error(c->stage.as<double>(), ex.as<double>(), exact_is_primitive ? 1 : 0, c->geo.as<double>(),
                                  c->qw.as<double>(), c->nq, c->K, c->Kp, c->gamma, part.as<double>(), nb,
                                  c->nf == 4 ? 1 : 0, c->stream),
                  "l2 error");
        std::vector<double> h((size_t)nb * 8);
        hip_check(hipMemcpyAsync(h.data(), part.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        double e2[4] = {0, 0, 0, 0}, n2[4] = {0, 0, 0, 0};
        for (int b = 0; b < nb; ++b)
            for (int v = 0; v < 4; ++v) {
                e2[v] += h[(size_t)b * 8 + v];
                n2[v] += h[(size_t)b * 8 + 4 + v];
            }
        for (int v = 0; v < 4; ++v) err[v] = n2[v] < 1e-14 ? std::sqrt(e2[v]) : std::sqrt(e2[v] / n2[v]);
    });
}

// Legacy-VTK output of an Euler state (VTKWriter::write_euler_solution and its helpers,
// S/utils/vtk_writer.cpp:72-176,232-311,323-489): every element subdivided at
// `resolution` (reference points (i/r, j/r), j outer, sub-triangles lower-left then
// upper-right), the DG solution evaluated there, primitive and derived fields plus the
// conserved ones as point data, ASCII, 10 significant digits in scientific notation.
dgx_status dgx_write_vtk(dgx_ctx* c, const double* u_or_null, const char* filename, double gamma, int32_t resolution) {
    return guarded([&] {
        if (!c || !filename) throw std::invalid_argument("null argument");
        if (resolution < 1) throw std::invalid_argument("resolution factor must be at least 1");
        const bool quad = c->nf == 4;
        const int K = c->K, np = c->np, r = resolution, n1 = r + 1;
        std::vector<double> u;
        if (!u_or_null) {
            u.resize((size_t)K * np * 4);
            if (dgx_get_state(c, u.data()) != DGX_OK) throw std::runtime_error(g_error);
            u_or_null = u.data();
        }
        // reference subdivision points and sub-cells (get_reference_subdivision_points /
        // _connectivity, S/utils/vtk_writer.cpp:326-424): triangles of the unit triangle, or quads
        // of [-1,1]^2, j outer
        std::vector<double> rp;
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < (quad ? n1 : n1 - j); ++i) {
                rp.push_back(quad ? -1.0 + 2.0 * i / r : double(i) / r);
                rp.push_back(quad ? -1.0 + 2.0 * j / r : double(j) / r);
            }
        const int npe = (int)rp.size() / 2;
        auto tri_index = [&](int i, int j) {
            int idx = 0;
            for (int jj = 0; jj < j; ++jj) idx += n1 - jj;
            return idx + i;
        };
        std::vector<int> sub;
        const int nvc = quad ? 4 : 3;  // vertices per sub-cell
        for (int j = 0; j < r; ++j)
            for (int i = 0; i < (quad ? r : r - j); ++i) {
                if (quad) {
                    sub.insert(sub.end(), {j * n1 + i, j * n1 + i + 1, (j + 1) * n1 + i + 1, (j + 1) * n1 + i});
                    continue;
                }
                sub.insert(sub.end(), {tri_index(i, j), tri_index(i + 1, j), tri_index(i, j + 1)});
                if (i < r - j - 1) sub.insert(sub.end(), {tri_index(i + 1, j), tri_index(i + 1, j + 1), tri_index(i, j + 1)});
            }
        const int nse = (int)sub.size() / nvc;
        std::vector<double> phi((size_t)npe * np);
        if (quad) {
            const QuadBasis B(c->p, true);
            for (int q = 0; q < npe; ++q) B.values(rp[2 * q], rp[2 * q + 1], &phi[(size_t)q * np]);
        } else {
            const TriBasis B(c->basis, c->p);
            for (int q = 0; q < npe; ++q) B.values(rp[2 * q], rp[2 * q + 1], &phi[(size_t)q * np]);
        }
        const size_t npts = (size_t)K * npe;
        std::vector<double> xy(npts * 2), U(npts * 4, 0.0);
        for (int e = 0; e < K; ++e) {
            const double* V = &c->h_verts[(size_t)c->dev_of(e) * 2 * c->nf];  // u is in the caller's order
            for (int q = 0; q < npe; ++q) {
                const double xi = rp[2 * q], eta = rp[2 * q + 1], n0 = 1.0 - xi - eta;
                const size_t k = (size_t)e * npe + q;
                if (quad) {
                    const QuadMap qm = quad_map(V, xi, eta);
                    xy[2 * k] = qm.x;
                    xy[2 * k + 1] = qm.y;
                } else {
                    xy[2 * k] = n0 * V[0] + xi * V[2] + eta * V[4];
                    xy[2 * k + 1] = n0 * V[1] + xi * V[3] + eta * V[5];
                }
                for (int i = 0; i < np; ++i)
                    for (int v = 0; v < 4; ++v) U[4 * k + v] += u_or_null[((size_t)e * np + i) * 4 + v] * phi[(size_t)q * np + i];
            }
        }
        std::ofstream f(std::string(filename) + ".vtk");
        if (!f.is_open()) throw std::runtime_error(std::string("Failed to open file: ") + filename + ".vtk");
        f << "# vtk DataFile Version 3.0\nDGFEM Euler Solution\nASCII\nDATASET UNSTRUCTURED_GRID\n";
        f << "POINTS " << npts << " double\n" << std::scientific << std::setprecision(10);
        for (size_t k = 0; k < npts; ++k) f << xy[2 * k] << " " << xy[2 * k + 1] << " 0.0\n";
        const size_t ncells = (size_t)K * nse;
        f << "CELLS " << ncells << " " << ncells * (1 + nvc) << "\n";
        for (int e = 0; e < K; ++e)
            for (int t = 0; t < nse; ++t) {
                const size_t o = (size_t)e * npe;
                f << nvc;
                for (int k = 0; k < nvc; ++k) f << " " << o + sub[nvc * t + k];
                f << "\n";
            }
        f << "CELL_TYPES " << ncells << "\n";
        for (size_t t = 0; t < ncells; ++t) f << (quad ? 9 : 5) << "\n";  // VTK_QUAD, VTK_TRIANGLE
        f << "POINT_DATA " << npts << "\n" << std::scientific << std::setprecision(10);
        auto field = [&](const char* name, auto fn) {
            f << "SCALARS " << name << " double 1\nLOOKUP_TABLE default\n";
            for (size_t k = 0; k < npts; ++k) {
                const double rho = U[4 * k], uu = U[4 * k + 1] / rho, vv = U[4 * k + 2] / rho;
                const double p = (gamma - 1.0) * (U[4 * k + 3] - 0.5 * rho * (uu * uu + vv * vv));
                f << fn(k, rho, uu, vv, p) << "\n";
            }
        };
        field("density", [](size_t, double rho, double, double, double) { return rho; });
        field("u_velocity", [](size_t, double, double uu, double, double) { return uu; });
        field("v_velocity", [](size_t, double, double, double vv, double) { return vv; });
        field("velocity_magnitude", [](size_t, double, double uu, double vv, double) { return std::sqrt(uu * uu + vv * vv); });
        field("pressure", [](size_t, double, double, double, double p) { return p; });
        field("temperature", [](size_t, double rho, double, double, double p) { return p / rho; });
        field("mach", [&](size_t, double rho, double uu, double vv, double p) {
            return std::sqrt(uu * uu + vv * vv) / std::sqrt(gamma * p / rho);
        });
        field("rho", [&](size_t k, double, double, double, double) { return U[4 * k]; });
        field("rho_u", [&](size_t k, double, double, double, double) { return U[4 * k + 1]; });
        field("rho_v", [&](size_t k, double, double, double, double) { return U[4 * k + 2]; });
        field("E", [&](size_t k, double, double, double, double) { return U[4 * k + 3]; });
        if (!f) throw std::runtime_error(std::string("Failed to write file: ") + filename + ".vtk");
    });
}

dgx_status dgx_project(dgx_ctx* c, const double* Uq) {
    return guarded([&] {
        if (!c || !Uq) throw std::invalid_argument("null argument");
        size_t bytes = (size_t)c->K * c->nq * 4 * sizeof(double);
        double U[4] = {0, 0, 0, 0};
        const size_t npts = (size_t)c->K * c->nq;
        for (size_t k = 0; k < npts; ++k)
            for (int v = 0; v < 4; ++v) U[v] += Uq[k * 4 + v];
        for (int v = 0; v < 4; ++v) U[v] /= double(npts);
        c->set_offset(U);
        std::vector<double> tmp;
        hip_check(hipMemcpyAsync(c->stage.p, c->rows_in(Uq, (size_t)c->nq * 4, tmp), bytes, hipMemcpyHostToDevice, c->stream),
                  "copy");
        if (c->nf == 4)
            hip_check(launch_project_quad(c->precision, c->np, c->nq, c->stage.as<double>(), c->u.p,
                                          c->tab64.as<double>(), c->geo.as<double>(), c->minve.as<double>(), c->K,
                                          c->Kp, c->off, c->stream),
                      "project");
        else
            hip_check(launch_project(c->precision, c->p, c->rule, c->stage.as<double>(), c->u.p, nullptr, c->K, c->Kp,
                                     c->pm.as<double>(), c->off, c->stream),
                      "project");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

dgx_status dgx_eval_volume(dgx_ctx* c, double* Uq) {
    return guarded([&] {
        if (!c || !Uq) throw std::invalid_argument("null argument");
        size_t bytes = (size_t)c->K * c->nq * 4 * sizeof(double);
        hip_check(launch_eval_n(c->precision, c->np, c->nq, c->u.p, c->tab64.as<double>(), c->K, c->Kp,
                                c->stage.as<double>(), c->uref, c->stream),
                  "eval");
        hip_check(hipMemcpyAsync(Uq, c->stage.p, bytes, hipMemcpyDeviceToHost, c->stream), "copy");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        c->rows_out(Uq, (size_t)c->nq * 4);
    });
}

dgx_status dgx_residual(dgx_ctx* c, const double* u, double* r) {
    return guarded([&] {
        if (!c || !r) throw std::invalid_argument("null argument");
        size_t bytes = (size_t)c->K * c->np * 4 * sizeof(double);
        const void* src = c->u.p;
        std::vector<double> tmp;
        if (u) {
            hip_check(hipMemcpyAsync(c->stage.p, c->rows_in(u, (size_t)c->np * 4, tmp), bytes, hipMemcpyHostToDevice, c->stream),
                      "copy");
            hip_check(launch_pack(c->precision, c->np, c->stage.as<double>(), c->u2.p, c->K, c->Kp, c->off, c->stream), "pack");
            src = c->u2.p;
        }
        c->stage_launch(src, src, c->u1.p, MODE_RESIDUAL, 0.0, 0.0, 0.0);
        hip_check(launch_unpack(c->precision, c->np, c->u1.p, c->stage.as<double>(), c->K, c->Kp, dgx_ctx::kNoOff, c->stream), "unpack");
        hip_check(hipMemcpyAsync(r, c->stage.p, bytes, hipMemcpyDeviceToHost, c->stream), "copy");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        c->rows_out(r, (size_t)c->np * 4);
    });
}

dgx_status dgx_apply_minv(dgx_ctx* c, double* r) {
    return guarded([&] {
        if (!c || !r) throw std::invalid_argument("null argument");
        size_t bytes = (size_t)c->K * c->np * 4 * sizeof(double);
        std::vector<double> tmp;
        hip_check(hipMemcpyAsync(c->stage.p, c->rows_in(r, (size_t)c->np * 4, tmp), bytes, hipMemcpyHostToDevice, c->stream),
                  "copy");
        if (c->nf == 4)
            hip_check(launch_minv_quad(c->np, c->minve.as<double>(), c->stage.as<double>(), c->K, c->Kp, c->stream), "minv");
        else
            hip_check(launch_apply_minv(c->precision, c->np, c->minv.as<double>(), c->geo.as<double>(),
                                        c->stage.as<double>(), c->K, c->Kp, c->stream),
                      "minv");
        hip_check(hipMemcpyAsync(r, c->stage.p, bytes, hipMemcpyDeviceToHost, c->stream), "copy");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        c->rows_out(r, (size_t)c->np * 4);
    });
}

dgx_status dgx_rk3_steps_async(dgx_ctx* c, double dt, int32_t nsteps) {
    return guarded([&] {
        if (!c) throw std::invalid_argument("null context");
        if (nsteps < 0 || !std::isfinite(dt)) throw std::invalid_argument("bad step count or dt");
        for (int s = 0; s < nsteps; ++s) c->rk3_step(dt);
    });
}

dgx_status dgx_synchronize(dgx_ctx* c) {
    return guarded([&] {
        if (!c) throw std::invalid_argument("null context");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

// The reference checks the state for NaN after every step and stops at the first step that
// has one (run_time_integration, S/solver/compressible_solver_base.cpp:279-301).  Here the steps
// run back to back and the state is checked once at the end; since a NaN coefficient stays NaN
// through every later stage (each element's update reads its own coefficients), a clean final
// state means no step had one.  Otherwise the run is replayed from a copy of its starting
// state one step at a time with the check after each; the stage kernels are deterministic, so
// the replay reproduces the same states bit for bit and stops at the reference's step.
dgx_status dgx_rk3_run(dgx_ctx* c, double dt, int32_t nsteps, int32_t* steps_done) {
    if (steps_done) *steps_done = 0;
    return guarded([&] {
        if (!c) throw std::invalid_argument("null context");
        if (nsteps < 0 || !std::isfinite(dt)) throw std::invalid_argument("bad step count or dt");
        if (nsteps == 0) return;
        const size_t sb = c->state_bytes();
        if (!c->ub.p) c->ub.alloc(sb);
        hip_check(hipMemcpyAsync(c->ub.p, c->u.p, sb, hipMemcpyDeviceToDevice, c->stream), "copy");
        for (int s = 0; s < nsteps; ++s) c->rk3_step(dt);
        if (c->nan_count() == 0.0) {
            if (steps_done) *steps_done = nsteps;
            return;
        }
        hip_check(hipMemcpyAsync(c->u.p, c->ub.p, sb, hipMemcpyDeviceToDevice, c->stream), "copy");
        for (int s = 1; s <= nsteps; ++s) {
            c->rk3_step(dt);
            if (c->nan_count() > 0.0) {
                if (steps_done) *steps_done = s;
                throw Fail(DGX_E_NONFINITE, "Solution contains NaN values at step " + std::to_string(s) +
                                                " of " + std::to_string(nsteps));
            }
        }
        throw Fail(DGX_E_DEVICE, "a NaN in the final state did not recur when the steps were replayed");
    });
}

dgx_status dgx_rk3_steps(dgx_ctx* c, double dt, int32_t nsteps) { return dgx_rk3_run(c, dt, nsteps, nullptr); }

dgx_status dgx_rk3_profile(dgx_ctx* c, double dt, int32_t nsteps, double* stage_ms, double* total_ms) {
    return guarded([&] {
        if (!c || nsteps <= 0) throw std::invalid_argument("bad arguments");
        // five warm-up steps, then events around nsteps back-to-back steps on the context stream:
        // the mean stage time is the total over the 3 nsteps launches (a per-launch event pair
        // would add its own record overhead to every launch)
        auto step = [&] {
            const double ca[3] = {0.0, 0.75, 1.0 / 3.0}, cb[3] = {1.0, 0.25, 2.0 / 3.0};
            void* in[3] = {c->u.p, c->u1.p, c->u2.p};
            void* outb[3] = {c->u1.p, c->u2.p, c->u1.p};
            for (int k = 0; k < 3; ++k) c->stage_launch(in[k], c->u.p, outb[k], MODE_RK, dt, ca[k], cb[k]);
            std::swap(c->u, c->u1);
        };
        for (int s = 0; s < 5; ++s) step();
        hip_check(hipEventRecord(c->ev0, c->stream), "record");
        for (int s = 0; s < nsteps; ++s) step();
        hip_check(hipEventRecord(c->ev1, c->stream), "record");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        float tot = 0.f;
        hip_check(hipEventElapsedTime(&tot, c->ev0, c->ev1), "elapsed");
        if (stage_ms) *stage_ms = tot / (3.0 * nsteps);
        if (total_ms) *total_ms = tot;
    });
}

dgx_status dgx_conservation_sums(dgx_ctx* c, double out[4]) {
    return guarded([&] {
        if (!c || !out) throw std::invalid_argument("null argument");
        std::vector<double> h = c->reduce(0);
        for (int v = 0; v < 4; ++v) {
            double s = 0.0;
            for (int b = 0; b < c->red_blocks(); ++b) s += h[b * 4 + v];
            out[v] = s;
        }
    });
}

dgx_status dgx_kinetic_energy(dgx_ctx* c, double* J) {
    return guarded([&] {
        if (!c || !J) throw std::invalid_argument("null argument");
        std::vector<double> h = c->reduce(1);
        double s = 0.0;
        for (int b = 0; b < c->red_blocks(); ++b) s += h[b * 4];
        *J = s;
    });
}

// ------------------------------------------------------------------------------------------------
// discrete adjoint
// ------------------------------------------------------------------------------------------------
namespace {

void adjoint_supported(const dgx_ctx* c) {
    if (!c) throw std::invalid_argument("null context");
    if (c->nf == 4) throw Fail(DGX_E_UNSUPPORTED, "the adjoint is built for triangles");
    if (c->physics != DGX_EULER) throw Fail(DGX_E_UNSUPPORTED, "the adjoint is built for the Euler residual only");
    if (c->has_slip) throw Fail(DGX_E_UNSUPPORTED, "the adjoint does not differentiate slip-wall ghost states");
    if (c->precision == DGX_FP32_BF16VOL)
        throw Fail(DGX_E_UNSUPPORTED, "the adjoint is built for the FP32 and FP64 operators");
}

// host [K][Np][4] -> device SoA in `dst`; states carry the storage offset, adjoints do not
void to_device(dgx_ctx* c, const double* h, void* dst, bool state) {
    size_t bytes = (size_t)c->K * c->np * 4 * sizeof(double);
    std::vector<double> tmp;
    hip_check(hipMemcpyAsync(c->stage.p, c->rows_in(h, (size_t)c->np * 4, tmp), bytes, hipMemcpyHostToDevice, c->stream),
              "copy");
    hip_check(launch_pack(c->precision, c->np, c->stage.as<double>(), dst, c->K, c->Kp,
                          state ? c->off : dgx_ctx::kNoOff, c->stream),
              "pack");
    hip_check(hipStreamSynchronize(c->stream), "sync");  // the host copy may be a temporary
}
void to_host(dgx_ctx* c, const void* src, double* h) {
    size_t bytes = (size_t)c->K * c->np * 4 * sizeof(double);
    hip_check(launch_unpack(c->precision, c->np, src, c->stage.as<double>(), c->K, c->Kp, dgx_ctx::kNoOff, c->stream), "unpack");
    hip_check(hipMemcpyAsync(h, c->stage.p, bytes, hipMemcpyDeviceToHost, c->stream), "copy");
    hip_check(hipStreamSynchronize(c->stream), "sync");
    c->rows_out(h, (size_t)c->np * 4);
}

void vjp(dgx_ctx* c, const void* uk, const void* lam, const void* add, const void* acc_in, void* acc_out, void* out,
         int mode, double dt, double ca, double cb) {
    VjpParams vp{};
    vp.uk = uk;
    vp.lam = lam;
    vp.add = add;
    vp.acc_in = acc_in;
    vp.acc_out = acc_out;
    vp.out = out;
    vp.geo = c->geo.as<double>();
    vp.nbr = c->nbr.as<int32_t>();
    vp.code = c->code.as<uint32_t>();
    vp.tab = mode == MODE_RESIDUAL ? c->tabR.p : c->tabL.p;
    vp.bc = c->bc.as<double>();
    vp.K = c->K;
    vp.Kp = c->Kp;
    vp.dt = dt;
    vp.ca = ca;
    vp.cb = cb;
    vp.gamma = c->gamma;
    vp.mode = mode;
    for (int v = 0; v < 4; ++v) vp.uref[v] = c->uref.v[v];
    vp.tabm = mode == MODE_RESIDUAL ? c->tabVR.p : c->tabVL.p;
    c->set_vjp_tile_params(vp, mode);
    hip_check(launch_vjp(c->precision, c->p, c->rule, vp, c->stream), "vjp kernel launch");
}

struct AdjointBuffers {
    std::vector<DevBuf> ckpt, stages[2];
    DevBuf l0, l1, l2, acc;
    hipEvent_t ev[5] = {};  // forward done; recomputation of stage set 0 / 1 done; reverse steps on set 0 / 1 done
    ~AdjointBuffers() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

// Reverse sweep of nsteps SSP-RK3 steps from the state in c->u with checkpoints every
// `every` steps; `lam` holds lambda_T on entry (device, in b.l0) and lambda_0 on exit.
// Segment j's stage states are recomputed from its checkpoint into stage set j mod 2 on the side
// stream while the reverse steps of segment j + 1 run on the context's stream from the other
// set: the recomputation of one segment overlaps the stage adjoints of the next, which are
// latency-bound launches that leave most of the GPU idle.  The forward pass keeps the last two
// segments' stage states in the sets as it goes.
// The checkpoint and stage-state buffers and the events of a sweep of nsteps steps.
void adjoint_buffers(dgx_ctx* c, AdjointBuffers& b, int nsteps, int every) {
    const size_t sb = c->state_bytes();
    const int nck = (nsteps + every - 1) / every;
    if (b.ckpt.size() < (size_t)nck) b.ckpt.resize(nck);
    for (int j = 0; j < nck; ++j) if (!b.ckpt[j].p) b.ckpt[j].alloc(sb);
    for (auto& set : b.stages) {
        if (set.size() < (size_t)(3 * every + 1)) set.resize(3 * every + 1);
        for (auto& s : set) if (!s.p) s.alloc(sb);
    }
    for (hipEvent_t& e : b.ev)
        if (!e) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
}

void rk3_adjoint(dgx_ctx* c, AdjointBuffers& b, double dt, int nsteps, int every, hipEvent_t mid = nullptr) {
    const size_t sb = c->state_bytes();
    const int nck = (nsteps + every - 1) / every;
    adjoint_buffers(c, b, nsteps, every);
    hipStream_t A = c->stream, B = c->side_stream();
    // Forward with checkpoints; the last two segments keep their stage states in the two stage
    // sets instead (the reverse sweep starts on them without recomputation).  A segment's
    // starting state lives where the sweep reads it -- its checkpoint, or entry 0 of its stage
    // set -- and the previous segment's last stage writes it there, so only the initial state is
    // copied.
    const int kept = std::min(nck, 2);
    auto start_of = [&](int j) -> void* { return j < nck - kept ? b.ckpt[j].p : b.stages[j & 1][0].p; };
    void* x = start_of(0);
    hip_check(hipMemcpyAsync(x, c->u.p, sb, hipMemcpyDeviceToDevice, A), "initial state");
    for (int j = 0; j < nck; ++j) {
        const int s0 = j * every, len = std::min(every, nsteps - s0);
        const bool keep = j >= nck - kept;
        std::vector<DevBuf>& st = b.stages[j & 1];
        for (int k = 0; k < len; ++k) {
            void *t1, *t2, *y;
            if (keep) {
                t1 = st[3 * k + 1].p;
                t2 = st[3 * k + 2].p;
                y = st[3 * k + 3].p;
            } else {
                // two of the context's state buffers other than x; the step's output reuses the
                // first stage's (it is read only by the second stage)
                void* pool[3] = {c->u.p, c->u1.p, c->u2.p};
                int n = 0;
                void* pick[2] = {nullptr, nullptr};
                for (void* q : pool)
                    if (q != x && n < 2) pick[n++] = q;
                t1 = pick[0];
                t2 = pick[1];
                y = t1;
            }
            if (k + 1 == len) y = j + 1 < nck ? start_of(j + 1) : c->u.p;
            c->stage_launch(x, x, t1, MODE_RK, dt, 0.0, 1.0);
            c->stage_launch(t1, x, t2, MODE_RK, dt, 0.75, 0.25);
            c->stage_launch(t2, x, y, MODE_RK, dt, 1.0 / 3.0, 2.0 / 3.0);
            x = y;
        }
    }
    if (mid) hip_check(hipEventRecord(mid, A), "record");
    hip_check(hipEventRecord(b.ev[0], A), "record");
    for (int set = 0; set < 2; ++set) hip_check(hipEventRecord(b.ev[1 + set], A), "record");  // the kept sets
    hip_check(hipStreamWaitEvent(B, b.ev[0], 0), "wait");
    // the state at step k of segment j: its start for k = 0, else entry 3k of its stage set
    auto state_at = [&](int j, int k) -> void* { return k == 0 ? start_of(j) : b.stages[j & 1][3 * k].p; };
    // stage states of segment j into its set, on the side stream, from its checkpoint
    bool used[2] = {false, false};  // a reverse step has read the set
    auto recompute = [&](int j) {
        const int set = j & 1, s0 = j * every, len = std::min(every, nsteps - s0);
        std::vector<DevBuf>& st = b.stages[set];
        if (used[set]) hip_check(hipStreamWaitEvent(B, b.ev[3 + set], 0), "wait");
        for (int k = 0; k < len; ++k) {
            void* v0 = state_at(j, k);
            void* v1 = st[3 * k + 1].p;
            void* v2 = st[3 * k + 2].p;
            c->stage_launch(v0, v0, v1, MODE_RK, dt, 0.0, 1.0, B);
            c->stage_launch(v1, v0, v2, MODE_RK, dt, 0.75, 0.25, B);
            if (k + 1 < len) c->stage_launch(v2, v0, st[3 * k + 3].p, MODE_RK, dt, 1.0 / 3.0, 2.0 / 3.0, B);
        }
        hip_check(hipEventRecord(b.ev[1 + set], B), "record");
    };
    for (int j = nck - 1; j >= 0; --j) {
        if (j > 0 && j - 1 < nck - kept) recompute(j - 1);
        const int set = j & 1, s0 = j * every, len = std::min(every, nsteps - s0);
        std::vector<DevBuf>& st = b.stages[set];
        hip_check(hipStreamWaitEvent(A, b.ev[1 + set], 0), "wait");
        for (int k = len - 1; k >= 0; --k) {
            void* v0 = state_at(j, k);
            void* v1 = st[3 * k + 1].p;
            void* v2 = st[3 * k + 2].p;
            vjp(c, v2, b.l0.p, nullptr, nullptr, b.acc.p, b.l1.p, MODE_RK, dt, 1.0 / 3.0, 2.0 / 3.0);
            vjp(c, v1, b.l1.p, nullptr, b.acc.p, b.acc.p, b.l2.p, MODE_RK, dt, 0.75, 0.25);
            vjp(c, v0, b.l2.p, b.acc.p, nullptr, nullptr, b.l0.p, MODE_RK, dt, 0.0, 1.0);
        }
        hip_check(hipEventRecord(b.ev[3 + set], A), "record");
        used[set] = true;
    }
}

}  // namespace

dgx_status dgx_residual_vjp(dgx_ctx* c, const double* u, const double* lam, double* out) {
    return guarded([&] {
        adjoint_supported(c);
        if (!lam || !out) throw std::invalid_argument("null argument");
        const void* src = c->u.p;
        if (u) { to_device(c, u, c->u2.p, true); src = c->u2.p; }
        to_device(c, lam, c->u1.p, false);
        DevBuf o;
        o.alloc(c->state_bytes());
        vjp(c, src, c->u1.p, nullptr, nullptr, nullptr, o.p, MODE_RESIDUAL, 0, 0, 0);
        to_host(c, o.p, out);
    });
}

dgx_status dgx_rhs_vjp(dgx_ctx* c, const double* u, const double* lam, double* out) {
    return guarded([&] {
        adjoint_supported(c);
        if (!lam || !out) throw std::invalid_argument("null argument");
        const void* src = c->u.p;
        if (u) { to_device(c, u, c->u2.p, true); src = c->u2.p; }
        to_device(c, lam, c->u1.p, false);
        DevBuf o;
        o.alloc(c->state_bytes());
        vjp(c, src, c->u1.p, nullptr, nullptr, nullptr, o.p, MODE_RHS, 0, 0, 0);
        to_host(c, o.p, out);
    });
}

dgx_status dgx_kinetic_energy_grad(dgx_ctx* c, double* J, double* grad) {
    return guarded([&] {
        if (!c) throw std::invalid_argument("null context");
        if (c->nf == 4) throw Fail(DGX_E_UNSUPPORTED, "the adjoint is built for triangles");
        if (J) {
            std::vector<double> h = c->reduce(1);
            double s = 0.0;
            for (int b = 0; b < c->red_blocks(); ++b) s += h[b * 4];
            *J = s;
        }
        if (grad) {
            TabLayout L(c->p, c->rule);
            hip_check(launch_ke_grad(c->precision, c->p, c->rule, c->u.p, c->tab64.as<double>(), c->qw.as<double>(),
                                     c->geo.as<double>(), c->u1.p, c->K, c->Kp, c->uref, c->stream),
                      "ke grad");
            to_host(c, c->u1.p, grad);
        }
    });
}

dgx_status dgx_rk3_vjp(dgx_ctx* c, double dt, int32_t nsteps, int32_t every, const double* lambda_T, double* lambda_0) {
    return guarded([&] {
        adjoint_supported(c);
        if (nsteps <= 0 || every <= 0 || !lambda_0) throw std::invalid_argument("bad adjoint arguments");
        AdjointBuffers b;
        for (DevBuf* x : {&b.l0, &b.l1, &b.l2, &b.acc}) x->alloc(c->state_bytes());
        // seed: lambda_T given, or the kinetic-energy gradient at u_T (computed after the forward)
        if (lambda_T) {
            to_device(c, lambda_T, b.l0.p, false);
            rk3_adjoint(c, b, dt, nsteps, every);
        } else {
            // forward once to reach u_T, seed, then restart from the saved initial state
            DevBuf u0;
            u0.alloc(c->state_bytes());
            hip_check(hipMemcpyAsync(u0.p, c->u.p, c->state_bytes(), hipMemcpyDeviceToDevice, c->stream), "copy");
            for (int s = 0; s < nsteps; ++s) c->rk3_step(dt);
            hip_check(launch_ke_grad(c->precision, c->p, c->rule, c->u.p, c->tab64.as<double>(), c->qw.as<double>(),
                                     c->geo.as<double>(), b.l0.p, c->K, c->Kp, c->uref, c->stream),
                      "ke grad");
            hip_check(hipMemcpyAsync(c->u.p, u0.p, c->state_bytes(), hipMemcpyDeviceToDevice, c->stream), "copy");
            rk3_adjoint(c, b, dt, nsteps, every);
        }
        to_host(c, b.l0.p, lambda_0);
    });
}

dgx_status dgx_rk3_vjp_profile(dgx_ctx* c, double dt, int32_t nsteps, int32_t every, double* fwd_ms,
                               double* ckpt_fwd_ms, double* rev_ms) {
    return guarded([&] {
        adjoint_supported(c);
        if (nsteps <= 0 || every <= 0) throw std::invalid_argument("bad adjoint arguments");
        AdjointBuffers b;
        for (DevBuf* x : {&b.l0, &b.l1, &b.l2, &b.acc}) x->alloc(c->state_bytes());
        DevBuf u0;
        u0.alloc(c->state_bytes());
        hip_check(hipMemcpyAsync(u0.p, c->u.p, c->state_bytes(), hipMemcpyDeviceToDevice, c->stream), "copy");
        // every buffer of the timed sweep allocated before it, and a warm-up sweep of three
        // segments (one recomputed) for the side stream and the code paths
        adjoint_buffers(c, b, nsteps, every);
        hip_check(launch_ke_grad(c->precision, c->p, c->rule, c->u.p, c->tab64.as<double>(), c->qw.as<double>(),
                                 c->geo.as<double>(), b.l0.p, c->K, c->Kp, c->uref, c->stream), "ke grad");
        rk3_adjoint(c, b, dt, std::min(nsteps, 3 * every), every);
        hip_check(hipMemcpyAsync(c->u.p, u0.p, c->state_bytes(), hipMemcpyDeviceToDevice, c->stream), "copy");
        // forward only
        hip_check(hipEventRecord(c->ev0, c->stream), "record");
        for (int s = 0; s < nsteps; ++s) c->rk3_step(dt);
        hip_check(hipEventRecord(c->ev1, c->stream), "record");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        float f = 0;
        hip_check(hipEventElapsedTime(&f, c->ev0, c->ev1), "elapsed");
        // adjoint (forward with checkpoints + recompute + reverse), seeded at u_T
        hip_check(launch_ke_grad(c->precision, c->p, c->rule, c->u.p, c->tab64.as<double>(), c->qw.as<double>(),
                                 c->geo.as<double>(), b.l0.p, c->K, c->Kp, c->uref, c->stream), "ke grad");
        hip_check(hipMemcpyAsync(c->u.p, u0.p, c->state_bytes(), hipMemcpyDeviceToDevice, c->stream), "copy");
        hipEvent_t mid = nullptr;
        hip_check(hipEventCreate(&mid), "hipEventCreate");
        hip_check(hipEventRecord(c->ev0, c->stream), "record");
        rk3_adjoint(c, b, dt, nsteps, every, mid);
        hip_check(hipEventRecord(c->ev1, c->stream), "record");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        float a = 0, r = 0;
        hip_check(hipEventElapsedTime(&a, c->ev0, mid), "elapsed");
        hip_check(hipEventElapsedTime(&r, mid, c->ev1), "elapsed");
        (void)hipEventDestroy(mid);
        hip_check(hipMemcpyAsync(c->u.p, u0.p, c->state_bytes(), hipMemcpyDeviceToDevice, c->stream), "copy");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        if (fwd_ms) *fwd_ms = f;
        if (ckpt_fwd_ms) *ckpt_fwd_ms = a;
        if (rev_ms) *rev_ms = r;
    });
}

dgx_status dgx_max_cfl(dgx_ctx* c, double dt, double* cfl) {
    return guarded([&] {
        if (!c || !cfl) throw std::invalid_argument("null argument");
        std::vector<double> h = c->reduce(2);
        double mx = 0.0, bad = 0.0;
        for (int b = 0; b < c->red_blocks(); ++b) { mx = std::max(mx, h[b * 4]); bad += h[b * 4 + 3]; }
        *cfl = bad > 0 ? INFINITY : dt * mx;
    });
}

dgx_status dgx_density_range(dgx_ctx* c, double out[2]) {
    return guarded([&] {
        if (!c || !out) throw std::invalid_argument("null argument");
        std::vector<double> h = c->reduce(2);
        double lo = INFINITY, hi = -INFINITY;
        for (int b = 0; b < c->red_blocks(); ++b) { lo = std::min(lo, h[b * 4 + 1]); hi = std::max(hi, h[b * 4 + 2]); }
        if (!std::isfinite(lo) || !std::isfinite(hi)) { out[0] = out[1] = NAN; return; }
        out[0] = lo;
        out[1] = hi;
    });
}

}  // extern "C"
