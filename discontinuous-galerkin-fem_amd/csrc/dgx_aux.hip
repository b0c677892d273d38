// Auxiliary device kernels: layout conversion between the host [K][Np][4] FP64 layout and the
// device SoA state, L2 projection of point values, evaluation at volume points, the inverse
// mass on host-layout data and deterministic FP64 reductions (conservation sums, kinetic
// energy, CFL / density range / finiteness).
#include "dgx_internal.hpp"

namespace dgx {
namespace {

template <typename Real>
__global__ void pack_kernel(const double* __restrict__ h, Real* __restrict__ u, int np, int K, int Kp, Off4 off) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= K) return;
    for (int k = 0; k < np * 4; ++k) {
        double x = h[(size_t)e * np * 4 + k];
        u[sidx(k >> 2, k & 3, e, Kp)] = Real(k < 4 ? x - off.v[k] : x);
    }
}

template <typename Real>
__global__ void unpack_kernel(const Real* __restrict__ u, double* __restrict__ h, int np, int K, int Kp, Off4 off) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= K) return;
    for (int k = 0; k < np * 4; ++k) {
        double x = double(u[sidx(k >> 2, k & 3, e, Kp)]);
        h[(size_t)e * np * 4 + k] = k < 4 ? x + off.v[k] : x;
    }
}

// u_i = sum_q Pm[q][i] U(x_q), Pm[q][i] = sum_j Minv_ref[i][j] w_q phi_j(x_q): the L2 projection
// M_e^-1 int phi U with affine |J| cancelling (project_initial_condition,
// S/solver/compressible_solver_base.cpp:63-104).  FP64 arithmetic.
template <typename Real>
__global__ void project_kernel(const double* __restrict__ Uq, Real* __restrict__ u,
                               const double* __restrict__ Pm, int np, int nq, int K, int Kp, Off4 off) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= K) return;
    for (int i = 0; i < np; ++i)
        for (int v = 0; v < 4; ++v) {
            double s = 0.0;
            for (int q = 0; q < nq; ++q) s += Pm[q * np + i] * Uq[((size_t)e * nq + q) * 4 + v];
            u[sidx(i, v, e, Kp)] = Real(i == 0 ? s - off.v[v] : s);
        }
}

template <typename Real>
__global__ void eval_kernel(const Real* __restrict__ u, const double* __restrict__ phiV,
                            double* __restrict__ Uq, int np, int nq, int K, int Kp, Off4 uref) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= K) return;
    for (int q = 0; q < nq; ++q)
        for (int v = 0; v < 4; ++v) {
            double s = uref.v[v];
            for (int i = 0; i < np; ++i) s += phiV[q * np + i] * double(u[sidx(i, v, e, Kp)]);
            Uq[((size_t)e * nq + q) * 4 + v] = s;
        }
}

// r <- M_ref^-1 r / |J| on host-layout data (apply_mass_inv, :54-61), FP64.
__global__ void minv_kernel(const double* __restrict__ minv, const double* __restrict__ geo,
                            double* __restrict__ r, int np, int K, int Kp) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= K) return;
    double rdet = 1.0 / geo[G_DET * Kp + e];
    double tmp[15 * 4];
    for (int k = 0; k < np * 4; ++k) tmp[k] = r[(size_t)e * np * 4 + k];
    for (int i = 0; i < np; ++i)
        for (int v = 0; v < 4; ++v) {
            double s = 0.0;
            for (int j = 0; j < np; ++j) s += minv[i * np + j] * tmp[j * 4 + v];
            r[(size_t)e * np * 4 + i * 4 + v] = s * rdet;
        }
}

constexpr int kRed = 256;
constexpr int kMaxOut = 4;

// Per-block reductions in FP64 with a fixed tree order (deterministic).  Output per block:
//   kind 0: conservation sums  sum_e |J| sum_i m_i u_i,v         (4 values, sum)
//   kind 1: kinetic energy     sum_e sum_q w_q |J| |m|^2/(2 rho)  (1 value, sum)
//   kind 2: [max (|u|+c)/h_min, rho min, rho max, #bad points]   (max, min, max, sum)
//   kind 3: [#NaN state entries, #infinite state entries]         (sum, sum)
template <typename Real>
__global__ __launch_bounds__(kRed) void reduce_kernel(int kind, const Real* __restrict__ u,
                                                      const double* __restrict__ tab64,
                                                      const double* __restrict__ geo,
                                                      const double* __restrict__ hmin,
                                                      const double* __restrict__ qw, int np, int nq,
                                                      int phiV_off, int mass_off, int K, int Kp,
                                                      double gamma, Off4 uref, Off4 off,
                                                      double* __restrict__ partial) {
    __shared__ double sh[kMaxOut][kRed];
    int e = blockIdx.x * kRed + threadIdx.x;
    double val[kMaxOut] = {0.0, 0.0, 0.0, 0.0};
    if (kind == 2) { val[1] = INFINITY; val[2] = -INFINITY; }
    if (e < K) {
        double det = geo[G_DET * Kp + e];
        if (kind == 0) {
            for (int v = 0; v < 4; ++v) val[v] = tab64[mass_off] * off.v[v];
            for (int i = 0; i < np; ++i) {
                double m = tab64[mass_off + i];
                for (int v = 0; v < 4; ++v) val[v] += m * double(u[sidx(i, v, e, Kp)]);
            }
            for (int v = 0; v < 4; ++v) val[v] *= det;
        } else if (kind == 3) {
            for (int k = 0; k < np * 4; ++k) {
                double x = double(u[sidx(k >> 2, k & 3, e, Kp)]);
                if (isnan(x)) val[0] += 1.0;
                else if (isinf(x)) val[1] += 1.0;
            }
        } else {
            double ws = 0.0;
            for (int q = 0; q < nq; ++q) {
                double U[4] = {uref.v[0], uref.v[1], uref.v[2], uref.v[3]};
                for (int i = 0; i < np; ++i) {
                    double ph = tab64[phiV_off + q * np + i];
                    for (int v = 0; v < 4; ++v) U[v] += ph * double(u[sidx(i, v, e, Kp)]);
                }
                if (kind == 1) {
                    val[0] += qw[q] * det * 0.5 * (U[1] * U[1] + U[2] * U[2]) / U[0];
                } else {
                    double rho = U[0];
                    if (isfinite(rho)) { val[1] = fmin(val[1], rho); val[2] = fmax(val[2], rho); }
                    if (!(rho > 0.0) || !isfinite(rho)) { val[3] += 1.0; continue; }
                    double uu = U[1] / rho, vv = U[2] / rho;
                    double p = (gamma - 1.0) * (U[3] - rho * 0.5 * (uu * uu + vv * vv));
                    if (!(p > 0.0) || !isfinite(p)) { val[3] += 1.0; continue; }
                    double c = sqrt(fmax(gamma * p / rho, 0.0));
                    ws = fmax(ws, sqrt(uu * uu + vv * vv) + c);
                }
            }
            if (kind == 2) val[0] = ws / hmin[e];
        }
    }
    for (int k = 0; k < kMaxOut; ++k) sh[k][threadIdx.x] = val[k];
    __syncthreads();
    for (int s = kRed / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int k = 0; k < kMaxOut; ++k) {
                double a = sh[k][threadIdx.x], b = sh[k][threadIdx.x + s];
                double r;
                if (kind == 2 && k == 0) r = fmax(a, b);
                else if (kind == 2 && k == 1) r = fmin(a, b);
                else if (kind == 2 && k == 2) r = fmax(a, b);
                else r = a + b;
                sh[k][threadIdx.x] = r;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < kMaxOut; ++k) partial[blockIdx.x * kMaxOut + k] = sh[k][0];
}

// Relative-L2 error of the primitive variables W = (rho, u, v, p) at the volume points
// (compute_variable_error of the reference's accuracy tests,
// T/solver/shear_flow_accuracy_test.cpp:32-79, X/taylor_green_test.cpp:44-86): per block, FP64
// partial sums of w_q |J| (W - W_exact)^2 (entries 0..3) and w_q |J| W_exact^2 (entries 4..7).
// Uq holds the numerical point values [K][nq][4]; Ex the exact ones, conserved or primitive.
__device__ __forceinline__ void to_primitive(const double* U, double g, double* W) {
    W[0] = U[0];
    W[1] = U[1] / U[0];
    W[2] = U[2] / U[0];
    W[3] = (g - 1.0) * (U[3] - 0.5 * U[0] * (W[1] * W[1] + W[2] * W[2]));
}
__global__ __launch_bounds__(kRed) void l2err_kernel(const double* __restrict__ Uq, const double* __restrict__ Ex,
                                                     int ex_prim, const double* __restrict__ geo,
                                                     const double* __restrict__ qw, int nq, int K, int Kp,
                                                     double gamma, int per_point, double* __restrict__ partial) {
    __shared__ double sh[8][kRed];
    const int e = blockIdx.x * kRed + threadIdx.x;
    double val[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (e < K) {
        const double det = per_point ? 0.0 : geo[G_DET * Kp + e];
        for (int q = 0; q < nq; ++q) {
            const size_t o = ((size_t)e * nq + q) * 4;
            double W[4], X[4];
            to_primitive(Uq + o, gamma, W);
            if (ex_prim)
                for (int v = 0; v < 4; ++v) X[v] = Ex[o + v];
            else
                to_primitive(Ex + o, gamma, X);
            const double dw = per_point ? geo[(size_t)(5 * q + 4) * Kp + e] : qw[q] * det;
            for (int v = 0; v < 4; ++v) {
                const double d = W[v] - X[v];
                val[v] += d * d * dw;
                val[4 + v] += X[v] * X[v] * dw;
            }
        }
    }
    for (int k = 0; k < 8; ++k) sh[k][threadIdx.x] = val[k];
    __syncthreads();
    for (int s = kRed / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 8; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 8; ++k) partial[blockIdx.x * 8 + k] = sh[k][0];
}

inline dim3 grid_for(int K, int b) { return dim3((K + b - 1) / b); }

}  // namespace

hipError_t launch_pack(int precision, int np, const double* h, void* u, int K, int Kp, Off4 off, hipStream_t s) {
    if (precision == DGX_FP64)
        hipLaunchKernelGGL(pack_kernel<double>, grid_for(K, 256), dim3(256), 0, s, h, (double*)u, np, K, Kp, off);
    else
        hipLaunchKernelGGL(pack_kernel<float>, grid_for(K, 256), dim3(256), 0, s, h, (float*)u, np, K, Kp, off);
    return hipGetLastError();
}

hipError_t launch_unpack(int precision, int np, const void* u, double* h, int K, int Kp, Off4 off, hipStream_t s) {
    if (precision == DGX_FP64)
        hipLaunchKernelGGL(unpack_kernel<double>, grid_for(K, 256), dim3(256), 0, s, (const double*)u, h, np, K, Kp, off);
    else
        hipLaunchKernelGGL(unpack_kernel<float>, grid_for(K, 256), dim3(256), 0, s, (const float*)u, h, np, K, Kp, off);
    return hipGetLastError();
}

hipError_t launch_project(int precision, int p, int rule, const double* Uq, void* u, const void*,
                          int K, int Kp, const double* Pm, Off4 off, hipStream_t s) {
    int np = np_of(p), nq = nq_of(p, rule);
    if (precision == DGX_FP64)
        hipLaunchKernelGGL(project_kernel<double>, grid_for(K, 256), dim3(256), 0, s, Uq, (double*)u, Pm, np, nq, K, Kp, off);
    else
        hipLaunchKernelGGL(project_kernel<float>, grid_for(K, 256), dim3(256), 0, s, Uq, (float*)u, Pm, np, nq, K, Kp, off);
    return hipGetLastError();
}

hipError_t launch_eval(int precision, int p, int rule, const void* u, const void* phiV64, int K,
                       int Kp, double* Uq, Off4 uref, hipStream_t s) {
    return launch_eval_n(precision, np_of(p), nq_of(p, rule), u, phiV64, K, Kp, Uq, uref, s);
}
hipError_t launch_eval_n(int precision, int np, int nq, const void* u, const void* phiV64, int K, int Kp, double* Uq,
                         Off4 uref, hipStream_t s) {
    const double* ph = static_cast<const double*>(phiV64);
    if (precision == DGX_FP64)
        hipLaunchKernelGGL(eval_kernel<double>, grid_for(K, 256), dim3(256), 0, s, (const double*)u, ph, Uq, np, nq, K, Kp, uref);
    else
        hipLaunchKernelGGL(eval_kernel<float>, grid_for(K, 256), dim3(256), 0, s, (const float*)u, ph, Uq, np, nq, K, Kp, uref);
    return hipGetLastError();
}

hipError_t launch_l2_error(const double* Uq, const double* Ex, int ex_prim, const double* geo, const double* qw, int nq,
                           int K, int Kp, double gamma, double* partial, int nblocks, int per_point, hipStream_t s) {
    hipLaunchKernelGGL(l2err_kernel, dim3(nblocks), dim3(kRed), 0, s, Uq, Ex, ex_prim, geo, qw, nq, K, Kp, gamma, per_point,
                       partial);
    return hipGetLastError();
}

hipError_t launch_apply_minv(int, int np, const double* minv, const double* geo, double* r, int K,
                             int Kp, hipStream_t s) {
    hipLaunchKernelGGL(minv_kernel, grid_for(K, 256), dim3(256), 0, s, minv, geo, r, np, K, Kp);
    return hipGetLastError();
}

hipError_t launch_reduce(int kind, int precision, int p, int rule, const void* u, const void* tab64,
                         const double* geo, const double* hmin, const double* qw, int K, int Kp,
                         double gamma, Off4 uref, Off4 off, double* partial, int nblocks, hipStream_t s) {
    int np = np_of(p), nq = nq_of(p, rule);
    TabLayout L(p, rule);
    const double* t = static_cast<const double*>(tab64);
    if (nblocks != (K + kRed - 1) / kRed) return hipErrorInvalidValue;
    if (precision == DGX_FP64)
        hipLaunchKernelGGL(reduce_kernel<double>, dim3(nblocks), dim3(kRed), 0, s, kind, (const double*)u, t, geo, hmin, qw,
                           np, nq, L.phiV, L.mass, K, Kp, gamma, uref, off, partial);
    else
        hipLaunchKernelGGL(reduce_kernel<float>, dim3(nblocks), dim3(kRed), 0, s, kind, (const float*)u, t, geo, hmin, qw,
                           np, nq, L.phiV, L.mass, K, Kp, gamma, uref, off, partial);
    return hipGetLastError();
}

}  // namespace dgx
