// Host-side builder of the MFMA stage kernel's lane-major tables (layout in dgx_mfma.hpp).
#include "dgx_mfma.hpp"

#include <cstdint>
#include <cstring>
#include <stdexcept>

namespace dgx {

namespace {

// Round to bfloat16, nearest-even (finite inputs).
uint16_t to_bf16(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

// Two bfloat16 values packed in one 32-bit word (low half first), stored in a float slot.
float pack_bf16(float lo, float hi) {
    const uint32_t w = uint32_t(to_bf16(lo)) | (uint32_t(to_bf16(hi)) << 16);
    float f;
    std::memcpy(&f, &w, 4);
    return f;
}

template <int P>
void build(const std::vector<double>& t, std::vector<float>& out, bool bf16) {
    using S = MfmaShape<P>;
    const TabLayout L(P, 0);
    const int np = S::NP, nq = S::NQ, nfq = S::NFQ;
    if ((int)t.size() != L.total) throw std::runtime_error("table size mismatch");
    out.assign(bf16 ? S::TOTAL_BF : S::TOTAL, 0.0f);
    auto at = [&](int table, int lane) -> float& { return out[(size_t)table * 64 + lane]; };
    for (int lane = 0; bf16 && lane < 64; ++lane) {
        const int i = lane % 16, g = lane / 16;
        int vp = mf_vpoint(S::RQ, nq, i / 4, i % 4);
        if (vp < 0) vp = 0;
        const int mrow = mf_mode(S::R, np, i / 4, i % 4);
        float e4[4], x4[4], y4[4];
        for (int tt = 0; tt < 4; ++tt) {
            const int m = mf_mode(S::R, np, g, tt);     // k = 4g + tt: mode slot (g, tt)
            const int q = mf_vpoint(S::RQ, nq, g, tt);  // or volume slot (g, tt)
            e4[tt] = m < 0 ? 0.0f : float(t[L.phiV + vp * np + m]);
            x4[tt] = (mrow < 0 || q < 0) ? 0.0f : float(t[L.PX + q * np + mrow]);
            y4[tt] = (mrow < 0 || q < 0) ? 0.0f : float(t[L.PY + q * np + mrow]);
        }
        for (int h = 0; h < 2; ++h) {
            out[(size_t)S::T_EB + h * 64 + lane] = pack_bf16(e4[2 * h], e4[2 * h + 1]);
            out[(size_t)S::T_XB + h * 64 + lane] = pack_bf16(x4[2 * h], x4[2 * h + 1]);
            out[(size_t)S::T_YB + h * 64 + lane] = pack_bf16(y4[2 * h], y4[2 * h + 1]);
        }
    }
    for (int lane = 0; lane < 64; ++lane) {
        const int i = lane % 16, k = lane / 16;
        // rows of an evaluation product are point slots; padding rows repeat point 0 so that
        // their (unused) fluxes stay finite
        int vp = mf_vpoint(S::RQ, nq, i / 4, i % 4);
        if (vp < 0) vp = 0;
        int fp = mf_fpoint(nfq, i / 4, i % 4);
        if (fp < 0) fp = 0;
        for (int s = 0; s < S::R; ++s) {
            const int m = mf_mode(S::R, np, k, s);
            at(S::T_EV + s, lane) = m < 0 ? 0.0f : float(t[L.phiV + vp * np + m]);
            at(S::T_EF + s, lane) = m < 0 ? 0.0f : float(t[L.phiF + fp * np + m]);
        }
        // rows of a projection / lift product are mode slots, k runs over point slots
        const int mrow = mf_mode(S::R, np, i / 4, i % 4);
        for (int r = 0; r < S::RQ; ++r) {
            const int q = mf_vpoint(S::RQ, nq, k, r);
            const bool ok = mrow >= 0 && q >= 0;
            at(S::T_PX + r, lane) = ok ? float(t[L.PX + q * np + mrow]) : 0.0f;
            at(S::T_PY + r, lane) = ok ? float(t[L.PY + q * np + mrow]) : 0.0f;
        }
        for (int r = 0; r < S::RF; ++r) {
            const int f = mf_fpoint(nfq, k, r);
            at(S::T_LF + r, lane) = (mrow >= 0 && f >= 0) ? float(t[L.LF + f * np + mrow]) : 0.0f;
        }
    }
    // neighbour traces: NT[(fn*2 + flip)][i][q] = phi_{fn, q'}(mode i), q' = nfq - 1 - q for a
    // reversed face
    float* nt = &out[(size_t)S::T_N * 64];
    for (int fn = 0; fn < 3; ++fn)
        for (int flip = 0; flip < 2; ++flip)
            for (int m = 0; m < np; ++m)
                for (int q = 0; q < nfq; ++q) {
                    const int qq = flip ? nfq - 1 - q : q;
                    nt[((fn * 2 + flip) * np + m) * nfq + q] = float(t[L.phiF + (fn * nfq + qq) * np + m]);
                }
}

template <int P>
void build_vjp(const std::vector<double>& t, std::vector<float>& out) {
    using S = MfmaVjpShape<P>;
    using F = MfmaShape<P>;
    const TabLayout L(P, 0);
    const int np = F::NP, nq = F::NQ, nfq = F::NFQ;
    if ((int)t.size() != L.total) throw std::runtime_error("table size mismatch");
    out.assign(S::TOTAL, 0.0f);
    auto at = [&](int table, int lane) -> float& { return out[(size_t)table * 64 + lane]; };
    for (int lane = 0; lane < 64; ++lane) {
        const int i = lane % 16, k = lane / 16;
        const int vq = mf_vpoint(S::RQ, nq, i / 4, i % 4), fq = mf_fpoint(nfq, i / 4, i % 4);
        const int vp = vq < 0 ? 0 : vq, fp = fq < 0 ? 0 : fq;
        for (int s = 0; s < S::R; ++s) {
            const int m = mf_mode(S::R, np, k, s);
            if (m < 0) continue;
            at(S::T_EV + s, lane) = float(t[L.phiV + vp * np + m]);
            at(S::T_XV + s, lane) = vq < 0 ? 0.0f : float(t[L.PX + vq * np + m]);
            at(S::T_YV + s, lane) = vq < 0 ? 0.0f : float(t[L.PY + vq * np + m]);
            at(S::T_EF + s, lane) = float(t[L.phiF + fp * np + m]);
            at(S::T_LE + s, lane) = fq < 0 ? 0.0f : float(t[L.LF + fq * np + m]);
        }
        const int mrow = mf_mode(S::R, np, i / 4, i % 4);
        for (int r = 0; r < S::RQ; ++r) {
            const int q = mf_vpoint(S::RQ, nq, k, r);
            at(S::T_VT + r, lane) = (mrow >= 0 && q >= 0) ? float(t[L.phiV + q * np + mrow]) : 0.0f;
        }
        for (int r = 0; r < S::RF; ++r) {
            const int f = mf_fpoint(nfq, k, r);
            at(S::T_FT + r, lane) = (mrow >= 0 && f >= 0) ? float(t[L.phiF + f * np + mrow]) : 0.0f;
        }
    }
    float* h = &out[(size_t)S::T_N * 64];
    for (int fp = 0; fp < 3 * nfq; ++fp)
        for (int m = 0; m < np; ++m) {
            float* x = &h[((size_t)fp * S::HV + m / 2) * 4 + 2 * (m & 1)];
            x[0] = float(t[L.phiF + fp * np + m]);
            x[1] = float(t[L.LF + fp * np + m]);
        }
}

// Populated face-gradient slot of (point fp, component c, mode i) and its two values T_ck,
// checked against the structure the kernel relies on (dgx_mfma.hpp).
struct FaceGradEntry {
    double v[2];
};
template <int P>
FaceGradEntry face_grad_entry(const std::vector<double>& t, const TabLayout& L, int fp, int c, int i, bool quirk) {
    using S = MfmaNsShape<P>;
    const int np = S::NP, nfq = S::NFQ, q = fp % nfq;
    auto fg = [&](int d, int k) { return t[L.FG + (fp * 8 + (c * 2 + d) * 2 + k) * np + i]; };
    const int e = (q * np + c * nfq * np) & 1;
    const int d = quirk ? (e ^ (i & 1)) : c;
    if (fg(1 - d, 0) != 0.0 || fg(1 - d, 1) != 0.0)
        throw std::runtime_error("face-gradient table does not have the expected populated slots");
    if (quirk && S::R % 2 != 0 && e != 0)
        throw std::runtime_error("face-gradient row parity is not uniform for an odd slot count");
    return {{fg(d, 0), fg(d, 1)}};
}

template <int P>
void build_ns(const std::vector<double>& t, bool trace, bool quirk, std::vector<float>& out) {
    using S = MfmaNsShape<P>;
    using F = MfmaShape<P>;
    const TabLayout L(P, 0);
    const int np = F::NP, nq = F::NQ, nfq = F::NFQ;
    if ((int)t.size() != L.total) throw std::runtime_error("table size mismatch");
    out.assign(trace ? S::TOTAL_TH : S::TOTAL_M, 0.0f);
    auto at = [&](int table, int lane) -> float& { return out[(size_t)table * 64 + lane]; };
    for (int lane = 0; lane < 64; ++lane) {
        const int i = lane % 16, k = lane / 16;
        int vp = mf_vpoint(S::RQ, nq, i / 4, i % 4), fp = mf_fpoint(nfq, i / 4, i % 4);
        if (vp < 0) vp = 0;  // padding rows repeat a valid point (finite states)
        if (fp < 0) fp = 0;
        const int mrow = mf_mode(S::R, np, i / 4, i % 4);
        for (int s = 0; s < S::R; ++s) {
            const int m = mf_mode(S::R, np, k, s);
            if (m < 0) continue;
            if (trace) {
                at(S::T_EF + s, lane) = float(t[L.phiF + fp * np + m]);
                for (int c = 0; c < 2; ++c) {
                    const FaceGradEntry x = face_grad_entry<P>(t, L, fp, c, m, quirk);
                    for (int k = 0; k < 2; ++k) at(S::T_FG + (c * 2 + k) * S::R + s, lane) = float(x.v[k]);
                }
            } else {
                at(S::M_EV + s, lane) = float(t[L.phiV + vp * np + m]);
                at(S::M_GX + s, lane) = float(t[L.DXV + vp * np + m]);
                at(S::M_GY + s, lane) = float(t[L.DYV + vp * np + m]);
            }
        }
        if (trace) continue;
        for (int r = 0; r < S::RQ; ++r) {
            const int q = mf_vpoint(S::RQ, nq, k, r);
            const bool ok = mrow >= 0 && q >= 0;
            at(S::M_PX + r, lane) = ok ? float(t[L.PX + q * np + mrow]) : 0.0f;
            at(S::M_PY + r, lane) = ok ? float(t[L.PY + q * np + mrow]) : 0.0f;
        }
        for (int r = 0; r < S::RF; ++r) {
            const int f = mf_fpoint(nfq, k, r);
            at(S::M_LF + r, lane) = (mrow >= 0 && f >= 0) ? float(t[L.LF + f * np + mrow]) : 0.0f;
        }
    }
    if (!trace) return;
    // external-face tables
    for (int fp = 0; fp < 3 * nfq; ++fp) {
        float* h = &out[S::TOTAL_T + (size_t)fp * S::HP * 4];
        for (int i = 0; i < np; ++i) {
            for (int c = 0; c < 2; ++c) {
                const FaceGradEntry x = face_grad_entry<P>(t, L, fp, c, i, quirk);
                h[i * 4 + 2 * c] = float(x.v[0]);
                h[i * 4 + 2 * c + 1] = float(x.v[1]);
            }
            h[np * 4 + i] = float(t[L.phiF + fp * np + i]);
        }
    }
}

}  // namespace

void build_mfma_ns_tables(int p, const std::vector<double>& tab, bool trace, bool quirk, std::vector<float>& out) {
    switch (p) {
    case 3: build_ns<3>(tab, trace, quirk, out); break;
    case 4: build_ns<4>(tab, trace, quirk, out); break;
    default: throw std::runtime_error("no MFMA Navier-Stokes tables for this order");
    }
}

void build_mfma_vjp_tables(int p, const std::vector<double>& tab, std::vector<float>& out) {
    switch (p) {
    case 3: build_vjp<3>(tab, out); break;
    case 4: build_vjp<4>(tab, out); break;
    default: throw std::runtime_error("no MFMA adjoint tables for this order");
    }
}

int mfma_table_size(int p) {
    switch (p) {
    case 1: return MfmaShape<1>::TOTAL;
    case 2: return MfmaShape<2>::TOTAL;
    case 3: return MfmaShape<3>::TOTAL;
    case 4: return MfmaShape<4>::TOTAL;
    default: return 0;
    }
}

void build_mfma_tables(int p, const std::vector<double>& tab, std::vector<float>& out, bool bf16) {
    switch (p) {
    case 1: build<1>(tab, out, bf16); break;
    case 2: build<2>(tab, out, bf16); break;
    case 3: build<3>(tab, out, bf16); break;
    case 4: build<4>(tab, out, bf16); break;
    default: throw std::runtime_error("no MFMA tables for this order");
    }
}

}  // namespace dgx
