// lrk_mlp.h -- FP64 CPU restatement of the learned linear-response kernel's
// network, its training loss and the loss gradient (test infrastructure: the
// checker the HIP kernels are compared against; never part of the product path).
//
// No reference counterpart: the reference has no semi-implicit scheme.  The
// architecture, the parameter layout and the loss are the ones include/fts.h
// documents (FTS_LRK_*):
//   x_j = (f_j - shift_j) scale_j   (j < FTS_LRK_FEATURES; x_7 = 0)
//   h1 = tanh(W1 x + b1), h2 = tanh(W2 h1 + b2), h3 = tanh(W3 h2 + b3)
//   z = W4 h3 + b4,  r_c = 4 tanh(z_c / 4)
// and the loss of one training record (moments m, E_c, S1_{c,b}, S2_{c,b}) is
//   sum_c sum_{b: S2 > 0, E_c > 0} (S2 / E_c) (e^{r_c(b, m)} - S1 / S2)^2.
// Pinned by central finite differences of the loss (gradient), by the closed
// form of a network whose only non-zero parameters are the output biases, and
// by the residual identity the record statistics satisfy.
#pragma once

#include <cmath>
#include <cstring>
#include <vector>

#include "../include/fts.h"

namespace oracle {

constexpr int LRK_H = FTS_LRK_HIDDEN;
constexpr int LRK_NIN = FTS_LRK_IN;
constexpr int LRK_NB = FTS_LRK_BINS;

struct LrkActs {
    double x[LRK_NIN];
    double h1[LRK_H], h2[LRK_H], h3[LRK_H];
    double z[FTS_LRK_OUT], r[FTS_LRK_OUT];
};

// Forward pass of one feature vector f[FTS_LRK_FEATURES].
inline void lrk_forward(const double* P, const double* f, LrkActs& a)
{
    for (int j = 0; j < LRK_NIN; ++j)
        a.x[j] = j < FTS_LRK_FEATURES ? (f[j] - P[FTS_LRK_OFF_SHIFT + j]) * P[FTS_LRK_OFF_SCALE + j] : 0.0;
    auto layer = [](const double* W, const double* b, const double* in, int nin, double* out) {
        for (int o = 0; o < LRK_H; ++o) {
            double s = b[o];
            for (int i = 0; i < nin; ++i) s += W[o * nin + i] * in[i];
            out[o] = std::tanh(s);
        }
    };
    layer(P + FTS_LRK_OFF_W1, P + FTS_LRK_OFF_B1, a.x, LRK_NIN, a.h1);
    layer(P + FTS_LRK_OFF_W2, P + FTS_LRK_OFF_B2, a.h1, LRK_H, a.h2);
    layer(P + FTS_LRK_OFF_W3, P + FTS_LRK_OFF_B3, a.h2, LRK_H, a.h3);
    for (int c = 0; c < FTS_LRK_OUT; ++c) {
        double s = P[FTS_LRK_OFF_B4 + c];
        for (int i = 0; i < LRK_H; ++i) s += P[FTS_LRK_OFF_W4 + c * LRK_H + i] * a.h3[i];
        a.z[c] = s;
        a.r[c] = 4.0 * std::tanh(0.25 * s);
    }
}

// Feature vector of bin b under moments m[FTS_LRK_MOMENTS].
inline void lrk_features(int b, const double* m, double* f)
{
    f[0] = (b + 0.5) / LRK_NB;
    for (int j = 0; j < FTS_LRK_MOMENTS; ++j) f[1 + j] = m[j];
}

// Backward pass: accumulates dL/dparams into G given dL/dz.
inline void lrk_backward(const double* P, const LrkActs& a, const double* dz, double* G)
{
    double d3[LRK_H], d2[LRK_H], d1[LRK_H];
    for (int c = 0; c < FTS_LRK_OUT; ++c) {
        G[FTS_LRK_OFF_B4 + c] += dz[c];
        for (int i = 0; i < LRK_H; ++i) G[FTS_LRK_OFF_W4 + c * LRK_H + i] += dz[c] * a.h3[i];
    }
    for (int i = 0; i < LRK_H; ++i) {
        double s = 0.0;
        for (int c = 0; c < FTS_LRK_OUT; ++c) s += P[FTS_LRK_OFF_W4 + c * LRK_H + i] * dz[c];
        d3[i] = s * (1.0 - a.h3[i] * a.h3[i]);
    }
    auto back = [&](int offW, int offB, const double* dout, const double* in, int nin, double* din, const double* hin) {
        for (int o = 0; o < LRK_H; ++o) {
            G[offB + o] += dout[o];
            for (int i = 0; i < nin; ++i) G[offW + o * nin + i] += dout[o] * in[i];
        }
        if (!din) return;
        for (int i = 0; i < nin; ++i) {
            double s = 0.0;
            for (int o = 0; o < LRK_H; ++o) s += P[offW + o * nin + i] * dout[o];
            din[i] = s * (1.0 - hin[i] * hin[i]);
        }
    };
    back(FTS_LRK_OFF_W3, FTS_LRK_OFF_B3, d3, a.h2, LRK_H, d2, a.h2);
    back(FTS_LRK_OFF_W2, FTS_LRK_OFF_B2, d2, a.h1, LRK_H, d1, a.h1);
    back(FTS_LRK_OFF_W1, FTS_LRK_OFF_B1, d1, a.x, LRK_NIN, nullptr, nullptr);
}

// Mean loss over n records and (if G) its gradient; shift and scale get no gradient.
inline double lrk_loss_grad(const double* P, const double* records, int n, double* G)
{
    if (G) std::memset(G, 0, sizeof(double) * FTS_LRK_NPARAMS);
    double loss = 0.0;
    const double inv_n = n > 0 ? 1.0 / n : 0.0;
    for (int k = 0; k < n; ++k) {
        const double* rec = records + static_cast<size_t>(k) * FTS_LRK_RECORD;
        for (int b = 0; b < LRK_NB; ++b) {
            double f[FTS_LRK_FEATURES];
            lrk_features(b, rec + FTS_LRK_REC_MOMENTS, f);
            LrkActs a;
            lrk_forward(P, f, a);
            double dz[FTS_LRK_OUT] = {0.0, 0.0};
            bool any = false;
            for (int c = 0; c < FTS_LRK_OUT; ++c) {
                const double E = rec[FTS_LRK_REC_E + c];
                const double S1 = rec[FTS_LRK_REC_S1 + c * LRK_NB + b];
                const double S2 = rec[FTS_LRK_REC_S2 + c * LRK_NB + b];
                if (!(E > 0.0) || !(S2 > 0.0)) continue;
                const double w = S2 / E, y = S1 / S2, er = std::exp(a.r[c]);
                loss += inv_n * w * (er - y) * (er - y);
                const double dr = inv_n * 2.0 * w * (er - y) * er;
                dz[c] = dr * (1.0 - 0.0625 * a.r[c] * a.r[c]);
                any = true;
            }
            if (G && any) lrk_backward(P, a, dz, G);
        }
    }
    return loss;
}

// Exact bin of a wave vector with integer spectral indices (nx, ny, nz) on an
// m1 x m2 x m3 grid of one spacing: |k| / k_ref = (2 / sqrt 3) sqrt(sum (n_d / m_d)^2)
// with k_ref = sqrt(3) pi / dL, so bin = floor(B |k| / k_ref) is the largest b with
// 3 b^2 m1^2 m2^2 m3^2 <= 4 B^2 (nx^2 m2^2 m3^2 + ny^2 m1^2 m3^2 + nz^2 m1^2 m2^2),
// clamped to B - 1.  Integer arithmetic: no rounding can move a k across a bin edge.
inline int lrk_bin(long long nx, long long ny, long long nz, long long m1, long long m2, long long m3)
{
    using u128 = unsigned __int128;
    const u128 a1 = static_cast<u128>(m1 * m1), a2 = static_cast<u128>(m2 * m2), a3 = static_cast<u128>(m3 * m3);
    const u128 rhs = static_cast<u128>(4) * LRK_NB * LRK_NB *
                     (static_cast<u128>(nx * nx) * a2 * a3 + static_cast<u128>(ny * ny) * a1 * a3 + static_cast<u128>(nz * nz) * a1 * a2);
    const u128 den = static_cast<u128>(3) * a1 * a2 * a3;
    int b = 0;
    while (b + 1 < LRK_NB && den * static_cast<u128>(b + 1) * static_cast<u128>(b + 1) <= rhs) ++b;
    return b;
}

} // namespace oracle
