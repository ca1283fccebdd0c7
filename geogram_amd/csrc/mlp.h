// Warm start of the optimal-transport Newton solver: a small MLP
// 6 -> 64 -> 64 -> 1 (leaky ReLU, FP32) mapping a seed's features to its
// Laguerre weight.  No counterpart in geogram.
//
// One wavefront evaluates rows one after another with lane j owning hidden
// unit j: its column of W1 and W2, and (backward) its row of W2, stay in
// registers for the whole tile; a layer is 64 readlane broadcasts and FMAs,
// the output a butterfly wave sum.  The backward pass accumulates the wave's
// gradient in registers and writes it as one partial per wave; partials are
// added in a fixed order, so training is bit-reproducible (and the CPU
// emulation reproduces it).
//
// Parameter layout (kMlpParams floats): W1[6][64], b1[64], W2[64][64] (row k =
// input unit k), b2[64], W3[64], b3.
#pragma once

#include "wave.h"

namespace vgc {

constexpr int kMlpIn = 6;
constexpr int kMlpHid = 64;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpIn * kMlpHid;
constexpr int kOffW2 = kOffB1 + kMlpHid;
constexpr int kOffB2 = kOffW2 + kMlpHid * kMlpHid;
constexpr int kOffW3 = kOffB2 + kMlpHid;
constexpr int kOffB3 = kOffW3 + kMlpHid;
constexpr int kMlpParams = kOffB3 + 1;  // 4673
constexpr float kLeak = 0.01f;

VGC_HD LF leaky(const LF& a) { return select(a > 0.0f, a, a * kLeak); }
VGC_HD LF leaky_d(const LF& a) { return select(a > 0.0f, LF(1.0f), LF(kLeak)); }

#if defined(VGC_EMU)
inline LF vfmaf(const LF& a, const LF& b, const LF& c) {
    LF r;
    for (int i = 0; i < kWave; ++i) r[i] = std::fma(a[i], b[i], c[i]);
    return r;
}
inline float wave_sumf(const LF& a) {
    LF x = a;
    for (int m = 32; m >= 1; m >>= 1) x = x + shfl_xor(x, m);
    return x[0];
}
#else
VGC_HD float vfmaf(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
VGC_HD float wave_sumf(float a) {
    float x = a;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) x = x + __shfl_xor(x, m);
    return x;
}
#endif

// Rows [r0, r1) of X (6 features each).  Forward: out[r] = MLP(X[r]).
// Backward (T != nullptr): adds to the wave's gradient partial G (kMlpParams
// floats, written by this call) the gradient of sum_r (y_r - T_r)^2 * scale,
// and returns the wave's loss sum (lane 0).
template <bool BACKWARD>
VGC_HD float mlp_rows(const float* __restrict__ Pm, const float* __restrict__ X,
                      const float* __restrict__ T, int r0, int r1, float scale,
                      float* __restrict__ out, float* __restrict__ G) {
    const LI lane = lane_id();
    LF w1[kMlpIn], w2c[kMlpHid], w2r[kMlpHid];
    for (int k = 0; k < kMlpIn; ++k) w1[k] = lds_ld(Pm + kOffW1 + k * kMlpHid, lane);
    const LF b1 = lds_ld(Pm + kOffB1, lane);
    for (int k = 0; k < kMlpHid; ++k) w2c[k] = lds_ld(Pm + kOffW2 + k * kMlpHid, lane);
    if (BACKWARD)
        for (int k = 0; k < kMlpHid; ++k) w2r[k] = lds_ld(Pm + kOffW2 + k, lane * kMlpHid);
    const LF b2 = lds_ld(Pm + kOffB2, lane);
    const LF w3 = lds_ld(Pm + kOffW3, lane);
    const float b3 = Pm[kOffB3];
    LF gw1[kMlpIn], gw2[kMlpHid], gb1 = LF(0.0f), gb2 = LF(0.0f), gw3 = LF(0.0f);
    float gb3 = 0.0f, loss = 0.0f;
    if (BACKWARD) {
        for (int k = 0; k < kMlpIn; ++k) gw1[k] = LF(0.0f);
        for (int k = 0; k < kMlpHid; ++k) gw2[k] = LF(0.0f);
    }
    for (int r = r0; r < r1; ++r) {
        float x[kMlpIn];
        for (int k = 0; k < kMlpIn; ++k) x[k] = X[(long long)r * kMlpIn + k];
        LF a1 = b1;
        for (int k = 0; k < kMlpIn; ++k) a1 = vfmaf(LF(x[k]), w1[k], a1);
        const LF h1 = leaky(a1);
        LF a2 = b2;
        for (int k = 0; k < kMlpHid; ++k) a2 = vfmaf(LF(readlane(h1, k)), w2c[k], a2);
        const LF h2 = leaky(a2);
        const float y = wave_sumf(h2 * w3) + b3;
        if (!BACKWARD) {
            if (lane0()) out[r] = y;
            continue;
        }
        const float e = y - T[r];
        loss += e * e;
        const float dy = 2.0f * e * scale;
        gw3 = vfmaf(LF(dy), h2, gw3);
        gb3 += dy;
        const LF dh2 = (LF(dy) * w3) * leaky_d(a2);
        for (int k = 0; k < kMlpHid; ++k) gw2[k] = vfmaf(LF(readlane(h1, k)), dh2, gw2[k]);
        gb2 = gb2 + dh2;
        LF s = LF(0.0f);
        for (int k = 0; k < kMlpHid; ++k) s = vfmaf(w2r[k], LF(readlane(dh2, k)), s);
        const LF dh1 = s * leaky_d(a1);
        for (int k = 0; k < kMlpIn; ++k) gw1[k] = vfmaf(LF(x[k]), dh1, gw1[k]);
        gb1 = gb1 + dh1;
    }
    if (BACKWARD) {
        for (int k = 0; k < kMlpIn; ++k) lds_st(G + kOffW1 + k * kMlpHid, lane, gw1[k], LB(true));
        lds_st(G + kOffB1, lane, gb1, LB(true));
        for (int k = 0; k < kMlpHid; ++k) lds_st(G + kOffW2 + k * kMlpHid, lane, gw2[k], LB(true));
        lds_st(G + kOffB2, lane, gb2, LB(true));
        lds_st(G + kOffW3, lane, gw3, LB(true));
        if (lane0()) G[kOffB3] = gb3;
    }
    return loss;
}

// Adam step on element i (bias corrections c1 = 1 - beta1^t, c2 = 1 - beta2^t)
VGC_SD void adam_elem(int i, float* p, const float* g, float* m, float* v, float lr, float b1,
                      float b2, float c1, float c2, float eps) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Features of seed i: position in the box, log relative target mass, log of
// the mean and max distance to its 16 nearest neighbours over the mean
// spacing h = (V / n)^(1/3).
VGC_SD void ot_features_row(int i, const double* xyz, const double* nu, const float* d2knn,
                            const double* box, double inv_cell, double h, float* f) {
    f[0] = float((xyz[3 * (long long)i] - box[0]) / (box[3] - box[0]));
    f[1] = float((xyz[3 * (long long)i + 1] - box[1]) / (box[4] - box[1]));
    f[2] = float((xyz[3 * (long long)i + 2] - box[2]) / (box[5] - box[2]));
    f[3] = float(log(nu[i] * inv_cell));
    double s = 0.0, mx = 0.0;
    for (int k = 0; k < 16; ++k) {
        double d = sqrt(double(d2knn[(long long)i * 16 + k]));
        s += d;
        mx = d > mx ? d : mx;
    }
    f[4] = float(log(s / 16.0 / h));
    f[5] = float(log(mx / h));
}

}  // namespace vgc
