// Learned MAC: features of a cluster pair and the network's forward pass,
// shared by the traversal (decisions), the sample generator and training.
#pragma once

#include "btk_device.hpp"

namespace btk {

constexpr int kMacIn = BTK_MAC_FEATURES;
constexpr int kMacHid = 64;
constexpr int kMacOut = BTK_MAC_OUTPUTS;
// parameter layout: W1[64][10], b1[64], W2[64][64], b2[64], W3[12][64], b3[12]
constexpr int kMacW1 = 0;
constexpr int kMacB1 = kMacW1 + kMacHid * kMacIn;
constexpr int kMacW2 = kMacB1 + kMacHid;
constexpr int kMacB2 = kMacW2 + kMacHid * kMacHid;
constexpr int kMacW3 = kMacB2 + kMacHid;
constexpr int kMacB3 = kMacW3 + kMacOut * kMacHid;
constexpr int kMacParams = kMacB3 + kMacOut;
static_assert(kMacParams == BTK_MAC_WEIGHTS, "parameter count");

// Output scale: network output = log10(relative error) / 4.
constexpr float kMacScale = 4.0f;
// Separation ratios (r_t + r_s)/d of the pairs the learned MAC decides: the
// range the training samples cover.
constexpr float kMacSepMin = 0.3f;
constexpr float kMacSepMax = 1.2f;

// Panels as the MAC sees them.
struct MacPanels {
    const double *lo0, *lo1, *lo2, *hi0, *hi1, *hi2, *radius;
    const int *count, *level;
};

// The pair's 10 features; false when the boxes overlap (one contains the
// other, or they share volume) or the centres coincide: such pairs are never
// approximated by the learned MAC.  Centre and distance expressions are the
// traversal's (interaction_list.cpp:141-154); g is the distance between the
// two boxes (0 when they touch).
__device__ __forceinline__ bool mac_features(int it, int is, const MacPanels &T, const MacPanels &S, int thr,
                                             float *f) {
    const bool overlap = T.lo0[it] < S.hi0[is] && S.lo0[is] < T.hi0[it] && T.lo1[it] < S.hi1[is] &&
                         S.lo1[is] < T.hi1[it] && T.lo2[it] < S.hi2[is] && S.lo2[is] < T.hi2[it];
    if (overlap) return false;
    const double ex = 0.5 * (S.lo0[is] + S.hi0[is]) - 0.5 * (T.lo0[it] + T.hi0[it]);
    const double ey = 0.5 * (S.lo1[is] + S.hi1[is]) - 0.5 * (T.lo1[it] + T.hi1[it]);
    const double ez = 0.5 * (S.lo2[is] + S.hi2[is]) - 0.5 * (T.lo2[it] + T.hi2[it]);
    const double dist = sqrt(ex * ex + ey * ey + ez * ez);
    if (!(dist > 0)) return false;
    const double gx = fmax(0.0, fmax(S.lo0[is] - T.hi0[it], T.lo0[it] - S.hi0[is]));
    const double gy = fmax(0.0, fmax(S.lo1[is] - T.hi1[it], T.lo1[it] - S.hi1[is]));
    const double gz = fmax(0.0, fmax(S.lo2[is] - T.hi2[it], T.lo2[it] - S.hi2[is]));
    const double gap = sqrt(gx * gx + gy * gy + gz * gz);
    const double rt = T.radius[it], rs = S.radius[is];
    const int nt = T.count[it], ns = S.count[is];
    (void)thr;
    f[0] = (float)fmin((rt + rs) / dist, 4.0);
    f[1] = (float)fmin(rt / dist, 4.0);
    f[2] = (float)fmin(rs / dist, 4.0);
    f[3] = log2f((float)nt + 1.0f) / 16.0f;
    f[4] = log2f((float)ns + 1.0f) / 16.0f;
    f[5] = (float)T.level[it] / 16.0f;
    f[6] = (float)S.level[is] / 16.0f;
    f[7] = (float)fmax(-2.0, fmin(2.0, log2(fmax(rt, 1e-300) / fmax(rs, 1e-300)) / 4.0));
    f[8] = (float)(gap / dist);
    f[9] = (float)(gap > 0 ? fmin((rt + rs) / gap, 8.0) / 8.0 : 1.0);
    return true;
}

// Outputs o0 .. o0+NO-1 of the network (ReLU hidden layers) for features f;
// w may be in global memory or LDS.
template <int NO>
__device__ __forceinline__ void mac_outputs(const float *w, const float *f, int o0, float *y) {
    float h1[kMacHid];  // fully unrolled loops keep it in registers
#pragma unroll
    for (int i = 0; i < kMacHid; i++) {
        float a = w[kMacB1 + i];
#pragma unroll
        for (int j = 0; j < kMacIn; j++) a = fmaf(w[kMacW1 + i * kMacIn + j], f[j], a);
        h1[i] = fmaxf(a, 0.0f);
    }
#pragma unroll
    for (int k = 0; k < NO; k++) y[k] = w[kMacB3 + o0 + k];
    for (int i = 0; i < kMacHid; i++) {
        float a = w[kMacB2 + i];
#pragma unroll
        for (int j = 0; j < kMacHid; j++) a = fmaf(w[kMacW2 + i * kMacHid + j], h1[j], a);
        a = fmaxf(a, 0.0f);
#pragma unroll
        for (int k = 0; k < NO; k++) y[k] = fmaf(w[kMacW3 + (o0 + k) * kMacHid + i], a, y[k]);
    }
}

// What the traversal needs to consult the network.
struct MacView {
    const float *w = nullptr;  // null: the theta MAC alone
    float log_budget = 0.0f;   // log10(budget) / 4, compared with the output
    int di = 0;                // degree index of the configured degree (4, 6, 8, 10)
    float cost[4] = {1.0f, 1.0f, 1.0f, 1.0f};  // relative cost of one PP, PC, CP, CC evaluation
    MacPanels T{}, S{};
};

// Proxy points per cluster at degree index di (degrees 4, 6, 8, 10).
__device__ __forceinline__ int mac_proxies(int di) {
    const int n1 = 5 + 2 * di;
    return n1 * n1 * n1;
}

// Predicted errors (network outputs) of the pair's PC, CP and CC
// approximations at every degree (output 4 * (type - 1) + degree index);
// false outside the trained domain: separated boxes (a positive gap, so no
// proxy point of one box meets one of the other) with separation ratio in
// [kMacSepMin, kMacSepMax).
__device__ __forceinline__ bool mac_predict(int it, int is, const MacView &m, const float *w, float *y) {
    float f[kMacIn];
    if (!mac_features(it, is, m.T, m.S, 0, f)) return false;
    if (!(f[8] > 0.0f) || f[0] < kMacSepMin || f[0] >= kMacSepMax) return false;
    mac_outputs<kMacOut>(w, f, 0, y);
    return true;
}

// The learned decision for a pair, as a code type | degree index << 2, or -1
// to leave the pair to the theta MAC.  type is the theta MAC's count-given
// type.  The candidates are the theta MAC's own choice when it accepts (type
// at the configured degree m.di, or PP), and every approximation at a degree
// up to m.di whose predicted error is within the budget: CC for any two
// panels, PC and CP where the panel on the point side holds at most thr
// points (a leaf, as in the reference's types); the pair takes the candidate
// of least cost, its kernel evaluations weighted by the type's cost per
// evaluation (m.cost; equal weights count evaluations).  A pair the theta
// MAC would refine is accepted only by an approximation cheaper than its
// direct sum.
__device__ __forceinline__ int mac_decide(int it, int is, const MacView &m, const float *w, int thr, int type,
                                          bool accepted) {
    const int ct = m.T.count[it], cs = m.S.count[is];
    const double nt = ct, ns = cs;
    const double direct = nt * ns * m.cost[0];
    auto cost = [&](int t, int d) {
        const double P = mac_proxies(d);
        return (t == 1 ? nt * P : (t == 2 ? P * ns : P * P)) * m.cost[t];
    };
    double best = direct;
    int code = -1;
    if (accepted) {
        code = type == 0 ? 0 : (type | (m.di << 2));
        best = type == 0 ? direct : cost(type, m.di);
    }
    const bool pc_ok = ct <= thr, cp_ok = cs <= thr;
    // the cheapest approximation must beat the best so far
    double low = cost(3, 0);
    if (pc_ok) low = fmin(low, cost(1, 0));
    if (cp_ok) low = fmin(low, cost(2, 0));
    if (!(low < best)) return code;
    float y[kMacOut];
    if (!mac_predict(it, is, m, w, y)) return code;
#pragma unroll
    for (int t = 1; t <= 3; t++) {
        if ((t == 1 && !pc_ok) || (t == 2 && !cp_ok)) continue;
        for (int d = 0; d <= m.di; d++) {
            const double c = cost(t, d);
            if (y[(t - 1) * 4 + d] <= m.log_budget && c < best) {
                best = c;
                code = t | (d << 2);
            }
        }
    }
    return code;
}

// The installed model as the traversal of (tt, ts) uses it; throws when the
// learned MAC cannot serve this configuration.
MacView mac_view(const btk_run_config &cfg, const Tree &tt, const Tree &ts);

}  // namespace btk
