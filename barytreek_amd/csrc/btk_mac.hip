// Learned MAC: training samples measured by the treecode itself, and the
// network's training (forward, backward and Adam kernels written here).
//
// A sample is a cluster pair (T, S) that the traversal visits.  Its targets
// are 64 points of T and its sources 1024 points of S (evenly spaced in the
// tree order, all of them when fewer), with charges U[-1,1] from a hash of the
// point index.  For degrees 4, 6, 8 and 10 the kernel forms the PC
// (targets from S's proxy weights), CP (S's points to T's proxies, then
// interpolation to the targets) and CC (S's proxy weights to T's proxies,
// then interpolation) approximations of the 1/r potential and its gradient
// kernel d/r^3 exactly as the treecode does (barycentric basis on the
// Chebyshev-Lobatto nodes, proxy points cos(pi/n*i), every component
// interpolated on its own), and their relative L2 error over the four
// components at the targets against the FP64 direct sum over the same
// points: 12 targets per sample, stored as log10 / 4 with log10 clamped to
// [-14, 0].
#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "btk_mac.hpp"

namespace btk {
namespace {

constexpr int TB = 256;
constexpr int kSampT = 64;    // targets per sample
constexpr int kSampS = 1024;  // sources per sample
inline int nblk(long long n) { return (int)((n + TB - 1) / TB); }

struct SampleTree {
    MacPanels p;
    const int *start;
    const double *sx, *sy, *sz;
};

struct Tables4 {
    NodeTables t[4];  // degrees 4, 6, 8, 10
};

// Charge of point i: U[-1,1) from a splitmix64 hash.
__device__ __forceinline__ double hash_charge(unsigned seed, int i) {
    unsigned long long z = ((unsigned long long)seed << 32) ^ (unsigned long long)(unsigned)i;
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

constexpr int kSampC = 4;  // potential and gradient components

// 1/r and d/r^3 (the potential+gradient kernel's components).
__device__ __forceinline__ void g_all(double dx, double dy, double dz, double *g) {
    const double inv = rsqrt_f64(dx * dx + dy * dy + dz * dz);
    const double h = inv * inv * inv;
    g[0] = inv;
    g[1] = h * dx;
    g[2] = h * dy;
    g[3] = h * dz;
}

// Shared memory of one sample (one workgroup).
struct SampleLds {
    double W[kMaxN1 * kMaxN1 * kMaxN1];            // S's proxy weights
    double Psi[kSampC][kMaxN1 * kMaxN1 * kMaxN1];  // T's proxy potentials, per component
    double LB[kSampT][3][kMaxN1];                  // basis of a chunk of 64 sources
    double LQ[kSampT];
    double LT[kSampT][3][kMaxN1];                  // basis of the targets
    double tx[kSampT], ty[kSampT], tz[kSampT], ex[kSampC][kSampT];
    double ap[3][kSampC][kSampT];                  // PC, CP, CC at the targets
    double red[4][kSampC][kSampT];
};

// Sampled point k of m out of a panel's cnt points.
__device__ __forceinline__ int pick(int start, int cnt, int m, int k) {
    return start + (int)(((long long)k * cnt) / m);
}

template <int N1>
__device__ void sample_degree(const SampleTree &tr, int T, int S, int nT, int nS, unsigned seed,
                              const NodeTables &tab, SampleLds &L) {
    constexpr int P = N1 * N1 * N1;
    const int tid = threadIdx.x;
    const int t = tid % kSampT, part = tid / kSampT;  // 4 parts per target
    const int cS = tr.p.count[S], cT = tr.p.count[T];
    const double sl[3] = {tr.p.lo0[S], tr.p.lo1[S], tr.p.lo2[S]}, sh[3] = {tr.p.hi0[S], tr.p.hi1[S], tr.p.hi2[S]};
    const double tl[3] = {tr.p.lo0[T], tr.p.lo1[T], tr.p.lo2[T]}, th[3] = {tr.p.hi0[T], tr.p.hi1[T], tr.p.hi2[T]};
    (void)cT;
    // S's proxy weights W_l = sum_s L_l(s) q_s (base_pots, upward_pass_impl.hpp:9-44)
    for (int l = tid; l < P; l += blockDim.x) L.W[l] = 0.0;
    for (int c0 = 0; c0 < nS; c0 += kSampT) {
        __syncthreads();
        if (tid < kSampT) {
            const int k = c0 + tid;
            if (k < nS) {
                const int i = pick(tr.start[S], cS, nS, k);
                bli_basis<N1>(tr.sx[i], sl[0], sh[0], tab.a, N1, L.LB[tid][0]);
                bli_basis<N1>(tr.sy[i], sl[1], sh[1], tab.a, N1, L.LB[tid][1]);
                bli_basis<N1>(tr.sz[i], sl[2], sh[2], tab.a, N1, L.LB[tid][2]);
                L.LQ[tid] = hash_charge(seed, i);
            } else {
                L.LQ[tid] = 0.0;
                for (int d = 0; d < 3; d++)
                    for (int b = 0; b < N1; b++) L.LB[tid][d][b] = 0.0;
            }
        }
        __syncthreads();
        const int nc = min(kSampT, nS - c0);
        for (int l = tid; l < P; l += blockDim.x) {
            const int m = l % N1, k = (l / N1) % N1, j = l / (N1 * N1);
            double a = 0.0;
            for (int s = 0; s < nc; s++) a = fma(L.LB[s][0][j] * L.LB[s][1][k] * L.LB[s][2][m], L.LQ[s], a);
            L.W[l] += a;
        }
    }
    // basis of the targets in T's box
    if (tid < nT) {
        bli_basis<N1>(L.tx[tid], tl[0], th[0], tab.a, N1, L.LT[tid][0]);
        bli_basis<N1>(L.ty[tid], tl[1], th[1], tab.a, N1, L.LT[tid][1]);
        bli_basis<N1>(L.tz[tid], tl[2], th[2], tab.a, N1, L.LT[tid][2]);
    }
    __syncthreads();
    // PC: targets from S's proxy points and weights
    {
        double a[kSampC] = {0.0, 0.0, 0.0, 0.0};
        if (t < nT) {
            for (int l = part; l < P; l += 4) {
                const int m = l % N1, k = (l / N1) % N1, j = l / (N1 * N1);
                double g[kSampC];
                g_all(L.tx[t] - proxy_node(tab.b[j], sl[0], sh[0]), L.ty[t] - proxy_node(tab.b[k], sl[1], sh[1]),
                      L.tz[t] - proxy_node(tab.b[m], sl[2], sh[2]), g);
                for (int c = 0; c < kSampC; c++) a[c] = fma(g[c], L.W[l], a[c]);
            }
        }
        for (int c = 0; c < kSampC; c++) L.red[part][c][t] = a[c];
        __syncthreads();
        if (tid < nT)
            for (int c = 0; c < kSampC; c++)
                L.ap[0][c][tid] = L.red[0][c][tid] + L.red[1][c][tid] + L.red[2][c][tid] + L.red[3][c][tid];
        __syncthreads();
    }
    // CP and CC: T's proxy values from S's points (CP) or proxies (CC), then
    // interpolation of each component to the targets (leaf_to_point)
    for (int kind = 1; kind <= 2; kind++) {
        for (int k = tid; k < P; k += blockDim.x) {
            const int m = k % N1, kk = (k / N1) % N1, j = k / (N1 * N1);
            const double px = proxy_node(tab.b[j], tl[0], th[0]);
            const double py = proxy_node(tab.b[kk], tl[1], th[1]);
            const double pz = proxy_node(tab.b[m], tl[2], th[2]);
            double a[kSampC] = {0.0, 0.0, 0.0, 0.0}, g[kSampC];
            if (kind == 1) {
                for (int s = 0; s < nS; s++) {
                    const int i = pick(tr.start[S], cS, nS, s);
                    g_all(px - tr.sx[i], py - tr.sy[i], pz - tr.sz[i], g);
                    const double q = hash_charge(seed, i);
                    for (int c = 0; c < kSampC; c++) a[c] = fma(g[c], q, a[c]);
                }
            } else {
                for (int l = 0; l < P; l++) {
                    const int lm = l % N1, lk = (l / N1) % N1, lj = l / (N1 * N1);
                    g_all(px - proxy_node(tab.b[lj], sl[0], sh[0]), py - proxy_node(tab.b[lk], sl[1], sh[1]),
                          pz - proxy_node(tab.b[lm], sl[2], sh[2]), g);
                    for (int c = 0; c < kSampC; c++) a[c] = fma(g[c], L.W[l], a[c]);
                }
            }
            for (int c = 0; c < kSampC; c++) L.Psi[c][k] = a[c];
        }
        __syncthreads();
        double a[kSampC] = {0.0, 0.0, 0.0, 0.0};
        if (t < nT) {
            for (int k = part; k < P; k += 4) {
                const int m = k % N1, kk = (k / N1) % N1, j = k / (N1 * N1);
                const double b = L.LT[t][0][j] * L.LT[t][1][kk] * L.LT[t][2][m];
                for (int c = 0; c < kSampC; c++) a[c] = fma(b, L.Psi[c][k], a[c]);
            }
        }
        for (int c = 0; c < kSampC; c++) L.red[part][c][t] = a[c];
        __syncthreads();
        if (tid < nT)
            for (int c = 0; c < kSampC; c++)
                L.ap[kind][c][tid] = L.red[0][c][tid] + L.red[1][c][tid] + L.red[2][c][tid] + L.red[3][c][tid];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(TB) k_mac_samples(int np, const int2 *__restrict__ pairs, SampleTree tr, int thr,
                                                    Tables4 tabs, unsigned seed, float *__restrict__ X,
                                                    float *__restrict__ Y) {
    __shared__ SampleLds L;
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    const int T = pairs[b].x, S = pairs[b].y;
    const int nT = min(kSampT, tr.p.count[T]), nS = min(kSampS, tr.p.count[S]);
    if (tid < nT) {
        const int i = pick(tr.start[T], tr.p.count[T], nT, tid);
        L.tx[tid] = tr.sx[i];
        L.ty[tid] = tr.sy[i];
        L.tz[tid] = tr.sz[i];
        double a[kSampC] = {0.0, 0.0, 0.0, 0.0}, g[kSampC];  // the direct sum over the sampled sources
        for (int s = 0; s < nS; s++) {
            const int k = pick(tr.start[S], tr.p.count[S], nS, s);
            g_all(L.tx[tid] - tr.sx[k], L.ty[tid] - tr.sy[k], L.tz[tid] - tr.sz[k], g);
            const double q = hash_charge(seed, k);
            for (int c = 0; c < kSampC; c++) a[c] = fma(g[c], q, a[c]);
        }
        for (int c = 0; c < kSampC; c++) L.ex[c][tid] = a[c];
    }
    if (tid == 0) {
        float f[kMacIn];
        mac_features(T, S, tr.p, tr.p, thr, f);
        for (int j = 0; j < kMacIn; j++) X[(size_t)b * kMacIn + j] = f[j];
    }
    __syncthreads();
    for (int di = 0; di < 4; di++) {
        switch (di) {
            case 0: sample_degree<5>(tr, T, S, nT, nS, seed, tabs.t[0], L); break;
            case 1: sample_degree<7>(tr, T, S, nT, nS, seed, tabs.t[1], L); break;
            case 2: sample_degree<9>(tr, T, S, nT, nS, seed, tabs.t[2], L); break;
            default: sample_degree<11>(tr, T, S, nT, nS, seed, tabs.t[3], L); break;
        }
        if (tid < 3) {
            double num = 0.0, den = 0.0;
            for (int c = 0; c < kSampC; c++)
                for (int t = 0; t < nT; t++) {
                    const double d = L.ap[tid][c][t] - L.ex[c][t];
                    num += d * d;
                    den += L.ex[c][t] * L.ex[c][t];
                }
            // log10 of the error, clamped to [-14, 0]: only errors near a
            // budget matter, and touching boxes give CC errors of 1e14
            const double e = den > 0 ? sqrt(num / den) : 1.0;
            Y[(size_t)b * kMacOut + tid * 4 + di] = (float)(fmin(0.0, log10(fmax(e, 1e-14))) / kMacScale);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Training: one thread per sample for the forward pass and the deltas, one
// thread per parameter for the weight gradients (sums over the batch).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(TB) k_mac_forward(int n, const float *__restrict__ X, const float *__restrict__ w,
                                                    float *__restrict__ H1, float *__restrict__ H2,
                                                    float *__restrict__ O) {
    __shared__ float sw[kMacParams];
    for (int i = threadIdx.x; i < kMacParams; i += TB) sw[i] = w[i];
    __syncthreads();
    const int b = blockIdx.x * TB + threadIdx.x;
    if (b >= n) return;
    float x[kMacIn], h1[kMacHid];
    for (int j = 0; j < kMacIn; j++) x[j] = X[(size_t)b * kMacIn + j];
    for (int i = 0; i < kMacHid; i++) {
        float a = sw[kMacB1 + i];
        for (int j = 0; j < kMacIn; j++) a = fmaf(sw[kMacW1 + i * kMacIn + j], x[j], a);
        h1[i] = fmaxf(a, 0.0f);
        if (H1) H1[(size_t)b * kMacHid + i] = h1[i];
    }
    float y[kMacOut];
    for (int o = 0; o < kMacOut; o++) y[o] = sw[kMacB3 + o];
    for (int i = 0; i < kMacHid; i++) {
        float a = sw[kMacB2 + i];
        for (int j = 0; j < kMacHid; j++) a = fmaf(sw[kMacW2 + i * kMacHid + j], h1[j], a);
        const float h2 = fmaxf(a, 0.0f);
        if (H2) H2[(size_t)b * kMacHid + i] = h2;
        for (int o = 0; o < kMacOut; o++) y[o] = fmaf(sw[kMacW3 + o * kMacHid + i], h2, y[o]);
    }
    for (int o = 0; o < kMacOut; o++) O[(size_t)b * kMacOut + o] = y[o];
}

// Loss L = mean over the batch and the 12 outputs of (O - Y)^2; deltas
// D3 = dL/dO, D2 = dL/d(pre-activation 2), D1 = dL/d(pre-activation 1).
__global__ void __launch_bounds__(TB) k_mac_delta(int n, const float *__restrict__ Y, const float *__restrict__ w,
                                                  const float *__restrict__ H1, const float *__restrict__ H2,
                                                  const float *__restrict__ O, float *__restrict__ D1,
                                                  float *__restrict__ D2, float *__restrict__ D3,
                                                  double *__restrict__ loss) {
    __shared__ float sw[kMacParams];
    __shared__ double sl[TB];
    for (int i = threadIdx.x; i < kMacParams; i += TB) sw[i] = w[i];
    __syncthreads();
    const int b = blockIdx.x * TB + threadIdx.x;
    double l = 0.0;
    if (b < n) {
        const float scale = 2.0f / ((float)n * kMacOut);
        float d3[kMacOut];
        for (int o = 0; o < kMacOut; o++) {
            const float r = O[(size_t)b * kMacOut + o] - Y[(size_t)b * kMacOut + o];
            l += (double)r * r;
            d3[o] = scale * r;
            D3[(size_t)b * kMacOut + o] = d3[o];
        }
        float d2[kMacHid];
        for (int i = 0; i < kMacHid; i++) {
            float a = 0.0f;
            for (int o = 0; o < kMacOut; o++) a = fmaf(sw[kMacW3 + o * kMacHid + i], d3[o], a);
            d2[i] = H2[(size_t)b * kMacHid + i] > 0.0f ? a : 0.0f;
            D2[(size_t)b * kMacHid + i] = d2[i];
        }
        for (int j = 0; j < kMacHid; j++) {
            float a = 0.0f;
            for (int i = 0; i < kMacHid; i++) a = fmaf(sw[kMacW2 + i * kMacHid + j], d2[i], a);
            D1[(size_t)b * kMacHid + j] = H1[(size_t)b * kMacHid + j] > 0.0f ? a : 0.0f;
        }
    }
    sl[threadIdx.x] = l;
    __syncthreads();
    for (int st = TB / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) sl[threadIdx.x] += sl[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, sl[0] / ((double)n * kMacOut));
}

// Gradient of parameter g: the batch sum of delta x input of its layer.
__global__ void __launch_bounds__(TB) k_mac_wgrad(int n, const float *__restrict__ X, const float *__restrict__ H1,
                                                  const float *__restrict__ H2, const float *__restrict__ D1,
                                                  const float *__restrict__ D2, const float *__restrict__ D3,
                                                  float *__restrict__ G) {
    const int g = blockIdx.x * TB + threadIdx.x;
    if (g >= kMacParams) return;
    const float *d, *in;
    int dw, iw, di, ii;
    if (g < kMacB1) {
        d = D1, dw = kMacHid, di = (g - kMacW1) / kMacIn, in = X, iw = kMacIn, ii = (g - kMacW1) % kMacIn;
    } else if (g < kMacW2) {
        d = D1, dw = kMacHid, di = g - kMacB1, in = nullptr, iw = 0, ii = 0;
    } else if (g < kMacB2) {
        d = D2, dw = kMacHid, di = (g - kMacW2) / kMacHid, in = H1, iw = kMacHid, ii = (g - kMacW2) % kMacHid;
    } else if (g < kMacW3) {
        d = D2, dw = kMacHid, di = g - kMacB2, in = nullptr, iw = 0, ii = 0;
    } else if (g < kMacB3) {
        d = D3, dw = kMacOut, di = (g - kMacW3) / kMacHid, in = H2, iw = kMacHid, ii = (g - kMacW3) % kMacHid;
    } else {
        d = D3, dw = kMacOut, di = g - kMacB3, in = nullptr, iw = 0, ii = 0;
    }
    float a = 0.0f;
    for (int b = 0; b < n; b++) a = fmaf(d[(size_t)b * dw + di], in ? in[(size_t)b * iw + ii] : 1.0f, a);
    G[g] = a;
}

__global__ void k_mac_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                           float *__restrict__ v, float lr, float b1, float b2, float c1, float c2) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= kMacParams) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

__global__ void k_mac_gather(int nb, const int *__restrict__ idx, const float *__restrict__ X,
                             const float *__restrict__ Y, float *__restrict__ Xb, float *__restrict__ Yb) {
    const int b = blockIdx.x * TB + threadIdx.x;
    if (b >= nb) return;
    const int i = idx[b];
    for (int j = 0; j < kMacIn; j++) Xb[(size_t)b * kMacIn + j] = X[(size_t)i * kMacIn + j];
    for (int o = 0; o < kMacOut; o++) Yb[(size_t)b * kMacOut + o] = Y[(size_t)i * kMacOut + o];
}

// Buffers of one batch.
struct MacBatch {
    DBuf<float> H1, H2, O, D1, D2, D3;
    DBuf<double> loss;
    explicit MacBatch(int n)
        : H1((size_t)n * kMacHid), H2((size_t)n * kMacHid), O((size_t)n * kMacOut), D1((size_t)n * kMacHid),
          D2((size_t)n * kMacHid), D3((size_t)n * kMacOut), loss(1) {}
};

// Loss and gradient of the weights w on (X, Y), n samples.
void batch_gradient(const float *w, const float *X, const float *Y, int n, MacBatch &B, float *G, hipStream_t s) {
    k_mac_forward<<<nblk(n), TB, 0, s>>>(n, X, w, B.H1.p, B.H2.p, B.O.p);
    BTK_LAUNCH_CHECK();
    BTK_HIP(hipMemsetAsync(B.loss.p, 0, sizeof(double), s));
    k_mac_delta<<<nblk(n), TB, 0, s>>>(n, Y, w, B.H1.p, B.H2.p, B.O.p, B.D1.p, B.D2.p, B.D3.p, B.loss.p);
    BTK_LAUNCH_CHECK();
    k_mac_wgrad<<<nblk(kMacParams), TB, 0, s>>>(n, X, B.H1.p, B.H2.p, B.D1.p, B.D2.p, B.D3.p, G);
    BTK_LAUNCH_CHECK();
}

std::vector<float> he_init(unsigned seed) {
    std::mt19937 rng(seed);
    std::vector<float> w(kMacParams, 0.0f);
    std::normal_distribution<float> n1(0.0f, std::sqrt(2.0f / kMacIn)), n2(0.0f, std::sqrt(2.0f / kMacHid)),
        n3(0.0f, std::sqrt(1.0f / kMacHid));
    for (int i = 0; i < kMacHid * kMacIn; i++) w[kMacW1 + i] = n1(rng);
    for (int i = 0; i < kMacHid * kMacHid; i++) w[kMacW2 + i] = n2(rng);
    for (int i = 0; i < kMacOut * kMacHid; i++) w[kMacW3 + i] = n3(rng);
    return w;
}

}  // namespace

// Created once and never destroyed: a static-storage DBuf would be freed at
// exit after the thread-local bookkeeping of the block cache is gone.
MacModel &mac_model() {
    static MacModel *m = new MacModel;
    return *m;
}

void mac_set_weights(const float *h, hipStream_t s) {
    MacModel &m = mac_model();
    if (!m.w.p) m.w.resize(kMacParams);
    BTK_HIP(hipMemcpyAsync(m.w.p, h, sizeof(float) * kMacParams, hipMemcpyHostToDevice, s));
    BTK_HIP(hipStreamSynchronize(s));
    m.ready = true;
}

void mac_get_weights(float *h, hipStream_t s) {
    MacModel &m = mac_model();
    BTK_REQUIRE(m.ready, "no learned MAC model is installed");
    BTK_HIP(hipMemcpyAsync(h, m.w.p, sizeof(float) * kMacParams, hipMemcpyDeviceToHost, s));
    BTK_HIP(hipStreamSynchronize(s));
}

MacView mac_view(const btk_run_config &cfg, const Tree &tt, const Tree &ts) {
    MacModel &m = mac_model();
    BTK_REQUIRE(m.ready, "the learned MAC is selected but no model is installed (btk_mac_train / set_weights)");
    BTK_REQUIRE(tt.dim == 3, "the learned MAC is 3D only");
    const int n = cfg.interp_degree;
    BTK_REQUIRE(n == 4 || n == 6 || n == 8 || n == 10, "the learned MAC needs interp_degree 4, 6, 8 or 10");
    MacView v;
    v.w = m.w.p;
    v.log_budget = (float)(std::log10(m.budget) / kMacScale);
    v.di = n / 2 - 2;
    for (int k = 0; k < 4; k++) v.cost[k] = m.cost[k];
    v.T = MacPanels{tt.lo[0].p, tt.lo[1].p, tt.lo[2].p, tt.hi[0].p, tt.hi[1].p, tt.hi[2].p, tt.radius.p,
                    tt.count.p, tt.level.p};
    v.S = MacPanels{ts.lo[0].p, ts.lo[1].p, ts.lo[2].p, ts.hi[0].p, ts.hi[1].p, ts.hi[2].p, ts.radius.p,
                    ts.count.p, ts.level.p};
    return v;
}

int mac_sample(const btk_run_config &cfg_in, int n, const double *x, const double *y, const double *z,
               int n_samples, unsigned seed, float *X, float *Y, int *pairs, hipStream_t s) {
    BTK_REQUIRE(cfg_in.dim == 3, "learned MAC samples are 3D");
    BTK_REQUIRE(n_samples >= 0, "n_samples must be >= 0");
    btk_run_config cfg = cfg_in;
    auto t = build_tree(3, cfg.fmm_cluster_thresh, x, y, z, n, s);
    std::vector<int2> visited;
    auto l = build_list(&cfg, *t, *t, s, false, &visited);
    std::vector<double> lo[3], hi[3], rad;
    std::vector<int> cnt(t->pc);
    for (int d = 0; d < 3; d++) {
        lo[d].resize(t->pc);
        hi[d].resize(t->pc);
        BTK_HIP(hipMemcpy(lo[d].data(), t->lo[d].p, sizeof(double) * t->pc, hipMemcpyDeviceToHost));
        BTK_HIP(hipMemcpy(hi[d].data(), t->hi[d].p, sizeof(double) * t->pc, hipMemcpyDeviceToHost));
    }
    rad.resize(t->pc);
    BTK_HIP(hipMemcpy(rad.data(), t->radius.p, sizeof(double) * t->pc, hipMemcpyDeviceToHost));
    BTK_HIP(hipMemcpy(cnt.data(), t->count.p, sizeof(int) * t->pc, hipMemcpyDeviceToHost));
    // candidate pairs: the learned MAC's domain, a positive gap between the
    // boxes and separation ratio in [kMacSepMin, kMacSepMax)
    std::vector<int2> cand;
    for (const int2 &p : visited) {
        const int a = p.x, b = p.y;
        if (cnt[a] == 0 || cnt[b] == 0) continue;
        double e2 = 0.0, g2 = 0.0;
        for (int d = 0; d < 3; d++) {
            const double e = 0.5 * (lo[d][b] + hi[d][b]) - 0.5 * (lo[d][a] + hi[d][a]);
            e2 += e * e;
            const double g = std::max(0.0, std::max(lo[d][b] - hi[d][a], lo[d][a] - hi[d][b]));
            g2 += g * g;
        }
        if (!(g2 > 0) || !(e2 > 0)) continue;
        const double sep = (rad[a] + rad[b]) / std::sqrt(e2);
        if (sep >= kMacSepMin && sep < kMacSepMax) cand.push_back(p);
    }
    std::mt19937 rng(seed);
    std::shuffle(cand.begin(), cand.end(), rng);
    const int m = (int)std::min<size_t>(cand.size(), (size_t)n_samples);
    if (m == 0) return 0;
    DBuf<int2> dp(m);
    BTK_HIP(hipMemcpy(dp.p, cand.data(), sizeof(int2) * m, hipMemcpyHostToDevice));
    if (pairs) BTK_HIP(hipMemcpyAsync(pairs, dp.p, sizeof(int2) * m, hipMemcpyDeviceToDevice, s));
    SampleTree tr{MacPanels{t->lo[0].p, t->lo[1].p, t->lo[2].p, t->hi[0].p, t->hi[1].p, t->hi[2].p, t->radius.p,
                            t->count.p, t->level.p},
                  t->start.p, t->sx.p, t->sy.p, t->sz.p};
    Tables4 tabs{{node_tables(4), node_tables(6), node_tables(8), node_tables(10)}};
    k_mac_samples<<<m, TB, 0, s>>>(m, dp.p, tr, cfg.fmm_cluster_thresh, tabs, seed, X, Y);
    BTK_LAUNCH_CHECK();
    BTK_HIP(hipStreamSynchronize(s));
    (void)l;
    return m;
}

void mac_forward(const float *X, int n, float *out, hipStream_t s) {
    MacModel &m = mac_model();
    BTK_REQUIRE(m.ready, "no learned MAC model is installed");
    if (n <= 0) return;
    k_mac_forward<<<nblk(n), TB, 0, s>>>(n, X, m.w.p, nullptr, nullptr, out);
    BTK_LAUNCH_CHECK();
}

float mac_gradients(const float *X, const float *Y, int n, float *grad, hipStream_t s) {
    MacModel &m = mac_model();
    BTK_REQUIRE(m.ready, "no learned MAC model is installed");
    BTK_REQUIRE(n > 0, "empty batch");
    MacBatch B(n);
    batch_gradient(m.w.p, X, Y, n, B, grad, s);
    double l = 0.0;
    BTK_HIP(hipMemcpyAsync(&l, B.loss.p, sizeof(double), hipMemcpyDeviceToHost, s));
    BTK_HIP(hipStreamSynchronize(s));
    return (float)l;
}

float mac_train(const float *X, const float *Y, int n, int epochs, int batch, float lr, unsigned seed, bool reinit,
                hipStream_t s) {
    BTK_REQUIRE(n > 0 && epochs > 0 && batch > 0, "n, epochs and batch must be positive");
    MacModel &m = mac_model();
    if (reinit || !m.ready) {
        const std::vector<float> w = he_init(seed);
        mac_set_weights(w.data(), s);
    }
    const int nb = std::min(batch, n);
    MacBatch B(nb);
    DBuf<float> Xb((size_t)nb * kMacIn), Yb((size_t)nb * kMacOut), G(kMacParams), M(kMacParams), V(kMacParams);
    BTK_HIP(hipMemsetAsync(M.p, 0, sizeof(float) * kMacParams, s));
    BTK_HIP(hipMemsetAsync(V.p, 0, sizeof(float) * kMacParams, s));
    std::vector<int> perm(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    DBuf<int> dperm(n);
    std::mt19937 rng(seed ^ 0x5bd1e995u);
    const float b1 = 0.9f, b2 = 0.999f;
    int step = 0;
    double last = 0.0;
    for (int ep = 0; ep < epochs; ep++) {
        std::shuffle(perm.begin(), perm.end(), rng);
        BTK_HIP(hipMemcpyAsync(dperm.p, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        double sum = 0.0;
        int batches = 0;
        for (int b0 = 0; b0 + nb <= n; b0 += nb) {  // whole batches only
            k_mac_gather<<<nblk(nb), TB, 0, s>>>(nb, dperm.p + b0, X, Y, Xb.p, Yb.p);
            BTK_LAUNCH_CHECK();
            batch_gradient(m.w.p, Xb.p, Yb.p, nb, B, G.p, s);
            step++;
            const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
            k_mac_adam<<<nblk(kMacParams), TB, 0, s>>>(m.w.p, G.p, M.p, V.p, lr, b1, b2, c1, c2);
            BTK_LAUNCH_CHECK();
            double l = 0.0;
            BTK_HIP(hipMemcpyAsync(&l, B.loss.p, sizeof(double), hipMemcpyDeviceToHost, s));
            BTK_HIP(hipStreamSynchronize(s));
            sum += l;
            batches++;
        }
        last = batches ? sum / batches : 0.0;
    }
    BTK_HIP(hipStreamSynchronize(s));
    return (float)last;
}

}  // namespace btk
