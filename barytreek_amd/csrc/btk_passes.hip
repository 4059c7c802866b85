// Upward pass (upward_pass.cpp:7-29, upward_pass_impl.hpp:9-187) and
// downward pass (downward_pass.cpp:7-41, downward_pass_impl.hpp:9-379).
//
// The reference scatters every contribution with an atomic add: P atomics
// per point into its leaf, P^2 per child into its parent, P^2 per child from
// its parent.  Here every output is owned by one workgroup:
//   anterpolation    one workgroup per leaf; each thread owns a row of the
//                    (n+1)^dim proxy grid and streams the leaf's points,
//                    whose 1-D basis values sit in LDS;
//   child->parent    one workgroup per parent gathers its children, the
//                    (n+1)^dim x (n+1)^dim interpolation applied as dim
//                    successive 1-D contractions ((n+1)^(dim+1) each);
//   parent->child    one workgroup per child, the transposed contractions;
//   leaf->point      one workgroup per leaf, one target per thread, the
//                    leaf's target-proxy values in LDS.
// The 1-D matrices are the reference's: the parent's basis evaluated at the
// child's mapped nodes (upward_pass_impl.hpp:189-199).
#include "btk_device.hpp"

namespace btk {
namespace {

constexpr int TB = 256;
inline int nblk(long long n) { return (int)((n + TB - 1) / TB); }

struct Boxes {
    const double *lo0, *lo1, *lo2, *hi0, *hi1, *hi2;
};
Boxes boxes(const Tree &t) { return Boxes{t.lo[0].p, t.lo[1].p, t.lo[2].p, t.hi[0].p, t.hi[1].p, t.hi[2].p}; }

__device__ __forceinline__ double blo(const Boxes &b, int d, int p) { return d == 0 ? b.lo0[p] : (d == 1 ? b.lo1[p] : b.lo2[p]); }
__device__ __forceinline__ double bhi(const Boxes &b, int d, int p) { return d == 0 ? b.hi0[p] : (d == 1 ? b.hi1[p] : b.hi2[p]); }

__global__ void k_gather(int n, const int *__restrict__ perm, const double *__restrict__ src, double *__restrict__ dst) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

// ---------------------------------------------------------------------------
// Anterpolation: W[leaf][(j,k,l)] = sum_i Lx_j(x_i) Ly_k(y_i) Lz_l(z_i) q_i
// (upward_pass_impl.hpp:133-159, 51-71).
// ---------------------------------------------------------------------------
template <int DIM, int N1>
__global__ void __launch_bounds__(TB)
    k_anterp(const int *__restrict__ leaves, const int *__restrict__ start, const int *__restrict__ count, Boxes bx,
             const double *__restrict__ sx, const double *__restrict__ sy, const double *__restrict__ sz,
             const double *__restrict__ q, NodeTables tabs, double *__restrict__ W) {
    constexpr int ROWS = DIM == 3 ? N1 * N1 : N1;
    constexpr int G = (TB / ROWS) > 0 ? (TB / ROWS) : 1;
    constexpr int P = ROWS * N1;
    constexpr int TILE = 64;
    __shared__ double Lb[DIM][TILE][N1];
    __shared__ double qs[TILE];
    __shared__ double red[G * P];
    const int leaf = leaves[blockIdx.x];
    const int s0 = start[leaf], cnt = count[leaf];
    const int g = threadIdx.x / ROWS;
    const int r = threadIdx.x % ROWS;
    double acc[N1];
#pragma unroll
    for (int l = 0; l < N1; l++) acc[l] = 0.0;
    double lo[DIM], hi[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        lo[d] = blo(bx, d, leaf);
        hi[d] = bhi(bx, d, leaf);
    }
    for (int t0 = 0; t0 < cnt; t0 += TILE) {
        const int np = min(TILE, cnt - t0);
        for (int task = threadIdx.x; task < DIM * np; task += TB) {
            const int d = task / np, p = task % np;
            const double *co = d == 0 ? sx : (d == 1 ? sy : sz);
            const double x = co[s0 + t0 + p];
            double b[N1];
            bli_basis<N1>(x, d == 0 ? lo[0] : (d == 1 ? lo[1] : lo[DIM - 1]),
                          d == 0 ? hi[0] : (d == 1 ? hi[1] : hi[DIM - 1]), tabs.a, N1, b);
#pragma unroll
            for (int i = 0; i < N1; i++) Lb[d][p][i] = b[i];
        }
        if ((int)threadIdx.x < np) qs[threadIdx.x] = q[s0 + t0 + threadIdx.x];
        __syncthreads();
        if (g < G) {
            if constexpr (DIM == 3) {
                const int j = r / N1, k = r % N1;
                for (int p = g; p < np; p += G) {
                    const double a = Lb[0][p][j] * Lb[1][p][k] * qs[p];
#pragma unroll
                    for (int l = 0; l < N1; l++) acc[l] = fma(a, Lb[2][p][l], acc[l]);
                }
            } else {
                const int j = r;
                for (int p = g; p < np; p += G) {
                    const double a = Lb[0][p][j] * qs[p];
#pragma unroll
                    for (int l = 0; l < N1; l++) acc[l] = fma(a, Lb[1][p][l], acc[l]);
                }
            }
        }
        __syncthreads();
    }
    if (g < G) {
#pragma unroll
        for (int l = 0; l < N1; l++) red[g * P + r * N1 + l] = acc[l];
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < P; idx += TB) {
        double s = 0.0;
        for (int gg = 0; gg < G; gg++) s += red[gg * P + idx];
        W[(size_t)leaf * P + idx] = s;
    }
}

// 1-D interpolation matrices of a (parent, child) pair:
// S[d][a][b] = parent basis a at the child's mapped node b.
template <int DIM>
__device__ void interp_matrices(const Boxes &bx, int parent, int child, int n1, const NodeTables &tabs,
                                double (*S)[kMaxN1][kMaxN1]) {
    for (int task = threadIdx.x; task < DIM * n1; task += TB) {
        const int d = task / n1, b = task % n1;
        const double node = proxy_node(tabs.b[b], blo(bx, d, child), bhi(bx, d, child));
        double v[kMaxN1];
        bli_basis<kMaxN1>(node, blo(bx, d, parent), bhi(bx, d, parent), tabs.a, n1, v);
#pragma unroll
        for (int a = 0; a < kMaxN1; a++)
            if (a < n1) S[d][a][b] = v[a];
    }
}

constexpr int kMaxP = kMaxN1 * kMaxN1 * kMaxN1;
constexpr int kOutPerThread = (kMaxP + TB - 1) / TB;

// Calls f(std::integral_constant<int, n1>{}) for n1 = 2 .. kMaxN1.
template <typename F>
void launch_by_n1(int n1, F &&f) {
    switch (n1) {
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 3: f(std::integral_constant<int, 3>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        case 5: f(std::integral_constant<int, 5>{}); break;
        case 6: f(std::integral_constant<int, 6>{}); break;
        case 7: f(std::integral_constant<int, 7>{}); break;
        case 8: f(std::integral_constant<int, 8>{}); break;
        case 9: f(std::integral_constant<int, 9>{}); break;
        case 10: f(std::integral_constant<int, 10>{}); break;
        case 11: f(std::integral_constant<int, 11>{}); break;
        default: throw Error("interp_degree must be between 1 and 10");
    }
}

// ---------------------------------------------------------------------------
// child -> parent (upward_pass_impl.hpp:170-213): one workgroup per parent.
// ---------------------------------------------------------------------------
template <int DIM, int N1>
__global__ void __launch_bounds__(TB)
    k_m2m(int ls, const int *__restrict__ child0, const int *__restrict__ count, Boxes bx, NodeTables tabs,
          double *__restrict__ W) {
    constexpr int NC = DIM == 3 ? 8 : 4;
    constexpr int n1 = N1;  // compile-time: index arithmetic by constants, inner loops unrolled
    const int p = ls + blockIdx.x;
    const int c0 = child0[p];
    if (c0 < 0) return;
    const int P = DIM == 3 ? n1 * n1 * n1 : n1 * n1;
    __shared__ double S[DIM][kMaxN1][kMaxN1];
    __shared__ double A[kMaxP], B[kMaxP];
    double acc[kOutPerThread];
#pragma unroll
    for (int k = 0; k < kOutPerThread; k++) acc[k] = 0.0;
    for (int c = 0; c < NC; c++) {
        const int ch = c0 + c;
        if (count[ch] == 0) continue;  // an empty child is a leaf with zero weights
        interp_matrices<DIM>(bx, p, ch, n1, tabs, S);
        for (int idx = threadIdx.x; idx < P; idx += TB) A[idx] = W[(size_t)ch * P + idx];
        __syncthreads();
        if constexpr (DIM == 3) {
            const int n2 = n1 * n1;
            for (int idx = threadIdx.x; idx < P; idx += TB) {  // (j2, k1, l1) <- sum_j1
                const int j2 = idx / n2, rem = idx % n2;
                double s = 0.0;
                for (int j1 = 0; j1 < n1; j1++) s = fma(S[0][j2][j1], A[j1 * n2 + rem], s);
                B[idx] = s;
            }
            __syncthreads();
            for (int idx = threadIdx.x; idx < P; idx += TB) {  // (j2, k2, l1) <- sum_k1
                const int j2 = idx / n2, k2 = (idx / n1) % n1, l1 = idx % n1;
                double s = 0.0;
                for (int k1 = 0; k1 < n1; k1++) s = fma(S[1][k2][k1], B[j2 * n2 + k1 * n1 + l1], s);
                A[idx] = s;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kOutPerThread; k++) {  // (j2, k2, l2) <- sum_l1
                const int idx = threadIdx.x + k * TB;
                if (idx < P) {
                    const int base = idx - idx % n1, l2 = idx % n1;
                    double s = 0.0;
                    for (int l1 = 0; l1 < n1; l1++) s = fma(S[2][l2][l1], A[base + l1], s);
                    acc[k] += s;
                }
            }
        } else {
            for (int idx = threadIdx.x; idx < P; idx += TB) {  // (j2, k1) <- sum_j1
                const int j2 = idx / n1, k1 = idx % n1;
                double s = 0.0;
                for (int j1 = 0; j1 < n1; j1++) s = fma(S[0][j2][j1], A[j1 * n1 + k1], s);
                B[idx] = s;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kOutPerThread; k++) {  // (j2, k2) <- sum_k1
                const int idx = threadIdx.x + k * TB;
                if (idx < P) {
                    const int j2 = idx / n1, k2 = idx % n1;
                    double s = 0.0;
                    for (int k1 = 0; k1 < n1; k1++) s = fma(S[1][k2][k1], B[j2 * n1 + k1], s);
                    acc[k] += s;
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < kOutPerThread; k++) {
        const int idx = threadIdx.x + k * TB;
        if (idx < P) W[(size_t)p * P + idx] = acc[k];
    }
}

// ---------------------------------------------------------------------------
// parent -> child (downward_pass_impl.hpp:118-190): one workgroup per child.
// Psi is [C][pc][P]; the child's values are added to.
// ---------------------------------------------------------------------------
template <int DIM, int N1>
__global__ void __launch_bounds__(TB)
    k_l2l(int ls, const int *__restrict__ parent, Boxes bx, NodeTables tabs, int ncomp, size_t comp_stride,
          double *__restrict__ Psi) {
    constexpr int n1 = N1;  // compile-time: index arithmetic by constants, inner loops unrolled
    const int ch = ls + blockIdx.x;
    const int p = parent[ch];
    const int P = DIM == 3 ? n1 * n1 * n1 : n1 * n1;
    __shared__ double S[DIM][kMaxN1][kMaxN1];
    __shared__ double A[kMaxP], B[kMaxP];
    interp_matrices<DIM>(bx, p, ch, n1, tabs, S);
    for (int c = 0; c < ncomp; c++) {
        const double *pp = Psi + c * comp_stride + (size_t)p * P;
        double *pc = Psi + c * comp_stride + (size_t)ch * P;
        for (int idx = threadIdx.x; idx < P; idx += TB) A[idx] = pp[idx];
        __syncthreads();
        if constexpr (DIM == 3) {
            const int n2 = n1 * n1;
            for (int idx = threadIdx.x; idx < P; idx += TB) {  // (j, k2, l2) <- sum_j2 S[j2][j]
                const int j = idx / n2, rem = idx % n2;
                double s = 0.0;
                for (int j2 = 0; j2 < n1; j2++) s = fma(S[0][j2][j], A[j2 * n2 + rem], s);
                B[idx] = s;
            }
            __syncthreads();
            for (int idx = threadIdx.x; idx < P; idx += TB) {  // (j, k, l2) <- sum_k2
                const int j = idx / n2, k = (idx / n1) % n1, l2 = idx % n1;
                double s = 0.0;
                for (int k2 = 0; k2 < n1; k2++) s = fma(S[1][k2][k], B[j * n2 + k2 * n1 + l2], s);
                A[idx] = s;
            }
            __syncthreads();
            for (int idx = threadIdx.x; idx < P; idx += TB) {  // (j, k, l) <- sum_l2
                const int base = idx - idx % n1, l = idx % n1;
                double s = 0.0;
                for (int l2 = 0; l2 < n1; l2++) s = fma(S[2][l2][l], A[base + l2], s);
                pc[idx] += s;
            }
        } else {
            for (int idx = threadIdx.x; idx < P; idx += TB) {  // (j, k2) <- sum_j2
                const int j = idx / n1, k2 = idx % n1;
                double s = 0.0;
                for (int j2 = 0; j2 < n1; j2++) s = fma(S[0][j2][j], A[j2 * n1 + k2], s);
                B[idx] = s;
            }
            __syncthreads();
            for (int idx = threadIdx.x; idx < P; idx += TB) {  // (j, k) <- sum_k2
                const int j = idx / n1, k = idx % n1;
                double s = 0.0;
                for (int k2 = 0; k2 < n1; k2++) s = fma(S[1][k2][k], B[j * n1 + k2], s);
                pc[idx] += s;
            }
        }
        __syncthreads();
    }
}

// out[c][perm[i]] += sorted[c][i]: results accumulated in the tree's sorted
// order (coalesced) moved to the caller's order once.  Each output element
// has one source element, so plain adds suffice.
__global__ void k_add_unsorted(int n, long long total, const int *__restrict__ perm,
                               const double *__restrict__ sorted, double *__restrict__ out) {
    const long long k = (long long)blockIdx.x * TB + threadIdx.x;
    if (k >= total) return;
    const long long c = k / n, i = k % n;
    out[c * n + perm[i]] += sorted[k];
}

// ---------------------------------------------------------------------------
// leaf -> point (downward_pass_impl.hpp:207-236): one workgroup per leaf.
// ---------------------------------------------------------------------------
// Leaf -> point: one workgroup per leaf, one target per thread; the leaf's
// potentials sit in LDS proxy by proxy (NC components of one proxy together,
// padded to CP), so each read feeds one independent FMA chain per component
// and every offset fits one base register's immediate.  Sum-factorised:
// (n+1)^dim + (n+1)^(dim-1) + ... FMAs per component and target.
template <int DIM, int N1, int NC>
__global__ void __launch_bounds__(TB)
    k_l2p(const int *__restrict__ leaves, const int *__restrict__ start, const int *__restrict__ count, Boxes bx,
          const double *__restrict__ sx, const double *__restrict__ sy, const double *__restrict__ sz,
          const int *__restrict__ perm, NodeTables tabs, size_t comp_stride, const double *__restrict__ Psi,
          double *__restrict__ out, int M) {
    constexpr int P = DIM == 3 ? N1 * N1 * N1 : N1 * N1;
    constexpr int CP = NC == 3 ? 4 : NC;
    __shared__ __align__(16) double ps[P][CP];
    const int leaf = leaves[blockIdx.x];
    const int s0 = start[leaf], cnt = count[leaf];
    if (cnt == 0) return;
    for (int idx = threadIdx.x; idx < P * NC; idx += TB) {
        const int c = idx / P, pp = idx % P;
        ps[pp][c] = Psi[c * comp_stride + (size_t)leaf * P + pp];
    }
    __syncthreads();
    const double lo0 = bx.lo0[leaf], hi0 = bx.hi0[leaf], lo1 = bx.lo1[leaf], hi1 = bx.hi1[leaf];
    for (int i = threadIdx.x; i < cnt; i += TB) {
        double Lx[N1], Ly[N1], Lz[N1];
        bli_basis<N1>(sx[s0 + i], lo0, hi0, tabs.a, N1, Lx);
        bli_basis<N1>(sy[s0 + i], lo1, hi1, tabs.a, N1, Ly);
        if constexpr (DIM == 3) bli_basis<N1>(sz[s0 + i], bx.lo2[leaf], bx.hi2[leaf], tabs.a, N1, Lz);
        const int o = perm ? perm[s0 + i] : s0 + i;
        double phi[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) phi[c] = 0.0;
        if constexpr (DIM == 3) {
#pragma unroll
            for (int j = 0; j < N1; j++) {
                double sj[NC];
#pragma unroll
                for (int c = 0; c < NC; c++) sj[c] = 0.0;
#pragma unroll
                for (int k = 0; k < N1; k++) {
                    double sk[NC];
#pragma unroll
                    for (int c = 0; c < NC; c++) sk[c] = 0.0;
#pragma unroll
                    for (int l = 0; l < N1; l++) {
                        const double *v = ps[(j * N1 + k) * N1 + l];
#pragma unroll
                        for (int c = 0; c < NC; c++) sk[c] = fma(v[c], Lz[l], sk[c]);
                    }
#pragma unroll
                    for (int c = 0; c < NC; c++) sj[c] = fma(sk[c], Ly[k], sj[c]);
                }
#pragma unroll
                for (int c = 0; c < NC; c++) phi[c] = fma(sj[c], Lx[j], phi[c]);
                // one slab of the grid at a time: the scheduler hoists LDS
                // reads only within it (registers, no spills)
                __builtin_amdgcn_sched_barrier(0);
            }
        } else {
#pragma unroll
            for (int j = 0; j < N1; j++) {
                double sj[NC];
#pragma unroll
                for (int c = 0; c < NC; c++) sj[c] = 0.0;
#pragma unroll
                for (int k = 0; k < N1; k++) {
                    const double *v = ps[j * N1 + k];
#pragma unroll
                    for (int c = 0; c < NC; c++) sj[c] = fma(v[c], Ly[k], sj[c]);
                }
#pragma unroll
                for (int c = 0; c < NC; c++) phi[c] = fma(sj[c], Lx[j], phi[c]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
#pragma unroll
        for (int c = 0; c < NC; c++) out[(size_t)c * M + o] += phi[c];
    }
}

template <int DIM>
void launch_anterp(int n1, int nleaves, const Tree &t, const double *q, const NodeTables &tabs, double *W,
                   hipStream_t s) {
#define BTK_ANTERP(N)                                                                                           \
    case N:                                                                                                     \
        k_anterp<DIM, N><<<nleaves, TB, 0, s>>>(t.leaves.p, t.start.p, t.count.p, boxes(t), t.sx.p, t.sy.p,     \
                                                t.sz.p, q, tabs, W);                                            \
        break;
    switch (n1) {
        BTK_ANTERP(2) BTK_ANTERP(3) BTK_ANTERP(4) BTK_ANTERP(5) BTK_ANTERP(6) BTK_ANTERP(7) BTK_ANTERP(8)
        BTK_ANTERP(9) BTK_ANTERP(10) BTK_ANTERP(11)
        default: throw Error("interp_degree must be between 1 and 10");
    }
#undef BTK_ANTERP
    BTK_LAUNCH_CHECK();
}

template <int DIM, int NC>
void launch_l2p_nc(int n1, const int *leaves, int nleaves, const Tree &t, const NodeTables &tabs, size_t stride,
                   const int *perm, const double *Psi, double *out, hipStream_t s) {
#define BTK_L2P(N)                                                                                              \
    case N:                                                                                                     \
        k_l2p<DIM, N, NC><<<nleaves, TB, 0, s>>>(leaves, t.start.p, t.count.p, boxes(t), t.sx.p, t.sy.p,        \
                                                 t.sz.p, perm, tabs, stride, Psi, out, t.n);                    \
        break;
    switch (n1) {
        BTK_L2P(2) BTK_L2P(3) BTK_L2P(4) BTK_L2P(5) BTK_L2P(6) BTK_L2P(7) BTK_L2P(8) BTK_L2P(9) BTK_L2P(10)
        BTK_L2P(11)
        default: throw Error("interp_degree must be between 1 and 10");
    }
#undef BTK_L2P
    BTK_LAUNCH_CHECK();
}

template <int DIM>
void launch_l2p(int n1, const int *leaves, int nleaves, const Tree &t, const NodeTables &tabs, int ncomp,
                size_t stride, const int *perm, const double *Psi, double *out, hipStream_t s) {
    switch (ncomp) {
        case 1: launch_l2p_nc<DIM, 1>(n1, leaves, nleaves, t, tabs, stride, perm, Psi, out, s); break;
        case 3: launch_l2p_nc<DIM, 3>(n1, leaves, nleaves, t, tabs, stride, perm, Psi, out, s); break;
        case 4: launch_l2p_nc<DIM, 4>(n1, leaves, nleaves, t, tabs, stride, perm, Psi, out, s); break;
        default: throw Error("leaf-to-point: 1, 3 or 4 components");
    }
}

// ---------------------------------------------------------------------------
// Change of interpolation degree, one workgroup per proxy array (a panel's
// charges, or one component of a panel's potentials): out = (A x A x A) in
// with A the nb x na 1-D matrix, applied as three 1-D contractions through
// LDS.  Layout (j * n + k) * n + l as the passes.
// ---------------------------------------------------------------------------
struct DegreeMatrix {
    double a[kMaxN1 * kMaxN1];  // row-major nb x na
};

__global__ void __launch_bounds__(TB)
    k_degree_change(int na, int nb, DegreeMatrix A, const double *__restrict__ in, double *__restrict__ out) {
    __shared__ double X[kMaxP], Y[kMaxP];
    const size_t b = blockIdx.x;
    const int a2 = na * na, Pa = a2 * na, Pb = nb * nb * nb;
    for (int i = threadIdx.x; i < Pa; i += TB) X[i] = in[b * Pa + i];
    __syncthreads();
    for (int i = threadIdx.x; i < nb * a2; i += TB) {  // Y[j2][k1][l1] = sum_j1 A[j2][j1] X[j1][k1][l1]
        const int j2 = i / a2, r = i % a2;
        double v = 0.0;
        for (int j1 = 0; j1 < na; j1++) v = fma(A.a[j2 * na + j1], X[j1 * a2 + r], v);
        Y[i] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb * nb * na; i += TB) {  // X[j2][k2][l1] = sum_k1 A[k2][k1] Y[j2][k1][l1]
        const int j2 = i / (nb * na), k2 = (i / na) % nb, l1 = i % na;
        double v = 0.0;
        for (int k1 = 0; k1 < na; k1++) v = fma(A.a[k2 * na + k1], Y[j2 * a2 + k1 * na + l1], v);
        X[i] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < Pb; i += TB) {  // out[j2][k2][l2] = sum_l1 A[l2][l1] X[j2][k2][l1]
        const int jk = i / nb, l2 = i % nb;
        double v = 0.0;
        for (int l1 = 0; l1 < na; l1++) v = fma(A.a[l2 * na + l1], X[jk * na + l1], v);
        out[b * Pb + i] = v;
    }
}

// The downward change of degree for every lower degree at once: one
// workgroup per (component, panel) interpolates the panel's potentials of
// each lower degree that has far entries on it to the top degree, sums them
// in registers and adds the sum into the top-degree array, which is read and
// written once instead of once per degree.
constexpr int kMaxLow = 3;  // lower degrees of a list (at most four in all)
struct DegreeSet {
    DegreeMatrix A[kMaxLow];        // nb x na[k], row-major
    const double *in[kMaxLow];      // potentials of lower degree k, C * pct arrays
    const int *goff[kMaxLow];       // that degree's group offsets: panels with entries
    int na[kMaxLow];
    int nlow;
};

// One lower degree's contribution, sizes known at compile time: X holds the
// panel's degree-(NA-1) potentials, A (NB x NA, in LDS) interpolates each
// axis; acc[r] += the degree-(NB-1) values of this thread's outputs.
template <int NA, int NB>
__device__ __forceinline__ void degree_add_one(const double *__restrict__ in, const double *A, double *X, double *Y,
                                               double *acc) {
    constexpr int a2 = NA * NA, Pa = a2 * NA, nb2 = NB * NB, Pb = nb2 * NB;
    constexpr int R = (Pb + TB - 1) / TB;
    for (int i = threadIdx.x; i < Pa; i += TB) X[i] = in[i];
    __syncthreads();
    for (int i = threadIdx.x; i < NB * a2; i += TB) {  // Y[j2][k1][l1] = sum_j1 A[j2][j1] X[j1][k1][l1]
        const int j2 = i / a2, r = i % a2;
        double v = 0.0;
#pragma unroll
        for (int j1 = 0; j1 < NA; j1++) v = fma(A[j2 * NA + j1], X[j1 * a2 + r], v);
        Y[i] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb2 * NA; i += TB) {  // X[j2][k2][l1] = sum_k1 A[k2][k1] Y[j2][k1][l1]
        const int j2 = i / (NB * NA), k2 = (i / NA) % NB, l1 = i % NA;
        double v = 0.0;
#pragma unroll
        for (int k1 = 0; k1 < NA; k1++) v = fma(A[k2 * NA + k1], Y[j2 * a2 + k1 * NA + l1], v);
        X[i] = v;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {  // sum_l1 A[l2][l1] X[j2][k2][l1] into this thread's outputs
        const int i = threadIdx.x + r * TB;
        if (i < Pb) {
            const int jk = i / NB, l2 = i % NB;
            double v = 0.0;
#pragma unroll
            for (int l1 = 0; l1 < NA; l1++) v = fma(A[l2 * NA + l1], X[jk * NA + l1], v);
            acc[r] += v;
        }
    }
    __syncthreads();  // X is refilled by the next degree
}

template <int NA, int NB>
__device__ __forceinline__ void degree_add_dispatch(int na, const double *in, const double *A, double *X, double *Y,
                                                    double *acc) {
    if constexpr (NA < NB) {
        if (na == NA)
            degree_add_one<NA, NB>(in, A, X, Y, acc);
        else
            degree_add_dispatch<NA + 1, NB>(na, in, A, X, Y, acc);
    }
}

template <int NB>
__global__ void __launch_bounds__(TB)
    k_degree_add(DegreeSet ds, double *__restrict__ out, int pct) {
    constexpr int Pb = NB * NB * NB;
    constexpr int R = (Pb + TB - 1) / TB;
    __shared__ double X[kMaxP], Y[kMaxP];
    __shared__ double sA[kMaxLow][kMaxN1 * kMaxN1];
    const size_t b = blockIdx.x;
    const int p = (int)(b % pct);
    for (int k = 0; k < ds.nlow; k++)
        for (int i = threadIdx.x; i < NB * ds.na[k]; i += TB) sA[k][i] = ds.A[k].a[i];
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; r++) acc[r] = 0.0;
    bool any = false;
    for (int k = 0; k < ds.nlow; k++) {
        if (ds.goff[k][p + 1] == ds.goff[k][p]) continue;  // the same for the whole workgroup
        any = true;
        const int na = ds.na[k];
        degree_add_dispatch<2, NB>(na, ds.in[k] + b * (size_t)(na * na * na), sA[k], X, Y, acc);
    }
    if (!any) return;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int i = threadIdx.x + r * TB;
        if (i < Pb) out[b * Pb + i] += acc[r];
    }
}

// A[i][k] = L_k(y_i): the degree-d Lagrange basis on the nodes cos(pi k / d)
// at the degree-e proxy nodes cos(pi / e * i), in the reference interval.
DegreeMatrix degree_matrix(int d, int e) {
    const NodeTables td = node_tables(d), te = node_tables(e);
    DegreeMatrix A{};
    for (int i = 0; i <= e; i++)
        for (int k = 0; k <= d; k++) {
            double v = 1.0;
            for (int j = 0; j <= d; j++)
                if (j != k) v *= (te.b[i] - td.a[j]) / (td.a[k] - td.a[j]);
            A.a[i * (d + 1) + k] = v;
        }
    return A;
}

// Row-major transpose of an (e+1) x (d+1) matrix.
DegreeMatrix transposed(const DegreeMatrix &A, int d, int e) {
    DegreeMatrix T{};
    for (int i = 0; i <= e; i++)
        for (int k = 0; k <= d; k++) T.a[k * (e + 1) + i] = A.a[i * (d + 1) + k];
    return T;
}

}  // namespace

void gather_sorted(const Tree &t, const double *orig, double *sorted, hipStream_t s) {
    if (t.n == 0) return;
    k_gather<<<nblk(t.n), TB, 0, s>>>(t.n, t.perm.p, orig, sorted);
    BTK_LAUNCH_CHECK();
}

void upward_pass(const btk_run_config &cfg, const Tree &t, const double *q_sorted, double *W, hipStream_t s) {
    const int n1 = cfg.interp_degree + 1;
    BTK_REQUIRE(n1 >= 2 && n1 <= kMaxN1, "interp_degree must be between 1 and 10");
    const NodeTables tabs = node_tables(cfg.interp_degree);
    if (t.dim == 3)
        launch_anterp<3>(n1, t.n_leaves, t, q_sorted, tabs, W, s);
    else
        launch_anterp<2>(n1, t.n_leaves, t, q_sorted, tabs, W, s);
    // child_to_parent, deepest level first (upward_pass.cpp:23-28)
    for (int lev = t.levels - 2; lev >= 0; lev--) {
        const int ls = t.level_start[lev], np = t.level_start[lev + 1] - ls;
        if (t.dim == 3)
            launch_by_n1(n1, [&](auto N) {
                k_m2m<3, decltype(N)::value><<<np, TB, 0, s>>>(ls, t.child0.p, t.count.p, boxes(t), tabs, W);
            });
        else
            launch_by_n1(n1, [&](auto N) {
                k_m2m<2, decltype(N)::value><<<np, TB, 0, s>>>(ls, t.child0.p, t.count.p, boxes(t), tabs, W);
            });
        BTK_LAUNCH_CHECK();
    }
}

void downward_pass(const btk_run_config &cfg, const btk_options &o, const Tree &t, double *Psi, double *out,
                   hipStream_t s, bool sorted_out) {
    const int n1 = cfg.interp_degree + 1;
    BTK_REQUIRE(n1 >= 2 && n1 <= kMaxN1, "interp_degree must be between 1 and 10");
    const int ncomp = components(t.dim, o.kernel);
    const NodeTables tabs = node_tables(cfg.interp_degree);
    const int P = ipow(n1, t.dim);
    const size_t stride = (size_t)t.pc * P;
    // parent_to_child, levels 0..L-2 (downward_pass.cpp:21-25); every panel of
    // level lev+1 is a child of a non-leaf panel of level lev
    for (int lev = 0; lev + 1 < t.levels; lev++) {
        const int ls = t.level_start[lev + 1], np = t.level_start[lev + 2] - ls;
        if (t.dim == 3)
            launch_by_n1(n1, [&](auto N) {
                k_l2l<3, decltype(N)::value><<<np, TB, 0, s>>>(ls, t.parent.p, boxes(t), tabs, ncomp, stride, Psi);
            });
        else
            launch_by_n1(n1, [&](auto N) {
                k_l2l<2, decltype(N)::value><<<np, TB, 0, s>>>(ls, t.parent.p, boxes(t), tabs, ncomp, stride, Psi);
            });
        BTK_LAUNCH_CHECK();
    }
    // leaf_to_point: every leaf, or only the deepest level's (downward_pass.cpp:26-28)
    const int nl = o.deepest_only ? t.n_leaves_deepest : t.n_leaves;
    const int *lv = t.leaves.p + (t.n_leaves - nl);
    if (nl == 0) return;
    const int *perm = sorted_out ? nullptr : t.perm.p;
    if (t.dim == 3)
        launch_l2p<3>(n1, lv, nl, t, tabs, ncomp, stride, perm, Psi, out, s);
    else
        launch_l2p<2>(n1, lv, nl, t, tabs, ncomp, stride, perm, Psi, out, s);
}

void add_unsorted(const Tree &t, int C, const double *sorted, double *out, hipStream_t s) {
    const long long total = (long long)C * t.n;
    if (total == 0) return;
    k_add_unsorted<<<nblk(total), TB, 0, s>>>(t.n, total, t.perm.p, sorted, out);
    BTK_LAUNCH_CHECK();
}

}  // namespace btk

namespace btk {

size_t proxy_offset(const List &l, int pc, int comps, int di) {
    size_t off = 0;
    for (int k = 0; k < di; k++) off += (size_t)comps * pc * ipow(l.deg[k] + 1, l.dim);
    return off;
}

void upward_all(const btk_run_config &cfg, const List &l, const Tree &t, const double *q_sorted, double *W,
                hipStream_t s) {
    const int top = l.ndeg - 1, n = l.deg[top];
    BTK_REQUIRE(n == cfg.interp_degree, "the list was built for another interp_degree");
    double *Wn = W + proxy_offset(l, t.pc, 1, top);
    upward_pass(cfg, t, q_sorted, Wn, s);
    if (top == 0 || t.pc == 0) return;
    BTK_REQUIRE(t.dim == 3, "several interpolation degrees: 3D only");
    for (int di = 0; di < top; di++) {
        // W_d[k] = sum_m W_n[m] L_k(y_m): A[k][m], degree d out, degree n in
        const int d = l.deg[di];
        const DegreeMatrix A = transposed(degree_matrix(d, n), d, n);
        k_degree_change<<<t.pc, TB, 0, s>>>(n + 1, d + 1, A, Wn, W + proxy_offset(l, t.pc, 1, di));
        BTK_LAUNCH_CHECK();
    }
}

void downward_all(const btk_run_config &cfg, const btk_options &o, const List &l, const Tree &t, double *Psi,
                  double *out, hipStream_t s, bool sorted_out) {
    const int top = l.ndeg - 1, n = l.deg[top];
    BTK_REQUIRE(n == cfg.interp_degree, "the list was built for another interp_degree");
    const int C = components(t.dim, o.kernel);
    double *Pn = Psi + proxy_offset(l, t.pc, C, top);
    if (top > 0 && t.pc > 0) {
        BTK_REQUIRE(t.dim == 3, "several interpolation degrees: 3D only");
        BTK_REQUIRE(top <= kMaxLow, "at most four interpolation degrees");
        BTK_REQUIRE(l.pct == t.pc, "the list's target panels are not this tree's");
        // Psi_n[m] += sum_d sum_k L_k(y_m) Psi_d[k]: degree n out, each lower
        // degree d in.  Only the far (CP, CC) entries of degree d write
        // Psi_d, so a panel without such entries skips that degree.
        DegreeSet ds{};
        ds.nlow = top;
        for (int di = 0; di < top; di++) {
            const int d = l.deg[di];
            ds.A[di] = degree_matrix(d, n);
            ds.in[di] = Psi + proxy_offset(l, t.pc, C, di);
            ds.goff[di] = l.goff.p + (size_t)(l.ndeg + di) * t.pc;
            ds.na[di] = d + 1;
        }
        switch (n + 1) {
#define BTK_DEGREE_ADD(N)                                                                                       \
    case N: k_degree_add<N><<<C * t.pc, TB, 0, s>>>(ds, Pn, t.pc); break;
            BTK_DEGREE_ADD(3) BTK_DEGREE_ADD(4) BTK_DEGREE_ADD(5) BTK_DEGREE_ADD(6) BTK_DEGREE_ADD(7)
            BTK_DEGREE_ADD(8) BTK_DEGREE_ADD(9) BTK_DEGREE_ADD(10) BTK_DEGREE_ADD(11)
#undef BTK_DEGREE_ADD
            default: throw Error("interp_degree must be between 2 and 10 with several degrees");
        }
        BTK_LAUNCH_CHECK();
    }
    downward_pass(cfg, o, t, Pn, out, s, sorted_out);
}

}  // namespace btk
