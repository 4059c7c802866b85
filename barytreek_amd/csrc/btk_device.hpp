// Device-side building blocks: barycentric basis, kernel evaluations.
//
// The library is compiled with -ffp-contract=off, so every expression below
// rounds exactly as written; fused multiply-adds are spelled out with fma()
// where a kernel wants them and the reference's rounding does not matter.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>

#include "btk_internal.hpp"

namespace btk {

constexpr int kMaxN1 = BTK_MAX_DEGREE + 1;
constexpr int kWave = 64;

// Barycentric Lagrange basis of degree n1-1 on the Chebyshev-Lobatto nodes of
// [lo, hi] at x (bli_impl.hpp:8-63): nodes ta[i]*r + o, weights 1/2, (-1)^i,
// (-1)^n/2, a node within DBL_EPSILON of x gives the unit vector.
template <int N1>
__device__ __forceinline__ void bli_basis(double x, double lo, double hi, const double *ta, int n1, double *out) {
    const double r = 0.5 * (hi - lo);
    const double o = 0.5 * (hi + lo);
    double c[N1];
    int hit = -1;
#pragma unroll
    for (int i = 0; i < N1; i++) {
        if (i < n1) {
            c[i] = ta[i] * r + o;
            if (hit < 0 && fabs(x - c[i]) < DBL_EPSILON) hit = i;
        }
    }
    if (hit >= 0) {
#pragma unroll
        for (int i = 0; i < N1; i++)
            if (i < n1) out[i] = (i == hit) ? 1.0 : 0.0;
        return;
    }
    double den = 0.0;
#pragma unroll
    for (int i = 0; i < N1; i++) {
        if (i < n1) {
            const double sgn = (i & 1) ? -1.0 : 1.0;
            const double w = (i == 0) ? 0.5 : ((i == n1 - 1) ? 0.5 * sgn : sgn);
            const double v = w / (x - c[i]);
            out[i] = v;
            den += v;
        }
    }
#pragma unroll
    for (int i = 0; i < N1; i++)
        if (i < n1) out[i] = out[i] / den;
}

// Mapped proxy node i of [lo, hi] (bli_impl.hpp:65-77): tb[i]*r + o.
__device__ __forceinline__ double proxy_node(double tb_i, double lo, double hi) {
    const double r = 0.5 * (hi - lo);
    const double o = 0.5 * (hi + lo);
    return tb_i * r + o;
}

// 1/sqrt(x) in FP64: hardware estimate refined by two Newton steps
// (relative error a few units in the last place).
__device__ __forceinline__ double rsqrt_f64(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-x, y * y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-x, y * y, 1.0);
    y = fma(0.5 * y, e, y);
    return y;
}

// 1/x in FP64: hardware estimate refined by two Newton steps (about one unit
// in the last place; the IEEE division costs three times the instructions).
__device__ __forceinline__ double recip_f64(double x) {
    double y = __builtin_amdgcn_rcp(x);
    y = fma(fma(-x, y, 1.0), y, y);
    y = fma(fma(-x, y, 1.0), y, y);
    return y;
}

// log(x) in FP64 for finite x > 0 (a few units in the last place).
// x = 2^e m with m in [sqrt(1/2), sqrt(2)), so log x = e ln 2 + log m and
// log m = 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 3 - 2 sqrt(2) < 0.1716;
// the Taylor series 2 (s + s^3/3 + ... + s^19/19) stops where the next term
// is below 2^-54 s.  ln 2 = kLn2Hi + kLn2Lo with kLn2Hi's low 21 bits zero,
// so e * kLn2Hi is exact for any exponent.
constexpr double kLn2Hi = 0.6931471803691238;      // 0x3fe62e42fee00000
constexpr double kLn2Lo = 1.9082149292705877e-10;  // ln 2 - kLn2Hi, rounded
__device__ __forceinline__ double log_f64(double x) {
    double m = __builtin_amdgcn_frexp_mant(x);  // [1/2, 1)
    int e = __builtin_amdgcn_frexp_exp(x);
    if (m < 0.70710678118654752440) {
        m = m + m;
        e -= 1;
    }
    const double f = m - 1.0;  // exact
    const double s = f * recip_f64(m + 1.0);
    const double z = s * s;
    double p = 1.0 / 19.0;
    p = fma(p, z, 1.0 / 17.0);
    p = fma(p, z, 1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, 1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, 1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, 1.0 / 3.0);
    const double s2 = s + s;
    const double lm = fma(s2 * z, p, s2);  // 2s + 2s z p
    const double de = (double)e;
    const double r = fma(de, kLn2Hi, fma(de, kLn2Lo, lm));
    return x > 0.0 ? r : -__builtin_inf();  // log 0 = -inf, as the C library's
}

// Table-driven FP64 log (Tang, "Table-driven implementation of the logarithm
// function in IEEE floating-point arithmetic", ACM TOMS 16(4), 1990) for the
// 2D FP64 kernels, where log is most of an evaluation.  x = 2^e m with m in
// [1, 2); entry j of kLogTab covers m in [1 + j/128, 1 + (j+1)/128) and holds
// inv = 1/c (c the interval's midpoint) rounded to 9 bits and L = -log(inv).
// r = m inv - 1 is formed by one fma (m inv is exact in its 63 bits), |r| <
// 0.006, and log x = e ln 2 + L + log1p(r) with log1p by its Taylor series to
// r^8 (remainder below 2^-60 relative).  22 instructions against the
// reduction above's 34; absolute error a few units of 2^-53 |log x| or of
// 2^-53 ln 2 (near x = 1, where the terms cancel).  Each workgroup fills
// its table with log_f64 (fill_log_tab) before use.
constexpr int kLogTab = 128;

__device__ __forceinline__ void fill_log_tab(double2 *tab, int tid, int nthreads) {
    for (int j = tid; j < kLogTab; j += nthreads) {
        const double c = 1.0 + (j + 0.5) / kLogTab;
        const double inv = rint(512.0 / c) / 512.0;
        tab[j] = make_double2(-log_f64(inv), inv);
    }
}

__device__ __forceinline__ double log_f64_tab(double x, const double2 *tab) {
    const double m = 2.0 * __builtin_amdgcn_frexp_mant(x);  // [1, 2)
    const int e = __builtin_amdgcn_frexp_exp(x) - 1;
    const int j = (__double2hiint(m) >> 13) & (kLogTab - 1);  // the top 7 bits of m's fraction
    const double2 t = tab[j];
    const double r = fma(m, t.y, -1.0);
    double p = fma(r, -1.0 / 8.0, 1.0 / 7.0);
    p = fma(p, r, -1.0 / 6.0);
    p = fma(p, r, 1.0 / 5.0);
    p = fma(p, r, -1.0 / 4.0);
    p = fma(p, r, 1.0 / 3.0);
    p = fma(p, r, -0.5);
    const double lp = fma(r * r, p, r);  // log1p(r)
    const double de = (double)e;
    const double res = fma(de, kLn2Hi, fma(de, kLn2Lo, t.x + lp));
    return x > 0.0 ? res : -__builtin_inf();
}

// ---------------------------------------------------------------------------
// Kernel evaluations.  acc[] collects the sums without the constant factor;
// finish() applies it.  Components:
//   POT        acc0 = sum q G0          (3D G0 = 1/R^(1/2), 2D G0 = log R)
//   REF_VEL3D  acc  = sum q d / R       (3 components)
//   POT_GRAD   acc0 as POT; acc1.. = sum q d / R^(3/2) (3D), sum q d / R (2D)
// with R = r^2 + eps^2 and d = t - s.
// ---------------------------------------------------------------------------
template <int DIM, int KERNEL>
struct KernelTraits {
    static constexpr int C = (KERNEL == BTK_KERNEL_POT) ? 1 : (KERNEL == BTK_KERNEL_REF_VEL3D ? 3 : 1 + DIM);
};

// d = t - s is passed in; EXCL skips the pair when R == eps^2, i.e. when
// |d|^2 vanishes at working precision (the self pair of coincident sets).
// LOGTAB: 2D log by the table ltab (log_f64_tab) instead of log_f64.
template <int DIM, int KERNEL, bool EXCL, bool LOGTAB = false>
__device__ __forceinline__ void eval_f64(double dx, double dy, double dz, double q, double eps2, double *acc,
                                         const double2 *ltab = nullptr) {
    if constexpr (DIM == 3) {
        const double R = fma(dx, dx, fma(dy, dy, fma(dz, dz, eps2)));
        const bool skip = EXCL && R == eps2;
        if constexpr (KERNEL == BTK_KERNEL_POT) {
            const double inv = skip ? 0.0 : rsqrt_f64(R);
            acc[0] = fma(q, inv, acc[0]);
        } else if constexpr (KERNEL == BTK_KERNEL_REF_VEL3D) {
            const double g = skip ? 0.0 : q * recip_f64(R);
            acc[0] = fma(dx, g, acc[0]);
            acc[1] = fma(dy, g, acc[1]);
            acc[2] = fma(dz, g, acc[2]);
        } else {
            const double inv = skip ? 0.0 : rsqrt_f64(R);
            const double qi = q * inv;
            acc[0] += qi;
            const double h = qi * inv * inv;
            acc[1] = fma(h, dx, acc[1]);
            acc[2] = fma(h, dy, acc[2]);
            acc[3] = fma(h, dz, acc[3]);
        }
    } else {
        (void)dz;
        const double R = fma(dx, dx, fma(dy, dy, eps2));
        const bool skip = EXCL && R == eps2;
        double lg = 0.0;
        if constexpr (LOGTAB)
            lg = skip ? 0.0 : log_f64_tab(R, ltab);
        else
            lg = skip ? 0.0 : log_f64(R);
        acc[0] = fma(q, lg, acc[0]);
        if constexpr (KERNEL == BTK_KERNEL_POT_GRAD) {
            const double g = skip ? 0.0 : q * recip_f64(R);
            acc[1] = fma(g, dx, acc[1]);
            acc[2] = fma(g, dy, acc[2]);
        }
    }
}

// Constant factors of each component (see eval_f64).
template <int DIM, int KERNEL>
__device__ __forceinline__ double comp_scale(int c, double gfc) {
    if constexpr (DIM == 3) {
        if constexpr (KERNEL == BTK_KERNEL_POT_GRAD) return c == 0 ? gfc : -gfc;
        return gfc;
    } else {
        if constexpr (KERNEL == BTK_KERNEL_POT_GRAD) return c == 0 ? gfc : 2.0 * gfc;
        return gfc;
    }
}

// ---------------------------------------------------------------------------
// FP32 evaluations (mixed precision).  The difference d = t - s is formed in
// FP64 and rounded once, so it keeps FP32 relative precision even for pairs
// much closer than their coordinates' magnitude; R = |d|^2 + eps^2, the
// hardware log2 / reciprocal / reciprocal square root and the products are
// FP32.  The 2D potential accumulates q log2 R, rescaled by ln 2 in
// comp_scale_f32.
// ---------------------------------------------------------------------------
// Kahan's compensated summation (Kahan, CACM 8(1), 1965): c carries the
// rounding error of the running sum a, whose value is a - c.
__device__ __forceinline__ void kahan_add(float &a, float &c, float term) {
    const float y = term - c;
    const float t = a + y;
    c = (t - a) - y;
    a = t;
}

// The FP32 kernel with Kahan-compensated sums (acc, comp per component).
template <int DIM, int KERNEL, bool EXCL>
__device__ __forceinline__ void eval_f32_kahan(float dx, float dy, float dz, float q, float eps2, float *acc,
                                               float *comp) {
    if constexpr (DIM == 3) {
        const float R = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
        const bool skip = EXCL && R == eps2;
        if constexpr (KERNEL == BTK_KERNEL_POT) {
            const float inv = skip ? 0.0f : __builtin_amdgcn_rsqf(R);
            kahan_add(acc[0], comp[0], q * inv);
        } else if constexpr (KERNEL == BTK_KERNEL_REF_VEL3D) {
            const float g = skip ? 0.0f : q * __builtin_amdgcn_rcpf(R);
            kahan_add(acc[0], comp[0], dx * g);
            kahan_add(acc[1], comp[1], dy * g);
            kahan_add(acc[2], comp[2], dz * g);
        } else {
            const float inv = skip ? 0.0f : __builtin_amdgcn_rsqf(R);
            const float qi = q * inv;
            kahan_add(acc[0], comp[0], qi);
            const float h = qi * inv * inv;
            kahan_add(acc[1], comp[1], h * dx);
            kahan_add(acc[2], comp[2], h * dy);
            kahan_add(acc[3], comp[3], h * dz);
        }
    } else {
        (void)dz;
        const float R = fmaf(dx, dx, fmaf(dy, dy, eps2));
        const bool skip = EXCL && R == eps2;
        const float lg = skip ? 0.0f : __builtin_amdgcn_logf(R);
        kahan_add(acc[0], comp[0], q * lg);
        if constexpr (KERNEL == BTK_KERNEL_POT_GRAD) {
            const float g = skip ? 0.0f : q * __builtin_amdgcn_rcpf(R);
            kahan_add(acc[1], comp[1], g * dx);
            kahan_add(acc[2], comp[2], g * dy);
        }
    }
}

template <int DIM, int KERNEL, bool EXCL>
__device__ __forceinline__ void eval_f32(float dx, float dy, float dz, float q, float eps2, float *acc) {
    if constexpr (DIM == 3) {
        const float R = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
        const bool skip = EXCL && R == eps2;
        if constexpr (KERNEL == BTK_KERNEL_POT) {
            const float inv = skip ? 0.0f : __builtin_amdgcn_rsqf(R);
            acc[0] = fmaf(q, inv, acc[0]);
        } else if constexpr (KERNEL == BTK_KERNEL_REF_VEL3D) {
            const float g = skip ? 0.0f : q * __builtin_amdgcn_rcpf(R);
            acc[0] = fmaf(dx, g, acc[0]);
            acc[1] = fmaf(dy, g, acc[1]);
            acc[2] = fmaf(dz, g, acc[2]);
        } else {
            const float inv = skip ? 0.0f : __builtin_amdgcn_rsqf(R);
            const float qi = q * inv;
            acc[0] += qi;
            const float h = qi * inv * inv;
            acc[1] = fmaf(h, dx, acc[1]);
            acc[2] = fmaf(h, dy, acc[2]);
            acc[3] = fmaf(h, dz, acc[3]);
        }
    } else {
        (void)dz;
        const float R = fmaf(dx, dx, fmaf(dy, dy, eps2));
        const bool skip = EXCL && R == eps2;
        const float lg = skip ? 0.0f : __builtin_amdgcn_logf(R);
        acc[0] = fmaf(q, lg, acc[0]);
        if constexpr (KERNEL == BTK_KERNEL_POT_GRAD) {
            const float g = skip ? 0.0f : q * __builtin_amdgcn_rcpf(R);
            acc[1] = fmaf(g, dx, acc[1]);
            acc[2] = fmaf(g, dy, acc[2]);
        }
    }
}

template <int DIM, int KERNEL>
__device__ __forceinline__ double comp_scale_f32(int c, double gfc) {
    if constexpr (DIM == 2) {
        if (c == 0) return gfc * 0.69314718055994530942;  // log R = ln 2 * log2 R
    }
    return comp_scale<DIM, KERNEL>(c, gfc);
}

}  // namespace btk
