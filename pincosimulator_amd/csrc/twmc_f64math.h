// twmc_f64math.h -- FP64 Box-Muller for the per-step Wiener increments of the
// FP64 kernels, written for the arguments Box-Muller takes here: u1 = (a + 1/2)
// 2^-32 and theta = b 2^-32 revolutions for 32-bit integers a, b.  General
// library log / sin / cos spend most of their instructions on argument ranges
// that never occur here (large angles, subnormals, infinities); these forms
// reduce exactly in integer arithmetic and evaluate short polynomials:
//   sin, cos: the 1/256 turn j nearest to b is removed from b itself; a
//     256-row table (twmc_f64tab.h) holds cos and sin of theta_j = 2 pi j/256;
//     the rest, x = (b - j 2^24) 2 pi 2^-32, |x| <= pi/256, enters by
//     sin x = x - x^3 ps(x^2), cos x - 1 = x^2 pc(x^2) and the angle sum;
//   ln: u1 = m 2^e from the exponent and significand bits of a + 1/2, m in
//     [1, 2); a 128-row table (twmc_f64tab.h) indexed by m's top 7 fraction
//     bits gives 2 c_j and L_j with ln m = L_j + ln(1 + r), r = m c_j - 1
//     exact and |r| <= 2^-7, ln(1 + r) = r + r^2 pq(r), plus e ln 2 with ln 2
//     split so that e ln2_hi is exact;
//   ps, pc (degree 2), pq (degree 6): Chebyshev fits over those ranges
//     (gen_f64_series.py), each adding less than 1e-17 to the relative error
//     of its function;
//   the radius sqrt(-2 ln u1) by Newton steps from the hardware reciprocal
//     square root: its argument (2e-10 < -2 ln u1 < 46) needs none of the
//     library square root's scaling for extreme exponents.
// N samples are evaluated side by side, so each series coefficient is formed
// once for all N.  Every multiply-add is an explicit fma and contraction is
// off, so the host build (used by the accuracy test) evaluates the same
// correctly rounded operations in the same order as the device build; only
// the Newton seed differs (a coarser one on the host, see rsq_seed).  Pure
// C++: no HIP header is needed to compile it on the host.
#ifndef TWMC_F64MATH_H
#define TWMC_F64MATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "twmc_f64tab.h"

#if defined(__HIPCC__)
#define TWMC_HD __host__ __device__
#else
#define TWMC_HD
#endif
#if defined(__clang__)
#define TWMC_NO_CONTRACT _Pragma("clang fp contract(off)")
#define TWMC_UNROLL _Pragma("unroll")
#else
#define TWMC_NO_CONTRACT
#define TWMC_UNROLL _Pragma("GCC unroll 16")
#endif

namespace twmc {
namespace f64 {

// 2 pi 2^-32: the double nearest 2 pi, scaled by a power of two
constexpr double kTurn32 = 6.283185307179586 / 4294967296.0;
// ln 2 = kLn2Hi + kLn2Lo to 2e-31; kLn2Hi has 44 significant bits
constexpr double kLn2Hi = 0.6931471805599472;
constexpr double kLn2Lo = -1.8641886737243033e-15;

// a b + c as one three-operand v_fma_f64 on the GPU.  The series' steps
// p = p z + k read a coefficient k that stays live for the next site; as the
// two-operand v_fmac_f64 the compiler would first copy k into the result
// register, one more instruction per step.
TWMC_HD inline double fma3(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
#else
  return fma(a, b, c);
#endif
}

TWMC_HD inline uint64_t bits(double x) {
  uint64_t u;
  memcpy(&u, &x, sizeof u);
  return u;
}

TWMC_HD inline double from_bits(uint64_t u) {
  double x;
  memcpy(&x, &u, sizeof x);
  return x;
}

// The significand of x > 0 scaled into [1/2, 1), exactly (one v_frexp_mant_f64).
TWMC_HD inline double frexp_mant(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_frexp_mant(x);
#else
  int e;
  return frexp(x, &e);
#endif
}

// Seed for the square root's Newton steps: the hardware's approximate
// 1/sqrt(x) on the GPU.  On the host the correctly rounded value cut to 21
// significant bits, a seed no better than the hardware's, so the accuracy test
// of the host build covers the refinement from a coarse start.
TWMC_HD inline double rsq_seed(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rsq(x);
#else
  return from_bits(bits(1.0 / sqrt(x)) & 0xffffffff00000000ull);
#endif
}

// sqrt(x) for normal x well inside the exponent range (no scaling): from
// y ~ 1/sqrt(x), g ~ sqrt(x) and h ~ 1/(2 sqrt(x)) are refined together
// (Goldschmidt: r = 1/2 - g h is the shared relative error, squared by one
// step), then g once more by the exact residual x - g^2.
TWMC_HD inline double sqrt_normal(double x) {
  TWMC_NO_CONTRACT
  const double y = rsq_seed(x);
  double g = x * y, h = 0.5 * y;
  const double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  return fma(fma(-g, g, x), h, g);
}

// The series coefficients, highest degree first: the top one, then the
// others (near the Taylor coefficients 1/7!, -1/5!, 1/3!; -1/6!, 1/4!, -1/2!;
// -1/8, 1/7, ..., -1/2, which the fit adjusts).  A kernel may hand the
// functions a copy held in vector registers for its whole run (pin), instead of
// having every use form its constant anew.
struct Series {
  double s_top, ks[2];  // ps
  double c_top, kc[2];  // pc
  double l_top, kl[6];  // pq
};
constexpr Series kSeries = {
    0.0001984120759012081,
    {-0.008333333333298177, 0.16666666666666666},
    -0.0013888832862873938,
    {0.041666666666350265, -0.5},
    -0.1250106818314722,
    {0.14286901154505832, -0.1666666663406768, 0.1999999996377896, -0.2500000000000025, 0.3333333333333361,
     -0.5}};

// On the GPU: makes the coefficients of k values the compiler cannot see
// through, so it keeps them in vector registers rather than forming each
// constant at each use.
TWMC_HD inline void pin(Series& k) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(k.s_top), "+v"(k.c_top), "+v"(k.l_top));
  for (int j = 0; j < 2; ++j) asm volatile("" : "+v"(k.ks[j]), "+v"(k.kc[j]));
  for (int j = 0; j < 6; ++j) asm volatile("" : "+v"(k.kl[j]));
#else
  (void)k;
#endif
}

// The tables (TWMC_F64_TABLES, 768 doubles): the ln rows (2 c_j, L_j) from
// kTabLog, the sin / cos rows (cos theta_j, sin theta_j) from kTabSinCos.  On
// the host the functions read them from kTables; device code hands them a copy
// in device memory.
constexpr int kTabLog = 0, kTabSinCos = 256, kTabSize = 768;
constexpr double kTables[kTabSize] = TWMC_F64_TABLES;

// Two doubles read together: a table row.
struct alignas(16) Row {
  double a, b;
};

// sin and cos of 2 pi b[i] 2^-32, i < N.
template <int N>
TWMC_HD inline void sincos_turn32(const uint32_t* b, double* s, double* c, const Series& k = kSeries,
                                  const double* tab = kTables) {
  TWMC_NO_CONTRACT
  const Row* rows = reinterpret_cast<const Row*>(tab + kTabSinCos);
  double x[N], z[N], ps[N], pc[N], cj[N], sj[N];
  TWMC_UNROLL
  for (int i = 0; i < N; ++i) {
    // b = j 2^24 + rem (mod 2^32): j the nearest 1/256 turn, rem in [-2^23, 2^23)
    const uint32_t j = (b[i] + 0x800000u) >> 24;
    const int32_t rem = int32_t(b[i] << 8) >> 8;
    x[i] = double(rem) * kTurn32;  // angle - theta_j
    z[i] = x[i] * x[i];
    const Row row = rows[j];
    cj[i] = row.a;
    sj[i] = row.b;
    ps[i] = k.s_top;
    pc[i] = k.c_top;
  }
  TWMC_UNROLL
  for (int j = 0; j < 2; ++j)
    TWMC_UNROLL
    for (int i = 0; i < N; ++i) {
      ps[i] = fma3(ps[i], z[i], k.ks[j]);
      pc[i] = fma3(pc[i], z[i], k.kc[j]);
    }
  TWMC_UNROLL
  for (int i = 0; i < N; ++i) {
    // sin x and cos x - 1, then the angle sum with theta_j
    const double sx = fma(-(x[i] * z[i]), ps[i], x[i]);
    const double cm = z[i] * pc[i];
    c[i] = fma(cj[i], cm, fma(-sj[i], sx, cj[i]));
    s[i] = fma(sj[i], cm, fma(cj[i], sx, sj[i]));
  }
}

// ln((a[i] + 1/2) 2^-32), in (-22.9, 0), i < N.  A row (2 c, L) of the ln
// table: r = m c - 1 and ln m = L + ln(1 + r), the exponent raised by one for
// m > sqrt 2, which L absorbs as -ln 2.
template <int N>
TWMC_HD inline void log_u32(const uint32_t* a, double* out, const Series& k = kSeries,
                            const double* tab = kTables) {
  TWMC_NO_CONTRACT
  const Row* rows = reinterpret_cast<const Row*>(tab + kTabLog);
  double r[N], lt[N], p[N];
  int e[N];
  TWMC_UNROLL
  for (int i = 0; i < N; ++i) {
    // a + 1/2 = m 2^(E - 1023) exactly, m in [1, 2) with at most 33 significant
    // bits; row j from its top 7 fraction bits; u = m 2^(E - 1055)
    const double v = double(a[i]) + 0.5;
    const uint32_t hi = uint32_t(bits(v) >> 32);
    const uint32_t j = (hi >> 13) & 127u;
    // E - 1055, plus one for rows j >= 53: adding 75 to j carries into E
    e[i] = int((hi + (75u << 13)) >> 20) - 1055;
    const Row row = rows[j];
    // m / 2 (the significand scaled into [1/2, 1)) times the row's 2 c
    r[i] = fma(frexp_mant(v), row.a, -1.0);  // exact: |r| <= 2^-7, m c has at most 52 bits
    lt[i] = row.b;
    p[i] = k.l_top;
  }
  // ln(1 + r) = r + r^2 pq(r)
  TWMC_UNROLL
  for (int j = 0; j < 6; ++j)
    TWMC_UNROLL
    for (int i = 0; i < N; ++i) p[i] = fma3(p[i], r[i], k.kl[j]);
  TWMC_UNROLL
  for (int i = 0; i < N; ++i) {
    const double t = lt[i] + fma(r[i] * r[i], p[i], r[i]);
    const double de = double(e[i]);
    out[i] = fma(de, kLn2Hi, fma(de, kLn2Lo, t));
  }
}

// N complex N(0, 1) samples from the Philox word pairs (a[i], b[i]):
// Re = sqrt(-2 ln u1) cos(2 pi theta), Im = sqrt(-2 ln u1) sin(2 pi theta).
template <int N>
TWMC_HD inline void box_muller(const uint32_t* a, const uint32_t* b, double* re, double* im,
                               const Series& k = kSeries, const double* tab = kTables) {
  TWMC_NO_CONTRACT
  double l[N], s[N], c[N];
  log_u32<N>(a, l, k, tab);
  sincos_turn32<N>(b, s, c, k, tab);
  TWMC_UNROLL
  for (int i = 0; i < N; ++i) {
    const double r = sqrt_normal(-2.0 * l[i]);
    re[i] = r * c[i];
    im[i] = r * s[i];
  }
}

}  // namespace f64
}  // namespace twmc

#endif
