// Accuracy of the FP64 Box-Muller pieces in csrc/twmc_f64math.h against x87
// long double (64-bit significand) references, in units in the last place of
// the double result.  Prints one line per function:
//   <name> <max ulp> <samples>
// and exits non-zero when a bound is exceeded or a value is not finite.
// Inputs: every 32-bit a / b of the form k * 4093 + (k * 2654435761 mod 4093)
// for k below 2^20 (1M values spread over the whole range), plus the edges
// 0, 1, 2^32 - 1, every power of two, the eighth turns, the sin / cos
// table's 1/256 turns and the midpoints between them (where the row switches)
// with their neighbours, a + 1/2 at and just above sqrt(2) 2^p, and at the ln
// table's row boundaries (1 + j/128) 2^p.
#include <math.h>
#include <stdint.h>
#include <stdio.h>

#include <vector>

#include "twmc_f64math.h"

static double ulp_err(double got, long double want) {
  if (!isfinite(got)) return INFINITY;
  const double w = (double)want;
  const double ulp = nextafter(fabs(w), INFINITY) - fabs(w);
  return (double)(fabsl((long double)got - want) / (long double)ulp);
}

int main() {
  std::vector<uint32_t> xs;
  for (uint64_t k = 0; k < (1u << 20); ++k)
    xs.push_back((uint32_t)(k * 4093u + (uint32_t)((k * 2654435761ull) % 4093u)));
  const uint32_t edges[] = {0u, 1u, 2u, 0xffffffffu, 0xfffffffeu, 0x7fffffffu, 0x80000000u, 0x80000001u};
  for (uint32_t e : edges) xs.push_back(e);
  for (int p = 0; p < 32; ++p) {
    const uint32_t v = 1u << p;
    xs.push_back(v);
    xs.push_back(v - 1u);
    xs.push_back(v + 1u);
  }
  for (uint32_t j = 0; j < 8; ++j)  // eighth turns
    for (int d = -2; d <= 2; ++d) xs.push_back(j * 0x20000000u + (uint32_t)d);
  for (uint32_t j = 0; j < 256; ++j)  // the sin / cos table's rows and the midpoints between them
    for (int d = -2; d <= 2; ++d) {
      xs.push_back(j * 0x1000000u + (uint32_t)d);
      xs.push_back(j * 0x1000000u + 0x800000u + (uint32_t)d);
    }
  for (uint32_t j = 0; j < 128; ++j)  // the ln table's row boundaries: a + 1/2 = (1 + j/128) 2^p
    for (int p = 8; p < 32; p += 3) {
      const uint32_t c = (uint32_t)((1.0 + j / 128.0) * (double)(1u << p));
      for (int d = -1; d <= 1; ++d) xs.push_back(c + (uint32_t)d);
    }
  for (int p = 1; p < 32; ++p) {  // ln's reduction boundary: a + 1/2 near sqrt(2) 2^p
    const uint32_t c = (uint32_t)(1.4142135623730951 * (double)(1u << p));
    for (int d = -3; d <= 3; ++d) xs.push_back(c + (uint32_t)d);
    for (uint32_t d = 0; d < 4000u; d += 37u) xs.push_back(c + d);
  }

  const long double two_pi = 6.283185307179586476925286766559L;
  double e_log = 0, e_sin = 0, e_cos = 0, e_bm = 0;
  for (uint32_t a : xs) {
    const long double u = ((long double)a + 0.5L) / 4294967296.0L;
    const long double want = logl(u);
    double got;
    twmc::f64::log_u32<1>(&a, &got);
    e_log = fmax(e_log, ulp_err(got, want));
  }
  for (uint32_t b : xs) {
    const long double ang = two_pi * ((long double)b / 4294967296.0L);
    double s, c;
    twmc::f64::sincos_turn32<1>(&b, &s, &c);
    // absolute error in ulps of 1/2 near the zeros of sin / cos, relative elsewhere
    const long double ws = sinl(ang), wc = cosl(ang);
    e_sin = fmax(e_sin, fabsl(ws) < 0.5L ? (double)(fabsl((long double)s - ws) / 0x1p-54L) : ulp_err(s, ws));
    e_cos = fmax(e_cos, fabsl(wc) < 0.5L ? (double)(fabsl((long double)c - wc) / 0x1p-54L) : ulp_err(c, wc));
  }
  for (size_t i = 0; i + 1 < xs.size(); i += 2) {
    // two samples side by side, as the kernels evaluate them
    uint32_t a[2], b[2];
    double re[2], im[2];
    for (int h = 0; h < 2; ++h) {
      a[h] = xs[i + h];
      b[h] = xs[((i + h) * 7919u + 13u) % xs.size()];
    }
    twmc::f64::box_muller<2>(a, b, re, im);
    for (int h = 0; h < 2; ++h) {
      const long double u = ((long double)a[h] + 0.5L) / 4294967296.0L;
      const long double r = sqrtl(-2.0L * logl(u));
      const long double ang = two_pi * ((long double)b[h] / 4294967296.0L);
      // absolute error in ulps of the radius: the sample's own scale
      const long double scale = 0x1p-52L * fmaxl(r, 1e-300L);
      e_bm = fmax(e_bm, (double)(fmaxl(fabsl(re[h] - r * cosl(ang)), fabsl(im[h] - r * sinl(ang))) / scale));
    }
  }
  printf("log_u32 %.3f %zu\n", e_log, xs.size());
  printf("sin_turn32 %.3f %zu\n", e_sin, xs.size());
  printf("cos_turn32 %.3f %zu\n", e_cos, xs.size());
  printf("box_muller %.3f %zu\n", e_bm, xs.size());
  const bool ok = e_log <= 2.0 && e_sin <= 2.0 && e_cos <= 2.0 && e_bm <= 4.0;
  return ok ? 0 : 1;
}
