// twmc_cplx.h -- small complex helpers and fixed-radix DFT butterflies used by
// the step kernels.  S = -1 is the forward transform (e^{-i}), S = +1 the
// backward one (e^{+i}); outputs are in natural order.
#ifndef TWMC_CPLX_H
#define TWMC_CPLX_H

#include <hip/hip_runtime.h>

namespace twmc {

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
  T re, im;
};

template <typename T>
__host__ __device__ __forceinline__ cpx<T> mk(T r, T i) {
  cpx<T> z;
  z.re = r;
  z.im = i;
  return z;
}
template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return mk(a.re + b.re, a.im + b.im); }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return mk(a.re - b.re, a.im - b.im); }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T>
__device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return mk(a.re * s, a.im * s); }
// multiply by S i
template <int S, typename T>
__device__ __forceinline__ cpx<T> mul_si(cpx<T> z) {
  return S < 0 ? mk(z.im, -z.re) : mk(-z.im, z.re);
}

template <int S, typename T>
__device__ __forceinline__ void dft2(cpx<T>* v) {
  cpx<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int S, typename T>
__device__ __forceinline__ void dft4(cpx<T>& v0, cpx<T>& v1, cpx<T>& v2, cpx<T>& v3) {
  cpx<T> t0 = cadd(v0, v2), t1 = csub(v0, v2);
  cpx<T> t2 = cadd(v1, v3), t3 = mul_si<S>(csub(v1, v3));
  v0 = cadd(t0, t2);
  v2 = csub(t0, t2);
  v1 = cadd(t1, t3);
  v3 = csub(t1, t3);
}

template <int S, typename T>
__device__ __forceinline__ void dft4(cpx<T>* v) { dft4<S>(v[0], v[1], v[2], v[3]); }

// radix 8 = two radix-4 halves combined with W8^q, W8 = exp(S 2 pi i / 8)
template <int S, typename T>
__device__ __forceinline__ void dft8(cpx<T>& v0, cpx<T>& v1, cpx<T>& v2, cpx<T>& v3, cpx<T>& v4,
                                     cpx<T>& v5, cpx<T>& v6, cpx<T>& v7) {
  const T c = T(0.70710678118654752440);
  dft4<S>(v0, v2, v4, v6);
  dft4<S>(v1, v3, v5, v7);
  // odd half times W8^q
  cpx<T> o1 = S < 0 ? mk(c * (v3.re + v3.im), c * (v3.im - v3.re)) : mk(c * (v3.re - v3.im), c * (v3.im + v3.re));
  cpx<T> o2 = mul_si<S>(v5);
  cpx<T> o3 = S < 0 ? mk(c * (v7.im - v7.re), -c * (v7.re + v7.im)) : mk(-c * (v7.re + v7.im), c * (v7.re - v7.im));
  cpx<T> e0 = v0, e1 = v2, e2 = v4, e3 = v6, o0 = v1;
  v0 = cadd(e0, o0);
  v4 = csub(e0, o0);
  v1 = cadd(e1, o1);
  v5 = csub(e1, o1);
  v2 = cadd(e2, o2);
  v6 = csub(e2, o2);
  v3 = cadd(e3, o3);
  v7 = csub(e3, o3);
}

template <int S, typename T>
__device__ __forceinline__ void dft8(cpx<T>* v) { dft8<S>(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]); }

template <int S, typename T>
__device__ __forceinline__ void dft3(cpx<T>* v) {
  const T s3 = T(S) * T(0.86602540378443864676);
  cpx<T> t = cadd(v[1], v[2]), d = csub(v[1], v[2]);
  cpx<T> m = mk(v[0].re - T(0.5) * t.re, v[0].im - T(0.5) * t.im);
  v[0] = cadd(v[0], t);
  v[1] = mk(m.re - s3 * d.im, m.im + s3 * d.re);
  v[2] = mk(m.re + s3 * d.im, m.im - s3 * d.re);
}

template <int S, typename T>
__device__ __forceinline__ void dft5(cpx<T>* v) {
  const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
  const T s1 = T(S) * T(0.95105651629515357212), s2 = T(S) * T(0.58778525229247312917);
  cpx<T> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
  cpx<T> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
  cpx<T> m1 = mk(v[0].re + c1 * a1.re + c2 * a2.re, v[0].im + c1 * a1.im + c2 * a2.im);
  cpx<T> m2 = mk(v[0].re + c2 * a1.re + c1 * a2.re, v[0].im + c2 * a1.im + c1 * a2.im);
  cpx<T> n1 = mk(s1 * b1.re + s2 * b2.re, s1 * b1.im + s2 * b2.im);
  cpx<T> n2 = mk(s2 * b1.re - s1 * b2.re, s2 * b1.im - s1 * b2.im);
  v[0] = mk(v[0].re + a1.re + a2.re, v[0].im + a1.im + a2.im);
  v[1] = mk(m1.re - n1.im, m1.im + n1.re);
  v[4] = mk(m1.re + n1.im, m1.im - n1.re);
  v[2] = mk(m2.re - n2.im, m2.im + n2.re);
  v[3] = mk(m2.re + n2.im, m2.im - n2.re);
}

}  // namespace twmc

#endif
