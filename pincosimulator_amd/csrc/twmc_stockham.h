// twmc_stockham.h -- LDS-resident mixed-radix Stockham FFT passes shared by
// the LDS step kernels (generic and batched sweep).  A workgroup transforms a
// batch of sequences held in LDS, ping-ponging between two buffers.
#ifndef TWMC_STOCKHAM_H
#define TWMC_STOCKHAM_H

#include "twmc_cplx.h"
#include "twmc_internal.h"

namespace twmc {

template <int S, typename T>
__device__ __forceinline__ cpx<T> twiddle(const cpx<T>* tw, int m) {
  cpx<T> w = tw[m];
  return S < 0 ? w : mk(w.re, -w.im);
}

// One Stockham pass of radix R over `batch` transforms of length L laid out
// with element stride es and batch stride bs.  Butterfly j reads
// in[j + r L/R], applies exp(S 2 pi i r (j mod Ns)/(Ns R)), and writes
// out[(j/Ns) Ns R + (j mod Ns) + q Ns].
template <int S, int R, typename T>
__device__ __forceinline__ void stockham_pass(const cpx<T>* __restrict__ src, cpx<T>* __restrict__ dst,
                                              int L, int Ns, int batch, int bs, int es,
                                              const cpx<T>* tw) {
  const int LR = L / R;
  const int nbf = batch * LR;
  const int tstep = L / (Ns * R);
  for (int w = threadIdx.x; w < nbf; w += blockDim.x) {
    int b = w / LR;
    int j = w - b * LR;
    int k = j % Ns;
    const cpx<T>* in = src + b * bs;
    cpx<T>* out = dst + b * bs;
    cpx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      v[r] = in[(j + r * LR) * es];
      if (r > 0 && k > 0) v[r] = cmul(v[r], twiddle<S>(tw, r * k * tstep));
    }
    if constexpr (R == 2) dft2<S>(v);
    if constexpr (R == 3) dft3<S>(v);
    if constexpr (R == 4) dft4<S>(v);
    if constexpr (R == 5) dft5<S>(v);
    if constexpr (R == 8) dft8<S>(v);
    int base = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int q = 0; q < R; ++q) out[(base + q * Ns) * es] = v[q];
  }
}

// A pass of any radix R, used for the prime factors above 5 (the length-R
// DFT as its direct sum).  The pass twiddle and the DFT root fold into one
// table entry: output q of butterfly j is
//     sum_r in[j + r L/R] exp(S 2 pi i ((r k tstep + (q r mod R) L/R) mod L) / L)
// with k = j mod Ns, tstep = L / (Ns R); r k tstep < L and (q r mod R) L/R < L,
// so one subtraction keeps the index inside the length-L table.
template <int S, typename T>
__device__ __forceinline__ void stockham_pass_any(const cpx<T>* __restrict__ src, cpx<T>* __restrict__ dst,
                                                  int L, int Ns, int R, int batch, int bs, int es,
                                                  const cpx<T>* tw) {
  const int LR = L / R;
  const int nbf = batch * LR;
  const int tstep = L / (Ns * R);
  for (int w = threadIdx.x; w < nbf; w += blockDim.x) {
    int b = w / LR;
    int j = w - b * LR;
    int k = j % Ns;
    const cpx<T>* in = src + b * bs;
    cpx<T>* out = dst + b * bs;
    const int base = (j / Ns) * Ns * R + k;
    for (int q = 0; q < R; ++q) {
      cpx<T> acc = in[j * es];
      int qr = 0;  // (q r) mod R
      for (int r = 1; r < R; ++r) {
        qr += q;
        if (qr >= R) qr -= R;
        int m = r * k * tstep + qr * LR;
        if (m >= L) m -= L;
        acc = cadd(acc, cmul(in[(j + r * LR) * es], twiddle<S>(tw, m)));
      }
      out[(base + q * Ns) * es] = acc;
    }
  }
}

// Full 1D transforms along one dimension; returns the buffer holding the result.
template <int S, typename T>
__device__ cpx<T>* fft_dim(cpx<T>* src, cpx<T>* dst, const RadixPlan& P, int batch, int bs, int es,
                           const cpx<T>* tw) {
  int Ns = 1;
  for (int ip = 0; ip < P.npass; ++ip) {
    int R = P.radix[ip];
    switch (R) {
      case 8: stockham_pass<S, 8>(src, dst, P.len, Ns, batch, bs, es, tw); break;
      case 4: stockham_pass<S, 4>(src, dst, P.len, Ns, batch, bs, es, tw); break;
      case 2: stockham_pass<S, 2>(src, dst, P.len, Ns, batch, bs, es, tw); break;
      case 3: stockham_pass<S, 3>(src, dst, P.len, Ns, batch, bs, es, tw); break;
      case 5: stockham_pass<S, 5>(src, dst, P.len, Ns, batch, bs, es, tw); break;
      default: stockham_pass_any<S>(src, dst, P.len, Ns, R, batch, bs, es, tw); break;
    }
    __syncthreads();
    cpx<T>* t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  return src;
}

}  // namespace twmc

#endif
