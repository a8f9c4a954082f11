// twmc_fused64.hip -- register-resident TWMC step kernel for a 64 x 64
// periodic lattice in FP32 (the 2D uniform-hopping configuration).
//
// One wavefront (64 lanes) owns one trajectory: lane x holds row x of beta
// (64 complex values) and a 64-value FFT working row, both in registers.
// Per step (src/TWMC/Solvers/TWMCBaseSolver.cpp:232-256):
//   row FFT (in registers, radix 8 x 8)  -> LDS transpose -> column FFT
//   -> K' = i (-2J)(cos kx + cos ky) dt/N  (the reference's K, /N and dt folded)
//   -> inverse column FFT -> LDS transpose -> inverse row FFT = dt h
//   -> drift + Philox/Box-Muller increment + update, all in registers.
// The forward transforms leave frequencies in digit-reversed register slots
// (slot 8a+b holds frequency a+8b); the inverse transforms take exactly that
// order back to natural order, so no reordering pass exists anywhere.
// The only LDS traffic is the two transposes (32 B per site-step); HBM is
// touched once per launch (load and store of beta) plus the frame's FP64
// ensemble sums.  A persistent grid walks the batch, one trajectory per wave
// at a time.
#include "twmc_cplx.h"
#include "twmc_internal.h"
#include "twmc_rng.h"

namespace twmc {

namespace {

constexpr int kL = 64;
constexpr int kPad = 65;  // LDS row stride (complex): conflict-free column reads

// cos(2 pi m / 64) for m = 0..16; the rest follow by symmetry
constexpr float kCq[17] = {
    1.000000000e+00f, 9.951847267e-01f, 9.807852804e-01f, 9.569403357e-01f, 9.238795325e-01f,
    8.819212643e-01f, 8.314696123e-01f, 7.730104534e-01f, 7.071067812e-01f, 6.343932842e-01f,
    5.555702330e-01f, 4.713967368e-01f, 3.826834324e-01f, 2.902846773e-01f, 1.950903220e-01f,
    9.801714033e-02f, 0.0f};

__host__ __device__ constexpr float cos64(int m) {
  m &= 63;
  return m <= 16 ? kCq[m] : m <= 32 ? -kCq[32 - m] : m <= 48 ? -kCq[m - 32] : kCq[64 - m];
}
__host__ __device__ constexpr float sin64(int m) {
  m &= 63;
  return m <= 16 ? kCq[16 - m] : m <= 32 ? kCq[m - 16] : m <= 48 ? -kCq[48 - m] : -kCq[m - 48];
}
__host__ __device__ constexpr int rev8(int s) { return (s >> 3) | ((s & 7) << 3); }

// z * exp(S 2 pi i m / 64) with m a compile-time constant after unrolling
template <int S>
__device__ __forceinline__ cpx<float> twmul(cpx<float> z, int m) {
  if (m == 0) return z;
  if (m == 16) return mul_si<S>(z);
  if (m == 32) return mk(-z.re, -z.im);
  if (m == 48) return mul_si<-S>(z);
  const float c = cos64(m), s = float(S) * sin64(m);
  return mk(z.re * c - z.im * s, z.re * s + z.im * c);
}

// 64-point DFT in registers.  S = -1: natural order in, digit-reversed out.
// S = +1: digit-reversed in, natural order out (the exact mirror).
template <int S>
__device__ __forceinline__ void fft64(cpx<float> (&z)[kL]) {
  if constexpr (S < 0) {
#pragma unroll
    for (int n2 = 0; n2 < 8; ++n2)
      dft8<S>(z[n2], z[8 + n2], z[16 + n2], z[24 + n2], z[32 + n2], z[40 + n2], z[48 + n2], z[56 + n2]);
#pragma unroll
    for (int k1 = 1; k1 < 8; ++k1)
#pragma unroll
      for (int n2 = 1; n2 < 8; ++n2) z[8 * k1 + n2] = twmul<S>(z[8 * k1 + n2], k1 * n2);
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1)
      dft8<S>(z[8 * k1], z[8 * k1 + 1], z[8 * k1 + 2], z[8 * k1 + 3], z[8 * k1 + 4], z[8 * k1 + 5],
              z[8 * k1 + 6], z[8 * k1 + 7]);
  } else {
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1)
      dft8<S>(z[8 * k1], z[8 * k1 + 1], z[8 * k1 + 2], z[8 * k1 + 3], z[8 * k1 + 4], z[8 * k1 + 5],
              z[8 * k1 + 6], z[8 * k1 + 7]);
#pragma unroll
    for (int k1 = 1; k1 < 8; ++k1)
#pragma unroll
      for (int n2 = 1; n2 < 8; ++n2) z[8 * k1 + n2] = twmul<S>(z[8 * k1 + n2], k1 * n2);
#pragma unroll
    for (int n2 = 0; n2 < 8; ++n2)
      dft8<S>(z[n2], z[8 + n2], z[16 + n2], z[24 + n2], z[32 + n2], z[40 + n2], z[48 + n2], z[56 + n2]);
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// CAPTURE_ONLY instantiates the frame-capture launch that ends a batch (no
// steps) under its own symbol, so profiles of the step kernel hold only
// full segments.
template <bool CAPTURE_ONLY>
__global__ void __launch_bounds__(64, 1) fused64_step_kernel(StepArgs g) {
  __shared__ cpx<float> T[kL * kPad];
  const int lane = threadIdx.x;  // row x for row phases, column t for column phases

  // uniform site parameters (the kernel is chosen only when they are uniform)
  const cpx<float> r0 = reinterpret_cast<const cpx<float>*>(g.rsl)[0];
  const cpx<float> u0 = reinterpret_cast<const cpx<float>*>(g.U)[0];
  const cpx<float> f0 = reinterpret_cast<const cpx<float>*>(g.F)[0];
  const float dtf = float(g.dt);
  const float Ur = dtf * u0.re, Ui = dtf * u0.im;         // dt * U
  // dt (rsl + i U (|b|^2 - 1)) = (A0 - Ui |b|^2) + i (B0 + Ur |b|^2)
  const float A0 = fmaf(dtf, r0.re, Ui), B0 = fmaf(dtf, r0.im, -Ur);
  const float fr = -dtf * f0.im, fi = dtf * f0.re;        // dt * i F
  const float sig = float(g.dt4);
  const float nl = -1.3862943611198906f * sig * sig;      // -2 ln2 sigma^2 (Box-Muller)
  // K = i kscale (cos kx + cos ky); kscale dt / N is applied in the update
  const float kfac = float(g.kscale * g.dt * g.inv_n);
  const float cy = cosf(6.283185307179586f * float(rev8(lane)) / 64.0f);
  double* part = g.partials + (size_t)blockIdx.x * kQ * 4096;

  for (int b = blockIdx.x; b < g.n_traj; b += gridDim.x) {
    const uint64_t traj = g.traj_ids[b];
    cpx<float>* gb = reinterpret_cast<cpx<float>*>(g.beta) + (size_t)b * 4096;
    cpx<float> beta[kL];
    // coalesced load through LDS: lane x finally holds row x
#pragma unroll
    for (int r = 0; r < kL; ++r) T[r * kPad + lane] = gb[r * kL + lane];
    __syncthreads();
#pragma unroll
    for (int y = 0; y < kL; ++y) beta[y] = T[lane * kPad + y];
    __syncthreads();

    const int iters = CAPTURE_ONLY ? 1 : g.n_steps;
    for (int s = 0; s < iters; ++s) {
      // values the compiler must not hoist out of the step loop (they would
      // pin 64 + 96 registers for the whole loop)
      float cyv = cy;
      uint32_t pbase = (uint32_t)lane * 32u;
      asm volatile("" : "+v"(cyv), "+v"(pbase));

      cpx<float> z[kL];
#pragma unroll
      for (int y = 0; y < kL; ++y) z[y] = beta[y];
      fft64<-1>(z);  // along y: slot q holds ky = rev8(q)
#pragma unroll
      for (int q = 0; q < kL; ++q) T[q * kPad + lane] = z[q];
      __syncthreads();
#pragma unroll
      for (int x = 0; x < kL; ++x) z[x] = T[lane * kPad + x];
      __syncthreads();
      fft64<-1>(z);  // along x: slot q holds kx = rev8(q), ky = rev8(lane)

      if (s == 0 && g.capture) {
        // Frame capture through LDS so the FP64 accumulation runs as a rolled
        // loop: partial sums are written in natural row-major order.
#pragma unroll
        for (int q = 0; q < kL; ++q) T[q * kPad + lane] = z[q];
        __syncthreads();
        for (int r = 0; r < kL; ++r) {  // kx = rev8(r), ky = rev8(lane)
          const cpx<float> v = T[r * kPad + lane];
          part[2 * 4096 + rev8(r) * kL + rev8(lane)] +=
              double(v.re) * double(v.re) + double(v.im) * double(v.im);
        }
        __syncthreads();
#pragma unroll
        for (int y = 0; y < kL; ++y) T[lane * kPad + y] = beta[y];
        __syncthreads();
        double m2 = 0.0, m4 = 0.0;
        cpx<float>* out = g.dump ? reinterpret_cast<cpx<float>*>(g.traj_out) +
                                       ((size_t)b * g.n_frames + g.frame) * 4096
                                 : nullptr;
        for (int r = 0; r < kL; ++r) {  // site (x = r, y = lane)
          const cpx<float> v = T[r * kPad + lane];
          const int i = r * kL + lane;
          const double nn = double(v.re) * double(v.re) + double(v.im) * double(v.im);
          part[i] += nn;
          part[4096 + i] += nn * nn;
          part[3 * 4096 + i] += double(v.re);
          part[4 * 4096 + i] += double(v.im);
          m2 += nn;
          m4 += nn * nn;
          if (out) out[i] = v;
        }
        __syncthreads();
        m2 = wave_sum(m2);
        m4 = wave_sum(m4);
        if (lane == 0) {
          g.traj_m2[(size_t)b * g.n_frames + g.frame] = m2 / 4096.0;
          g.traj_m4[(size_t)b * g.n_frames + g.frame] = m4 / 4096.0;
        }
      }
      if (CAPTURE_ONLY) break;

#pragma unroll
      for (int q = 0; q < kL; ++q) {
        const float c = cos64(rev8(q)) + cyv;
        z[q] = mk(-c * z[q].im, c * z[q].re);
      }
      fft64<+1>(z);  // back along x: slot x natural
#pragma unroll
      for (int x = 0; x < kL; ++x) T[lane * kPad + x] = z[x];
      __syncthreads();
#pragma unroll
      for (int q = 0; q < kL; ++q) z[q] = T[q * kPad + lane];
      __syncthreads();
      fft64<+1>(z);  // back along y: kfac z[y] = dt h(x = lane, y)

      const uint32_t step = uint32_t(g.step0 + (uint64_t)s);
#pragma unroll
      for (int p = 0; p < kL / 2; ++p) {
        const u32x4 w = philox4x32_10(make_counter(traj, TAG_WIENER, step, pbase + (uint32_t)p),
                                      g.key0, g.key1);
        float rs[2], cs[2], sn[2];
        box_muller_f32_scaled(w.x, w.y, nl, rs[0], cs[0], sn[0]);
        box_muller_f32_scaled(w.z, w.w, nl, rs[1], cs[1], sn[1]);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int y = 2 * p + h;
          const cpx<float> v = beta[y];
          const float m = fmaf(v.im, v.im, v.re * v.re);
          const float cr = fmaf(-Ui, m, A0), ci = fmaf(Ur, m, B0);
          // increment = noise + dt iF + dt h + drift, all small, added to beta once
          float dr = fmaf(rs[h], cs[h], fr);
          dr = fmaf(kfac, z[y].re, dr);
          dr = fmaf(cr, v.re, dr);
          dr = fmaf(-ci, v.im, dr);
          float di = fmaf(rs[h], sn[h], fi);
          di = fmaf(kfac, z[y].im, di);
          di = fmaf(cr, v.im, di);
          di = fmaf(ci, v.re, di);
          beta[y] = mk(v.re + dr, v.im + di);
        }
      }
    }

#pragma unroll
    for (int y = 0; y < kL; ++y) T[lane * kPad + y] = beta[y];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kL; ++r) gb[r * kL + lane] = T[r * kPad + lane];
    __syncthreads();
  }
}

// Frame capture of the 64 x 64 stencil path.  Per trajectory: the site means
// of |beta|^2 and |beta|^4 and, if asked, the frame dump; and |DFT beta|^2
// summed over this wave's trajectories in FP64 registers (after the two
// register FFTs lane holds ky = rev8(lane), slot q holds kx = rev8(q)),
// written once per wave, in natural k order, as the k-space row of partial
// slot blockIdx.x.  The per-site real-space sums come from the site-major
// kernel beside it, so no partial row is read back here: every sum has a
// fixed trajectory order and repeated runs agree bit for bit.
__global__ void __launch_bounds__(64, 1) capture64_kernel(StepArgs g) {
  __shared__ cpx<float> T[kL * kPad];
  const int lane = threadIdx.x;
  double acc[kL];
#pragma unroll
  for (int q = 0; q < kL; ++q) acc[q] = 0.0;

  for (int b = blockIdx.x; b < g.n_traj; b += gridDim.x) {
    const cpx<float>* gb = reinterpret_cast<const cpx<float>*>(g.beta) + (size_t)b * 4096;
    // coalesced load through LDS: row r, column lane
#pragma unroll
    for (int r = 0; r < kL; ++r) T[r * kPad + lane] = gb[r * kL + lane];
    __syncthreads();
    if (g.dump) {
      cpx<float>* out = reinterpret_cast<cpx<float>*>(g.traj_out) + ((size_t)b * g.n_frames + g.frame) * 4096;
#pragma unroll
      for (int r = 0; r < kL; ++r) out[r * kL + lane] = T[r * kPad + lane];
    }
    cpx<float> z[kL];  // lane x holds row x
#pragma unroll
    for (int y = 0; y < kL; ++y) z[y] = T[lane * kPad + y];
    __syncthreads();

    double m2 = 0.0, m4 = 0.0;
#pragma unroll
    for (int y = 0; y < kL; ++y) {
      const double nn = double(z[y].re) * double(z[y].re) + double(z[y].im) * double(z[y].im);
      m2 += nn;
      m4 += nn * nn;
    }
    m2 = wave_sum(m2);
    m4 = wave_sum(m4);
    if (lane == 0) {
      g.traj_m2[(size_t)b * g.n_frames + g.frame] = m2 / 4096.0;
      g.traj_m4[(size_t)b * g.n_frames + g.frame] = m4 / 4096.0;
    }

    fft64<-1>(z);  // along y: slot q holds ky = rev8(q)
#pragma unroll
    for (int q = 0; q < kL; ++q) T[q * kPad + lane] = z[q];
    __syncthreads();
#pragma unroll
    for (int x = 0; x < kL; ++x) z[x] = T[lane * kPad + x];
    __syncthreads();
    fft64<-1>(z);  // along x: slot q holds kx = rev8(q), ky = rev8(lane)
#pragma unroll
    for (int q = 0; q < kL; ++q) acc[q] += double(z[q].re) * double(z[q].re) + double(z[q].im) * double(z[q].im);
  }
  double* part = g.partials + (size_t)blockIdx.x * kQ * 4096 + 2 * 4096;
#pragma unroll
  for (int q = 0; q < kL; ++q) part[rev8(q) * kL + rev8(lane)] = acc[q];
}

}  // namespace

hipError_t launch_capture64(const StepArgs& a, int grid, hipStream_t s) {
  hipLaunchKernelGGL(capture64_kernel, dim3(grid), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t capture64_occupancy(int* blocks_per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void*)capture64_kernel, 64, 0);
}

bool fused64_supported(int nx, int ny, int hopping, int precision, int param_stride) {
  return nx == 64 && ny == 64 && hopping == 0 && precision == 32 && param_stride == 0;
}

hipError_t launch_fused64(const StepArgs& a, int grid, hipStream_t s) {
  if (a.n_steps > 0)
    hipLaunchKernelGGL(fused64_step_kernel<false>, dim3(grid), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL(fused64_step_kernel<true>, dim3(grid), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t fused64_occupancy(int* blocks_per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void*)fused64_step_kernel<false>, 64, 0);
}

// the kernel writes its partial sums in natural order
void fused64_perms(int* perm_site, int* perm_k) {
  for (int j = 0; j < 4096; ++j) perm_site[j] = perm_k[j] = j;
}

}  // namespace twmc
