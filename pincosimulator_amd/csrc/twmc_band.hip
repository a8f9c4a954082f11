// twmc_band.hip -- register-resident TWMC step kernel for matrix hopping with
// a banded coupling matrix (chains, ladders, the Lieb chain) in FP32.
//
// The reference's Lieb solver applies h_j = -i sum_i beta_i C_ij with a sparse
// product (TWMCLiebSolver.cpp:259-267).  When every nonzero C_ij has
// |i - j| <= W (cyclically, when N = 64 S), that product is a (2W+1)-point
// stencil along the site index, and a dense or LDS-staged product only moves
// zeros.  One wavefront owns one trajectory: lane l holds sites l S .. l S + S - 1
// in registers (S even; N padded to 64 S, the padding sites have zero
// coefficients and are never stored), together with the coefficients
// dt C_(j+d) j of its sites and its dt-scaled site parameters.  The W sites on
// either side of a lane's block come from lanes l - 1 and l + 1 by whole-wave
// DPP rotations.  Per step: the hopping sums of all owned sites from the old
// state, then per site pair one Philox4x32-10 call with Box-Muller and the
// update, the same FMA chain as the other FP32 kernels.
#include "twmc_cplx.h"
#include "twmc_internal.h"
#include "twmc_rng.h"

namespace twmc {

namespace {

// lane l reads lane l - 1 (DPP wave_ror:1) / lane l + 1 (wave_rol:1), wrapping at 64
__device__ __forceinline__ float from_prev_lane(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x13c, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_next_lane(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x134, 0xf, 0xf, true));
}

// the Lieb chain's bonds as offsets d of the coupled site j + d, by the site
// index mod 3 (A_i = 3i, B_i = 3i + 1, C_i = 3i + 2; TWMCLiebSolver.cpp:52-85):
// A-B (+1); B-A (-1), B-C (+1), B-C of the previous cell (-2); C-B (-1),
// C-B of the next cell (+2)
__host__ __device__ constexpr bool lieb_bond(int r, int d) {
  return r == 0 ? d == 1 : r == 1 ? (d == -1 || d == 1 || d == -2) : (d == -1 || d == 2);
}

// LIEB (S = 6, W = 2): each lane's six sites are two whole cells, so the
// offsets without a bond are known per register and their multiplications
// by zero are left out (6 of the 10 complex products per cell remain)
template <int S, int W, bool FT, bool LIEB>
__global__ void __launch_bounds__(64) band_step_kernel(StepArgs g) {
  static_assert(!LIEB || (S == 6 && W == 2), "the Lieb pattern needs two cells per lane");
  constexpr int D = 2 * W + 1;
  const int lane = threadIdx.x;
  const int n = g.n;
  const int base = lane * S;
  const float dt = float(g.dt), sig = float(g.dt4);
  const float nl = -1.3862943611198906f * sig * sig;  // -2 ln2 sigma^2 (Box-Muller)

  // this lane's coefficients and site tables; the site parameters are shared
  // by all trajectories (the kernel is chosen only then):
  //   dt (rsl + i U (|b|^2 - 1)) = (A0 - Ui |b|^2) + i (B0 + Ur |b|^2), dt i F = (fr, fi)
  float cf[S][D], A0[S], B0[S], Ur[S], Ui[S], fr[S], fi[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int j = base + s;
#pragma unroll
    for (int d = 0; d < D; ++d) cf[s][d] = g.band_coef[(size_t)j * D + d];
    const bool live = j < n;
    const cpx<float> r = live ? reinterpret_cast<const cpx<float>*>(g.rsl)[j] : mk(0.f, 0.f);
    const cpx<float> u = live ? reinterpret_cast<const cpx<float>*>(g.U)[j] : mk(0.f, 0.f);
    const cpx<float> f = live ? reinterpret_cast<const cpx<float>*>(g.F)[j] : mk(0.f, 0.f);
    A0[s] = dt * (r.re + u.im);
    B0[s] = dt * (r.im - u.re);
    Ur[s] = dt * u.re;
    Ui[s] = dt * u.im;
    fr[s] = -dt * f.im;
    fi[s] = dt * f.re;
  }

  for (int b = blockIdx.x; b < g.n_traj; b += gridDim.x) {
    const uint64_t traj = g.traj_ids[b];
    cpx<float>* gb = reinterpret_cast<cpx<float>*>(g.beta) + (size_t)b * n;
    float br[S], bi[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int j = base + s;
      const cpx<float> v = j < n ? gb[j] : mk(0.f, 0.f);
      br[s] = v.re;
      bi[s] = v.im;
    }

    if (g.capture) {
      // per-trajectory site means and the optional frame dump; the per-site
      // ensemble sums come from the site-major kernel launched before this one
      double m2 = 0.0, m4 = 0.0;
      cpx<float>* out =
          g.dump ? reinterpret_cast<cpx<float>*>(g.traj_out) + ((size_t)b * g.n_frames + g.frame) * n : nullptr;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const int j = base + s;
        if (j < n) {
          const double nn = double(br[s]) * double(br[s]) + double(bi[s]) * double(bi[s]);
          m2 += nn;
          m4 += nn * nn;
          if (out) out[j] = mk(br[s], bi[s]);
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        m2 += __shfl_xor(m2, o, 64);
        m4 += __shfl_xor(m4, o, 64);
      }
      if (lane == 0) {
        g.traj_m2[(size_t)b * g.n_frames + g.frame] = m2 / double(n);
        g.traj_m4[(size_t)b * g.n_frames + g.frame] = m4 / double(n);
      }
    }

    int ft_id = 0;
    for (int st = 0; st < g.n_steps; ++st) {
      const uint32_t step = uint32_t(g.step0 + (uint64_t)st);
      // site-pair base of this lane, opaque so the Philox rounds stay in the loop
      uint32_t pbase = (uint32_t)(base / 2);
      asm volatile("" : "+v"(pbase));
      float dr[S], di[S];
      if constexpr (FT) {
        // F(t) before this step (TWMCLiebSolver.cpp:252-256), as dt i F
        const float fdt = float(drive_interval(g.ft_times, g.ft_n, g.t0 + double(st) * g.dt, ft_id));
        const cpx<float>* fv = reinterpret_cast<const cpx<float>*>(g.ft_vals) + (size_t)ft_id * n;
        const cpx<float>* fs = reinterpret_cast<const cpx<float>*>(g.ft_slopes) + (size_t)ft_id * n;
#pragma unroll
        for (int s = 0; s < S; ++s) {
          const int j = base + s;
          const cpx<float> v = j < n ? fv[j] : mk(0.f, 0.f), sl = j < n ? fs[j] : mk(0.f, 0.f);
          dr[s] = -dt * fmaf(fdt, sl.im, v.im);
          di[s] = dt * fmaf(fdt, sl.re, v.re);
        }
      } else {
#pragma unroll
        for (int s = 0; s < S; ++s) {
          dr[s] = fr[s];
          di[s] = fi[s];
        }
      }
      // old values of the W sites before and after this lane's block
      float lre[W], lim[W], rre[W], rim[W];
#pragma unroll
      for (int k = 0; k < W; ++k) {
        lre[k] = from_prev_lane(br[S - W + k]);
        lim[k] = from_prev_lane(bi[S - W + k]);
        rre[k] = from_next_lane(br[k]);
        rim[k] = from_next_lane(bi[k]);
      }
      // dt h_j = -i sum_d dt C_(j+d) j beta_(j+d): (sum c Im beta, -sum c Re beta)
#pragma unroll
      for (int s = 0; s < S; ++s)
#pragma unroll
        for (int d = -W; d <= W; ++d) {
          if (LIEB && !lieb_bond(s % 3, d)) continue;
          const int t = s + d;
          const float vre = t < 0 ? lre[W + t] : t >= S ? rre[t - S] : br[t];
          const float vim = t < 0 ? lim[W + t] : t >= S ? rim[t - S] : bi[t];
          dr[s] = fmaf(cf[s][d + W], vim, dr[s]);
          di[s] = fmaf(-cf[s][d + W], vre, di[s]);
        }
      // noise and update
#pragma unroll
      for (int p = 0; p < S / 2; ++p) {
        const u32x4 w = philox4x32_10(make_counter(traj, TAG_WIENER, step, pbase + (uint32_t)p), g.key0, g.key1);
        float rn[2], cs[2], sn[2];
        box_muller_f32_scaled(w.x, w.y, nl, rn[0], cs[0], sn[0]);
        box_muller_f32_scaled(w.z, w.w, nl, rn[1], cs[1], sn[1]);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int s = 2 * p + h;
          const float x = br[s], y = bi[s];
          const float mq = fmaf(y, y, x * x);
          const float cr = fmaf(-Ui[s], mq, A0[s]), ci = fmaf(Ur[s], mq, B0[s]);
          // increment = noise + dt iF + dt h + drift, all small, added to beta once
          float ur = fmaf(rn[h], cs[h], dr[s]);
          ur = fmaf(cr, x, ur);
          ur = fmaf(-ci, y, ur);
          float ui = fmaf(rn[h], sn[h], di[s]);
          ui = fmaf(cr, y, ui);
          ui = fmaf(ci, x, ui);
          br[s] = x + ur;
          bi[s] = y + ui;
        }
      }
    }

#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int j = base + s;
      if (j < n) gb[j] = mk(br[s], bi[s]);
    }
  }
}

template <int S, int W>
hipError_t launch_sw(const StepArgs& a, int grid, hipStream_t s) {
  if constexpr (S == 6 && W == 2) {
    if (a.band_lieb) {
      if (a.ft_n > 0)
        hipLaunchKernelGGL((band_step_kernel<6, 2, true, true>), dim3(grid), dim3(64), 0, s, a);
      else
        hipLaunchKernelGGL((band_step_kernel<6, 2, false, true>), dim3(grid), dim3(64), 0, s, a);
      return hipGetLastError();
    }
  }
  if (a.ft_n > 0)
    hipLaunchKernelGGL((band_step_kernel<S, W, true, false>), dim3(grid), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL((band_step_kernel<S, W, false, false>), dim3(grid), dim3(64), 0, s, a);
  return hipGetLastError();
}

template <int S>
hipError_t launch_s(const StepArgs& a, int grid, hipStream_t s) {
  switch (a.band_w) {
    case 1: return launch_sw<S, 1>(a, grid, s);
    case 2: return launch_sw<S, 2>(a, grid, s);
    default: return launch_sw<S, 3>(a, grid, s);
  }
}

}  // namespace

int band_sites_per_lane(int n) { return 2 * ((n + 127) / 128); }

int band_width(const double* coupling, int n) {
  if (n < 1 || n > kBandMaxSites) return -1;
  const int S = band_sites_per_lane(n);
  const bool cyclic = 64 * S == n;
  int w = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      if (coupling[(size_t)i * n + j] == 0.0) continue;
      int d = i > j ? i - j : j - i;
      if (cyclic) d = d < n - d ? d : n - d;
      w = d > w ? d : w;
    }
  return w;
}

bool band_supported(int n, int hopping, int precision, int param_stride, int width) {
  return hopping == 1 && precision == 32 && param_stride == 0 && n >= 1 && n <= kBandMaxSites && width >= 0 &&
         width <= kBandMaxWidth && width <= band_sites_per_lane(n);
}

void band_coefficients(const double* coupling, int n, int w, double dt, float* out) {
  const int S = band_sites_per_lane(n), D = 2 * w + 1;
  const bool cyclic = 64 * S == n;
  for (int j = 0; j < 64 * S; ++j)
    for (int d = -w; d <= w; ++d) {
      int i = j + d;
      double c = 0.0;
      if (j < n) {
        if (cyclic) i = (i % n + n) % n;
        if (i >= 0 && i < n) c = coupling[(size_t)i * n + j];
      }
      out[(size_t)j * D + d + w] = float(dt * c);
    }
}

bool band_lieb_pattern(const float* coef, int n, int w) {
  const int S = band_sites_per_lane(n);
  if (S != 6 || w != 2) return false;
  for (int j = 0; j < 64 * S; ++j)
    for (int d = -w; d <= w; ++d)
      if (!lieb_bond(j % 3, d) && coef[(size_t)j * (2 * w + 1) + d + w] != 0.f) return false;
  return true;
}

hipError_t launch_band(const StepArgs& a, int grid, hipStream_t s) {
  switch (band_sites_per_lane(a.n)) {
    case 2: return launch_s<2>(a, grid, s);
    case 4: return launch_s<4>(a, grid, s);
    case 6: return launch_s<6>(a, grid, s);
    default: return launch_s<8>(a, grid, s);
  }
}

// occupancy of the instantiation launch_band runs for (n, w, lieb) (the
// FT = false variant, the larger of the two)
template <int S>
const void* band_kernel_for(int w, int lieb) {
  if constexpr (S == 6) {
    if (lieb && w == 2) return (const void*)band_step_kernel<6, 2, false, true>;
  }
  switch (w) {
    case 1: return (const void*)band_step_kernel<S, 1, false, false>;
    case 2: return (const void*)band_step_kernel<S, 2, false, false>;
    default: return (const void*)band_step_kernel<S, 3, false, false>;
  }
}

hipError_t band_occupancy(int n, int w, int lieb, int* blocks_per_cu) {
  const void* f = nullptr;
  switch (band_sites_per_lane(n)) {
    case 2: f = band_kernel_for<2>(w, lieb); break;
    case 4: f = band_kernel_for<4>(w, lieb); break;
    case 6: f = band_kernel_for<6>(w, lieb); break;
    default: f = band_kernel_for<8>(w, lieb); break;
  }
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, 64, 0);
}

}  // namespace twmc
