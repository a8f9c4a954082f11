// twmc_stencil64.hip -- register-resident TWMC step kernel for a 64 x 64
// periodic lattice in FP32 with the hopping applied as its exact real-space
// form.
//
// The reference applies the hopping in k-space (TWMCBaseSolver.cpp:80-88,
// :235-242): h = IDFT(K . DFT(beta)) / N with
// K = -2iJ([nx != 1] cos(2 pi kx / nx) + [ny != 1] cos(2 pi ky / ny)).
// Because e^{ik} + e^{-ik} = 2 cos k, that product is, term by term, the
// nearest-neighbour sum
//     h(x, y) = -iJ (beta(x-1, y) + beta(x+1, y) + beta(x, y-1) + beta(x, y+1))
// with periodic wrap -- the same linear map, without the FFT round trip (the
// two agree to rounding, ~1e-16 relative in FP64).  This kernel evaluates that
// form: one wavefront owns one trajectory, lane x holds row x (64 complex
// values, most in registers, the last few in the wave's LDS slice);
// y-neighbours are neighbouring sites of the row, x-neighbours come from
// lanes x-1 and x+1 (whole-wave DPP rotations).  Per step and site: the
// 4-term sum, the local drift, one Philox4x32-10 call per site pair with
// Box-Muller, the update.  Frames are captured by separate launches (the
// register-FFT capture kernel for the per-trajectory means and the k-space
// sums, the site-major kernel for the per-site sums), so this loop carries no
// capture code.
#include "twmc_cplx.h"
#include "twmc_internal.h"
#include "twmc_rng.h"

namespace twmc {

namespace {

constexpr int kL = 64;

// Whole-wave rotations by one lane (DPP wave_rol:1 / wave_ror:1): lane x
// reads lane x + 1 or x - 1, wrapping at 64.  Used as the first operand of an
// add, the compiler folds each into a v_add_f32_dpp, so the x-neighbours cost
// no LDS traffic and no separate instruction.
__device__ __forceinline__ float rot_up(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x134, 0xf, 0xf, true));
}
__device__ __forceinline__ float rot_down(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x13c, 0xf, 0xf, true));
}

// UI0: Im U = 0 (a real interaction), so the real part of the drift factor is
// the constant A0 and its per-site FMA drops out.  R: sites y < R of the row
// live in registers, the other kL - R in LDS ([y - R][lane], 512 bytes per
// site and wave, conflict-free); with R = 46 the kernel fits four waves per
// SIMD, so a batch of 16 trajectories per SIMD runs in four full rounds
template <bool UI0, int R>
__global__ void __launch_bounds__(64, R < kL ? 4 : 2) stencil64_step_kernel(StepArgs g) {
  extern __shared__ __align__(16) unsigned char smem[];
  cpx<float>* const row_lds = reinterpret_cast<cpx<float>*>(smem);  // [kL - R][64]
  const int lane = threadIdx.x;  // row x

  // uniform site parameters (the kernel is chosen only when they are uniform)
  const cpx<float> r0 = reinterpret_cast<const cpx<float>*>(g.rsl)[0];
  const cpx<float> u0 = reinterpret_cast<const cpx<float>*>(g.U)[0];
  const cpx<float> f0 = reinterpret_cast<const cpx<float>*>(g.F)[0];
  const float dtf = float(g.dt);
  const float Ur = dtf * u0.re, Ui = dtf * u0.im;   // dt * U
  // dt (rsl + i U (|b|^2 - 1)) = (A0 - Ui |b|^2) + i (B0 + Ur |b|^2)
  const float A0 = fmaf(dtf, r0.re, Ui), B0 = fmaf(dtf, r0.im, -Ur);
  const float fr = -dtf * f0.im, fi = dtf * f0.re;  // dt * i F
  const float sig = float(g.dt4);
  const float nl = -1.3862943611198906f * sig * sig;  // -2 ln2 sigma^2
  // dt h = -i J dt S: the FFT form's -2J (kscale) times the 1/2 of each cosine
  const float jdt = float(-0.5 * g.kscale * g.dt);

  for (int b = blockIdx.x; b < g.n_traj; b += gridDim.x) {
    const uint64_t traj = g.traj_ids[b];
    cpx<float>* gb = reinterpret_cast<cpx<float>*>(g.beta) + (size_t)b * 4096 + (size_t)lane * kL;
    cpx<float> beta[R];
    // site y of this lane's row: a register below R, else its LDS slot
    auto at = [&](int y) __attribute__((always_inline)) {
      return y < R ? beta[y < R ? y : 0] : row_lds[(y - R) * 64 + lane];
    };
    auto put = [&](int y, cpx<float> v) __attribute__((always_inline)) {
      if (y < R)
        beta[y < R ? y : 0] = v;
      else
        row_lds[(y - R) * 64 + lane] = v;
    };
    // lane x reads its own row: 512 contiguous bytes, 16-byte loads
#pragma unroll
    for (int y = 0; y < kL; y += 2) {
      const float4 v = *reinterpret_cast<const float4*>(gb + y);
      put(y, mk(v.x, v.y));
      put(y + 1, mk(v.z, v.w));
    }

    for (int s = 0; s < g.n_steps; ++s) {
      // opaque per step so the compiler keeps the Philox rounds inside the loop
      uint32_t pbase = (uint32_t)lane * 32u;
      asm volatile("" : "+v"(pbase));
      const uint32_t step = uint32_t(g.step0 + (uint64_t)s);
      const cpx<float> first = at(0);  // old value of the y + 1 wrap of y = 63
      // y-neighbour sum of the next site, formed from old values before the
      // site between them is overwritten (so each new value can take its old
      // value's register)
      const cpx<float> last = at(kL - 1), second = at(1);
      float syr = last.re + second.re, syi = last.im + second.im;
      // LDS sites read for a y-sum two sites ahead are kept for their own turn
      cpx<float> ahead[2] = {second, second};
#pragma unroll
      for (int p = 0; p < kL / 2; ++p) {
        const u32x4 w = philox4x32_10(make_counter(traj, TAG_WIENER, step, pbase + (uint32_t)p), g.key0, g.key1);
        float rs[2], cs[2], sn[2];
        box_muller_f32_scaled(w.x, w.y, nl, rs[0], cs[0], sn[0]);
        box_muller_f32_scaled(w.z, w.w, nl, rs[1], cs[1], sn[1]);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int y = 2 * p + h;
          const cpx<float> v = (y >= R && y >= 2) ? ahead[y & 1] : at(y);
          // every lane reads its neighbours' old beta[y] before any lane writes it
          const float sr = rot_down(v.re) + (rot_up(v.re) + syr);
          const float si = rot_down(v.im) + (rot_up(v.im) + syi);
          if (y < kL - 1) {
            const cpx<float> y2 = (y + 2 == kL) ? first : (y + 2 == kL - 1) ? last : at(y + 2 < kL ? y + 2 : 0);
            if (y + 2 >= R && y + 2 < kL) ahead[y & 1] = y2;
            syr = v.re + y2.re;
            syi = v.im + y2.im;
          }
          const float m = fmaf(v.im, v.im, v.re * v.re);
          const float cr = UI0 ? A0 : fmaf(-Ui, m, A0), ci = fmaf(Ur, m, B0);
          // increment = noise + dt iF + drift + hopping, all small, added to beta once
          float dr = fmaf(rs[h], cs[h], fr);
          dr = fmaf(cr, v.re, dr);
          dr = fmaf(-ci, v.im, dr);
          dr = fmaf(jdt, si, dr);
          float di = fmaf(rs[h], sn[h], fi);
          di = fmaf(cr, v.im, di);
          di = fmaf(ci, v.re, di);
          di = fmaf(-jdt, sr, di);
          put(y, mk(v.re + dr, v.im + di));
        }
      }
    }

#pragma unroll
    for (int y = 0; y < kL; y += 2) {
      const cpx<float> v0 = at(y), v1 = at(y + 1);
      *reinterpret_cast<float4*>(gb + y) = make_float4(v0.re, v0.im, v1.re, v1.im);
    }
  }
}

}  // namespace

bool stencil64_supported(int nx, int ny, int hopping, int precision, int param_stride) {
  return nx == 64 && ny == 64 && hopping == 0 && precision == 32 && param_stride == 0;
}

// sites per row kept in registers (the rest in LDS)
constexpr int kRegSites = 46;
constexpr size_t kRowLds = (size_t)(kL - kRegSites) * 64 * sizeof(cpx<float>);

hipError_t launch_stencil64(const StepArgs& a, int grid, hipStream_t s) {
  if (a.u_real)
    hipLaunchKernelGGL((stencil64_step_kernel<true, kRegSites>), dim3(grid), dim3(64), kRowLds, s, a);
  else
    hipLaunchKernelGGL((stencil64_step_kernel<false, kRegSites>), dim3(grid), dim3(64), kRowLds, s, a);
  return hipGetLastError();
}

hipError_t stencil64_occupancy(int* blocks_per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void*)stencil64_step_kernel<false, kRegSites>,
                                                      64, kRowLds);
}

}  // namespace twmc
