// twmc_rng.h -- counter-based Wiener increments for the TWMC kernels.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) keyed by the 64-bit run
// seed, with the counter
//     { step, site/2, traj & 0xffffffff, (tag << 24) | ((traj >> 32) & 0xffffff) }
// so every increment is a pure function of (seed, trajectory id, step, site):
// results do not depend on batch composition, launch shape or resume point.
// The words multiplied in the first round (step and trajectory) are uniform
// across a wavefront, so that round runs on the scalar unit.
// Site 2p takes output words (x0, x1), site 2p+1 takes (x2, x3); each pair is
// turned into one complex N(0,1) sample (Re, Im independent) by Box-Muller:
//     u1 = (a + 1/2) 2^-32,  theta = b 2^-32 (revolutions),
//     Re = sqrt(-2 ln u1) cos(2 pi theta),  Im = sqrt(-2 ln u1) sin(2 pi theta).
// The reference draws these from std::mt19937 + normal_distribution
// (src/TWMC/Solvers/TWMCBaseSolver.cpp:119-120, :285-291); its Re/Im draw order
// is unspecified there, so this layout fixes one.
#ifndef TWMC_RNG_H
#define TWMC_RNG_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace twmc {

enum : uint32_t {
  TAG_WIENER = 0,
  TAG_INIT = 1,
  TAG_U = 2,
  TAG_OMEGA = 3,
  TAG_F = 4,
  TAG_BINIT = 5,
  TAG_BSIGMA = 6,
  TAG_BOND = 7,
  TAG_GAUSS = 8  // Gaussian warm-start sample, counter pair = k-mode pair
};

struct u32x4 {
  uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ void mulhilo32(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  uint64_t p = uint64_t(a) * uint64_t(b);
  hi = uint32_t(p >> 32);
  lo = uint32_t(p);
}

// a ^ b ^ c; on the GPU one v_bitop3_b32 (truth table 0x96) instead of two
// XORs.  Used from the third round on, where every operand is per-lane; the
// first two rounds keep plain XORs so that their uniform parts stay scalar.
__host__ __device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
  return a ^ b ^ c;
#endif
}

// Ten Philox rounds; the key is bumped by the Weyl constants between rounds.
__host__ __device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo32(0xD2511F53u, c.x, hi0, lo0);
    mulhilo32(0xCD9E8D57u, c.z, hi1, lo1);
    u32x4 n;
    n.x = r < 2 ? (hi1 ^ c.y ^ k0) : xor3(hi1, c.y, k0);
    n.y = lo1;
    n.z = r < 2 ? (hi0 ^ c.w ^ k1) : xor3(hi0, c.w, k1);
    n.w = lo0;
    c = n;
  }
  return c;
}

// philox4x32_10 for counters whose step (c.x) and trajectory words (c.z, c.w)
// are the same across the wave and only the pair (c.y) differs per lane, as
// make_counter builds them in a kernel that steps one trajectory per wave or
// workgroup.  The same rounds, with each XOR of round 2 and 3 grouped so that
// its two wave-uniform operands combine first on the scalar unit: a vector
// three-input XOR can read only one scalar operand, and would need a copy.
// Which words are uniform, by round (U uniform, V per lane):
//   in:  x U, y V, z U, w U      after 0: x V, y U, z U, w U
//   after 1: x U, y U, z V, w V  after 2: x V, y V, z V, w U  (then all V)
__host__ __device__ __forceinline__ u32x4 philox4x32_10_wave(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo32(0xD2511F53u, c.x, hi0, lo0);
    mulhilo32(0xCD9E8D57u, c.z, hi1, lo1);
    u32x4 n;
    if (r < 2) {
      n.x = hi1 ^ c.y ^ k0;
      n.z = hi0 ^ c.w ^ k1;
    } else if (r == 2) {
      n.x = hi1 ^ (c.y ^ k0);  // c.y, k0 uniform
      n.z = c.w ^ (hi0 ^ k1);  // hi0, k1 uniform
    } else if (r == 3) {
      n.x = xor3(hi1, c.y, k0);
      n.z = hi0 ^ (c.w ^ k1);  // c.w, k1 uniform
    } else {
      n.x = xor3(hi1, c.y, k0);
      n.z = xor3(hi0, c.w, k1);
    }
    n.y = lo1;
    n.w = lo0;
    c = n;
  }
  return c;
}

// Round r (0-based) of philox4x32_10 alone, with that round's key
// (k0 + r 0x9E3779B9, k1 + r 0xBB67AE85), for callers that spread the ten
// rounds over other work; rounds 0..9 in order give philox4x32_10 exactly.
__host__ __device__ __forceinline__ void philox_round(u32x4& c, uint32_t k0, uint32_t k1, int r) {
  const uint32_t kr0 = k0 + uint32_t(r) * 0x9E3779B9u, kr1 = k1 + uint32_t(r) * 0xBB67AE85u;
  uint32_t hi0, lo0, hi1, lo1;
  mulhilo32(0xD2511F53u, c.x, hi0, lo0);
  mulhilo32(0xCD9E8D57u, c.z, hi1, lo1);
  u32x4 n;
  n.x = r < 2 ? (hi1 ^ c.y ^ kr0) : xor3(hi1, c.y, kr0);
  n.y = lo1;
  n.z = r < 2 ? (hi0 ^ c.w ^ kr1) : xor3(hi0, c.w, kr1);
  n.w = lo0;
  c = n;
}

__host__ __device__ __forceinline__ u32x4 make_counter(uint64_t traj, uint32_t tag, uint32_t step, uint32_t pair) {
  u32x4 c;
  c.x = step;
  c.y = pair;
  c.z = uint32_t(traj & 0xffffffffu);
  c.w = (tag << 24) | uint32_t((traj >> 32) & 0xffffffu);
  return c;
}

// FP64 Box-Muller with accurate libm/ocml log, sqrt, cos, sin.
__host__ __device__ __forceinline__ void box_muller_f64(uint32_t a, uint32_t b, double& re, double& im) {
  const double two_m32 = 2.3283064365386963e-10;
  double u1 = (double(a) + 0.5) * two_m32;
  double th = double(b) * two_m32;
  double r = sqrt(-2.0 * log(u1));
  double ang = 6.283185307179586 * th;
  re = r * cos(ang);
  im = r * sin(ang);
}

// FP32 Box-Muller on the hardware transcendental units: v_log_f32 (log2),
// v_sqrt_f32, and v_sin_f32 / v_cos_f32, which take revolutions, so the
// 2 pi factor costs nothing.  u1 lies in (0, 1]; u1 = 1 gives r = 0.
__device__ __forceinline__ void box_muller_f32(uint32_t a, uint32_t b, float& re, float& im) {
  const float two_m32 = 2.3283064365386963e-10f;
  const float two_m33 = 1.1641532182693481e-10f;
  float u1 = fmaf(float(a), two_m32, two_m33);
  float th = float(b) * two_m32;
  float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));
  re = r * __builtin_amdgcn_cosf(th);
  im = r * __builtin_amdgcn_sinf(th);
}

// Box-Muller in the form the FP32 update kernels fold into their FMA chains:
// rs = sigma r = sqrt(nl log2(u1)) with nl = -2 ln2 sigma^2 (so sigma costs no
// instruction), and cos / sin of the angle.  u1 is mapped as in
// box_muller_f32.  The angle is the top 23 bits of b as the float 1.f in
// [1, 2) -- one v_alignbit_b32 -- since v_sin_f32 / v_cos_f32 take revolutions
// and drop the integer part; it differs from b 2^-32 by less than 2^-23 turns.
__device__ __forceinline__ void box_muller_f32_scaled(uint32_t a, uint32_t b, float nl, float& rs, float& c,
                                                      float& s) {
  const float two_m32 = 2.3283064365386963e-10f;
  const float two_m33 = 1.1641532182693481e-10f;
  const float u1 = fmaf(float(a), two_m32, two_m33);
  rs = __builtin_amdgcn_sqrtf(nl * __builtin_amdgcn_logf(u1));
  const float th = __builtin_bit_cast(float, __builtin_amdgcn_alignbit(0x7fu, b, 9u));
  c = __builtin_amdgcn_cosf(th);
  s = __builtin_amdgcn_sinf(th);
}

}  // namespace twmc

#endif
