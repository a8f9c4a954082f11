// twmc_mfma.hip -- matrix-hopping TWMC step on the matrix cores.
//
// For general (disordered, Lieb, long-range) hopping the reference computes
// h = -i (beta . C) per trajectory with a sparse product
// (src/TWMC/Solvers/TWMCLiebSolver.cpp:259-267).  Here the hopping of a block
// of 32 trajectories is one GEMM per step,
//     Y = C^T [Re beta_1 .. Re beta_32 | Im beta_1 .. Im beta_32]   (N x 64),
// on v_mfma_f32_32x32x16_bf16 with FP32 accumulation.  FP32 operands are
// split as x = x_hi + x_lo (both BF16) and the product is
//     A_hi X_hi + A_hi X_lo + A_lo X_hi,
// which keeps about 16 significant bits (the dropped A_lo X_lo term and the
// BF16 rounding of the low parts are ~2^-16 relative).
//
// One workgroup (4 waves) owns 32 trajectories for a whole segment.  Wave w
// owns TPW row tiles of 32 sites; its accumulators hold Y for those sites and
// both column tiles (Re, Im), so each lane has Re and Im of the same
// (site, trajectory) and the Euler update runs in the GEMM epilogue, in
// registers, on the FP32 state that never leaves the lane.  Each step the
// state is re-staged in LDS as the BF16 hi/lo B operand; A = C^T (hi/lo) is
// read from L2 (it is shared by every workgroup).
//
// A = C^T is stored zero-padded to 4 TPW 32 rows and in fragment order
// (mfma_swizzle), so each wave's A fragment load is one contiguous 1 KiB read.
//
// MFMA lane layout (gfx950, checked on the device): A lane l holds row l%32,
// k = (l/32)*8 + 0..7; B lane l holds column l%32, same k; result register r
// of lane l is row (r/4)*8 + (l/32)*4 + r%4, column l%32.
#include <cstring>

#include "twmc_cplx.h"
#include "twmc_internal.h"
#include "twmc_rng.h"

namespace twmc {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kTB = 32;  // trajectories per workgroup

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// two FP32 values -> packed BF16 high parts (round to nearest even, one
// v_cvt_pk_bf16_f32) and packed BF16 parts of the remainders, so that
// x = hi + lo to ~2^-16 relative
__device__ __forceinline__ void split_bf16(float x0, float x1, uint32_t& hi, uint32_t& lo) {
  const f32x2 v = {x0, x1};
  hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
  const f32x2 d = {x0 - __uint_as_float(hi << 16), x1 - __uint_as_float(hi & 0xffff0000u)};
  lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(d, bf16x2));
}

template <int TPW, bool OVL, bool CAPTURE_ONLY>
__global__ void __launch_bounds__(256, 1) mfma_step_kernel(StepArgs g) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int np = g.np, n = g.n;
  const int ldk = np + 8;  // padded row: conflict-free 16-byte column reads
  uint16_t* Xhi = reinterpret_cast<uint16_t*>(smem);
  uint16_t* Xlo = Xhi + 64 * ldk;
  cpx<float>* Prsl = reinterpret_cast<cpx<float>*>(Xlo + 64 * ldk);
  cpx<float>* PU = Prsl + np;
  cpx<float>* PF = PU + np;
  double* red = reinterpret_cast<double*>(PF + np);  // [4 waves][32 traj][2]

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int col = lane & 31, half = lane >> 5;
  const int tiles = np / 32;  // wave w owns row tiles w*TPW .. w*TPW+TPW-1 that are < tiles
  // site tables pre-scaled by dt for the epilogue's FMA chain:
  //   dt (rsl + i U (|b|^2 - 1)) = (A - dt U.im |b|^2) + i (B + dt U.re |b|^2),
  //   Prsl = (A, B) = (dt (rsl.re + U.im), dt (rsl.im - U.re)), PU = dt U, PF = dt i F
  const float dt = float(g.dt), sig = float(g.dt4);
  const float nl = -1.3862943611198906f * sig * sig;  // -2 ln2 sigma^2 (Box-Muller)
  for (int i = tid; i < np; i += blockDim.x) {
    const bool in = i < n;
    const cpx<float> r = in ? reinterpret_cast<const cpx<float>*>(g.rsl)[i] : mk(0.f, 0.f);
    const cpx<float> u = in ? reinterpret_cast<const cpx<float>*>(g.U)[i] : mk(0.f, 0.f);
    const cpx<float> f = in ? reinterpret_cast<const cpx<float>*>(g.F)[i] : mk(0.f, 0.f);
    Prsl[i] = mk(dt * (r.re + u.im), dt * (r.im - u.re));
    PU[i] = mk(dt * u.re, dt * u.im);
    PF[i] = mk(-dt * f.im, dt * f.re);
  }
  __syncthreads();

  const uint16_t* Ahi = reinterpret_cast<const uint16_t*>(g.a_hi);
  const uint16_t* Alo = reinterpret_cast<const uint16_t*>(g.a_lo);
  double* part = g.partials + (size_t)blockIdx.x * kQ * n;
  const int nblk = (g.n_traj + kTB - 1) / kTB;

  for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const int j = blk * kTB + col;
    const bool valid = j < g.n_traj;
    const uint64_t traj = valid ? g.traj_ids[j] : 0;
    cpx<float>* gb = reinterpret_cast<cpx<float>*>(g.beta) + (size_t)(valid ? j : 0) * n;
    // first site of this lane's tiles, opaque to the compiler: the per-site
    // addresses and bounds of the load, capture and write-back are then
    // formed in the block, not hoisted out of the block loop and held
    // across the steps
    int ib = wave * TPW * 32 + half * 4;
    asm volatile("" : "+v"(ib));

    // FP32 state in the MFMA result layout: [tile][register]
    float sre[TPW][16], sim[TPW][16];
#pragma unroll
    for (int t = 0; t < TPW; ++t)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int i = ib + t * 32 + (r >> 2) * 8 + (r & 3);
        cpx<float> v = (valid && wave * TPW + t < tiles && i < n) ? gb[i] : mk(0.f, 0.f);
        sre[t][r] = v.re;
        sim[t][r] = v.im;
      }

    // frame capture of the state at the start of the segment, before the
    // step loop (inside it, its per-site addresses would be held across the
    // steps)
    if (g.capture) {
      double m2 = 0.0, m4 = 0.0;
      cpx<float>* out = (g.dump && valid)
                            ? reinterpret_cast<cpx<float>*>(g.traj_out) + ((size_t)j * g.n_frames + g.frame) * n
                            : nullptr;
#pragma unroll
      for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int i = ib + t * 32 + (r >> 2) * 8 + (r & 3);
          const bool use = valid && wave * TPW + t < tiles && i < n;
          const double xr = use ? double(sre[t][r]) : 0.0, xi = use ? double(sim[t][r]) : 0.0;
          const double nn = xr * xr + xi * xi;
          m2 += nn;
          m4 += nn * nn;
          if (out && use) out[i] = mk(sre[t][r], sim[t][r]);
          // per-site sums over the block's 32 trajectories: the 32 lanes of
          // this half-wave hold the same site for different trajectories
          double a = nn, b = nn * nn, c = xr, d = xi;
#pragma unroll
          for (int o = 1; o < 32; o <<= 1) {
            a += __shfl_xor(a, o, 64);
            b += __shfl_xor(b, o, 64);
            c += __shfl_xor(c, o, 64);
            d += __shfl_xor(d, o, 64);
          }
          if (col == 0 && wave * TPW + t < tiles && i < n) {
            part[i] += a;
            part[n + i] += b;
            part[3 * n + i] += c;
            part[4 * n + i] += d;
          }
        }
      // lanes col and col+32 hold the same trajectory
      m2 += __shfl_xor(m2, 32, 64);
      m4 += __shfl_xor(m4, 32, 64);
      if (half == 0) {
        red[(wave * 32 + col) * 2] = m2;
        red[(wave * 32 + col) * 2 + 1] = m4;
      }
      __syncthreads();
      if (tid < 32 && blk * kTB + tid < g.n_traj) {
        double a2 = 0.0, a4 = 0.0;
        for (int w = 0; w < 4; ++w) {
          a2 += red[(w * 32 + tid) * 2];
          a4 += red[(w * 32 + tid) * 2 + 1];
        }
        const size_t o = (size_t)(blk * kTB + tid) * g.n_frames + g.frame;
        g.traj_m2[o] = a2 / double(n);
        g.traj_m4[o] = a4 / double(n);
      }
      __syncthreads();
    }
    if (CAPTURE_ONLY) continue;

    const int iters = g.n_steps;
    int ft_id = 0;
    for (int s = 0; s < iters; ++s) {
      // stage the state as the BF16 hi/lo B operand: column col (Re) and 32+col (Im)
#pragma unroll
      for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (wave * TPW + t >= tiles) continue;
          const int i0 = (wave * TPW + t) * 32 + q * 8 + half * 4;
          uint32_t rh[2], rl[2], ih[2], il[2];
#pragma unroll
          for (int e = 0; e < 2; ++e) {
            split_bf16(sre[t][4 * q + 2 * e], sre[t][4 * q + 2 * e + 1], rh[e], rl[e]);
            split_bf16(sim[t][4 * q + 2 * e], sim[t][4 * q + 2 * e + 1], ih[e], il[e]);
          }
          *reinterpret_cast<uint2*>(&Xhi[col * ldk + i0]) = make_uint2(rh[0], rh[1]);
          *reinterpret_cast<uint2*>(&Xlo[col * ldk + i0]) = make_uint2(rl[0], rl[1]);
          *reinterpret_cast<uint2*>(&Xhi[(32 + col) * ldk + i0]) = make_uint2(ih[0], ih[1]);
          *reinterpret_cast<uint2*>(&Xlo[(32 + col) * ldk + i0]) = make_uint2(il[0], il[1]);
        }
      __syncthreads();

      if (g.ft_n > 0) {
        // F(t) before this step into the site table (TWMCLiebSolver.cpp:252-256);
        // the epilogue reads it after the barrier that follows the GEMM (the
        // overlapped step, which reads it during the GEMM, after one of its
        // own), and the previous epilogue finished before the staging barrier
        const float fdt = float(drive_interval(g.ft_times, g.ft_n, g.t0 + double(s) * g.dt, ft_id));
        const cpx<float>* fv = reinterpret_cast<const cpx<float>*>(g.ft_vals) + (size_t)ft_id * n;
        const cpx<float>* fs = reinterpret_cast<const cpx<float>*>(g.ft_slopes) + (size_t)ft_id * n;
        for (int i = tid; i < n; i += blockDim.x)
          PF[i] = mk(-dt * fmaf(fdt, fs[i].im, fv[i].im), dt * fmaf(fdt, fs[i].re, fv[i].re));
      }

      if constexpr (OVL) {
        // np = 128 TPW: 8 TPW k blocks and 8 TPW site pairs per lane.  The
        // part of the epilogue that does not depend on Y = C^T beta -- the
        // pair's Philox draw, Box-Muller and the drift and drive terms of the
        // increment -- runs during the GEMM, one site pair per k block, its
        // pieces placed between the block's MFMAs (fenced, so the order is
        // kept and the matrix cores run under the VALU work); after the GEMM
        // only dt h is added.
        constexpr int NKB = 8 * TPW, kPieces = 9, kSlots = 6 * TPW;
        if (g.ft_n > 0) __syncthreads();  // the F(t) table written above is read during the GEMM
        const uint32_t step = uint32_t(g.step0 + (uint64_t)s);
        // opaque to the compiler: the per-pair Philox rounds, the table reads
        // and the A addresses are formed in the step, not hoisted out of it
        uint32_t pbase = (uint32_t)(wave * TPW * 16 + half * 2);
        asm volatile("" : "+v"(pbase));
        int zo = 0;
        asm volatile("" : "+s"(zo));
        const size_t abase = ((size_t)(wave * TPW) * NKB * 64 + lane) * 8 + zo;
        auto aload = [&](const uint16_t* A, int t, int kb) {
          return *reinterpret_cast<const bf16x8*>(&A[abase + ((size_t)t * NKB + kb) * 512]);
        };
        auto bload = [&](const uint16_t* X, int c, int kb) {
          return *reinterpret_cast<const bf16x8*>(&X[c * ldk + kb * 16 + half * 8 + zo]);
        };
        float pre_r[TPW][16], pre_i[TPW][16];
        f32x16 are[TPW], aim[TPW];
#pragma unroll
        for (int t = 0; t < TPW; ++t)
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            are[t][r] = 0.f;
            aim[t][r] = 0.f;
          }
        bf16x8 ah[TPW], al[TPW], bq[4];
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
          ah[t] = aload(Ahi, t, 0);
          al[t] = aload(Alo, t, 0);
        }
        bq[0] = bload(Xhi, col, 0);
        bq[1] = bload(Xhi, 32 + col, 0);
        bq[2] = bload(Xlo, col, 0);
        bq[3] = bload(Xlo, 32 + col, 0);
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb) {
          // operands of the next k block, loaded one block ahead
          bf16x8 ahn[TPW], aln[TPW], bn[4];
          if (kb + 1 < NKB) {
#pragma unroll
            for (int t = 0; t < TPW; ++t) {
              ahn[t] = aload(Ahi, t, kb + 1);
              aln[t] = aload(Alo, t, kb + 1);
            }
            bn[0] = bload(Xhi, col, kb + 1);
            bn[1] = bload(Xhi, 32 + col, kb + 1);
            bn[2] = bload(Xlo, col, kb + 1);
            bn[3] = bload(Xlo, 32 + col, kb + 1);
          }
          // site pair kb: row tile tp, result registers 4 q + 2 pr + 0/1
          const int tp = kb >> 3, q = (kb >> 1) & 3, pr = kb & 1;
          u32x4 w;
          float rn[2], cs[2], sn[2];
          auto piece = [&](int pc) __attribute__((always_inline)) {
            if (pc == 0) w = make_counter(traj, TAG_WIENER, step, pbase + (uint32_t)(tp * 16 + q * 4 + pr));
            // each piece's results pass through an empty volatile asm, which
            // the fences order: pure arithmetic alone could be moved across
            // them
            if (pc < 5) {
              philox_round(w, g.key0, g.key1, 2 * pc);
              philox_round(w, g.key0, g.key1, 2 * pc + 1);
              asm volatile("" : "+v"(w.x), "+v"(w.y), "+v"(w.z), "+v"(w.w));
            } else if (pc < 7) {
              const int h = pc - 5;
              box_muller_f32_scaled(h ? w.z : w.x, h ? w.w : w.y, nl, rn[h], cs[h], sn[h]);
              asm volatile("" : "+v"(rn[h]), "+v"(cs[h]), "+v"(sn[h]));
            } else {
              // increment without dt h: noise + dt iF + drift (TWMCLiebSolver.cpp:259-264)
              const int h = pc - 7, r = 4 * q + 2 * pr + h;
              const int i = (wave * TPW + tp) * 32 + q * 8 + half * 4 + 2 * pr + h;
              const float br = sre[tp][r], bi = sim[tp][r];
              const cpx<float> ab = Prsl[i + zo], u = PU[i + zo], f = PF[i + zo];
              const float mq = fmaf(bi, bi, br * br);
              const float cr = fmaf(-u.im, mq, ab.re), ci = fmaf(u.re, mq, ab.im);
              float dr = fmaf(rn[h], cs[h], f.re);
              dr = fmaf(cr, br, dr);
              pre_r[tp][r] = fmaf(-ci, bi, dr);
              float di = fmaf(rn[h], sn[h], f.im);
              di = fmaf(cr, bi, di);
              pre_i[tp][r] = fmaf(ci, br, di);
              asm volatile("" : "+v"(pre_r[tp][r]), "+v"(pre_i[tp][r]));
            }
          };
#pragma unroll
          for (int sl = 0; sl < kSlots; ++sl) {
#pragma unroll
            for (int pc = (sl * kPieces + kSlots - 1) / kSlots; pc < ((sl + 1) * kPieces + kSlots - 1) / kSlots; ++pc)
              piece(pc);
            // MFMA sl: product wm = sl / TPW of row tile tt = sl % TPW, in the
            // order hi.hi Re, hi.hi Im, hi.lo Re, hi.lo Im, lo.hi Re, lo.hi Im
            // (consecutive MFMAs on different accumulators)
            const int wm = sl / TPW, tt = sl % TPW;
            const bf16x8 a = wm < 4 ? ah[tt] : al[tt];
            const bf16x8 b = bq[wm < 4 ? wm : wm - 4];
            if (wm & 1)
              aim[tt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, aim[tt], 0, 0, 0);
            else
              are[tt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, are[tt], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
          }
          if (kb + 1 < NKB) {
#pragma unroll
            for (int t = 0; t < TPW; ++t) {
              ah[t] = ahn[t];
              al[t] = aln[t];
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) bq[e] = bn[e];
          }
        }
        __syncthreads();  // the next step overwrites the B operand
        // dt h, h = -i Y (h_re = Im Y, h_im = -Re Y), and the update
#pragma unroll
        for (int t = 0; t < TPW; ++t)
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const int i = (wave * TPW + t) * 32 + (r >> 2) * 8 + half * 4 + (r & 3) + zo;
            if (i < n) {
              sre[t][r] += fmaf(dt, aim[t][r], pre_r[t][r]);
              sim[t][r] += fmaf(-dt, are[t][r], pre_i[t][r]);
            }
          }
        continue;
      }
      f32x16 are[TPW], aim[TPW];
#pragma unroll
      for (int t = 0; t < TPW; ++t) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          are[t][r] = 0.f;
          aim[t][r] = 0.f;
        }
      }
      // A has 4 TPW 32 zero-padded rows, so every wave runs all TPW tiles
      // without a guard; A fragments are double-buffered one 16-wide k block
      // ahead of the MFMAs that use them (np / 16 is even)
      // A is stored in fragment order (one contiguous 1 KiB block per row tile
      // and 16-wide k block, lane-major), so every fragment load is coalesced
      const int nkb = np / 16;
      const size_t abase = ((size_t)(wave * TPW) * nkb * 64 + lane) * 8;
      auto afrag = [&](int t, int kblk) { return abase + ((size_t)t * nkb + kblk) * 512; };
      bf16x8 ah0[TPW], al0[TPW], ah1[TPW], al1[TPW];
#pragma unroll
      for (int t = 0; t < TPW; ++t) {
        ah0[t] = *reinterpret_cast<const bf16x8*>(&Ahi[afrag(t, 0)]);
        al0[t] = *reinterpret_cast<const bf16x8*>(&Alo[afrag(t, 0)]);
      }
      for (int kb = 0; kb < np; kb += 32) {
        // k block kb: buffer 0; prefetch kb + 16 into buffer 1
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
          ah1[t] = *reinterpret_cast<const bf16x8*>(&Ahi[afrag(t, kb / 16 + 1)]);
          al1[t] = *reinterpret_cast<const bf16x8*>(&Alo[afrag(t, kb / 16 + 1)]);
        }
        {
          const int ko = kb + half * 8;
          const bf16x8 brh = *reinterpret_cast<const bf16x8*>(&Xhi[col * ldk + ko]);
          const bf16x8 brl = *reinterpret_cast<const bf16x8*>(&Xlo[col * ldk + ko]);
          const bf16x8 bih = *reinterpret_cast<const bf16x8*>(&Xhi[(32 + col) * ldk + ko]);
          const bf16x8 bil = *reinterpret_cast<const bf16x8*>(&Xlo[(32 + col) * ldk + ko]);
#pragma unroll
          for (int t = 0; t < TPW; ++t) {
            are[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah0[t], brh, are[t], 0, 0, 0);
            aim[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah0[t], bih, aim[t], 0, 0, 0);
            are[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah0[t], brl, are[t], 0, 0, 0);
            aim[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah0[t], bil, aim[t], 0, 0, 0);
            are[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al0[t], brh, are[t], 0, 0, 0);
            aim[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al0[t], bih, aim[t], 0, 0, 0);
          }
        }
        // k block kb + 16: buffer 1; prefetch kb + 32 (the last block reloads itself)
        const int kn = kb + 32 < np ? kb + 32 : kb + 16;
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
          ah0[t] = *reinterpret_cast<const bf16x8*>(&Ahi[afrag(t, kn / 16)]);
          al0[t] = *reinterpret_cast<const bf16x8*>(&Alo[afrag(t, kn / 16)]);
        }
        {
          const int ko = kb + 16 + half * 8;
          const bf16x8 brh = *reinterpret_cast<const bf16x8*>(&Xhi[col * ldk + ko]);
          const bf16x8 brl = *reinterpret_cast<const bf16x8*>(&Xlo[col * ldk + ko]);
          const bf16x8 bih = *reinterpret_cast<const bf16x8*>(&Xhi[(32 + col) * ldk + ko]);
          const bf16x8 bil = *reinterpret_cast<const bf16x8*>(&Xlo[(32 + col) * ldk + ko]);
#pragma unroll
          for (int t = 0; t < TPW; ++t) {
            are[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah1[t], brh, are[t], 0, 0, 0);
            aim[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah1[t], bih, aim[t], 0, 0, 0);
            are[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah1[t], brl, are[t], 0, 0, 0);
            aim[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah1[t], bil, aim[t], 0, 0, 0);
            are[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al1[t], brh, are[t], 0, 0, 0);
            aim[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al1[t], bih, aim[t], 0, 0, 0);
          }
        }
      }
      __syncthreads();  // the next step overwrites the B operand

      // epilogue: h = -i Y; drift, increment and update (TWMCLiebSolver.cpp:259-267)
      const uint32_t step = uint32_t(g.step0 + (uint64_t)s);
      // site-pair base of this lane; opaque to the compiler so the per-pair
      // second Philox round is not hoisted out of the step loop
      uint32_t pbase = (uint32_t)(wave * TPW * 16 + half * 2);
      asm volatile("" : "+v"(pbase));
#pragma unroll
      for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
          for (int pr = 0; pr < 2; ++pr) {
            if (wave * TPW + t >= tiles) continue;
            const int i0 = (wave * TPW + t) * 32 + q * 8 + half * 4 + 2 * pr;  // even site
            const u32x4 w = philox4x32_10(make_counter(traj, TAG_WIENER, step, pbase + (uint32_t)(t * 16 + q * 4 + pr)),
                                          g.key0, g.key1);
            float rn[2], cs[2], sn[2];
            box_muller_f32_scaled(w.x, w.y, nl, rn[0], cs[0], sn[0]);
            box_muller_f32_scaled(w.z, w.w, nl, rn[1], cs[1], sn[1]);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              const int r = 4 * q + 2 * pr + h;
              const int i = i0 + h;
              if (i < n) {
                const float br = sre[t][r], bi = sim[t][r];
                const cpx<float> ab = Prsl[i], u = PU[i], f = PF[i];
                const float mq = fmaf(bi, bi, br * br);
                const float cr = fmaf(-u.im, mq, ab.re), ci = fmaf(u.re, mq, ab.im);
                // increment = noise + dt iF + dt h + drift, h = -i Y (h_re = Im Y, h_im = -Re Y),
                // all small, added to beta once
                float dr = fmaf(rn[h], cs[h], f.re);
                dr = fmaf(dt, aim[t][r], dr);
                dr = fmaf(cr, br, dr);
                dr = fmaf(-ci, bi, dr);
                float di = fmaf(rn[h], sn[h], f.im);
                di = fmaf(-dt, are[t][r], di);
                di = fmaf(cr, bi, di);
                di = fmaf(ci, br, di);
                sre[t][r] = br + dr;
                sim[t][r] = bi + di;
              }
            }
          }
    }

    if (!CAPTURE_ONLY && valid) {
#pragma unroll
      for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int i = ib + t * 32 + (r >> 2) * 8 + (r & 3);
          if (wave * TPW + t < tiles && i < n) gb[i] = mk(sre[t][r], sim[t][r]);
        }
    }
  }
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// The step on v_mfma_f32_16x16x32_bf16 for np = 128 TPW (np = 128, 256,
// 384): 8 waves (two per SIMD, so one wave's MFMAs run while the other waits
// for operands or at a barrier), 32 trajectories per workgroup.  Wave w owns
// the TPW 16-site row tiles w TPW .. w TPW + TPW - 1 and all four 16-column
// tiles: Re and Im of trajectories 0-15 (ct = 0, 1) and 16-31 (ct = 2, 3).
// Lane l holds column l % 16 and rows 4 (l / 16) + r, r < 4, of each result
// tile, so Re and Im of the same (site, trajectory) meet in one lane.  A
// (16-row tiles, 32-wide k blocks, fragment order) streams from L2 one k block
// ahead; B (the state as BF16 hi/lo) is staged in LDS every step.  A lane owns
// 4 TPW site pairs per step, one per k block: each k block's MFMAs carry that
// pair's Philox draw, Box-Muller and drift/drive terms (fenced pieces); after
// the GEMM only dt h is added.
template <int TPW, bool CAPTURE_ONLY>
__global__ void __launch_bounds__(512, 1) mfma16_step_kernel(StepArgs g) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NP = 128 * TPW, NKB = NP / 32, LDK = NP + 8;  // LDK: padded row, conflict-free reads
  const int n = g.n;
  uint16_t* Xhi = reinterpret_cast<uint16_t*>(smem);
  uint16_t* Xlo = Xhi + 64 * LDK;
  cpx<float>* Prsl = reinterpret_cast<cpx<float>*>(Xlo + 64 * LDK);
  cpx<float>* PU = Prsl + NP;
  cpx<float>* PF = PU + NP;
  double* red = reinterpret_cast<double*>(PF + NP);  // [8 waves][32 traj][2]

  // the wave index is uniform: made scalar, the operand addresses become a
  // scalar base plus one per-lane offset instead of a register pair per load
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int col = lane & 15, rg = lane >> 4;
  const float dt = float(g.dt), sig = float(g.dt4);
  const float nl = -1.3862943611198906f * sig * sig;  // -2 ln2 sigma^2 (Box-Muller)
  // site tables pre-scaled by dt, as in mfma_step_kernel
  for (int i = tid; i < NP; i += blockDim.x) {
    const bool in = i < n;
    const cpx<float> r = in ? reinterpret_cast<const cpx<float>*>(g.rsl)[i] : mk(0.f, 0.f);
    const cpx<float> u = in ? reinterpret_cast<const cpx<float>*>(g.U)[i] : mk(0.f, 0.f);
    const cpx<float> f = in ? reinterpret_cast<const cpx<float>*>(g.F)[i] : mk(0.f, 0.f);
    Prsl[i] = mk(dt * (r.re + u.im), dt * (r.im - u.re));
    PU[i] = mk(dt * u.re, dt * u.im);
    PF[i] = mk(-dt * f.im, dt * f.re);
  }
  __syncthreads();

  const uint16_t* Ahi = reinterpret_cast<const uint16_t*>(g.a_hi);
  const uint16_t* Alo = reinterpret_cast<const uint16_t*>(g.a_lo);
  double* part = g.partials + (size_t)blockIdx.x * kQ * n;
  const int nblk = (g.n_traj + kTB - 1) / kTB;

  for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    int jt[2];
    bool valid[2];
    uint64_t traj[2];
    cpx<float>* gb[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      jt[q] = blk * kTB + q * 16 + col;
      valid[q] = jt[q] < g.n_traj;
      traj[q] = valid[q] ? g.traj_ids[jt[q]] : 0;
      gb[q] = reinterpret_cast<cpx<float>*>(g.beta) + (size_t)(valid[q] ? jt[q] : 0) * n;
    }
    // first site of this lane's rows, opaque to the compiler (per-site
    // addresses are formed in the block, not held across the steps)
    int ib = wave * TPW * 16 + rg * 4;
    asm volatile("" : "+v"(ib));

    // FP32 state in the result layout: [tile][trajectory group][register]
    float sre[TPW][2][4], sim[TPW][2][4];
#pragma unroll
    for (int t = 0; t < TPW; ++t)
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = ib + t * 16 + r;
          const cpx<float> v = (valid[q] && i < n) ? gb[q][i] : mk(0.f, 0.f);
          sre[t][q][r] = v.re;
          sim[t][q][r] = v.im;
        }

    if (g.capture) {
      double m2[2] = {0.0, 0.0}, m4[2] = {0.0, 0.0};
#pragma unroll
      for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = ib + t * 16 + r;
          double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const bool use = valid[q] && i < n;
            const double xr = use ? double(sre[t][q][r]) : 0.0, xi = use ? double(sim[t][q][r]) : 0.0;
            const double nn = xr * xr + xi * xi;
            m2[q] += nn;
            m4[q] += nn * nn;
            if (g.dump && use)
              reinterpret_cast<cpx<float>*>(g.traj_out)[((size_t)jt[q] * g.n_frames + g.frame) * n + i] =
                  mk(sre[t][q][r], sim[t][q][r]);
            a += nn;
            b += nn * nn;
            c += xr;
            d += xi;
          }
          // per-site sums over the block's 32 trajectories: the 16 lanes of a
          // row group hold the same site for trajectories q 16 + l % 16
#pragma unroll
          for (int o = 1; o < 16; o <<= 1) {
            a += __shfl_xor(a, o, 64);
            b += __shfl_xor(b, o, 64);
            c += __shfl_xor(c, o, 64);
            d += __shfl_xor(d, o, 64);
          }
          if (col == 0 && i < n) {
            part[i] += a;
            part[n + i] += b;
            part[3 * n + i] += c;
            part[4 * n + i] += d;
          }
        }
      // the four row groups hold the same trajectories
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        m2[q] += __shfl_xor(m2[q], 16, 64);
        m2[q] += __shfl_xor(m2[q], 32, 64);
        m4[q] += __shfl_xor(m4[q], 16, 64);
        m4[q] += __shfl_xor(m4[q], 32, 64);
        if (rg == 0) {
          red[(wave * 32 + q * 16 + col) * 2] = m2[q];
          red[(wave * 32 + q * 16 + col) * 2 + 1] = m4[q];
        }
      }
      __syncthreads();
      if (tid < 32 && blk * kTB + tid < g.n_traj) {
        double a2 = 0.0, a4 = 0.0;
        for (int w = 0; w < 8; ++w) {
          a2 += red[(w * 32 + tid) * 2];
          a4 += red[(w * 32 + tid) * 2 + 1];
        }
        const size_t o = (size_t)(blk * kTB + tid) * g.n_frames + g.frame;
        g.traj_m2[o] = a2 / double(n);
        g.traj_m4[o] = a4 / double(n);
      }
      __syncthreads();
    }
    if (CAPTURE_ONLY) continue;

    int ft_id = 0;
    for (int s = 0; s < g.n_steps; ++s) {
      // stage the state as the BF16 hi/lo B operand: columns q 16 + l % 16 (Re)
      // and 32 + q 16 + l % 16 (Im), this lane's four consecutive sites
#pragma unroll
      for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int k0 = (wave * TPW + t) * 16 + rg * 4;
          uint32_t rh[2], rl[2], ih[2], il[2];
#pragma unroll
          for (int e = 0; e < 2; ++e) {
            split_bf16(sre[t][q][2 * e], sre[t][q][2 * e + 1], rh[e], rl[e]);
            split_bf16(sim[t][q][2 * e], sim[t][q][2 * e + 1], ih[e], il[e]);
          }
          const int cr = q * 16 + col, ci = 32 + q * 16 + col;
          *reinterpret_cast<uint2*>(&Xhi[cr * LDK + k0]) = make_uint2(rh[0], rh[1]);
          *reinterpret_cast<uint2*>(&Xlo[cr * LDK + k0]) = make_uint2(rl[0], rl[1]);
          *reinterpret_cast<uint2*>(&Xhi[ci * LDK + k0]) = make_uint2(ih[0], ih[1]);
          *reinterpret_cast<uint2*>(&Xlo[ci * LDK + k0]) = make_uint2(il[0], il[1]);
        }
      if (g.ft_n > 0) {
        // F(t) before this step into the site table (TWMCLiebSolver.cpp:252-256),
        // read during the GEMM after the barrier below; the previous step's
        // readers finished before the barrier that ended its GEMM
        const float fdt = float(drive_interval(g.ft_times, g.ft_n, g.t0 + double(s) * g.dt, ft_id));
        const cpx<float>* fv = reinterpret_cast<const cpx<float>*>(g.ft_vals) + (size_t)ft_id * n;
        const cpx<float>* fs = reinterpret_cast<const cpx<float>*>(g.ft_slopes) + (size_t)ft_id * n;
        for (int i = tid; i < n; i += blockDim.x)
          PF[i] = mk(-dt * fmaf(fdt, fs[i].im, fv[i].im), dt * fmaf(fdt, fs[i].re, fv[i].re));
      }
      __syncthreads();

      constexpr int kPieces = 9, kSlots = 12 * TPW;
      const uint32_t step = uint32_t(g.step0 + (uint64_t)s);
      // opaque to the compiler: the per-pair Philox rounds, the table reads and
      // the operand addresses are formed in the step, not hoisted out of it
      uint32_t pbase = (uint32_t)(wave * TPW * 8 + rg * 2);
      asm volatile("" : "+v"(pbase));
      int zo = 0;
      asm volatile("" : "+s"(zo));
      const uint16_t* Ahw = Ahi + (size_t)(wave * TPW) * NKB * 512 + zo;  // scalar
      const uint16_t* Alw = Alo + (size_t)(wave * TPW) * NKB * 512 + zo;
      auto aload = [&](const uint16_t* A, int t, int kb) {
        return *reinterpret_cast<const bf16x8*>(&(A == Ahi ? Ahw : Alw)[((size_t)t * NKB + kb) * 512 + lane * 8]);
      };
      auto bload = [&](const uint16_t* X, int ct, int kb) {
        const int c = (ct & 1) * 32 + (ct >> 1) * 16 + col;
        return *reinterpret_cast<const bf16x8*>(&X[c * LDK + kb * 32 + rg * 8 + zo]);
      };
      f32x4 acc[TPW][4];
#pragma unroll
      for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[t][ct] = f32x4{0.f, 0.f, 0.f, 0.f};
      bf16x8 ah[TPW], al[TPW], bh, bl;
#pragma unroll
      for (int t = 0; t < TPW; ++t) {
        ah[t] = aload(Ahi, t, 0);
        al[t] = aload(Alo, t, 0);
      }
      bh = bload(Xhi, 0, 0);
      bl = bload(Xlo, 0, 0);
#pragma unroll
      for (int kb = 0; kb < NKB; ++kb) {
        // A of the next k block, one block ahead
        bf16x8 ahn[TPW], aln[TPW];
        if (kb + 1 < NKB) {
#pragma unroll
          for (int t = 0; t < TPW; ++t) {
            ahn[t] = aload(Ahi, t, kb + 1);
            aln[t] = aload(Alo, t, kb + 1);
          }
        }
        // site pair kb: row tile tp, trajectory group qp, pair pr of the lane's four sites
        const int tp = kb >> 2, qp = (kb >> 1) & 1, pr = kb & 1;
        u32x4 w;
        float rn[2], cs[2], sn[2];
        auto piece = [&](int pc) __attribute__((always_inline)) {
          if (pc == 0) w = make_counter(traj[qp], TAG_WIENER, step, pbase + (uint32_t)(tp * 8 + pr));
          if (pc < 5) {
            philox_round(w, g.key0, g.key1, 2 * pc);
            philox_round(w, g.key0, g.key1, 2 * pc + 1);
            asm volatile("" : "+v"(w.x), "+v"(w.y), "+v"(w.z), "+v"(w.w));
          } else if (pc < 7) {
            const int h = pc - 5;
            box_muller_f32_scaled(h ? w.z : w.x, h ? w.w : w.y, nl, rn[h], cs[h], sn[h]);
            asm volatile("" : "+v"(rn[h]), "+v"(cs[h]), "+v"(sn[h]));
          } else {
            // increment without dt h: noise + dt iF + drift (TWMCLiebSolver.cpp:259-264)
            const int h = pc - 7, r = 2 * pr + h;
            const int i = (wave * TPW + tp) * 16 + rg * 4 + r;
            const float br = sre[tp][qp][r], bi = sim[tp][qp][r];
            const cpx<float> ab = Prsl[i + zo], u = PU[i + zo], f = PF[i + zo];
            const float mq = fmaf(bi, bi, br * br);
            const float cr = fmaf(-u.im, mq, ab.re), ci = fmaf(u.re, mq, ab.im);
            float dr = fmaf(rn[h], cs[h], f.re);
            dr = fmaf(cr, br, dr);
            dr = fmaf(-ci, bi, dr);
            float di = fmaf(rn[h], sn[h], f.im);
            di = fmaf(cr, bi, di);
            di = fmaf(ci, br, di);
            // added to the state now (the B operand of this step is already
            // staged); dt h follows after the GEMM
            sre[tp][qp][r] = br + dr;
            sim[tp][qp][r] = bi + di;
            asm volatile("" : "+v"(sre[tp][qp][r]), "+v"(sim[tp][qp][r]));
          }
        };
        // 12 TPW MFMAs: column tile ct outer (its B operand loaded one tile
        // ahead), then the products hi.hi, hi.lo, lo.hi, then the row tiles
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
          bf16x8 bhn = bh, bln = bl;
          if (ct + 1 < 4) {
            bhn = bload(Xhi, ct + 1, kb);
            bln = bload(Xlo, ct + 1, kb);
          } else if (kb + 1 < NKB) {
            bhn = bload(Xhi, 0, kb + 1);
            bln = bload(Xlo, 0, kb + 1);
          }
#pragma unroll
          for (int pm = 0; pm < 3; ++pm)
#pragma unroll
            for (int t = 0; t < TPW; ++t) {
              const int sl = (ct * 3 + pm) * TPW + t;
#pragma unroll
              for (int pc = (sl * kPieces + kSlots - 1) / kSlots; pc < ((sl + 1) * kPieces + kSlots - 1) / kSlots;
                   ++pc)
                piece(pc);
              const bf16x8 a = pm < 2 ? ah[t] : al[t];
              const bf16x8 b = pm == 1 ? bl : bh;
              acc[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t][ct], 0, 0, 0);
              __builtin_amdgcn_sched_barrier(0);
            }
          bh = bhn;
          bl = bln;
        }
        if (kb + 1 < NKB) {
#pragma unroll
          for (int t = 0; t < TPW; ++t) {
            ah[t] = ahn[t];
            al[t] = aln[t];
          }
        }
      }
      __syncthreads();  // the next step overwrites the B operand
      // dt h, h = -i Y (h_re = Im Y, h_im = -Re Y), and the update
#pragma unroll
      for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = (wave * TPW + t) * 16 + rg * 4 + r + zo;
            if (i < n) {
              sre[t][q][r] = fmaf(dt, acc[t][2 * q + 1][r], sre[t][q][r]);
              sim[t][q][r] = fmaf(-dt, acc[t][2 * q][r], sim[t][q][r]);
            }
          }
    }

    if (!CAPTURE_ONLY) {
#pragma unroll
      for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = ib + t * 16 + r;
            if (valid[q] && i < n) gb[q][i] = mk(sre[t][q][r], sim[t][q][r]);
          }
    }
  }
}

template <int TPW>
hipError_t launch16_t(const StepArgs& a, int grid, size_t lds, hipStream_t s) {
  const void* f = a.n_steps > 0 ? (const void*)mfma16_step_kernel<TPW, false>
                                : (const void*)mfma16_step_kernel<TPW, true>;
  hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  if (a.n_steps > 0)
    hipLaunchKernelGGL((mfma16_step_kernel<TPW, false>), dim3(grid), dim3(512), lds, s, a);
  else
    hipLaunchKernelGGL((mfma16_step_kernel<TPW, true>), dim3(grid), dim3(512), lds, s, a);
  return hipGetLastError();
}

template <int TPW, bool OVL>
hipError_t launch_t(const StepArgs& a, int grid, size_t lds, hipStream_t s) {
  const void* f = a.n_steps > 0 ? (const void*)mfma_step_kernel<TPW, OVL, false>
                                : (const void*)mfma_step_kernel<TPW, false, true>;
  hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  if (a.n_steps > 0)
    hipLaunchKernelGGL((mfma_step_kernel<TPW, OVL, false>), dim3(grid), dim3(256), lds, s, a);
  else
    hipLaunchKernelGGL((mfma_step_kernel<TPW, false, true>), dim3(grid), dim3(256), lds, s, a);
  return hipGetLastError();
}

// row tiles per wave, and whether the step overlaps the epilogue with the
// GEMM: that needs one k block per site pair of a lane, np = 128 TPW (the
// 384-site Lieb chain is np = 384, TPW = 3), and TPW <= 3 for the registers
void mfma_shape(int np, int* tpw, bool* ovl) {
  *tpw = (np / 32 + 3) / 4;
  *ovl = *tpw <= 3 && np == 128 * *tpw;
}

}  // namespace

int mfma_padded(int n) { return (n + 31) / 32 * 32; }

// the 16x16x32 step (8 waves) for np = 128, 256, 384; 32x32x16 otherwise
// (at np = 512 the 16x16x32 step's four row tiles per wave spill registers)
bool mfma_use16(int np) { return np % 128 == 0 && np <= 384; }

// rows of the padded A operand: 8 waves x TPW tiles x 16, or 4 waves x TPW tiles x 32
int mfma_rows(int np) { return mfma_use16(np) ? np : 4 * ((np / 32 + 3) / 4) * 32; }

// row-major rows x np -> fragment order.  16x16x32: for 16-row tile T, 32-wide
// k block b and lane l, the 8 consecutive elements A[16 T + l % 16][32 b +
// 8 (l / 16) + e]; 32x32x16: for 32-row tile T, 16-wide k block b, the 8
// elements A[32 T + l % 32][16 b + 8 (l / 32) + e]
void mfma_swizzle(const double* a, int rows, int np, double* out) {
  const bool w16 = mfma_use16(np);
  const int tr = w16 ? 16 : 32, kw = w16 ? 32 : 16, nkb = np / kw;
  for (int T = 0; T < rows / tr; ++T)
    for (int b = 0; b < nkb; ++b)
      for (int l = 0; l < 64; ++l)
        for (int e = 0; e < 8; ++e)
          out[(((size_t)T * nkb + b) * 64 + l) * 8 + e] =
              a[(size_t)(T * tr + l % tr) * np + b * kw + (l / tr) * 8 + e];
}

size_t mfma_lds_bytes(int np) {
  const int waves = mfma_use16(np) ? 8 : 4;
  return (size_t)2 * 64 * (np + 8) * 2 + (size_t)3 * np * 8 + (size_t)waves * 32 * 2 * 8;
}

int mfma_block(int np) { return mfma_use16(np) ? 512 : 256; }

bool mfma_supported(int n, int hopping, int precision, int param_stride) {
  int np = mfma_padded(n);
  return hopping == 1 && precision == 32 && param_stride == 0 && np / 32 <= 16 && mfma_lds_bytes(np) <= 160 * 1024;
}

hipError_t launch_mfma(const StepArgs& a, int grid, hipStream_t s) {
  const size_t lds = mfma_lds_bytes(a.np);
  if (mfma_use16(a.np)) {
    switch (a.np / 128) {
      case 1: return launch16_t<1>(a, grid, lds, s);
      case 2: return launch16_t<2>(a, grid, lds, s);
      default: return launch16_t<3>(a, grid, lds, s);
    }
  }
  int tpw;
  bool ovl;
  mfma_shape(a.np, &tpw, &ovl);
  switch (tpw) {
    case 1: return ovl ? launch_t<1, true>(a, grid, lds, s) : launch_t<1, false>(a, grid, lds, s);
    case 2: return ovl ? launch_t<2, true>(a, grid, lds, s) : launch_t<2, false>(a, grid, lds, s);
    case 3: return ovl ? launch_t<3, true>(a, grid, lds, s) : launch_t<3, false>(a, grid, lds, s);
    default: return launch_t<4, false>(a, grid, lds, s);
  }
}

hipError_t mfma_occupancy(int np, int* blocks_per_cu) {
  const size_t lds = mfma_lds_bytes(np);
  if (mfma_use16(np)) {
    const int t = np / 128;
    const void* f = t == 1   ? (const void*)mfma16_step_kernel<1, false>
                    : t == 2 ? (const void*)mfma16_step_kernel<2, false>
                             : (const void*)mfma16_step_kernel<3, false>;
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, 512, lds);
  }
  int tpw;
  bool ovl;
  mfma_shape(np, &tpw, &ovl);
  const void* f = tpw == 1   ? (ovl ? (const void*)mfma_step_kernel<1, true, false> : (const void*)mfma_step_kernel<1, false, false>)
                  : tpw == 2 ? (ovl ? (const void*)mfma_step_kernel<2, true, false> : (const void*)mfma_step_kernel<2, false, false>)
                  : tpw == 3 ? (ovl ? (const void*)mfma_step_kernel<3, true, false> : (const void*)mfma_step_kernel<3, false, false>)
                             : (const void*)mfma_step_kernel<4, false, false>;
  hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, 256, lds);
}

void mfma_split(const double* a, size_t count, uint16_t* hi, uint16_t* lo) {
  auto rne = [](float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
  };
  auto back = [](uint16_t b) {
    uint32_t u = uint32_t(b) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
  };
  for (size_t k = 0; k < count; ++k) {
    float f = (float)a[k];
    hi[k] = rne(f);
    lo[k] = rne(f - back(hi[k]));
  }
}

}  // namespace twmc
