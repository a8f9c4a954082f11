// twmc_generic.hip -- LDS-resident TWMC step kernel for a lattice of any
// size (mixed-radix FFT: radices 2, 3, 4, 5, 8 and a direct-sum pass for each
// larger prime factor) and for matrix hopping, FP32 or FP64.  One workgroup
// advances one trajectory at a time (up to four, for lattices of up to 64
// sites); a persistent grid walks the batch.  This kernel is the FP64 path
// and the fallback for shapes the register-resident kernels do not cover.
//
// Per step (src/TWMC/Solvers/TWMCBaseSolver.cpp:232-256):
//   beta -> LDS, forward DFT (rows, then columns; Stockham, ping-pong),
//   K product, backward DFT, /N, drift + Philox increment, update in registers.
// With FFT hopping the steps run as nearest-neighbour launches (NN below, the
// same linear map without the transforms) and a stored frame as a
// capture-only launch (generic_capture_kernel), whose FFT gives the k-space
// sums.
// Matrix hopping (src/TWMC/Solvers/TWMCLiebSolver.cpp:259-267):
//   h_j = -i sum_i beta_i C_ij over the stored nonzeros of column j.
#include <algorithm>

#include "twmc_cplx.h"
#include "twmc_f64math.h"
#include "twmc_internal.h"
#include "twmc_rng.h"
#include "twmc_stockham.h"

namespace twmc {

namespace {

constexpr int kBlock = 256;

template <int S, typename T>
__device__ cpx<T>* fft2d(cpx<T>* a, cpx<T>* b, const StepArgs& g, const cpx<T>* twx, const cpx<T>* twy) {
  // rows: length ny, contiguous; then columns: length nx, stride ny
  cpx<T>* cur = fft_dim<S>(a, b, g.py, g.nx, g.ny, 1, twy);
  if (g.nx > 1) {
    cpx<T>* other = (cur == a) ? b : a;
    cur = fft_dim<S>(cur, other, g.px, g.ny, 1, g.ny, twx);
  }
  return cur;
}

// the FP64 Box-Muller's ln and sin / cos tables (twmc_f64math.h)
alignas(16) __device__ const double kF64TabDev[f64::kTabSize] = TWMC_F64_TABLES;

// the step's Wiener increments of a site pair from one Philox draw: FP64 in
// the range-specialised forms of twmc_f64math.h, both sites side by side; FP32
// on the hardware transcendental units
template <typename T>
__device__ __forceinline__ void box_muller_pair(const u32x4& x, T (&re)[2], T (&im)[2], const f64::Series& k,
                                                const double* ltab) {
  if constexpr (sizeof(T) == 8) {
    const uint32_t a[2] = {x.x, x.z}, b[2] = {x.y, x.w};
    f64::box_muller<2>(a, b, re, im, k, ltab);
  } else {
    box_muller_f32(x.x, x.y, re[0], im[0]);
    box_muller_f32(x.z, x.w, re[1], im[1]);
  }
}

__device__ __forceinline__ double warp_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// NN: a step launch with the hopping in its nearest-neighbour form and no
// frame capture (the capture runs as its own launch): the kernel then needs
// one LDS copy of the state (two with BLOCK 1024 and a split layout, see kDB)
// and none of the FFT's registers.
// UNIF (with NN): the site tables are uniform (StepArgs::site_uniform): their
// first entries are read once, into registers, instead of per site and step.
// LAY (with NN): the LDS layout of the staged state.  0: site order.  1 (ny
// even): split -- even sites in the first half, odd sites in the second, a
// pair's two sites in one row, neighbour indices by a division and wrap tests.
// 2 (ny a power of two >= 2, nx = 1 or a power of two >= 2): split, the
// neighbour indices as masks.  Plus 4 with a split layout when nx = 1 (no x
// neighbours: known at compile time, so no select of the x terms).
// BLOCK: threads per workgroup (kBlock; 64 or 128 for the step launch of a
// lattice of up to 128 or 256 sites; 1024 for the FP64 site-uniform step
// launch of 2049 to 4096 sites: two pairs per thread fit 128 registers, four
// waves per SIMD in the one workgroup per CU that the LDS state allows).
// TPB: trajectories per workgroup (64-thread step launches of lattices of up
// to 64 sites: 2, or 4 for up to 32 sites), each with BLOCK / TPB lanes and
// its own LDS copy of the state.
template <typename T, int PPT, bool NN = false, bool UNIF = false, int LAY = 0, int BLOCK = kBlock, int TPB = 1>
__global__ void __launch_bounds__(BLOCK) generic_step_kernel(StepArgs g) {
  static_assert(!UNIF || NN, "site-uniform launches are nearest-neighbour launches");
  static_assert(LAY == 0 || NN, "the split layouts are nearest-neighbour launches");
  static_assert(TPB == 1 || (NN && PPT == 1 && BLOCK == 64), "packed trajectories: 64-thread step launches");
  constexpr int kLanes = BLOCK / TPB;  // threads per trajectory
  const int tg = TPB > 1 ? (int)threadIdx.x / kLanes : 0;  // this thread's trajectory in the workgroup
  const int tl = TPB > 1 ? (int)threadIdx.x % kLanes : (int)threadIdx.x;
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = g.n;
  double* red = reinterpret_cast<double*>(smem);  // 8 doubles: per-wave partial sums
  cpx<T>* const base0 = reinterpret_cast<cpx<T>*>(smem + 64);
  cpx<T>* bufA = base0 + tg * n;
  cpx<T>* bufB = bufA + n;
  cpx<T>* twx = bufB + n;
  cpx<T>* twy = twx + g.nx;
  T* kcx = reinterpret_cast<T*>(twy + g.ny);
  T* kcy = kcx + g.nx;

  if constexpr (!NN) {
    for (int i = threadIdx.x; i < g.nx; i += blockDim.x) {
      twx[i] = reinterpret_cast<const cpx<T>*>(g.twx)[i];
      kcx[i] = reinterpret_cast<const T*>(g.kcx)[i];
    }
    for (int i = threadIdx.x; i < g.ny; i += blockDim.x) {
      twy[i] = reinterpret_cast<const cpx<T>*>(g.twy)[i];
      kcy[i] = reinterpret_cast<const T*>(g.kcy)[i];
    }
    __syncthreads();
  }

  const int npairs = (n + 1) / 2;
  // NN FP64: the Box-Muller series coefficients held in registers for the run
  f64::Series bmk = f64::kSeries;
  if constexpr (NN && sizeof(T) == 8) f64::pin(bmk);
  // FP64 Box-Muller tables: an LDS copy after the state for the
  // nearest-neighbour launch, else read from device memory
  // DB (the 1024-thread FP64 launch on a power-of-two lattice, one workgroup
  // per CU): two LDS copies of the state, each step reading one and writing
  // the new values into the other, so one barrier per step instead of two
  constexpr int kLay = LAY & 3;        // 0 site order, 1 / 2 split (see LAY)
  constexpr bool kSplit = kLay != 0;  // split layout, pairs within rows
  constexpr bool kXn = (LAY & 4) == 0;  // split layouts: the lattice has x neighbours (nx > 1)
  constexpr bool kDB = NN && kSplit && BLOCK == 1024;
  constexpr bool kBatch = kSplit && (UNIF || sizeof(T) == 4);  // neighbour reads batched
  cpx<T>* const bufD = bufA + n;
  const double* ltab = kF64TabDev;
  if constexpr (NN && sizeof(T) == 8) {
    double* lw = reinterpret_cast<double*>(base0 + (kDB ? 2 * n : TPB * n));
    for (int i = threadIdx.x; i < f64::kTabSize; i += blockDim.x) lw[i] = kF64TabDev[i];
    ltab = lw;
    __syncthreads();
  }
  double* part = g.partials + (size_t)blockIdx.x * kQ * n;
  constexpr bool kStencil = sizeof(T) == 4 || NN;
  // x = floor(i / ny) as umulhi(i, ceil(2^32 / ny)): exact for i ny < 2^32
  const uint32_t ymagic = g.ny > 1 ? (uint32_t)((0x100000000ull + (uint64_t)g.ny - 1) / (uint64_t)g.ny) : 0u;
  // pp / (ny / 2) as umulhi(pp, pmagic) for the split layout's row of a pair
  const uint32_t pmagic = g.ny > 1 ? (uint32_t)((0x100000000ull + (uint64_t)(g.ny >> 1) - 1) / (uint64_t)(g.ny >> 1)) : 0u;
  const T jval = T(-0.5 * g.kscale);  // J
  const T dt = T(g.dt), dt4 = T(g.dt4), kscale = T(g.kscale);
  const T djv = T(-0.5 * g.kscale * g.dt);  // dt J

  for (int b0 = blockIdx.x * TPB; b0 < g.n_traj; b0 += gridDim.x * TPB) {
    // a group past the batch's end steps the last trajectory again and stores
    // nothing, so that every thread reaches every barrier
    const bool live = b0 + tg < g.n_traj;
    const int b = live ? b0 + tg : g.n_traj - 1;
    const uint64_t traj = g.traj_ids[b];
    cpx<T>* gbeta = reinterpret_cast<cpx<T>*>(g.beta) + (size_t)b * n;
    const size_t poff = (size_t)(g.param_stride ? b : 0) * n;
    const cpx<T>* rsl = reinterpret_cast<const cpx<T>*>(g.rsl) + poff;
    const cpx<T>* Up = reinterpret_cast<const cpx<T>*>(g.U) + poff;
    const cpx<T>* Fp = reinterpret_cast<const cpx<T>*>(g.F) + poff;
    // UNIF: the site constants pre-scaled by dt, dt rsl, dt U and dt i F
    cpx<T> rd = mk(T(0), T(0)), ud = rd, fd = rd;
    T rd1 = T(1);
    if constexpr (UNIF) {
      const cpx<T> r0 = rsl[0], u0 = Up[0], f0 = Fp[0];
      rd = mk(dt * r0.re, dt * r0.im);
      ud = mk(dt * u0.re, dt * u0.im);
      fd = mk(-dt * f0.im, dt * f0.re);
      rd1 = T(1) + rd.re;
    }

    cpx<T> beta[2 * PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      int p = tl + kLanes * j;
      if (p < npairs) {
        beta[2 * j] = gbeta[2 * p];
        if (2 * p + 1 < n) beta[2 * j + 1] = gbeta[2 * p + 1];
      }
    }

    const int iters = g.n_steps > 0 ? g.n_steps : (g.capture ? 1 : 0);
    int ft_id = 0;
    if constexpr (kDB) {
#pragma unroll
      for (int j = 0; j < PPT; ++j) {
        const int p = tl + kLanes * j;
        if (p < npairs) {
          bufA[p] = beta[2 * j];
          bufA[npairs + p] = beta[2 * j + 1];
        }
      }
      __syncthreads();
    }
    for (int s = 0; s < iters; ++s) {
      // stage beta in LDS.  Split layout: even sites in the first half, odd sites in the
      // second (site 2p at p, 2p + 1 at n/2 + p), so a wave's stores and
      // neighbour reads step 16 bytes from lane to lane (no bank conflicts).
      // DB: the state of this step is already in nbuf, the new one goes to nxt
      cpx<T>* const nbuf = (kDB && (s & 1)) ? bufD : bufA;
      cpx<T>* const nxt = (kDB && (s & 1)) ? bufA : bufD;
#pragma unroll
      for (int j = 0; j < PPT && !kDB; ++j) {
        int p = tl + kLanes * j;
        if (p < npairs) {
          if constexpr (kSplit) {
            bufA[p] = beta[2 * j];
            bufA[npairs + p] = beta[2 * j + 1];
          } else {
            bufA[2 * p] = beta[2 * j];
            if (2 * p + 1 < n) bufA[2 * p + 1] = beta[2 * j + 1];
          }
        }
      }
      if constexpr (!kDB) __syncthreads();
      cpx<T>* cur = bufA;
      const bool cap = !NN && (s == 0 && g.capture);
      // FP32 (and the FP64 step launch, NN) uses the FFT form's exact
      // nearest-neighbour equivalent and transforms only for a frame's k-space
      // sums; an FP64 launch with steps and FFT hopping without NN applies the
      // FFT form every step (the host issues steps of FP64 FFT-hopping
      // batches as NN launches, captures as capture-only launches of this one)
      if (g.hopping == 0 && (!kStencil || cap)) cur = fft2d<-1>(bufA, bufB, g, twx, twy);

      if (cap) {
        // ensemble sums of the state before this step (a stored frame)
        if (g.hopping == 0) {
          for (int i = threadIdx.x; i < n; i += blockDim.x) {
            cpx<T> f = cur[i];
            part[2 * n + i] += double(f.re) * double(f.re) + double(f.im) * double(f.im);
          }
        }
        double m2 = 0.0, m4 = 0.0;
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
          int p = tl + kLanes * j;
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            int i = 2 * p + h;
            if (p < npairs && i < n) {
              cpx<T> z = beta[2 * j + h];
              double nn = double(z.re) * double(z.re) + double(z.im) * double(z.im);
              part[i] += nn;
              part[n + i] += nn * nn;
              part[3 * n + i] += double(z.re);
              part[4 * n + i] += double(z.im);
              m2 += nn;
              m4 += nn * nn;
              if (g.dump) {
                cpx<T>* out = reinterpret_cast<cpx<T>*>(g.traj_out) +
                              ((size_t)b * g.n_frames + g.frame) * n;
                out[i] = z;
              }
            }
          }
        }
        m2 = warp_sum(m2);
        m4 = warp_sum(m4);
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        if (lane == 0) {
          red[2 * wave] = m2;
          red[2 * wave + 1] = m4;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
          double a2 = 0.0, a4 = 0.0;
          for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            a2 += red[2 * w];
            a4 += red[2 * w + 1];
          }
          g.traj_m2[(size_t)b * g.n_frames + g.frame] = a2 / double(n);
          g.traj_m4[(size_t)b * g.n_frames + g.frame] = a4 / double(n);
        }
        __syncthreads();
      }
      if (s >= g.n_steps) break;

      if (NN || (g.hopping == 0 && kStencil)) {
        if (cap) {
          // the transform overwrote bufA: stage beta again for the neighbour sums
#pragma unroll
          for (int j = 0; j < PPT; ++j) {
            int p = tl + kLanes * j;
            if (p < npairs) {
              bufA[2 * p] = beta[2 * j];
              if (2 * p + 1 < n) bufA[2 * p + 1] = beta[2 * j + 1];
            }
          }
          __syncthreads();
        }
        cur = bufA;
      } else if (!NN && g.hopping == 0) {
        // K product: K = i kscale (fx cos kx + fy cos ky) (TWMCBaseSolver.cpp:80-88, :240)
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
          int kx = i / g.ny, ky = i - kx * g.ny;
          T kim = kscale * (kcx[kx] + kcy[ky]);
          cpx<T> f = cur[i];
          cur[i] = mk(-kim * f.im, kim * f.re);
        }
        __syncthreads();
        cpx<T>* other = (cur == bufA) ? bufB : bufA;
        cur = fft2d<+1>(cur, other, g, twx, twy);
      } else if (!NN) {
        // h = -i (beta . C) over the nonzeros of each column (TWMCLiebSolver.cpp:262)
        const T* vals = reinterpret_cast<const T*>(g.csc_val);
        for (int jcol = threadIdx.x; jcol < n; jcol += blockDim.x) {
          cpx<T> acc = mk(T(0), T(0));
          for (int e = g.csc_ptr[jcol]; e < g.csc_ptr[jcol + 1]; ++e) {
            cpx<T> z = bufA[g.csc_row[e]];
            T c = vals[e];
            acc = mk(acc.re + z.re * c, acc.im + z.im * c);
          }
          bufB[jcol] = mk(acc.im, -acc.re);  // -i acc
        }
        __syncthreads();
        cur = bufB;
      }

      // drift + increment + update (TWMCBaseSolver.cpp:245-253)
      const uint32_t step = uint32_t(g.step0 + (uint64_t)s);
      // zero the compiler cannot see through: the sites' neighbour addresses are
      // formed in each step instead of being hoisted and held across the loop
      int opq = 0;
      if constexpr (NN) asm volatile("" : "+s"(opq));
      // F(t) before this step (TWMCLiebSolver.cpp:252-256)
      T fdt = T(0);
      const cpx<T>* fv = Fp;
      const cpx<T>* fs = nullptr;
      if (g.ft_n > 0) {
        fdt = T(drive_interval(g.ft_times, g.ft_n, g.t0 + double(s) * g.dt, ft_id));
        fv = reinterpret_cast<const cpx<T>*>(g.ft_vals) + (size_t)ft_id * n;
        fs = reinterpret_cast<const cpx<T>*>(g.ft_slopes) + (size_t)ft_id * n;
      }
#pragma unroll
      for (int j = 0; j < PPT; ++j) {
        int p = tl + kLanes * j;
        if (p < npairs) {
          // split layout: the pair's sites 2p, 2p + 1 share a row (ny even) and are each
          // other's y neighbours; their other y neighbours and their x
          // neighbours are read from the split LDS layout.
          // Both sites' new values are formed before either is stored, so beta
          // holds the pair's state before this step throughout.
          cpx<T> znew[2], dz[2];
          T c1[2];
          int nbl = 0, nbr = 0, nbu = 0, nbd = 0;
          if constexpr (kSplit) {
            // in pairs (yh pairs per row): the left neighbour of 2p is the odd
            // site of pair rb + q - 1, the right neighbour of 2p + 1 the even
            // site of pair rb + q + 1 (q the pair's place in its row, mod yh);
            // the x neighbours are the pairs p +- yh (mod n/2), both sites
            const int pp = p + opq, yh = g.ny >> 1;
            if constexpr (kLay == 2) {
              const int q = pp & (yh - 1), rb = pp - q, hm = npairs - 1;
              nbl = npairs + rb + ((q - 1) & (yh - 1));
              nbr = rb + ((q + 1) & (yh - 1));
              nbu = (pp + yh) & hm;
              nbd = (pp - yh) & hm;
            } else {
              const int q = pp - (int)__umulhi((uint32_t)pp, pmagic) * yh;
              nbl = npairs + pp - 1 + (q == 0 ? yh : 0);
              nbr = pp + 1 - (q == yh - 1 ? yh : 0);
              nbu = pp + yh >= npairs ? pp + yh - npairs : pp + yh;
              nbd = pp < yh ? pp - yh + npairs : pp - yh;
            }
          }
          // split layout (kBatch): the pair's six neighbour reads issued
          // together (they need only the addresses), then the Philox rounds,
          // and only then their use -- one LDS round trip, under the integer
          // work (the scheduling barrier keeps the compiler from pulling the
          // sums, and so the waits, up to each read).  Not for FP64 with
          // site-varying parameters, where the held values would cost a wave
          // per SIMD; those read at the use below.
          T nsr[2] = {T(0), T(0)}, nsi[2] = {T(0), T(0)};
          cpx<T> nl, nr, nu[2], nd[2];
          if constexpr (kBatch) {
            nl = nbuf[nbl];
            nr = nbuf[nbr];
            if constexpr (kXn) {
              nu[0] = nbuf[nbu];
              nd[0] = nbuf[nbd];
              nu[1] = nbuf[npairs + nbu];
              nd[1] = nbuf[npairs + nbd];
            }
          }
          // the trajectory (one per workgroup) and step are wave-uniform,
          // unless the workgroup packs several trajectories
          u32x4 x;
          if constexpr (TPB > 1)
            x = philox4x32_10(make_counter(traj, TAG_WIENER, step, (uint32_t)p), g.key0, g.key1);
          else
            x = philox4x32_10_wave(make_counter(traj, TAG_WIENER, step, (uint32_t)p), g.key0, g.key1);
          if constexpr (kBatch) {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              // y + 1 and y - 1 neighbours, then x + 1 and x - 1
              const cpx<T> a = h == 0 ? beta[2 * j + 1] : nr;
              const cpx<T> c = h == 0 ? nl : beta[2 * j];
              nsr[h] = a.re + c.re;
              nsi[h] = a.im + c.im;
              if constexpr (kXn) {
                nsr[h] += nu[h].re + nd[h].re;
                nsi[h] += nu[h].im + nd[h].im;
              }
            }
          }
          T xr[2], xi[2];
          box_muller_pair<T>(x, xr, xi, bmk, ltab);
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            int i = 2 * p + h + opq;
            if (kSplit || i < n) {  // split: n even
              cpx<T> z = beta[2 * j + h];
              cpx<T> hh = mk(T(0), T(0));  // -i J S, or the FFT form's product
              if constexpr (!NN) hh = cur[i];
              T sr = T(0), si = T(0);
              if (NN || g.hopping == 0) {
                if constexpr (kStencil) {
                  // h = -i J S, S the sum over the periodic nearest neighbours
                  // (a dimension of length 1 has none): IDFT(K DFT beta)/N exactly
                  if constexpr (kBatch) {
                    sr = nsr[h];
                    si = nsi[h];
                  } else if constexpr (kSplit) {
                    // y + 1 and y - 1 neighbours, then x + 1 and x - 1
                    const cpx<T> a = h == 0 ? beta[2 * j + 1] : nbuf[nbr];
                    const cpx<T> c = h == 0 ? nbuf[nbl] : beta[2 * j];
                    sr = a.re + c.re;
                    si = a.im + c.im;
                    if constexpr (kXn) {
                      const cpx<T> d = nbuf[h * npairs + nbu];
                      const cpx<T> e = nbuf[h * npairs + nbd];
                      sr += d.re + e.re;
                      si += d.im + e.im;
                    }
                  } else {
                  const int x = g.ny > 1 ? (int)__umulhi((uint32_t)i, ymagic) : i;
                  const int y = i - x * g.ny;
                  if (g.ny > 1) {
                    const cpx<T> a = bufA[x * g.ny + (y + 1 == g.ny ? 0 : y + 1)];
                    const cpx<T> c = bufA[x * g.ny + (y == 0 ? g.ny - 1 : y - 1)];
                    sr = a.re + c.re;
                    si = a.im + c.im;
                  }
                  if (g.nx > 1) {
                    const cpx<T> a = bufA[(x + 1 == g.nx ? 0 : x + 1) * g.ny + y];
                    const cpx<T> c = bufA[(x == 0 ? g.nx - 1 : x - 1) * g.ny + y];
                    sr += a.re + c.re;
                    si += a.im + c.im;
                  }
                  }
                  hh = mk(jval * si, -jval * sr);
                } else {
                  hh = mk(hh.re / T(n), hh.im / T(n));
                }
              }
              if constexpr (UNIF) {
                // the same update with dt folded into the constants: the
                // increment dt (c z + h + i F) + dt4 xi as one FMA chain per
                // component, added to beta once; dt h = dt J (S_im, -S_re)
                const T q = fma(z.re, z.re, fma(z.im, z.im, T(-1)));
                const T ci = fma(ud.re, q, rd.im);
                if constexpr (sizeof(T) == 8) {
                  // FP64: z + c z = (1 + c) z with 1 + rd.re formed once (its
                  // rounding, 2^-53, is 1e-13 of dt rsl), the (1 + c) z term
                  // last, so the new value still takes one rounding at |z|
                  const T cr1 = fma(-ud.im, q, rd1);
                  T dr = fma(dt4, xr[h], fd.re);
                  dr = fma(djv, si, dr);
                  dr = fma(-ci, z.im, dr);
                  T di = fma(dt4, xi[h], fd.im);
                  di = fma(-djv, sr, di);
                  di = fma(ci, z.re, di);
                  if constexpr (kSplit) {
                    // applied after the pair's other site has read this one
                    c1[h] = cr1;
                    dz[h] = mk(dr, di);
                  } else {
                    z = mk(f64::fma3(cr1, z.re, dr), f64::fma3(cr1, z.im, di));
                  }
                } else {
                  const T cr = fma(-ud.im, q, rd.re);
                  T dr = fma(dt4, xr[h], fd.re);
                  dr = fma(djv, si, dr);
                  dr = fma(cr, z.re, dr);
                  dr = fma(-ci, z.im, dr);
                  T di = fma(dt4, xi[h], fd.im);
                  di = fma(-djv, sr, di);
                  di = fma(cr, z.im, di);
                  di = fma(ci, z.re, di);
                  z = mk(z.re + dr, z.im + di);
                }
              } else {
                cpx<T> r = rsl[i], u = Up[i], f = fv[i];
                if (fs) f = mk(f.re + fdt * fs[i].re, f.im + fdt * fs[i].im);
                T q = z.re * z.re + z.im * z.im - T(1);
                cpx<T> c = mk(r.re + (-u.im) * q, r.im + u.re * q);
                cpx<T> cb = cmul(c, z);
                cpx<T> sdr = mk(cb.re + hh.re + (-f.im), cb.im + hh.im + f.re);
                cpx<T> a = mk(sdr.re * dt, sdr.im * dt);
                z = mk(z.re + a.re + dt4 * xr[h], z.im + a.im + dt4 * xi[h]);
              }
              if constexpr (kSplit)
                znew[h] = z;
              else
                beta[2 * j + h] = z;
            }
          }
          if constexpr (kSplit) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              if constexpr (sizeof(T) == 8 && UNIF) {
                // three-operand FMAs: the new value can take the old one's
                // registers (the accumulator form would need a copy)
                const cpx<T> z = beta[2 * j + h];
                beta[2 * j + h] = mk(f64::fma3(c1[h], z.re, dz[h].re), f64::fma3(c1[h], z.im, dz[h].im));
              } else {
                beta[2 * j + h] = znew[h];
              }
            }
            if constexpr (kDB) {
              nxt[p] = beta[2 * j];
              nxt[npairs + p] = beta[2 * j + 1];
            }
          }
        }
      }
      __syncthreads();
    }

#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      int p = tl + kLanes * j;
      if (live && p < npairs) {
        gbeta[2 * p] = beta[2 * j];
        if (2 * p + 1 < n) gbeta[2 * p + 1] = beta[2 * j + 1];
      }
    }
  }
}

// Capture-only launch (no steps: a stored frame of a batch whose steps run as
// nearest-neighbour launches, or a batch's closing frame), for lattices of up
// to 4096 sites.  BLOCK threads: 1024 for 1025 to 4096 sites (one or two site
// pairs per thread, 16 waves for the FFT passes of a workgroup whose two LDS
// state copies may be the only one on its CU), else 256 (one or two pairs).
// Per trajectory, as the step kernel's capture does: the site means of
// |beta|^2 and |beta|^4, the frame dump and, with FFT hopping, |DFT beta|^2.
// The per-site sums over the workgroup's trajectories stay in registers (each
// thread owns fixed sites: the pairs tid + BLOCK j and the k-space sites
// tid + BLOCK j) and are added to the workgroup's partial row once at the
// end, so every sum takes the same trajectory order as a per-trajectory update
// of that row; the state is read and not written back.
// LDS bytes before the state in the capture kernel: two doubles per wave
constexpr int capture_red_bytes(int block) { return block > 256 ? block / 4 : 64; }

template <typename T, int PPT, int BLOCK>
__global__ void __launch_bounds__(BLOCK) generic_capture_kernel(StepArgs g) {
  static_assert(PPT <= 2, "register sums: at most 4 sites per thread");
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = g.n, npairs = (n + 1) / 2, tid = threadIdx.x;
  double* red = reinterpret_cast<double*>(smem);  // per-wave partial sums
  cpx<T>* const bufA = reinterpret_cast<cpx<T>*>(smem + capture_red_bytes(BLOCK));
  cpx<T>* const bufB = bufA + n;
  cpx<T>* const twx = bufB + n;
  cpx<T>* const twy = twx + g.nx;
  for (int i = tid; i < g.nx; i += BLOCK) twx[i] = reinterpret_cast<const cpx<T>*>(g.twx)[i];
  for (int i = tid; i < g.ny; i += BLOCK) twy[i] = reinterpret_cast<const cpx<T>*>(g.twy)[i];

  // [site h of pair j][|z|^2, |z|^4, Re z, Im z] and the k-space sites tid + j BLOCK
  double accr[2 * PPT][4], acck[2 * PPT];
#pragma unroll
  for (int j = 0; j < 2 * PPT; ++j) {
    acck[j] = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) accr[j][q] = 0.0;
  }
  const int wave = tid >> 6, lane = tid & 63;

  for (int b = blockIdx.x; b < g.n_traj; b += gridDim.x) {
    const cpx<T>* gbeta = reinterpret_cast<const cpx<T>*>(g.beta) + (size_t)b * n;
    cpx<T>* out = g.dump ? reinterpret_cast<cpx<T>*>(g.traj_out) + ((size_t)b * g.n_frames + g.frame) * n : nullptr;
    double m2 = 0.0, m4 = 0.0;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      const int p = tid + BLOCK * j;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int i = 2 * p + h;
        if (p < npairs && i < n) {
          const cpx<T> z = gbeta[i];
          bufA[i] = z;
          const double nn = double(z.re) * double(z.re) + double(z.im) * double(z.im);
          accr[2 * j + h][0] += nn;
          accr[2 * j + h][1] += nn * nn;
          accr[2 * j + h][2] += double(z.re);
          accr[2 * j + h][3] += double(z.im);
          m2 += nn;
          m4 += nn * nn;
          if (out) out[i] = z;
        }
      }
    }
    m2 = warp_sum(m2);
    m4 = warp_sum(m4);
    if (lane == 0) {
      red[2 * wave] = m2;
      red[2 * wave + 1] = m4;
    }
    __syncthreads();  // the staged state and the wave sums
    if (tid == 0) {
      double a2 = 0.0, a4 = 0.0;
      for (int w = 0; w < BLOCK / 64; ++w) {
        a2 += red[2 * w];
        a4 += red[2 * w + 1];
      }
      g.traj_m2[(size_t)b * g.n_frames + g.frame] = a2 / double(n);
      g.traj_m4[(size_t)b * g.n_frames + g.frame] = a4 / double(n);
    }
    if (g.hopping == 0) {
      const cpx<T>* cur = fft2d<-1>(bufA, bufB, g, twx, twy);
#pragma unroll
      for (int j = 0; j < 2 * PPT; ++j) {
        const int i = tid + BLOCK * j;
        if (i < n) {
          const cpx<T> f = cur[i];
          acck[j] += double(f.re) * double(f.re) + double(f.im) * double(f.im);
        }
      }
    }
    __syncthreads();  // the next trajectory overwrites the LDS state and the wave sums
  }

  double* part = g.partials + (size_t)blockIdx.x * kQ * n;
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    const int p = tid + BLOCK * j;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int i = 2 * p + h;
      if (p < npairs && i < n) {
        part[i] += accr[2 * j + h][0];
        part[n + i] += accr[2 * j + h][1];
        part[3 * n + i] += accr[2 * j + h][2];
        part[4 * n + i] += accr[2 * j + h][3];
      }
    }
  }
  if (g.hopping == 0) {
#pragma unroll
    for (int j = 0; j < 2 * PPT; ++j) {
      const int i = tid + BLOCK * j;
      if (i < n) part[2 * n + i] += acck[j];
    }
  }
}

// ---------------------------------------------------------------- init
// NoisyMatrix realisation of parameter `which` at site i for trajectory traj
// (src/Base/NoisyMatrix.cpp:118-160), FP64.
__device__ void gen_param(const InitArgs& a, int which, uint32_t tag, uint64_t traj, int i,
                          double& re, double& im) {
  re = a.base[which][2 * i];
  im = a.base[which][2 * i + 1];
  if (a.noise[which] == 0) return;
  u32x4 x = philox4x32_10(make_counter(traj, tag, 0u, (uint32_t)(i >> 1)), a.key0, a.key1);
  uint32_t w0 = (i & 1) ? x.z : x.x, w1 = (i & 1) ? x.w : x.y;
  double zr, zi;
  if (a.noise[which] == 2) {
    box_muller_f64(w0, w1, zr, zi);
  } else {
    const double two_m32 = 2.3283064365386963e-10;
    zr = (double(w0) + 0.5) * two_m32;
    zi = (double(w1) + 0.5) * two_m32;
  }
  re += zr * a.amp[which][2 * i];
  im += zi * a.amp[which][2 * i + 1];
}

template <typename T>
__global__ void init_kernel(InitArgs a) {
  const size_t total = (size_t)a.n_traj * a.n;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * blockDim.x) {
    int b = (int)(idx / a.n);
    int i = (int)(idx - (size_t)b * a.n);
    uint64_t traj = a.traj_ids[b];
    cpx<T>* beta = reinterpret_cast<cpx<T>*>(a.beta) + (size_t)b * a.n;
    if (a.mode == 0) {
      double br, bi, sr, si;
      gen_param(a, 3, TAG_BINIT, traj, i, br, bi);
      gen_param(a, 4, TAG_BSIGMA, traj, i, sr, si);
      u32x4 x = philox4x32_10(make_counter(traj, TAG_INIT, 0u, (uint32_t)(i >> 1)), a.key0, a.key1);
      double zr, zi;
      if (i & 1) box_muller_f64(x.z, x.w, zr, zi); else box_muller_f64(x.x, x.y, zr, zi);
      // beta_init + sigma (.) xi, complex product (TWMCBaseSolver.cpp:126-134)
      double pr = sr * zr - si * zi, pi = sr * zi + si * zr;
      beta[i] = mk(T(br + pr), T(bi + pi));
    } else if (a.mode == 1) {
      beta[i] = mk(T(a.base[3][2 * i]), T(a.base[3][2 * i + 1]));
    } else if (a.mode == 2) {
      const double* bp = a.beta_prev + 2 * ((size_t)b * a.n + i);
      beta[i] = mk(T(bp[0]), T(bp[1]));
    }  // mode 3: gauss_init_kernel writes the state
    if (a.per_traj_params || b == 0) {
      size_t off = (size_t)(a.per_traj_params ? b : 0) * a.n + i;
      double orr, oi, ur, ui, fr, fi;
      gen_param(a, 0, TAG_OMEGA, traj, i, orr, oi);
      gen_param(a, 1, TAG_U, traj, i, ur, ui);
      gen_param(a, 2, TAG_F, traj, i, fr, fi);
      double rr, ri;
      if (a.onsite == 0) {
        // rsl = -i omega - gamma/2 (TWMCBaseSolver.cpp:74)
        rr = oi - a.gamma / 2.0;
        ri = -orr;
      } else {
        // delta = -E + detuning; rsl = -i delta - gamma/2 (TWMCLiebSolver.cpp:44-45)
        double dr = -orr + a.detuning, di = -oi;
        rr = di - a.gamma / 2.0;
        ri = -dr;
      }
      reinterpret_cast<cpx<T>*>(a.rsl)[off] = mk(T(rr), T(ri));
      reinterpret_cast<cpx<T>*>(a.U)[off] = mk(T(ur), T(ui));
      reinterpret_cast<cpx<T>*>(a.F)[off] = mk(T(fr), T(fi));
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) gauss_init_kernel(GaussArgs g) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = g.n;
  cpx<T>* bufA = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* bufB = bufA + n;
  cpx<T>* twx = bufB + n;
  cpx<T>* twy = twx + g.nx;
  for (int i = threadIdx.x; i < g.nx; i += blockDim.x) twx[i] = reinterpret_cast<const cpx<T>*>(g.twx)[i];
  for (int i = threadIdx.x; i < g.ny; i += blockDim.x) twy[i] = reinterpret_cast<const cpx<T>*>(g.twy)[i];
  __syncthreads();
  const T inv_n = T(1.0 / double(n));
  for (int b = blockIdx.x; b < g.n_traj; b += gridDim.x) {
    const uint64_t traj = g.traj_ids[b];
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
      const u32x4 x = philox4x32_10(make_counter(traj, TAG_GAUSS, 0u, (uint32_t)(k >> 1)), g.key0, g.key1);
      double zr, zi;
      if (k & 1) box_muller_f64(x.z, x.w, zr, zi); else box_muller_f64(x.x, x.y, zr, zi);
      double re = g.kamp[k] * zr, im = g.kamp[k] * zi;
      if (k == 0) {
        re += double(n) * g.mean_re;
        im += double(n) * g.mean_im;
      }
      bufA[k] = mk(T(re), T(im));
    }
    __syncthreads();
    // inverse transform: rows (length ny), then columns (length nx)
    cpx<T>* cur = fft_dim<+1>(bufA, bufB, g.py, g.nx, g.ny, 1, twy);
    if (g.nx > 1) {
      cpx<T>* other = (cur == bufA) ? bufB : bufA;
      cur = fft_dim<+1>(cur, other, g.px, g.ny, 1, g.ny, twx);
    }
    cpx<T>* out = reinterpret_cast<cpx<T>*>(g.beta) + (size_t)b * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) out[i] = mk(cur[i].re * inv_n, cur[i].im * inv_n);
    __syncthreads();  // the next trajectory overwrites bufA
  }
}

// Per-site ensemble sums of one frame in site-major order: thread i of chunk
// c sums |beta_i|^2, |beta_i|^4, Re beta_i and Im beta_i over the trajectories
// of chunk c (a fixed range in a fixed order, so repeated runs agree bit for
// bit) and writes rows 0, 1, 3 and 4 of partial slot c.  Consecutive threads
// read consecutive sites of one trajectory: the state is read once, coalesced.
template <typename T>
__global__ void __launch_bounds__(256) site_sums_kernel(SiteSumArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int c = blockIdx.y;
  const int b0 = c * a.chunk, b1 = min(a.n_traj, b0 + a.chunk);
  const cpx<T>* beta = reinterpret_cast<const cpx<T>*>(a.beta);
  double s2 = 0.0, s4 = 0.0, sr = 0.0, si = 0.0;
  for (int b = b0; b < b1; ++b) {
    const cpx<T> v = beta[(size_t)b * a.n + i];
    const double re = double(v.re), im = double(v.im), nn = re * re + im * im;
    s2 += nn;
    s4 += nn * nn;
    sr += re;
    si += im;
  }
  double* part = a.partials + (size_t)c * kQ * a.n;
  part[i] = s2;
  part[a.n + i] = s4;
  part[3 * a.n + i] = sr;
  part[4 * a.n + i] = si;
  if (a.zero_k) part[2 * a.n + i] = 0.0;
}

__global__ void reduce_kernel(ReduceArgs a) {
  const int total = kQ * a.n;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x) {
    const int q = idx / a.n, j = idx - q * a.n;
    const int ns = (q == 2 && a.slots_k > 0) ? a.slots_k : a.slots;
    double s = 0.0;
    if (a.keep) {
      // read-only: 16 loads in flight per thread, the sum still in slot order
      const double* in = a.partials_in + idx;
      int sl = 0;
      for (; sl + 16 <= ns; sl += 16) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = in[(size_t)(sl + u) * total];
#pragma unroll
        for (int u = 0; u < 16; ++u) s += v[u];
      }
      for (; sl < ns; ++sl) s += in[(size_t)sl * total];
    } else {
      for (int sl = 0; sl < ns; ++sl) {
        size_t o = (size_t)sl * total + idx;
        s += a.partials_in[o];
        a.partials[o] = 0.0;
      }
    }
    const int* perm = (q == 2) ? a.perm_k : a.perm_site;
    int dst = perm ? perm[j] : j;
    a.obs[q * a.n + dst] = s;
  }
}

template <typename T, int PPT>
hipError_t launch_generic_t(const StepArgs& a, int grid, int block, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL((generic_step_kernel<T, PPT>), dim3(grid), dim3(block), lds, s, a);
  return hipGetLastError();
}

// lds: the step kernel's (generic_lds_bytes), which holds the capture kernel's
// state, twiddles and its 64 bytes of wave sums
template <typename T, int PPT, int BLOCK>
hipError_t launch_generic_capture_t(const StepArgs& a, int grid, size_t lds, hipStream_t s) {
  const void* f = (const void*)generic_capture_kernel<T, PPT, BLOCK>;
  lds += capture_red_bytes(BLOCK) - 64;
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((generic_capture_kernel<T, PPT, BLOCK>), dim3(grid), dim3(BLOCK), lds, s, a);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_generic_p(const StepArgs& a, int grid, int block, size_t lds, hipStream_t s) {
  int ppt = generic_pairs_per_thread(a.n, block);
  // a capture without steps: the capture kernel, its sums in registers
  if (a.n_steps == 0 && a.capture && block == kBlock && a.n <= 4096) {
    if (a.n > 2048) return launch_generic_capture_t<T, 2, 1024>(a, grid, lds, s);
    if (a.n > 1024) return launch_generic_capture_t<T, 1, 1024>(a, grid, lds, s);
    return ppt == 1 ? launch_generic_capture_t<T, 1, kBlock>(a, grid, lds, s)
                    : launch_generic_capture_t<T, 2, kBlock>(a, grid, lds, s);
  }
  if (lds > 65536) {
    hipError_t e;
    switch (ppt) {
      case 1: e = hipFuncSetAttribute((const void*)generic_step_kernel<T, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); break;
      case 2: e = hipFuncSetAttribute((const void*)generic_step_kernel<T, 2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); break;
      case 8: e = hipFuncSetAttribute((const void*)generic_step_kernel<T, 8>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); break;
      default: e = hipFuncSetAttribute((const void*)generic_step_kernel<T, 32>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); break;
    }
    if (e != hipSuccess) return e;
  }
  switch (ppt) {
    case 1: return launch_generic_t<T, 1>(a, grid, block, lds, s);
    case 2: return launch_generic_t<T, 2>(a, grid, block, lds, s);
    case 8: return launch_generic_t<T, 8>(a, grid, block, lds, s);
    default: return launch_generic_t<T, 32>(a, grid, block, lds, s);
  }
}

template <typename T, bool UNIF, int LAY>
const void* generic_nn_ptr_l(int ppt, int block, int tpb) {
  if constexpr (sizeof(T) == 8 && UNIF)
    if (block == 1024) return (const void*)generic_step_kernel<T, 2, true, UNIF, LAY, 1024>;
  if (block == 64)
    return tpb == 4   ? (const void*)generic_step_kernel<T, 1, true, UNIF, LAY, 64, 4>
           : tpb == 2 ? (const void*)generic_step_kernel<T, 1, true, UNIF, LAY, 64, 2>
                      : (const void*)generic_step_kernel<T, 1, true, UNIF, LAY, 64>;
  if (block == 128) return (const void*)generic_step_kernel<T, 1, true, UNIF, LAY, 128>;
  return ppt == 1   ? (const void*)generic_step_kernel<T, 1, true, UNIF, LAY>
         : ppt == 2 ? (const void*)generic_step_kernel<T, 2, true, UNIF, LAY>
         : ppt == 8 ? (const void*)generic_step_kernel<T, 8, true, UNIF, LAY>
                    : (const void*)generic_step_kernel<T, 32, true, UNIF, LAY>;
}

template <typename T, bool UNIF>
const void* generic_nn_ptr_u(int ppt, int lay, int block, int tpb) {
  switch (lay) {
    case 6: return generic_nn_ptr_l<T, UNIF, 6>(ppt, block, tpb);
    case 5: return generic_nn_ptr_l<T, UNIF, 5>(ppt, block, tpb);
    case 2: return generic_nn_ptr_l<T, UNIF, 2>(ppt, block, tpb);
    case 1: return generic_nn_ptr_l<T, UNIF, 1>(ppt, block, tpb);
    default: return generic_nn_ptr_l<T, UNIF, 0>(ppt, block, tpb);
  }
}

// the LDS layout of a nearest-neighbour step launch (LAY): split when a pair's
// two sites share a row, with mask indices on power-of-two lattices
int nn_layout(int nx, int ny) {
  auto p2 = [](int v) { return v >= 2 && (v & (v - 1)) == 0; };
  const int lay = (p2(ny) && (nx == 1 || p2(nx))) ? 2 : (ny >= 2 && ny % 2 == 0) ? 1 : 0;
  return (lay != 0 && nx == 1) ? lay | 4 : lay;
}

const void* generic_nn_ptr(int precision, int ppt, bool unif, int lay, int block, int tpb) {
  if (precision == 64)
    return unif ? generic_nn_ptr_u<double, true>(ppt, lay, block, tpb)
                : generic_nn_ptr_u<double, false>(ppt, lay, block, tpb);
  return unif ? generic_nn_ptr_u<float, true>(ppt, lay, block, tpb)
              : generic_nn_ptr_u<float, false>(ppt, lay, block, tpb);
}

// threads per workgroup of a nearest-neighbour step launch: one or two waves
// for lattices of up to 128 or 256 sites (a pair per thread, every wave busy),
// 1024 for FP64 site-uniform lattices of 2049 to 4096 sites, else kBlock
int generic_nn_block(int precision, int n, bool unif) {
  const int np = (n + 1) / 2;
  if (np <= 64) return 64;
  if (np <= 128) return 128;
  return (precision == 64 && unif && generic_pairs_per_thread(n, 1024) == 2) ? 1024 : kBlock;
}

// trajectories per workgroup of a nearest-neighbour step launch (TPB)
int nn_tpb(int n) {
  const int np = (n + 1) / 2;
  return np <= 16 ? 4 : np <= 32 ? 2 : 1;
}

}  // namespace

// LDS of a nearest-neighbour step launch: the per-wave sums, one copy of the
// state per trajectory of the workgroup (two for the double-buffered
// 1024-thread FP64 launch) and, FP64, the Box-Muller tables
size_t generic_nn_lds_bytes(int nx, int ny, int precision, bool site_uniform) {
  const int n = nx * ny;
  const bool db = generic_nn_block(precision, n, site_uniform) == 1024 && nn_layout(nx, ny) != 0;
  return 64 + (size_t)n * (precision == 64 ? 16 : 8) * (db ? 2 : nn_tpb(n)) +
         (precision == 64 ? sizeof(double) * f64::kTabSize : 0);
}

int generic_nn_traj_per_block(int n) { return nn_tpb(n); }

hipError_t generic_nn_occupancy(int precision, int nx, int ny, size_t lds, bool site_uniform,
                                int* blocks_per_cu) {
  const int block = generic_nn_block(precision, nx * ny, site_uniform);
  const int ppt = generic_pairs_per_thread(nx * ny, block);
  const void* f = generic_nn_ptr(precision, ppt, site_uniform, nn_layout(nx, ny), block, nn_tpb(nx * ny));
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, block, lds);
}

hipError_t launch_generic_nn(const StepArgs& a, int precision, int grid, size_t lds, hipStream_t s) {
  const int block = generic_nn_block(precision, a.n, a.site_uniform != 0);
  const int ppt = generic_pairs_per_thread(a.n, block);
  const void* f = generic_nn_ptr(precision, ppt, a.site_uniform != 0, nn_layout(a.nx, a.ny), block, nn_tpb(a.n));
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  void* args[] = {const_cast<StepArgs*>(&a)};
  return hipLaunchKernel(f, dim3(grid), dim3(block), args, lds, s);
}

int generic_pairs_per_thread(int n, int block) {
  int np = (n + 1) / 2;
  int need = (np + block - 1) / block;
  if (need <= 1) return 1;
  if (need <= 2) return 2;
  if (need <= 8) return 8;
  if (need <= 32) return 32;
  return -1;
}

size_t generic_lds_bytes(int nx, int ny, int precision) {
  size_t es = precision == 64 ? 16 : 8;
  size_t rs = precision == 64 ? 8 : 4;
  size_t n = (size_t)nx * ny;
  return 64 + 2 * n * es + (size_t)(nx + ny) * es + (size_t)(nx + ny) * rs;
}

hipError_t generic_occupancy(int precision, int n, int block, size_t lds, int* blocks_per_cu) {
  int ppt = generic_pairs_per_thread(n, block);
  const void* f;
  if (precision == 64) {
    f = ppt == 1 ? (const void*)generic_step_kernel<double, 1>
        : ppt == 2 ? (const void*)generic_step_kernel<double, 2>
        : ppt == 8 ? (const void*)generic_step_kernel<double, 8>
                   : (const void*)generic_step_kernel<double, 32>;
  } else {
    f = ppt == 1 ? (const void*)generic_step_kernel<float, 1>
        : ppt == 2 ? (const void*)generic_step_kernel<float, 2>
        : ppt == 8 ? (const void*)generic_step_kernel<float, 8>
                   : (const void*)generic_step_kernel<float, 32>;
  }
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, block, lds);
}

hipError_t launch_generic(const StepArgs& a, int precision, int grid, int block, size_t lds,
                          hipStream_t s) {
  if (precision == 64) return launch_generic_p<double>(a, grid, block, lds, s);
  return launch_generic_p<float>(a, grid, block, lds, s);
}

hipError_t launch_init(const InitArgs& a, int precision, hipStream_t s) {
  size_t total = (size_t)a.n_traj * a.n;
  int block = 256;
  size_t grid = (total + block - 1) / block;
  if (grid > 65536) grid = 65536;
  if (grid < 1) grid = 1;
  if (precision == 64)
    hipLaunchKernelGGL(init_kernel<double>, dim3((unsigned)grid), dim3(block), 0, s, a);
  else
    hipLaunchKernelGGL(init_kernel<float>, dim3((unsigned)grid), dim3(block), 0, s, a);
  return hipGetLastError();
}

size_t gauss_lds_bytes(int nx, int ny, int precision) {
  const size_t es = precision == 64 ? 16 : 8;
  return (2 * (size_t)nx * ny + nx + ny) * es;
}

hipError_t launch_gauss_init(const GaussArgs& a, int precision, int grid, hipStream_t s) {
  const size_t lds = gauss_lds_bytes(a.nx, a.ny, precision);
  const void* f = precision == 64 ? (const void*)gauss_init_kernel<double> : (const void*)gauss_init_kernel<float>;
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  if (precision == 64)
    hipLaunchKernelGGL(gauss_init_kernel<double>, dim3((unsigned)grid), dim3(256), lds, s, a);
  else
    hipLaunchKernelGGL(gauss_init_kernel<float>, dim3((unsigned)grid), dim3(256), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_reduce(const ReduceArgs& a, hipStream_t s) {
  int total = kQ * a.n;
  int block = 64;  // one output per thread: small blocks spread the few waves over every CU
  int grid = (total + block - 1) / block;
  hipLaunchKernelGGL(reduce_kernel, dim3(grid), dim3(block), 0, s, a);
  return hipGetLastError();
}

int site_sum_chunks(uint64_t n_traj, int max_chunks, int* chunk) {
  const int want = (int)std::max<uint64_t>(1, std::min<uint64_t>(n_traj, (uint64_t)max_chunks));
  *chunk = (int)((n_traj + want - 1) / want);
  return (int)((n_traj + *chunk - 1) / *chunk);
}

hipError_t launch_site_sums(const SiteSumArgs& a, int precision, hipStream_t s) {
  const int chunks = (a.n_traj + a.chunk - 1) / a.chunk;
  const dim3 grid((a.n + 255) / 256, chunks);
  if (precision == 64)
    hipLaunchKernelGGL(site_sums_kernel<double>, grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(site_sums_kernel<float>, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace twmc
