// twmc_sweep.hip -- a parameter sweep of many uneven lattices (1D chains and
// 2D squares of different sizes, each point with its own F, detuning, ...)
// advanced by ONE persistent launch per frame.
//
// Work items are (point, block of trajectories).  Workgroups take items from a
// global counter in decreasing order of cost (largest lattices first), so the
// uneven items balance across the chip.  A step item holds tb trajectories of
// one point stacked as the rows of one array (a chain's trajectory is one
// row, a 2D point's nx rows), so small lattices still fill a workgroup; the
// frame capture takes 2D points one trajectory per item (its FFT runs along
// both dimensions) and chains as stacked rows, transformed along the rows
// only.  tb per lattice shape is chosen by twmc_sweep_autotune, which times
// the candidates on the device.  A step applies the hopping as the exact
// nearest-neighbour form of the reference's FFT product (the x- and
// y-neighbours read from the LDS copy of the state), then the same Philox
// increments and Euler update as the single-system kernels, so every
// trajectory of a sweep equals the same trajectory run alone (to FP32
// rounding); the Stockham FFT runs only in the frame capture, for the
// k-space bins.
// The frame loop (counter reset + step kernel per frame) can be captured once
// into a hipGraph and launched as one graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/twmc_hip.h"
#include "twmc_cplx.h"
#include "twmc_internal.h"
#include "twmc_rng.h"
#include "twmc_stockham.h"

namespace twmc {

namespace {

constexpr int kSweepBlock = 256;
constexpr int kSweepMaxSites = 4096;  // sites per work item (LDS and registers)
constexpr int kPPT = kSweepMaxSites / 2 / kSweepBlock;
constexpr int kBins = 32;  // radial |k| bins: 0 is k = 0, 1..31 split (0, pi sqrt 2]
constexpr int kMaxRows = 256;  // stacked trajectories per work item
constexpr int kMaxNx = 64;     // rows of a 2D point
constexpr int kMaxNy = 512;    // row length (chain length or 2D row)

struct SweepPointDev {
  int nx, ny, tb, fft_x;  // fft_x = 1: 2D point (rows are lattice rows), 0: stacked 1D trajectories
  RadixPlan px, py;
  int tw_x, tw_y, kc_x, kc_y;  // offsets into the table pools
  float rsl_re, rsl_im, u_re, u_im, f_re, f_im;
  float dt4, kscale;
  int n_traj;
  uint64_t traj_id0;
  uint64_t state_off;  // complex elements
  uint64_t m_off;      // offset of [n_traj][n_frames] per-trajectory means
  int bin_off;         // offset of the mode -> radial bin table (one entry per mode of one trajectory)
  int index;           // point index (offset of its [n_frames][kBins + 2] k-space / mean-field sums)
};

struct SweepItemDev {
  int point, t0;
};

struct SweepArgs {
  const SweepPointDev* points;
  const SweepItemDev* items;
  int n_items;
  unsigned int* counter;
  const cpx<float>* tw_pool;
  const float* kc_pool;
  cpx<float>* beta;
  double* traj_m2;
  double* traj_m4;
  const unsigned char* bin_pool;  // mode -> bin
  double* ksum;                   // [points][n_frames][kBins + 2]: sum of |DFT beta|^2 per bin, sum of beta
  float dt;
  uint32_t key0, key1;
  uint64_t step0;
  int n_steps, capture, frame, n_frames;
  const uint64_t* traj_ids;  // one-point batches of the main API: ids by index (else traj_id0 + index)
};

// clears the work-queue counters of the capture and the step launch
__global__ void sweep_reset_kernel(unsigned int* counter) {
  if (threadIdx.x < 2 && blockIdx.x == 0) counter[threadIdx.x] = 0u;
}

// One work item advanced g.n_steps Euler-Maruyama steps (the step launch).
// The item is ntr trajectories of nxt rows each, stacked: 2D points have
// nxt = nx lattice rows and x-hopping (XHOP), chains one row each and no
// x-hopping; MULTI = false is the single-trajectory case (one trajectory id
// for the whole item).  Thread t owns the site pairs p = t + 256 j, sites 2p
// and 2p + 1 of one row (rows have even length), so the pair's two values are
// each other's y-neighbours; the other y-neighbours, and the x-neighbours as
// the 16-byte pairs of the rows above and below in the same trajectory, come
// from the LDS copy of the state.  The hopping h = -i J S, S the sum over the
// periodic nearest neighbours, is the exact real-space form of the reference's
// IDFT(K DFT(beta)) / N (TWMCBaseSolver.cpp:80-88, :235-242); a length-2
// dimension gives the one neighbour twice, as the FFT form does.
template <bool XHOP, bool MULTI>
__device__ __forceinline__ void step_item(const SweepArgs& g, const SweepPointDev& P, const SweepItemDev& item,
                                          cpx<float>* bufA, cpx<float> (&beta)[2 * kPPT], int nxt, int ny,
                                          int npairs) {
  // neighbour offsets and Philox counters of the owned pairs, fixed for the item
  int lft[kPPT], rgt[kPPT], up[kPPT], dn[kPPT];
  uint32_t pid[kPPT];
  uint64_t trj[kPPT];
  const uint64_t traj0 = P.traj_id0 + (uint64_t)item.t0;
#pragma unroll
  for (int j = 0; j < kPPT; ++j) {
    const int p = threadIdx.x + kSweepBlock * j;
    const int row = (2 * p) / ny, col = 2 * p - row * ny;
    const int tr = MULTI ? row / nxt : 0, r0 = tr * nxt, r = row - r0;
    lft[j] = row * ny + (col == 0 ? ny - 1 : col - 1);
    rgt[j] = row * ny + (col + 2 == ny ? 0 : col + 2);
    up[j] = ((r0 + (r == 0 ? nxt - 1 : r - 1)) * ny + col) / 2;  // pair index
    dn[j] = ((r0 + (r + 1 == nxt ? 0 : r + 1)) * ny + col) / 2;
    pid[j] = (uint32_t)(p - tr * (nxt * ny / 2));
    trj[j] = g.traj_ids ? g.traj_ids[item.t0 + tr] : traj0 + (uint64_t)tr;
  }
  // step constants: dt (rsl + i U (|b|^2 - 1)) = (A0 - Ui |b|^2) + i (B0 + Ur |b|^2),
  // dt iF = (fr, fi), dt h = -i jdt S, sigma^2 folded into the Box-Muller log multiplier
  const float dt = g.dt;
  const float Ur = dt * P.u_re, Ui = dt * P.u_im;
  const float A0 = fmaf(dt, P.rsl_re, Ui), B0 = fmaf(dt, P.rsl_im, -Ur);
  const float fr = -dt * P.f_im, fi = dt * P.f_re;
  const float jdt = dt * (-0.5f * P.kscale);  // dt J
  const float nl = -1.3862943611198906f * P.dt4 * P.dt4;
  float4* pairs = reinterpret_cast<float4*>(bufA);

  for (int s = 0; s < g.n_steps; ++s) {
#pragma unroll
    for (int j = 0; j < kPPT; ++j) {
      const int p = threadIdx.x + kSweepBlock * j;
      if (p < npairs) pairs[p] = make_float4(beta[2 * j].re, beta[2 * j].im, beta[2 * j + 1].re, beta[2 * j + 1].im);
    }
    __syncthreads();
    const uint32_t step = uint32_t(g.step0 + (uint64_t)s);
#pragma unroll
    for (int j = 0; j < kPPT; ++j) {
      const int p = threadIdx.x + kSweepBlock * j;
      if (p < npairs) {
        const u32x4 w = philox4x32_10(make_counter(trj[j], TAG_WIENER, step, pid[j]), g.key0, g.key1);
        float rs[2], cs[2], sn[2];
        box_muller_f32_scaled(w.x, w.y, nl, rs[0], cs[0], sn[0]);
        box_muller_f32_scaled(w.z, w.w, nl, rs[1], cs[1], sn[1]);
        const cpx<float> z0 = beta[2 * j], z1 = beta[2 * j + 1];
        const cpx<float> L = bufA[lft[j]], R = bufA[rgt[j]];
        float sr[2], si[2];
        sr[0] = z1.re + L.re;
        si[0] = z1.im + L.im;
        sr[1] = R.re + z0.re;
        si[1] = R.im + z0.im;
        if (XHOP) {
          const float4 u = pairs[up[j]], d = pairs[dn[j]];
          sr[0] += d.x + u.x;
          si[0] += d.y + u.y;
          sr[1] += d.z + u.z;
          si[1] += d.w + u.w;
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const cpx<float> z = h ? z1 : z0;
          const float mq = fmaf(z.im, z.im, z.re * z.re);
          const float cr = fmaf(-Ui, mq, A0), ci = fmaf(Ur, mq, B0);
          // increment = noise + dt iF + drift + hopping, all small, added to beta once
          float dr = fmaf(rs[h], cs[h], fr);
          dr = fmaf(cr, z.re, dr);
          dr = fmaf(-ci, z.im, dr);
          dr = fmaf(jdt, si[h], dr);
          float di = fmaf(rs[h], sn[h], fi);
          di = fmaf(cr, z.im, di);
          di = fmaf(ci, z.re, di);
          di = fmaf(-jdt, sr[h], di);
          beta[2 * j + h] = mk(z.re + dr, z.im + di);
        }
      }
    }
    __syncthreads();
  }
}

// Built for three waves per SIMD (three 256-thread workgroups per CU, at most
// 168 VGPRs; a 512-point sweep frame 42.5 -> 34.8 ms against two workgroups
// per CU at 204 VGPRs; four per CU spill and lose on stacked 2D items).
template <bool CAPTURE_ONLY>
__global__ void __launch_bounds__(kSweepBlock, 3) sweep_step_kernel(SweepArgs g) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* red = reinterpret_cast<double*>(smem);  // [8]
  cpx<float>* bufA = reinterpret_cast<cpx<float>*>(smem + 64);
  cpx<float>* bufB = bufA + kSweepMaxSites;
  cpx<float>* twx = bufB + kSweepMaxSites;    // [kMaxNx]
  cpx<float>* twy = twx + kMaxNx;             // [kMaxNy]
  float* kcx = reinterpret_cast<float*>(twy + kMaxNy);  // [kMaxNx]
  float* kcy = kcx + kMaxNx;                  // [kMaxNy]
  __shared__ int s_item;
  __shared__ double s_bins[kBins + 2];

  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = (int)atomicAdd(g.counter, 1u);
    __syncthreads();
    const int it = s_item;
    if (it >= g.n_items) break;  // every workgroup reaches this once the queue is drained
    const SweepItemDev item = g.items[it];
    const SweepPointDev P = g.points[item.point];
    // trajectories in the item: the step launch stacks P.tb of every point,
    // the capture launch takes 2D points one trajectory at a time
    const int ntr = (CAPTURE_ONLY && P.fft_x) ? 1 : min(P.tb, P.n_traj - item.t0);
    const int nxt = P.fft_x ? P.nx : 1;        // rows per trajectory
    const int rows_traj = P.fft_x ? 1 : ntr;  // chains: one trajectory per row
    const int nx = nxt * ntr;                  // rows of the item array
    const int ny = P.ny;
    const int m = nx * ny;
    const int npairs = m / 2;  // ny is even for every supported size
    if (CAPTURE_ONLY) {  // FFT tables, for the frame's k-space sums only
      for (int i = threadIdx.x; i < ny; i += blockDim.x) {
        twy[i] = g.tw_pool[P.tw_y + i];
        kcy[i] = g.kc_pool[P.kc_y + i];
      }
      if (P.fft_x)
        for (int i = threadIdx.x; i < P.nx; i += blockDim.x) {
          twx[i] = g.tw_pool[P.tw_x + i];
          kcx[i] = g.kc_pool[P.kc_x + i];
        }
    }
    __syncthreads();

    // the item's state: this thread's pairs p = t + 256 j in registers
    cpx<float>* gbeta = g.beta + P.state_off + (size_t)item.t0 * nxt * ny;
    auto load_state = [&](cpx<float> (&beta)[2 * kPPT]) __attribute__((always_inline)) {
#pragma unroll
      for (int j = 0; j < kPPT; ++j) {
        const int p = threadIdx.x + kSweepBlock * j;
        if (p < npairs) {
          beta[2 * j] = gbeta[2 * p];
          beta[2 * j + 1] = gbeta[2 * p + 1];
        }
      }
    };

    if constexpr (!CAPTURE_ONLY) {
      // the step launch: no capture, neither bufB nor the FFT tables.  Each
      // kind of item loads, steps and stores its own copy of the state, so no
      // register array is live across the three bodies (one shared copy made
      // the register allocator spill ~26 values per lane at this occupancy)
      auto run = [&](auto xhop, auto multi, int rows) __attribute__((always_inline)) {
        cpx<float> beta[2 * kPPT];
        load_state(beta);
        step_item<decltype(xhop)::value, decltype(multi)::value>(g, P, item, bufA, beta, rows, ny, npairs);
#pragma unroll
        for (int j = 0; j < kPPT; ++j) {
          const int p = threadIdx.x + kSweepBlock * j;
          if (p < npairs) {
            gbeta[2 * p] = beta[2 * j];
            gbeta[2 * p + 1] = beta[2 * j + 1];
          }
        }
      };
      if (!P.fft_x)
        run(std::false_type{}, std::true_type{}, 1);
      else if (ntr == 1)
        run(std::true_type{}, std::false_type{}, nxt);
      else
        run(std::true_type{}, std::true_type{}, nxt);
    } else {
      cpx<float> beta[2 * kPPT];
      load_state(beta);
      // the capture launch records the frame: per-trajectory means, mean
      // field and k-space bins of the state
#pragma unroll
      for (int j = 0; j < kPPT; ++j) {
        const int p = threadIdx.x + kSweepBlock * j;
        if (p < npairs) {
          bufA[2 * p] = beta[2 * j];
          bufA[2 * p + 1] = beta[2 * j + 1];
        }
      }
      __syncthreads();
      {
        // per-trajectory site means of |beta|^2 and |beta|^4 (FP64), summed
        // in a fixed order so that repeated runs agree bit for bit; for
        // stacked chains every row of the item is a trajectory
        if (threadIdx.x < kBins + 2) s_bins[threadIdx.x] = 0.0;
        if (P.fft_x) {
          double a2 = 0.0, a4 = 0.0;
#pragma unroll
          for (int j = 0; j < kPPT; ++j) {
            const int p = threadIdx.x + kSweepBlock * j;
            if (p < npairs)
#pragma unroll
              for (int h = 0; h < 2; ++h) {
                const cpx<float> z = beta[2 * j + h];
                const double nn = double(z.re) * double(z.re) + double(z.im) * double(z.im);
                a2 += nn;
                a4 += nn * nn;
              }
          }
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) {
            a2 += __shfl_xor(a2, o, 64);
            a4 += __shfl_xor(a4, o, 64);
          }
          if ((threadIdx.x & 63) == 0) {
            red[2 * (threadIdx.x >> 6)] = a2;
            red[2 * (threadIdx.x >> 6) + 1] = a4;
          }
          __syncthreads();
          if (threadIdx.x == 0) {
            double s2 = 0, s4 = 0;
            for (int w = 0; w < kSweepBlock / 64; ++w) {
              s2 += red[2 * w];
              s4 += red[2 * w + 1];
            }
            const size_t o = P.m_off + (size_t)item.t0 * g.n_frames + g.frame;
            g.traj_m2[o] = s2 / double(m);
            g.traj_m4[o] = s4 / double(m);
          }
        } else {
          for (int r = threadIdx.x; r < rows_traj; r += blockDim.x) {
            double s2 = 0, s4 = 0;
            for (int c = 0; c < ny; ++c) {
              const cpx<float> z = bufA[r * ny + c];
              const double nn = double(z.re) * double(z.re) + double(z.im) * double(z.im);
              s2 += nn;
              s4 += nn * nn;
            }
            const size_t o = P.m_off + (size_t)(item.t0 + r) * g.n_frames + g.frame;
            g.traj_m2[o] = s2 / double(ny);
            g.traj_m4[o] = s4 / double(ny);
          }
        }
        // ensemble sum of beta over all sites (mean field)
        double br = 0.0, bi = 0.0;
#pragma unroll
        for (int j = 0; j < kPPT; ++j) {
          const int p = threadIdx.x + kSweepBlock * j;
          if (p < npairs)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              br += double(beta[2 * j + h].re);
              bi += double(beta[2 * j + h].im);
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
          br += __shfl_xor(br, o, 64);
          bi += __shfl_xor(bi, o, 64);
        }
        __syncthreads();  // s_bins cleared, bufA reads done before the FFT overwrites it
        if ((threadIdx.x & 63) == 0) {
          atomicAdd(&s_bins[kBins], br);
          atomicAdd(&s_bins[kBins + 1], bi);
        }
      }

      {
        // k-space populations: forward FFT, rows (length ny) then columns
        // (length nx) for 2D points; |DFT beta|^2 per radial bin, summed over
        // this item's trajectories
        // the radix plans are read where they lie in global memory: indexing
        // the local copy's plan arrays would put the whole copy in scratch
        const SweepPointDev& Pg = g.points[item.point];
        cpx<float>* cur = fft_dim<-1>(bufA, bufB, Pg.py, nx, ny, 1, twy);
        if (P.fft_x) {
          cpx<float>* other = (cur == bufA) ? bufB : bufA;
          cur = fft_dim<-1>(cur, other, Pg.px, ny, 1, ny, twx);
        }
        for (int i = threadIdx.x; i < m; i += blockDim.x) {
          const cpx<float> f = cur[i];
          const int mode = P.fft_x ? i : i % ny;
          atomicAdd(&s_bins[g.bin_pool[P.bin_off + mode]], double(f.re) * double(f.re) + double(f.im) * double(f.im));
        }
        __syncthreads();
        if (threadIdx.x < kBins + 2)
          atomicAdd(&g.ksum[((size_t)P.index * g.n_frames + g.frame) * (kBins + 2) + threadIdx.x], s_bins[threadIdx.x]);
      }
    }
  }
}

// vacuum-type start beta_0 = beta_init + sigma xi (TWMCBaseSolver.cpp:126-134)
__global__ void sweep_init_kernel(const SweepPointDev* points, int n_points, const double* binit, cpx<float>* beta,
                                  uint32_t key0, uint32_t key1) {
  for (int p = 0; p < n_points; ++p) {
    const SweepPointDev P = points[p];
    const int n = P.nx * P.ny;
    const size_t total = (size_t)P.n_traj * n;
    const double br = binit[4 * p], bi = binit[4 * p + 1], sr = binit[4 * p + 2], si = binit[4 * p + 3];
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
      const int t = (int)(idx / n), i = (int)(idx - (size_t)t * n);
      const u32x4 w = philox4x32_10(make_counter(P.traj_id0 + t, TAG_INIT, 0u, (uint32_t)(i >> 1)), key0, key1);
      double zr, zi;
      if (i & 1) box_muller_f64(w.z, w.w, zr, zi); else box_muller_f64(w.x, w.y, zr, zi);
      beta[P.state_off + idx] = mk(float(br + (sr * zr - si * zi)), float(bi + (sr * zi + si * zr)));
    }
  }
}

// dynamic LDS: the capture launch holds bufA, bufB and the FFT tables; the
// step launch only bufA (so more workgroups fit on a CU)
size_t sweep_lds_bytes(bool capture) {
  return capture ? 64 + 2 * kSweepMaxSites * 8 + (kMaxNx + kMaxNy) * (8 + 4) : 64 + kSweepMaxSites * 8;
}

}  // namespace
}  // namespace twmc

using namespace twmc;

namespace {

thread_local std::string g_sweep_error;

struct SweepCtx {
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  double dt = 0.0;
  uint64_t frame_steps = 1, seed = 0;
  std::vector<twmc_sweep_point> pts;
  std::vector<SweepPointDev> dev_pts;
  std::vector<int> tb;  // trajectories per work item, per point
  std::vector<cpx<float>> tw_pool;
  std::vector<float> kc_pool;
  std::vector<unsigned char> bin_pool;
  uint64_t total_traj = 0, total_sites = 0;
  void *d_points = nullptr, *d_items = nullptr, *d_counter = nullptr, *d_tw = nullptr, *d_kc = nullptr,
       *d_beta = nullptr, *d_m2 = nullptr, *d_m4 = nullptr, *d_binit = nullptr, *d_bins = nullptr,
       *d_ksum = nullptr;
  size_t ksum_cap = 0;
  size_t m_cap = 0;
  int n_items = 0, n_items_cap = 0, grid = 0, grid_cap = 0;
  void* d_items_cap = nullptr;  // the capture launch's items: 2D points one trajectory each
  twmc_stats stats{};
};

int sfail(SweepCtx* c, int code, const std::string& m) {
  if (c) c->err = m;
  g_sweep_error = m;
  return code;
}

#define SCHK(c, expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) return sfail(c, TWMC_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

int add_tables(SweepCtx* c, int L, int& tw_off, int& kc_off) {
  tw_off = (int)c->tw_pool.size();
  kc_off = (int)c->kc_pool.size();
  for (int m = 0; m < L; ++m) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * m / L;
    c->tw_pool.push_back(mk(float(cosl(a)), float(sinl(a))));
    c->kc_pool.push_back(L == 1 ? 0.0f : float(cos(2.0 * M_PI / double(L) * double(m))));
  }
  return 0;
}

// radial bin of every mode of an nx x ny lattice: bin 0 is k = 0; the other
// modes go to 1 + floor(|k| / w) (at most 31) with w = pi sqrt 2 / 31 and
// |k| = 2 pi sqrt((kx'/nx)^2 + (ky'/ny)^2), kx' = min(kx, nx - kx)
int add_bins(SweepCtx* c, int nx, int ny) {
  int off = (int)c->bin_pool.size();
  const double w = M_PI * std::sqrt(2.0) / 31.0;
  for (int kx = 0; kx < nx; ++kx)
    for (int ky = 0; ky < ny; ++ky) {
      int ax = std::min(kx, nx - kx), ay = std::min(ky, ny - ky);
      if (ax == 0 && ay == 0) {
        c->bin_pool.push_back(0);
        continue;
      }
      double k = 2.0 * M_PI * std::sqrt(double(ax) * ax / (double(nx) * nx) + double(ay) * ay / (double(ny) * ny));
      int b = 1 + std::min(30, (int)std::floor(k / w));
      c->bin_pool.push_back((unsigned char)b);
    }
  return off;
}

// lays out points and the two work-item lists (largest first) for the
// current tb: the step launch stacks tb trajectories of every point per item,
// the capture launch takes 2D points one trajectory per item (chains as in the
// step launch)
int build_items(SweepCtx* c) {
  for (int pass = 0; pass < 2; ++pass) {
    const bool cap = pass == 1;
    std::vector<std::pair<double, SweepItemDev>> items;
    for (size_t p = 0; p < c->pts.size(); ++p) {
      const SweepPointDev& P = c->dev_pts[p];
      const int step = (cap && P.fft_x) ? 1 : c->tb[p];
      for (int t0 = 0; t0 < P.n_traj; t0 += step) {
        const int rows = std::min(step, P.n_traj - t0);
        const double cost = double(rows) * P.nx * P.ny * std::log2(double(P.nx * P.ny) + 1.0);
        items.push_back({cost, SweepItemDev{(int)p, t0}});
      }
    }
    std::stable_sort(items.begin(), items.end(), [](auto& a, auto& b) { return a.first > b.first; });
    std::vector<SweepItemDev> flat;
    for (auto& it : items) flat.push_back(it.second);
    void*& d = cap ? c->d_items_cap : c->d_items;
    (cap ? c->n_items_cap : c->n_items) = (int)flat.size();
    if (d) (void)hipFree(d);
    d = nullptr;
    SCHK(c, hipMalloc(&d, flat.size() * sizeof(SweepItemDev)));
    SCHK(c, hipMemcpy(d, flat.data(), flat.size() * sizeof(SweepItemDev), hipMemcpyHostToDevice));
  }
  for (size_t p = 0; p < c->pts.size(); ++p) c->dev_pts[p].tb = c->tb[p];
  SCHK(c, hipMemcpy(c->d_points, c->dev_pts.data(), c->dev_pts.size() * sizeof(SweepPointDev),
                    hipMemcpyHostToDevice));
  return TWMC_OK;
}

// one segment: reset the queues, the frame capture (if any), then the steps;
// each persistent launch drains its own queue counter
hipError_t launch_segment(SweepCtx* c, const SweepArgs& a) {
  hipLaunchKernelGGL(sweep_reset_kernel, dim3(1), dim3(64), 0, c->stream, (unsigned int*)c->d_counter);
  if (a.capture) {
    SweepArgs ca = a;
    ca.n_steps = 0;
    ca.counter = (unsigned int*)c->d_counter;
    ca.items = (const SweepItemDev*)c->d_items_cap;
    ca.n_items = c->n_items_cap;
    hipLaunchKernelGGL(sweep_step_kernel<true>, dim3(c->grid_cap), dim3(kSweepBlock), sweep_lds_bytes(true), c->stream,
                       ca);
  }
  if (a.n_steps > 0) {
    SweepArgs sa = a;
    sa.capture = 0;
    sa.counter = (unsigned int*)c->d_counter + 1;
    hipLaunchKernelGGL(sweep_step_kernel<false>, dim3(c->grid), dim3(kSweepBlock), sweep_lds_bytes(false), c->stream,
                       sa);
  }
  return hipGetLastError();
}

SweepArgs base_args(SweepCtx* c) {
  SweepArgs a{};
  a.points = (const SweepPointDev*)c->d_points;
  a.items = (const SweepItemDev*)c->d_items;
  a.n_items = c->n_items;
  a.counter = (unsigned int*)c->d_counter;
  a.tw_pool = (const cpx<float>*)c->d_tw;
  a.kc_pool = (const float*)c->d_kc;
  a.beta = (cpx<float>*)c->d_beta;
  a.traj_m2 = (double*)c->d_m2;
  a.traj_m4 = (double*)c->d_m4;
  a.bin_pool = (const unsigned char*)c->d_bins;
  a.ksum = (double*)c->d_ksum;
  a.dt = (float)c->dt;
  a.key0 = uint32_t(c->seed & 0xffffffffu);
  a.key1 = uint32_t(c->seed >> 32);
  return a;
}

int init_state(SweepCtx* c) {
  hipLaunchKernelGGL(sweep_init_kernel, dim3(1024), dim3(256), 0, c->stream, (const SweepPointDev*)c->d_points,
                     (int)c->pts.size(), (const double*)c->d_binit, (cpx<float>*)c->d_beta,
                     uint32_t(c->seed & 0xffffffffu), uint32_t(c->seed >> 32));
  SCHK(c, hipGetLastError());
  return TWMC_OK;
}

}  // namespace

namespace twmc {

// The main API's FP32 FFT-hopping batches of one site-uniform system run
// their steps on this kernel as a single point: the state layout is the
// same (trajectories contiguous, each row-major), each step item stacks tb
// trajectories, and the trajectory ids come from the batch's id array.  The
// batch's frames stay with the LDS kernel (its capture-only launch), which
// forms the per-site and per-mode sums of that API.
bool stack_supported(int nx, int ny) {
  return ny % 2 == 0 && ny <= kMaxNy && nx <= kMaxNx && nx * ny <= kSweepMaxSites && (nx == 1 || ny > 1);
}

hipError_t stack_prepare(StackPlan& sp, int nx, int ny, int n_traj, const double rsl[2], const double u[2],
                         const double f[2], double dt4, double kscale) {
  stack_release(sp);
  SweepPointDev P{};
  P.nx = nx;
  P.ny = ny;
  P.fft_x = nx > 1 ? 1 : 0;
  P.tb = std::max(1, std::min({n_traj, 2048 / (nx * ny), kMaxRows}));
  P.rsl_re = float(rsl[0]);
  P.rsl_im = float(rsl[1]);
  P.u_re = float(u[0]);
  P.u_im = float(u[1]);
  P.f_re = float(f[0]);
  P.f_im = float(f[1]);
  P.dt4 = float(dt4);
  P.kscale = float(kscale);
  P.n_traj = n_traj;
  std::vector<SweepItemDev> items;
  for (int t0 = 0; t0 < n_traj; t0 += P.tb) items.push_back(SweepItemDev{0, t0});
  sp.n_items = (int)items.size();
  sp.tb = P.tb;
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)sweep_step_kernel<false>, kSweepBlock,
                                                     sweep_lds_bytes(false));
  sp.grid = std::max(1, std::min(sp.n_items, std::max(1, per_cu) * cus));
  if (e == hipSuccess) e = hipMalloc(&sp.d_point, sizeof(SweepPointDev));
  if (e == hipSuccess) e = hipMalloc(&sp.d_items, items.size() * sizeof(SweepItemDev));
  if (e == hipSuccess) e = hipMalloc(&sp.d_counter, 2 * sizeof(unsigned int));
  if (e == hipSuccess) e = hipMemcpy(sp.d_point, &P, sizeof(SweepPointDev), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(sp.d_items, items.data(), items.size() * sizeof(SweepItemDev), hipMemcpyHostToDevice);
  if (e != hipSuccess) stack_release(sp);
  return e;
}

hipError_t stack_launch(const StackPlan& sp, const StepArgs& a, hipStream_t s) {
  SweepArgs g{};
  g.points = (const SweepPointDev*)sp.d_point;
  g.items = (const SweepItemDev*)sp.d_items;
  g.n_items = sp.n_items;
  g.counter = (unsigned int*)sp.d_counter;
  g.beta = (cpx<float>*)a.beta;
  g.dt = float(a.dt);
  g.key0 = a.key0;
  g.key1 = a.key1;
  g.step0 = a.step0;
  g.n_steps = a.n_steps;
  g.traj_ids = a.traj_ids;
  hipLaunchKernelGGL(sweep_reset_kernel, dim3(1), dim3(64), 0, s, (unsigned int*)sp.d_counter);
  hipLaunchKernelGGL(sweep_step_kernel<false>, dim3(sp.grid), dim3(kSweepBlock), sweep_lds_bytes(false), s, g);
  return hipGetLastError();
}

void stack_release(StackPlan& sp) {
  void* bufs[] = {sp.d_point, sp.d_items, sp.d_counter};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  sp.d_point = sp.d_items = sp.d_counter = nullptr;
  sp.n_items = 0;
}

}  // namespace twmc

extern "C" {

const char* twmc_sweep_last_error(const twmc_sweep_ctx* ctx) {
  return ctx ? reinterpret_cast<const SweepCtx*>(ctx)->err.c_str() : g_sweep_error.c_str();
}

int twmc_sweep_create(const twmc_sweep_point* pts, uint64_t n_points, double dt, uint64_t frame_steps,
                      uint64_t seed, int device, twmc_sweep_ctx** ctx_out) {
  if (!pts || !ctx_out || n_points == 0) return sfail(nullptr, TWMC_ERR_INVALID, "empty sweep");
  if (!(dt > 0.0) || frame_steps < 1) return sfail(nullptr, TWMC_ERR_INVALID, "dt and frame_steps must be positive");
  // one launch advances one frame; its step count is an int kernel argument
  if (frame_steps > (uint64_t(1) << 30)) return sfail(nullptr, TWMC_ERR_INVALID, "frame_steps above 2^30");
  *ctx_out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return sfail(nullptr, TWMC_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= ndev) return sfail(nullptr, TWMC_ERR_INVALID, "device index out of range");
  SweepCtx* c = new SweepCtx();
  c->device = device;
  c->dt = dt;
  c->frame_steps = frame_steps;
  c->seed = seed;
  c->pts.assign(pts, pts + n_points);
  std::map<std::pair<int, int>, std::pair<int, int>> tables;  // length -> (tw, kc) offsets
  std::map<std::pair<int, int>, int> bin_tables;                // shape -> bin table offset
  auto table_for = [&](int L) {
    auto it = tables.find({L, 0});
    if (it != tables.end()) return it->second;
    int a, b;
    add_tables(c, L, a, b);
    tables[{L, 0}] = {a, b};
    return std::make_pair(a, b);
  };
  uint64_t state_off = 0, m_off = 0;
  for (uint64_t p = 0; p < n_points; ++p) {
    const twmc_sweep_point& q = pts[p];
    int nx = q.nx, ny = q.ny;
    if (nx > 1 && ny == 1) std::swap(nx, ny);  // chains are 1 x N
    SweepPointDev P{};
    P.nx = nx;
    P.ny = ny;
    P.fft_x = nx > 1 ? 1 : 0;
    if (nx * ny > kSweepMaxSites || ny > kMaxNy || nx > kMaxNx || (ny % 2) != 0 || q.n_traj == 0 ||
        q.n_traj > (uint64_t)INT32_MAX || !make_radix_plan(ny, P.py) || !make_radix_plan(nx, P.px)) {
      delete c;
      return sfail(nullptr, TWMC_ERR_UNSUPPORTED,
                   "sweep point " + std::to_string(p) + ": up to 4096 sites, 64 rows, rows of even length <= 512");
    }
    auto ty = table_for(ny);
    auto tx = table_for(nx);
    P.tw_y = ty.first;
    P.kc_y = ty.second;
    P.tw_x = tx.first;
    P.kc_x = tx.second;
    // rsl = -i omega - gamma/2 (TWMCBaseSolver.cpp:74)
    P.rsl_re = float(q.omega[1] - q.gamma / 2.0);
    P.rsl_im = float(-q.omega[0]);
    P.u_re = float(q.U[0]);
    P.u_im = float(q.U[1]);
    P.f_re = float(q.F[0]);
    P.f_im = float(q.F[1]);
    P.dt4 = float(std::sqrt(q.gamma * dt / 4.0));
    P.kscale = float(-2.0 * q.J);
    P.n_traj = (int)q.n_traj;
    P.traj_id0 = q.traj_id0;
    P.state_off = state_off;
    P.m_off = m_off;
    P.index = (int)p;
    {
      auto bt = bin_tables.find({nx, ny});
      if (bt == bin_tables.end()) bt = bin_tables.emplace(std::make_pair(nx, ny), add_bins(c, nx, ny)).first;
      P.bin_off = bt->second;
    }
    state_off += (uint64_t)q.n_traj * nx * ny;
    m_off += q.n_traj;  // times n_frames at run time
    c->dev_pts.push_back(P);
    // default stacking: fill a workgroup with about 2048 sites
    c->tb.push_back(std::max(1, std::min({(int)q.n_traj, 2048 / (nx * ny), kMaxRows})));
  }
  c->total_traj = m_off;
  c->total_sites = state_off;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&c->d_points, c->dev_pts.size() * sizeof(SweepPointDev));
  if (e == hipSuccess) e = hipMalloc(&c->d_counter, 256);
  if (e == hipSuccess) e = hipMalloc(&c->d_tw, c->tw_pool.size() * sizeof(cpx<float>));
  if (e == hipSuccess) e = hipMalloc(&c->d_kc, c->kc_pool.size() * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&c->d_beta, state_off * sizeof(cpx<float>));
  std::vector<double> binit(4 * n_points);
  for (uint64_t p = 0; p < n_points; ++p) {
    binit[4 * p] = pts[p].beta_init[0];
    binit[4 * p + 1] = pts[p].beta_init[1];
    binit[4 * p + 2] = pts[p].beta_sigma[0];
    binit[4 * p + 3] = pts[p].beta_sigma[1];
  }
  if (e == hipSuccess) e = hipMalloc(&c->d_binit, binit.size() * 8);
  if (e == hipSuccess) e = hipMalloc(&c->d_bins, c->bin_pool.size());
  if (e == hipSuccess) e = hipMemcpy(c->d_bins, c->bin_pool.data(), c->bin_pool.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->d_tw, c->tw_pool.data(), c->tw_pool.size() * sizeof(cpx<float>), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->d_kc, c->kc_pool.data(), c->kc_pool.size() * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->d_binit, binit.data(), binit.size() * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipFuncSetAttribute((const void*)sweep_step_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sweep_lds_bytes(false));
  if (e == hipSuccess) e = hipFuncSetAttribute((const void*)sweep_step_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sweep_lds_bytes(true));
  int per_cu = 0, per_cu_cap = 0, cus = 0;
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)sweep_step_kernel<false>, kSweepBlock, sweep_lds_bytes(false));
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_cap, (const void*)sweep_step_kernel<true>, kSweepBlock, sweep_lds_bytes(true));
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess) {
    std::string m = std::string("sweep setup: ") + hipGetErrorString(e);
    twmc_sweep_destroy(reinterpret_cast<twmc_sweep_ctx*>(c));
    return sfail(nullptr, TWMC_ERR_HIP, m);
  }
  c->grid = std::max(1, per_cu) * cus;
  c->grid_cap = std::max(1, per_cu_cap) * cus;
  int rc = build_items(c);
  if (rc) {
    std::string m = c->err;
    twmc_sweep_destroy(reinterpret_cast<twmc_sweep_ctx*>(c));
    return sfail(nullptr, rc, m);
  }
  *ctx_out = reinterpret_cast<twmc_sweep_ctx*>(c);
  return TWMC_OK;
}

void twmc_sweep_destroy(twmc_sweep_ctx* ctx) {
  if (!ctx) return;
  SweepCtx* c = reinterpret_cast<SweepCtx*>(ctx);
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void* bufs[] = {c->d_points, c->d_items, c->d_items_cap, c->d_counter, c->d_tw, c->d_kc, c->d_beta, c->d_m2, c->d_m4, c->d_binit,
                  c->d_bins, c->d_ksum};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int twmc_sweep_set_batching(twmc_sweep_ctx* ctx, const int32_t* tb) {
  if (!ctx || !tb) return sfail(nullptr, TWMC_ERR_INVALID, "null argument");
  SweepCtx* c = reinterpret_cast<SweepCtx*>(ctx);
  for (size_t p = 0; p < c->pts.size(); ++p) {
    const SweepPointDev& P = c->dev_pts[p];
    const int v = tb[p];
    if (v < 1 || v > kMaxRows || (int64_t)v * P.nx * P.ny > kSweepMaxSites)
      return sfail(c, TWMC_ERR_INVALID, "stacking exceeds 4096 sites or 256 trajectories");
    c->tb[p] = std::min(v, P.n_traj);
  }
  return build_items(c);
}

int twmc_sweep_get_batching(twmc_sweep_ctx* ctx, int32_t* tb) {
  if (!ctx || !tb) return sfail(nullptr, TWMC_ERR_INVALID, "null argument");
  SweepCtx* c = reinterpret_cast<SweepCtx*>(ctx);
  for (size_t p = 0; p < c->tb.size(); ++p) tb[p] = c->tb[p];
  return TWMC_OK;
}

// Runs the sweep from the vacuum start; frames follow the reference loop with
// store_initial.  use_graph captures the frame loop into one hipGraph.
int twmc_sweep_run(twmc_sweep_ctx* ctx, double t_start, double t_end, int32_t use_graph, double* traj_m2, double* traj_m4,
                   double* ksum, uint64_t* n_frames_out, double* kernel_ms) {
  if (!ctx) return sfail(nullptr, TWMC_ERR_INVALID, "null argument");
  SweepCtx* c = reinterpret_cast<SweepCtx*>(ctx);
  SCHK(c, hipSetDevice(c->device));
  // frame schedule (TWMCBaseSolver.cpp:221-266), store_initial = true
  double dt_obs = c->dt * (double)c->frame_steps;
  uint64_t nf = (uint64_t)(size_t)nearbyint((t_end - t_start) / dt_obs) + 1;
  std::vector<uint64_t> cap{0};
  uint64_t i_frame = 1, i_step = 0;
  double t = t_start;
  while (t <= t_end) {
    t += c->dt;
    i_step++;
    if ((i_step % c->frame_steps == 0) && i_frame < nf) {
      cap.push_back(i_step);
      i_frame++;
      if (i_frame == nf) break;
    }
  }
  const uint64_t frames = i_frame;
  size_t need = (size_t)c->total_traj * frames;
  if (need > c->m_cap) {
    if (c->d_m2) (void)hipFree(c->d_m2);
    if (c->d_m4) (void)hipFree(c->d_m4);
    c->d_m2 = c->d_m4 = nullptr;
    SCHK(c, hipMalloc(&c->d_m2, need * 8));
    SCHK(c, hipMalloc(&c->d_m4, need * 8));
    c->m_cap = need;
  }
  const size_t kneed = c->pts.size() * frames * (kBins + 2);
  if (kneed > c->ksum_cap) {
    if (c->d_ksum) (void)hipFree(c->d_ksum);
    c->d_ksum = nullptr;
    SCHK(c, hipMalloc(&c->d_ksum, kneed * 8));
    c->ksum_cap = kneed;
  }
  SCHK(c, hipMemsetAsync(c->d_ksum, 0, kneed * 8, c->stream));  // before any graph capture
  // per-trajectory means are laid out [point block][traj][frame]
  std::vector<SweepPointDev> pts = c->dev_pts;
  for (auto& P : pts) P.m_off *= frames;
  SCHK(c, hipMemcpyAsync(c->d_points, pts.data(), pts.size() * sizeof(SweepPointDev), hipMemcpyHostToDevice, c->stream));
  int rc = init_state(c);
  if (rc) return rc;
  SCHK(c, hipStreamSynchronize(c->stream));

  // events and graph of this run, released on every return path
  struct RunResources {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~RunResources() {
      if (exec) (void)hipGraphExecDestroy(exec);
      if (graph) (void)hipGraphDestroy(graph);
      if (e1) (void)hipEventDestroy(e1);
      if (e0) (void)hipEventDestroy(e0);
    }
  } res;
  SCHK(c, hipEventCreate(&res.e0));
  SCHK(c, hipEventCreate(&res.e1));
  hipEvent_t e0 = res.e0, e1 = res.e1;
  if (use_graph) SCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  else SCHK(c, hipEventRecord(e0, c->stream));
  for (size_t k = 0; k < cap.size(); ++k) {
    uint64_t end = (k + 1 < cap.size()) ? cap[k + 1] : i_step;
    SweepArgs a = base_args(c);
    a.step0 = cap[k];
    a.n_steps = (int)(end - cap[k]);
    a.capture = 1;
    a.frame = (int)k;
    a.n_frames = (int)frames;
    hipError_t le = launch_segment(c, a);
    if (le != hipSuccess) {
      if (use_graph) {
        hipGraph_t g2;
        (void)hipStreamEndCapture(c->stream, &g2);
        if (g2) (void)hipGraphDestroy(g2);
      }
      return sfail(c, TWMC_ERR_HIP, std::string("sweep launch: ") + hipGetErrorString(le));
    }
  }
  if (use_graph) {
    SCHK(c, hipStreamEndCapture(c->stream, &res.graph));
    SCHK(c, hipGraphInstantiate(&res.exec, res.graph, nullptr, nullptr, 0));
    SCHK(c, hipEventRecord(e0, c->stream));
    SCHK(c, hipGraphLaunch(res.exec, c->stream));
  }
  SCHK(c, hipEventRecord(e1, c->stream));
  SCHK(c, hipStreamSynchronize(c->stream));
  float ms = 0.f;
  SCHK(c, hipEventElapsedTime(&ms, e0, e1));
  c->stats.step_kernel_ms += ms;
  c->stats.step_launches += cap.size();
  c->stats.site_steps += 0;
  for (size_t p = 0; p < c->pts.size(); ++p)
    c->stats.site_steps += (uint64_t)c->dev_pts[p].n_traj * c->dev_pts[p].nx * c->dev_pts[p].ny * i_step;
  c->stats.grid = c->grid;
  c->stats.block = kSweepBlock;
  c->stats.lds_bytes = (int)sweep_lds_bytes(false);
  if (kernel_ms) *kernel_ms = ms;
  if (n_frames_out) *n_frames_out = frames;
  if (traj_m2) SCHK(c, hipMemcpy(traj_m2, c->d_m2, need * 8, hipMemcpyDeviceToHost));
  if (traj_m4) SCHK(c, hipMemcpy(traj_m4, c->d_m4, need * 8, hipMemcpyDeviceToHost));
  if (ksum) SCHK(c, hipMemcpy(ksum, c->d_ksum, kneed * 8, hipMemcpyDeviceToHost));
  return TWMC_OK;
}

int twmc_sweep_mode_bins(twmc_sweep_ctx* ctx, uint64_t point, uint8_t* bins) {
  if (!ctx || !bins) return sfail(nullptr, TWMC_ERR_INVALID, "null argument");
  SweepCtx* c = reinterpret_cast<SweepCtx*>(ctx);
  if (point >= c->pts.size()) return sfail(c, TWMC_ERR_INVALID, "point index out of range");
  const SweepPointDev& P = c->dev_pts[point];
  std::memcpy(bins, &c->bin_pool[P.bin_off], (size_t)P.nx * P.ny);
  return TWMC_OK;
}

// Times each candidate stacking per lattice shape on a short run (steps steps
// of every point of that shape) and keeps the fastest per site-step.
int twmc_sweep_autotune(twmc_sweep_ctx* ctx, int32_t steps, int32_t* tb_out) {
  if (!ctx) return sfail(nullptr, TWMC_ERR_INVALID, "null argument");
  SweepCtx* c = reinterpret_cast<SweepCtx*>(ctx);
  if (steps < 1) steps = 20;
  std::map<std::pair<int, int>, std::vector<size_t>> by_shape;  // (rows, row length) -> points
  for (size_t p = 0; p < c->pts.size(); ++p) by_shape[{c->dev_pts[p].nx, c->dev_pts[p].ny}].push_back(p);
  std::vector<int32_t> best = c->tb;
  for (auto& kv : by_shape) {
    const int sites = kv.first.first * kv.first.second;
    if (2 * sites > kSweepMaxSites) {  // one trajectory per item is the only choice
      for (size_t p : kv.second) best[p] = 1;
      continue;
    }
    double best_rate = -1.0;
    for (int cand = 1; cand * sites <= kSweepMaxSites && cand <= kMaxRows; cand *= 2) {
      if (cand > c->dev_pts[kv.second[0]].n_traj && cand > 1) break;
      std::vector<int32_t> tb = c->tb;
      for (size_t p : kv.second) tb[p] = cand;
      // time only this shape: a context holding just these points
      std::vector<twmc_sweep_point> sub;
      for (size_t p : kv.second) sub.push_back(c->pts[p]);
      twmc_sweep_ctx* tctx = nullptr;
      int rc = twmc_sweep_create(sub.data(), sub.size(), c->dt, (uint64_t)steps, c->seed, c->device, &tctx);
      if (rc) return sfail(c, rc, twmc_sweep_last_error(nullptr));
      std::vector<int32_t> subtb(sub.size(), cand);
      rc = twmc_sweep_set_batching(tctx, subtb.data());
      double ms = 0.0;
      if (!rc) rc = twmc_sweep_run(tctx, 0.0, c->dt * steps, 0, nullptr, nullptr, nullptr, nullptr, &ms);
      uint64_t ss = reinterpret_cast<SweepCtx*>(tctx)->stats.site_steps;
      twmc_sweep_destroy(tctx);
      if (rc) return sfail(c, rc, "autotune run failed");
      double rate = ms > 0 ? double(ss) / ms : 0.0;
      if (rate > best_rate) {
        best_rate = rate;
        for (size_t p : kv.second) best[p] = cand;
      }
    }
  }
  int rc = twmc_sweep_set_batching(ctx, best.data());
  if (rc) return rc;
  if (tb_out)
    for (size_t p = 0; p < best.size(); ++p) tb_out[p] = best[p];
  return TWMC_OK;
}

int twmc_sweep_query(twmc_sweep_ctx* ctx, twmc_stats* st) {
  if (!ctx || !st) return sfail(nullptr, TWMC_ERR_INVALID, "null argument");
  *st = reinterpret_cast<SweepCtx*>(ctx)->stats;
  return TWMC_OK;
}

}  // extern "C"
