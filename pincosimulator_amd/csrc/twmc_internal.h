// twmc_internal.h -- launch interfaces shared between the C ABI layer
// (twmc_capi.hip) and the kernel translation units.
#ifndef TWMC_INTERNAL_H
#define TWMC_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace twmc {

// Number of accumulated per-site ensemble quantities per frame:
// 0: |beta|^2   1: |beta|^4   2: |DFT(beta)_k|^2   3: Re beta   4: Im beta
constexpr int kQ = 5;
constexpr int kMaxRadixPasses = 16;

// Factorisation of one FFT length into radices 8, 4, 2, 3, 5 (in that order),
// then the remaining prime factors in ascending order (direct-sum passes).
struct RadixPlan {
  int len;
  int npass;
  int radix[kMaxRadixPasses];
};

// Fills P for a length len >= 1; false only when the factors need more than
// kMaxRadixPasses passes (no length below 2^16 does).
inline bool make_radix_plan(int len, RadixPlan& P) {
  P.len = len;
  P.npass = 0;
  if (len < 1) return false;
  int m = len;
  auto take = [&](int r) {
    while (m % r == 0) {
      if (P.npass >= kMaxRadixPasses) return false;
      P.radix[P.npass++] = r;
      m /= r;
    }
    return true;
  };
  for (int r : {8, 4, 2, 3, 5})
    if (!take(r)) return false;
  for (int r = 7; r * r <= m; r += 2)
    if (!take(r)) return false;
  if (m > 1) {
    if (P.npass >= kMaxRadixPasses) return false;
    P.radix[P.npass++] = m;
  }
  return true;
}

// Everything a step kernel needs; plain data passed by value.
struct StepArgs {
  int nx, ny, n;             // lattice and site count
  int hopping;               // 0: FFT, 1: sparse matrix (CSC)
  int n_traj;                // trajectories in the batch
  int param_stride;          // 0: shared site parameters, n: per-trajectory
  double dt, dt4, inv_n;     // dt, sqrt(gamma dt / 4), 1 / N
  double kscale;             // -2 J  (K = i kscale (fx cos kx + fy cos ky))
  uint32_t key0, key1;       // Philox key (seed)
  uint64_t step0;            // absolute index of the segment's first step
  int n_steps;               // steps in this segment
  int capture;               // accumulate the state before the first step
  int frame;                 // frame index of that capture
  int n_frames;              // frames in the batch (trajectory dump stride)
  int dump;                  // write the captured state to traj_out
  RadixPlan px, py;          // plans for the x (length nx) and y (length ny) FFTs
  const uint64_t* traj_ids;  // [n_traj]
  void* beta;                // [n_traj][n] complex (T)
  const void* rsl;           // [sets][n] complex (T): -i omega - gamma/2 etc.
  const void* U;             // [sets][n] complex (T)
  const void* F;             // [sets][n] complex (T)
  const void* twx;           // [nx] complex (T): exp(-2 pi i m / nx)
  const void* twy;           // [ny] complex (T)
  const void* kcx;           // [nx] (T): [nx != 1] cos(2 pi i / nx)
  const void* kcy;           // [ny] (T)
  const int* csc_ptr;        // hopping matrix, column pointers [n+1]
  const int* csc_row;        // row index of each nonzero
  const void* csc_val;       // value (T)
  double* partials;          // [slots][kQ][n] slot-private FP64 ensemble sums
  double* traj_m2;           // [n_traj][n_frames]
  double* traj_m4;           // [n_traj][n_frames]
  void* traj_out;            // [n_traj][n_frames][n] complex (T)
  // MFMA matrix hopping: A = C^T split into BF16 high and low parts, [np][np]
  const void* a_hi;
  const void* a_lo;
  int np;                    // n padded to a multiple of 32
  int u_real;                // 1: Im U = 0 at every site, in every trajectory
  int site_uniform;          // 1: rsl, U and F are the same at every site and in every trajectory, F static
  const float* band_coef;    // banded kernel: [64 S][2W + 1] dt C_(j+d) j
  int band_w;                // banded kernel: half-width W
  int band_lieb;             // banded kernel: the coupling has the Lieb chain's pattern (S = 6, W = 2)
  // time-dependent drive: F(t) = ft_vals[id] + (t - ft_times[id]) ft_slopes[id]
  // with id the last table time below t (F = ft_vals[0] up to ft_times[0]);
  // t = t0 + s dt before step s of this launch, in FP64
  int ft_n;                  // table times, 0: static F
  double t0;
  const double* ft_times;    // [ft_n]
  const void* ft_vals;       // [ft_n][n] complex (T)
  const void* ft_slopes;     // [ft_n][n] complex (T), last row zero
};

// Interval of the F(t) table used at time t and the offset t - times[id]
// (NoisyMatrix::GetAtTimeWithSuggestion, NoisyMatrix.cpp:252-264).  `id`
// carries the previous interval: time only grows within a launch.
__device__ __forceinline__ double drive_interval(const double* times, int nt, double t, int& id) {
  if (t <= times[0]) {
    id = 0;
    return 0.0;
  }
  while (id + 1 < nt && times[id + 1] < t) ++id;
  return t - times[id];
}

struct InitArgs {
  int n, n_traj, mode;       // mode: 0 settings, 1 fixed point, 2 previous
  int per_traj_params;       // write per-trajectory rsl/U/F
  int onsite;                // 0: rsl = -i omega - gamma/2; 1: -i (detuning - E) - gamma/2
  double gamma, detuning;
  uint32_t key0, key1;
  const uint64_t* traj_ids;
  // site-resolved parameter bases and noise amplitudes (FP64, interleaved)
  const double* base[5];     // omega/E, U, F, beta_init, beta_sigma
  const double* amp[5];
  int noise[5];
  const double* beta_prev;   // mode 2
  void* beta;                // out [n_traj][n] (T)
  void* rsl;                 // out [sets][n] (T)
  void* U;
  void* F;
};

// Gaussian Wigner sample around a predicted steady state: per trajectory,
// DFT(beta)_k = N mean [k = 0] + kamp[k] (xi_re + i xi_im), beta = IDFT / N.
struct GaussArgs {
  int n, nx, ny, n_traj;
  RadixPlan px, py;
  const void* twx;           // [nx] complex (T): exp(-2 pi i m / nx)
  const void* twy;           // [ny] complex (T)
  const double* kamp;        // [n] standard deviation per quadrature of DFT(beta)_k
  double mean_re, mean_im;
  uint32_t key0, key1;
  const uint64_t* traj_ids;
  void* beta;                // out [n_traj][n] (T)
};

struct ReduceArgs {
  int n, slots;
  const double* partials_in;  // [slots][kQ][n]
  double* partials;           // same buffer, cleared after reading unless keep
  double* obs;                // [kQ][n] for one frame
  const int* perm_site;       // native index -> natural index, or null
  const int* perm_k;
  int slots_k;  // slots of the k-space row (q = 2) when > 0, else slots
  int keep;     // 1: the producers overwrite their rows, no clearing
};

// per-site real-space sums in site-major order (rows 0, 1, 3, 4 of the slots);
// at most this many trajectory chunks (partial slots) per frame
constexpr int kSiteSumChunks = 128;
struct SiteSumArgs {
  const void* beta;  // [n_traj][n] complex, precision of the batch
  double* partials;  // [chunks][kQ][n]
  int n, n_traj, chunk;
  int zero_k;        // 1: also write zeros to row 2 (no k-space sums)
};
hipError_t launch_site_sums(const SiteSumArgs& a, int precision, hipStream_t s);
// chunks of at most max_chunks for n_traj trajectories; *chunk = trajectories per chunk
int site_sum_chunks(uint64_t n_traj, int max_chunks, int* chunk);
hipError_t launch_capture64(const StepArgs& a, int grid, hipStream_t s);
hipError_t capture64_occupancy(int* blocks_per_cu);

hipError_t launch_init(const InitArgs& a, int precision, hipStream_t s);
hipError_t launch_gauss_init(const GaussArgs& a, int precision, int grid, hipStream_t s);
size_t gauss_lds_bytes(int nx, int ny, int precision);
hipError_t launch_generic(const StepArgs& a, int precision, int grid, int block, size_t lds,
                          hipStream_t s);
size_t generic_lds_bytes(int nx, int ny, int precision);
int generic_pairs_per_thread(int n, int block);
// the LDS kernel as a step launch with the hopping in its nearest-neighbour
// form and no capture (FFT hopping): one LDS copy of the state
size_t generic_nn_lds_bytes(int nx, int ny, int precision, bool site_uniform);
// trajectories per workgroup of the nearest-neighbour step launch (1, 2 or 4):
// its grid covers ceil(B / that) workgroups
int generic_nn_traj_per_block(int n);
hipError_t generic_nn_occupancy(int precision, int nx, int ny, size_t lds, bool site_uniform,
                                int* blocks_per_cu);
hipError_t launch_generic_nn(const StepArgs& a, int precision, int grid, size_t lds, hipStream_t s);
hipError_t generic_occupancy(int precision, int n, int block, size_t lds, int* blocks_per_cu);
hipError_t launch_reduce(const ReduceArgs& a, hipStream_t s);

// register-resident 64x64 FP32 kernel
bool fused64_supported(int nx, int ny, int hopping, int precision, int param_stride);
hipError_t launch_fused64(const StepArgs& a, int grid, hipStream_t s);
hipError_t fused64_occupancy(int* blocks_per_cu);
void fused64_perms(int* perm_site, int* perm_k);  // [4096] each
bool stencil64_supported(int nx, int ny, int hopping, int precision, int param_stride);
hipError_t launch_stencil64(const StepArgs& a, int grid, hipStream_t s);
hipError_t stencil64_occupancy(int* blocks_per_cu);

// matrix hopping on MFMA (FP32 state, BF16 x 3 products, FP32 accumulation)
bool mfma_supported(int n, int hopping, int precision, int param_stride);
int mfma_padded(int n);
size_t mfma_lds_bytes(int np);
hipError_t launch_mfma(const StepArgs& a, int grid, hipStream_t s);

// one-point batches of the main API on the sweep kernel: FP32, FFT hopping,
// site-uniform shared parameters, no F(t)
struct StackPlan {
  void* d_point = nullptr;
  void* d_items = nullptr;
  void* d_counter = nullptr;
  int n_items = 0, grid = 0, tb = 1;
};
bool stack_supported(int nx, int ny);
hipError_t stack_prepare(StackPlan& sp, int nx, int ny, int n_traj, const double rsl[2], const double u[2],
                         const double f[2], double dt4, double kscale);
hipError_t stack_launch(const StackPlan& sp, const StepArgs& a, hipStream_t s);
void stack_release(StackPlan& sp);

// banded coupling (matrix hopping with every nonzero C_ij at |i - j| <= W,
// cyclic when N = 64 S): one wave per trajectory, S sites per lane
constexpr int kBandMaxSites = 512;
constexpr int kBandMaxWidth = 3;
int band_sites_per_lane(int n);
int band_width(const double* coupling, int n);  // W, or -1 when n is out of range
bool band_supported(int n, int hopping, int precision, int param_stride, int width);
void band_coefficients(const double* coupling, int n, int w, double dt, float* out);  // [64 S][2w+1]
// true when the [64 S][2w+1] table has nonzeros only where a Lieb chain has
// bonds (A: +1; B: -2, -1, +1; C: -1, +2, by site index mod 3) and S = 6, W = 2
bool band_lieb_pattern(const float* coef, int n, int w);
hipError_t launch_band(const StepArgs& a, int grid, hipStream_t s);
hipError_t band_occupancy(int n, int w, int lieb, int* blocks_per_cu);
hipError_t mfma_occupancy(int np, int* blocks_per_cu);
// splits A (row-major np x np, FP64) into BF16 hi and lo parts (raw 16-bit)
void mfma_split(const double* a, size_t count, uint16_t* hi, uint16_t* lo);
int mfma_rows(int np);  // rows of the zero-padded A operand
int mfma_block(int np);  // threads per workgroup of the MFMA step kernel
void mfma_swizzle(const double* a, int rows, int np, double* out);  // row-major -> fragment order

}  // namespace twmc

#endif
