// twmc_capi.hip -- C ABI of libtwmc_hip.so (declared in include/twmc_hip.h).
//
// A context holds one parameter set on one GPU (the reference's
// TWMCSystemData plus a Solver after Setup, src/TWMC/Solvers/TWMCBaseSolver.cpp:24-92);
// a batch is the vector of tasks of one Solver::Compute call
// (TWMCBaseSolver.cpp:94-283), run as a sequence of segments: one launch of
// the step kernel per stored frame, each followed by the reduction of the
// frame's ensemble sums.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/twmc_hip.h"
#include "twmc_internal.h"
#include "twmc_rng.h"

using namespace twmc;

// Device memory of twmc_advance_noalloc beyond the caller's buffers (the
// site-shared parameter tables and the per-chunk partial sums): one
// fixed-size array, of which every GPU the library runs on holds its own copy
// until the process ends.  Host code uses it only through the per-device
// address from hipGetSymbolAddress (g_scratch_addr), never through its name.
static __device__ __align__(256) unsigned char scratch[TWMC_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_addr;  // device address of `scratch` per device index

thread_local std::string g_last_error;

struct Segment {
  int capture;     // accumulate the state before the first step
  int frame;       // frame index of that capture
  uint64_t steps;  // steps in the segment
};

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (b == 0) return hipSuccess;
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) bytes = b;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

struct Ctx {
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  twmc_params p{};
  int n = 0;
  int precision = 64;
  size_t es = 16;  // bytes per complex element
  int kernel = TWMC_KERNEL_GENERIC;
  int noisy_params = 0;
  double dt4 = 0.0;
  RadixPlan px{}, py{};
  // host copies of the site-resolved inputs (2N doubles each)
  std::vector<double> base[5], amp[5];
  int noise[5] = {0, 0, 0, 0, 0};
  // device-resident constants
  DevBuf d_base[5], d_amp[5], d_twx, d_twy, d_kcx, d_kcy, d_cscp, d_cscr, d_cscv, d_perm_site,
      d_perm_k, d_ahi, d_alo, d_kamp, d_ft_times, d_ft_vals, d_ft_slopes, d_band;
  int ft_n = 0;  // F(t) table times (0: static F)
  int np = 0;  // padded site count of the MFMA kernel
  int band_w = -1;  // half-width of the coupling matrix (matrix hopping), -1: not banded
  int band_lieb = 0;  // the banded coupling has the Lieb chain's pattern
  StackPlan stack;  // TWMC_KERNEL_STACK: the batch's point and work items on the sweep kernel
  // per-batch buffers
  DevBuf d_ids, d_beta, d_rsl, d_U, d_F, d_partials, d_obs, d_m2, d_m4, d_traj, d_prev;
  int csc_nnz = 0;
  // batch state
  bool batch_active = false;
  twmc_batch_desc desc{};
  uint64_t n_traj = 0;
  uint64_t n_frames = 0, n_steps = 0;
  double t_final = 0.0;
  int truncated = 0;
  std::vector<Segment> segs;
  size_t next_seg = 0;
  uint64_t steps_done = 0;
  // completion of the batch on the device, read by other threads
  // (twmc_progress): one event recorded after each segment's launches, the
  // steps done at that point, and how many of them have fired
  mutable std::mutex prog_mu;
  std::vector<hipEvent_t> seg_ev;
  std::vector<uint64_t> seg_steps;
  size_t seg_recorded = 0;
  mutable size_t seg_completed = 0;
  int grid = 0, block = 256;
  int cap_grid = 0;  // stencil64: grid of the capture launch (the 64x64 FFT kernel's)
  int grid_nn = 0;   // LDS kernel, FFT hopping: grid of the nearest-neighbour step launch
  size_t lds_nn = 0;
  size_t lds = 0;
  // timing
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_step, ev_red;
  twmc_stats stats{};
};

int fail(Ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  g_last_error = msg;
  return code;
}

int hip_fail(Ctx* c, hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return fail(c, TWMC_ERR_HIP, m);
}

#define HIPCHK(c, expr)                                 \
  do {                                                  \
    hipError_t e_ = (expr);                             \
    if (e_ != hipSuccess) return hip_fail(c, e_, #expr); \
  } while (0)

// Grid of a kernel that runs one trajectory per single-wave workgroup in a
// grid-stride loop and is bound by instruction issue.  With w waves per SIMD
// (up to what fits) the batch runs in rounds; a round with a waves active on
// a SIMD costs about a issue periods, but no less than a latency floor of
// 2.4 (measured: two register-resident 64x64 waves keep the main VALU 84.5%
// busy), so a partly filled last round costs the floor, not nothing.  Take
// the w with the smallest total, the larger w on ties.  16384 trajectories on
// 1024 SIMDs: three waves per SIMD (5 full rounds and a one-wave tail, 17.4)
// rather than two (8 rounds at the floor, 19.2) -- 2% faster measured;
// 8192 trajectories with up to four waves: four (2 rounds, 8).
int balanced_wave_grid(size_t B, int per_cu, int cus) {
  const size_t simds = 4 * (size_t)cus;
  const int wmax = std::max(1, per_cu / 4);
  auto round_cost = [](size_t active) { return std::max<size_t>(5 * active, 12); };  // x5: floor 2.4
  size_t best_cost = SIZE_MAX, best = simds;
  for (int w = 1; w <= wmax; ++w) {
    const size_t g = (size_t)w * simds;
    const size_t full = B / g, rest = B - full * g;
    const size_t cost = full * round_cost((size_t)w) + (rest ? round_cost((rest + simds - 1) / simds) : 0);
    if (cost <= best_cost) {
      best_cost = cost;
      best = g;
    }
  }
  return (int)std::max<size_t>(1, std::min(best, B));
}

void expand_matrix(const twmc_matrix& m, int n, std::vector<double>& base, std::vector<double>& amp,
                   int& noise) {
  base.assign(2 * (size_t)n, 0.0);
  amp.assign(2 * (size_t)n, 0.0);
  for (int i = 0; i < n; ++i) {
    base[2 * i] = m.base ? m.base[2 * i] : m.value[0];
    base[2 * i + 1] = m.base ? m.base[2 * i + 1] : m.value[1];
  }
  noise = m.noise_type;
  if (noise != TWMC_NOISE_NONE) {
    for (int i = 0; i < n; ++i) {
      amp[2 * i] = m.amp ? m.amp[2 * i] : m.amp_value[0];
      amp[2 * i + 1] = m.amp ? m.amp[2 * i + 1] : m.amp_value[1];
    }
  }
}

template <typename T>
std::vector<unsigned char> pack_complex(const std::vector<double>& v) {
  std::vector<unsigned char> out(v.size() * sizeof(T));
  T* o = reinterpret_cast<T*>(out.data());
  for (size_t i = 0; i < v.size(); ++i) o[i] = T(v[i]);
  return out;
}

// Host-to-device copy that has finished when it returns: the sources are
// temporaries of the caller (packed tables, the caller's trajectory ids and
// previous states) that may be released as soon as the call returns.
hipError_t upload(DevBuf& b, const void* src, size_t bytes, hipStream_t s) {
  hipError_t e = b.ensure(bytes);
  if (e != hipSuccess || bytes == 0) return e;
  e = hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e;
}

// The reference's time loop (TWMCBaseSolver.cpp:221-274) run on the host to
// fix the segment plan: frames at i_step % frame_steps == 0, loop while
// t <= t_end with t accumulated in FP64, break after the last frame.
void plan_segments(Ctx* c) {
  const twmc_batch_desc& d = c->desc;
  uint64_t fs = c->p.frame_steps;
  double dt_obs = c->p.dt * (double)fs;
  uint64_t nf = (uint64_t)(size_t)nearbyint((d.t_end - d.t_start) / dt_obs);
  if (d.store_initial) nf++;
  std::vector<uint64_t> cap;  // step index of each capture
  uint64_t i_frame = 0, i_step = 0;
  if (d.store_initial && i_frame < nf) {
    cap.push_back(0);
    i_frame++;
  }
  double t = d.t_start;
  while (t <= d.t_end) {
    t += c->p.dt;
    i_step++;
    if ((i_step % fs == 0) && i_frame < nf) {
      cap.push_back(i_step);
      i_frame++;
      if (i_frame == nf) break;
    }
  }
  c->truncated = i_frame < nf ? 1 : 0;
  c->n_frames = i_frame;
  c->n_steps = i_step;
  c->t_final = t;
  // one segment per stored frame: it starts at the capture step and runs to
  // the next one; steps before the first capture form a segment without one
  c->segs.clear();
  std::vector<std::pair<uint64_t, int>> starts;
  if (cap.empty() || cap[0] > 0) starts.push_back({0, -1});
  for (size_t k = 0; k < cap.size(); ++k) starts.push_back({cap[k], (int)k});
  for (size_t k = 0; k < starts.size(); ++k) {
    uint64_t end = (k + 1 < starts.size()) ? starts[k + 1].first : i_step;
    Segment sg{starts[k].second >= 0 ? 1 : 0, starts[k].second, end - starts[k].first};
    if (sg.capture || sg.steps > 0) c->segs.push_back(sg);
  }
}

// rsl, U and F the same at every site and in every trajectory, F static: the
// tables the init kernel forms are then uniform
bool site_uniform(const Ctx* c) {
  if (c->noisy_params || c->ft_n > 0) return false;
  for (int k = 0; k < 3; ++k)
    for (int i = 1; i < c->n; ++i)
      if (c->base[k][2 * i] != c->base[k][0] || c->base[k][2 * i + 1] != c->base[k][1]) return false;
  return true;
}

int choose_kernel(Ctx* c) {
  const twmc_params& p = c->p;
  // the register-resident kernel takes site-uniform, trajectory-shared parameters
  bool uniform = true;
  for (int k = 0; k < 3; ++k)
    for (int i = 1; i < c->n && uniform; ++i)
      uniform = c->base[k][2 * i] == c->base[k][0] && c->base[k][2 * i + 1] == c->base[k][1];
  bool fused_ok = fused64_supported(p.nx, p.ny, p.hopping, c->precision,
                                    (c->noisy_params || !uniform) ? 1 : 0) && c->ft_n == 0;
  size_t lds = generic_lds_bytes(p.nx, p.ny, c->precision);
  bool generic_ok = generic_pairs_per_thread(c->n, 256) > 0 && lds <= 160 * 1024 &&
                    (p.hopping == TWMC_HOP_MATRIX || (c->px.npass >= 0 && c->py.npass >= 0));
  bool mfma_ok = mfma_supported(c->n, p.hopping, c->precision, c->noisy_params ? 1 : 0);
  const bool band_ok = band_supported(c->n, p.hopping, c->precision, c->noisy_params ? 1 : 0, c->band_w);
  const bool stack_ok = p.hopping == TWMC_HOP_FFT && c->precision == 32 && !c->noisy_params && uniform &&
                        c->ft_n == 0 && stack_supported(p.nx, p.ny) && generic_ok;
  const bool stencil_ok = stencil64_supported(p.nx, p.ny, p.hopping, c->precision,
                                             (c->noisy_params || !uniform) ? 1 : 0) && c->ft_n == 0;
  int k = p.kernel;
  if (k == TWMC_KERNEL_STENCIL64 && !stencil_ok)
    return fail(c, TWMC_ERR_UNSUPPORTED,
                "stencil 64x64 kernel needs nx = ny = 64, FFT hopping, FP32 and site-uniform, disorder-free parameters");
  // AUTO: the 64x64 nearest-neighbour form first (the same linear map as the
  // reference's FFT hopping, about twice as fast as the register FFT), then
  // the register FFT; other uniform FP32 lattices of more than 64 sites on the
  // stacked LDS stencil (up to 64 sites the LDS kernel, which packs 2 or 4
  // trajectories per wave, is faster); for matrix hopping the banded stencil
  // when the coupling is banded (a dense product would mostly move zeros),
  // else the MFMA matrix product; the LDS kernel for everything else
  if (k == TWMC_KERNEL_AUTO)
    k = stencil_ok ? TWMC_KERNEL_STENCIL64
        : fused_ok ? TWMC_KERNEL_FUSED64
        : (stack_ok && c->n > 64) ? TWMC_KERNEL_STACK
        : band_ok  ? TWMC_KERNEL_BAND
        : mfma_ok  ? TWMC_KERNEL_MFMA
                   : TWMC_KERNEL_GENERIC;
  if (k == TWMC_KERNEL_STACK && !stack_ok)
    return fail(c, TWMC_ERR_UNSUPPORTED,
                "stacked kernel needs FFT hopping, FP32, site-uniform disorder-free parameters, static F, an even "
                "row length <= 512, at most 64 rows and 4096 sites");
  if (k == TWMC_KERNEL_BAND && !band_ok)
    return fail(c, TWMC_ERR_UNSUPPORTED,
                "banded kernel needs matrix hopping with nonzeros at |i - j| <= 3 (cyclic when N = 64 S), FP32, at "
                "most 512 sites and no per-trajectory disorder");
  if (k == TWMC_KERNEL_FUSED64 && !fused_ok)
    return fail(c, TWMC_ERR_UNSUPPORTED,
                "fused 64x64 kernel needs nx = ny = 64, FFT hopping, FP32 and site-uniform, disorder-free parameters");
  if (k == TWMC_KERNEL_GENERIC && !generic_ok)
    return fail(c, TWMC_ERR_UNSUPPORTED, "lattice too large for the LDS-resident kernel");
  if (k == TWMC_KERNEL_MFMA && !mfma_ok)
    return fail(c, TWMC_ERR_UNSUPPORTED,
                "MFMA kernel needs matrix hopping, FP32, at most 512 sites and no per-trajectory disorder");
  c->kernel = k;
  return TWMC_OK;
}

hipError_t timed(Ctx* c, std::vector<std::pair<hipEvent_t, hipEvent_t>>& evs, hipStream_t s, bool begin,
                 hipEvent_t& e) {
  if (!c->timing) return hipSuccess;
  hipError_t r = hipSuccess;
  if (begin) {
    r = hipEventCreate(&e);
    if (r == hipSuccess) r = hipEventRecord(e, s);
  } else {
    hipEvent_t e1 = nullptr;
    r = hipEventCreate(&e1);
    if (r == hipSuccess) r = hipEventRecord(e1, s);
    if (r == hipSuccess) evs.push_back({e, e1});
  }
  return r;
}

// stencil64 segment: the capture (if any) is the register-FFT capture kernel
// (per-trajectory means, dump, k-space sums per wave), the site-major kernel
// (per-site real-space sums per trajectory chunk) and the reduction of both
// partial sets, then the steps
int advance_stencil(Ctx* c, const StepArgs& a, hipStream_t s) {
  const int n = c->n;
  if (a.capture) {
    StepArgs ca = a;
    ca.n_steps = 0;
    hipEvent_t r0 = nullptr;
    HIPCHK(c, timed(c, c->ev_red, s, true, r0));
    HIPCHK(c, launch_capture64(ca, c->cap_grid, s));
    SiteSumArgs ss{};
    ss.beta = a.beta;
    ss.partials = static_cast<double*>(c->d_partials.p);
    ss.n = n;
    ss.n_traj = (int)c->n_traj;
    const int chunks = site_sum_chunks(c->n_traj, kSiteSumChunks, &ss.chunk);
    HIPCHK(c, launch_site_sums(ss, c->precision, s));
    ReduceArgs r{};
    r.n = n;
    r.slots = chunks;
    r.slots_k = c->cap_grid;
    r.keep = 1;
    r.partials_in = static_cast<const double*>(c->d_partials.p);
    r.partials = static_cast<double*>(c->d_partials.p);
    r.obs = static_cast<double*>(c->d_obs.p) + (size_t)a.frame * kQ * n;
    HIPCHK(c, launch_reduce(r, s));
    HIPCHK(c, timed(c, c->ev_red, s, false, r0));
  }
  if (a.n_steps > 0) {
    hipEvent_t e0 = nullptr;
    HIPCHK(c, timed(c, c->ev_step, s, true, e0));
    HIPCHK(c, launch_stencil64(a, c->grid, s));
    HIPCHK(c, timed(c, c->ev_step, s, false, e0));
    c->stats.site_steps += (uint64_t)a.n_steps * (uint64_t)n * c->n_traj;
  }
  return TWMC_OK;
}

// banded segment: at a capture the site-major sums read the frame state
// first, then the step launch records the per-trajectory means (and the
// dump) before stepping, then the reduction
int advance_band(Ctx* c, const StepArgs& a, hipStream_t s) {
  const int n = c->n;
  int chunks = 0;
  hipEvent_t r0 = nullptr, e0 = nullptr;
  if (a.capture) {
    HIPCHK(c, timed(c, c->ev_red, s, true, r0));
    SiteSumArgs ss{};
    ss.beta = a.beta;
    ss.partials = static_cast<double*>(c->d_partials.p);
    ss.n = n;
    ss.n_traj = (int)c->n_traj;
    ss.zero_k = 1;
    chunks = site_sum_chunks(c->n_traj, kSiteSumChunks, &ss.chunk);
    HIPCHK(c, launch_site_sums(ss, c->precision, s));
    HIPCHK(c, timed(c, c->ev_red, s, false, r0));
  }
  HIPCHK(c, timed(c, c->ev_step, s, true, e0));
  HIPCHK(c, launch_band(a, c->grid, s));
  HIPCHK(c, timed(c, c->ev_step, s, false, e0));
  c->stats.site_steps += (uint64_t)a.n_steps * (uint64_t)n * c->n_traj;
  if (a.capture) {
    ReduceArgs r{};
    r.n = n;
    r.slots = chunks;
    r.keep = 1;
    r.partials_in = static_cast<const double*>(c->d_partials.p);
    r.partials = static_cast<double*>(c->d_partials.p);
    r.obs = static_cast<double*>(c->d_obs.p) + (size_t)a.frame * kQ * n;
    HIPCHK(c, timed(c, c->ev_red, s, true, r0));
    HIPCHK(c, launch_reduce(r, s));
    HIPCHK(c, timed(c, c->ev_red, s, false, r0));
  }
  return TWMC_OK;
}

// stacked segment: a frame is the LDS kernel's capture-only launch and the
// reduction (the per-site and per-mode sums), then the steps on the sweep kernel
int advance_stack(Ctx* c, const StepArgs& a, hipStream_t s) {
  const int n = c->n;
  hipEvent_t r0 = nullptr, e0 = nullptr;
  if (a.capture) {
    StepArgs ca = a;
    ca.n_steps = 0;
    HIPCHK(c, timed(c, c->ev_red, s, true, r0));
    HIPCHK(c, launch_generic(ca, c->precision, c->grid, c->block, c->lds, s));
    ReduceArgs r{};
    r.n = n;
    r.slots = c->grid;
    r.partials_in = static_cast<const double*>(c->d_partials.p);
    r.partials = static_cast<double*>(c->d_partials.p);
    r.obs = static_cast<double*>(c->d_obs.p) + (size_t)a.frame * kQ * n;
    HIPCHK(c, launch_reduce(r, s));
    HIPCHK(c, timed(c, c->ev_red, s, false, r0));
  }
  if (a.n_steps > 0) {
    HIPCHK(c, timed(c, c->ev_step, s, true, e0));
    HIPCHK(c, stack_launch(c->stack, a, s));
    HIPCHK(c, timed(c, c->ev_step, s, false, e0));
    c->stats.site_steps += (uint64_t)a.n_steps * (uint64_t)n * c->n_traj;
  }
  return TWMC_OK;
}

// The address of this device's copy of `scratch`, looked up once per device
// (the device must be current).
hipError_t scratch_setup(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((size_t)device < g_scratch_addr.size() && g_scratch_addr[device]) return hipSuccess;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (g_scratch_addr.size() <= (size_t)device) g_scratch_addr.resize(device + 1, nullptr);
  g_scratch_addr[device] = ptr;
  return hipSuccess;
}

void* scratch_address(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return (size_t)device < g_scratch_addr.size() ? g_scratch_addr[device] : nullptr;
}

// Launch geometry of the step kernel (and of the 64x64 capture kernel) for a
// batch of B trajectories; allocates nothing.
struct Geometry {
  int grid = 0, block = 256, cap_grid = 0;
  size_t lds = 0;
  int grid_nn = 0;  // LDS kernel with FFT hopping: the step launch without the FFT's second buffer
  size_t lds_nn = 0;
};

int step_geometry(Ctx* c, size_t B, Geometry& geo) {
  int cus = 0;
  HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
  int per_cu = 1;
  size_t units = B;  // work items of the persistent grid
  if (c->kernel == TWMC_KERNEL_FUSED64) {
    geo.block = 64;
    geo.lds = 0;
    HIPCHK(c, fused64_occupancy(&per_cu));
  } else if (c->kernel == TWMC_KERNEL_STENCIL64) {
    geo.block = 64;
    geo.lds = 0;
    int cap_per_cu = 0;
    HIPCHK(c, capture64_occupancy(&cap_per_cu));
    if (cap_per_cu < 1) return fail(c, TWMC_ERR_UNSUPPORTED, "capture kernel does not fit on a CU");
    geo.cap_grid = (int)std::min<size_t>(B, (size_t)cap_per_cu * cus);
    HIPCHK(c, stencil64_occupancy(&per_cu));
  } else if (c->kernel == TWMC_KERNEL_BAND) {
    geo.block = 64;
    geo.lds = 0;
    HIPCHK(c, band_occupancy(c->n, c->band_w, c->band_lieb, &per_cu));
  } else if (c->kernel == TWMC_KERNEL_MFMA) {
    geo.block = mfma_block(c->np);
    geo.lds = mfma_lds_bytes(c->np);
    units = (B + 31) / 32;
    HIPCHK(c, mfma_occupancy(c->np, &per_cu));
  } else {
    geo.block = 256;
    geo.lds = generic_lds_bytes(c->p.nx, c->p.ny, c->precision);
    HIPCHK(c, generic_occupancy(c->precision, c->n, geo.block, geo.lds, &per_cu));
    if (c->p.hopping == TWMC_HOP_FFT) {
      int per_cu_nn = 0;
      geo.lds_nn = generic_nn_lds_bytes(c->p.nx, c->p.ny, c->precision, site_uniform(c));
      HIPCHK(c, generic_nn_occupancy(c->precision, c->p.nx, c->p.ny, geo.lds_nn, site_uniform(c), &per_cu_nn));
      if (per_cu_nn < 1) return fail(c, TWMC_ERR_UNSUPPORTED, "step kernel does not fit on a CU");
      const size_t tpb = (size_t)generic_nn_traj_per_block(c->n);
      geo.grid_nn = (int)std::min<size_t>((B + tpb - 1) / tpb, (size_t)per_cu_nn * cus);
    }
  }
  if (per_cu < 1) return fail(c, TWMC_ERR_UNSUPPORTED, "step kernel does not fit on a CU");
  if (c->kernel == TWMC_KERNEL_STENCIL64 || c->kernel == TWMC_KERNEL_FUSED64 || c->kernel == TWMC_KERNEL_BAND)
    geo.grid = balanced_wave_grid(B, per_cu, cus);
  else
    geo.grid = (int)std::min<size_t>(units, (size_t)per_cu * cus);
  return TWMC_OK;
}

// The step arguments every segment shares: lattice, parameters, plans and
// the context's device tables; the batch's buffers are set by the caller.
StepArgs shared_step_args(const Ctx* c) {
  StepArgs a{};
  a.nx = c->p.nx;
  a.ny = c->p.ny;
  a.n = c->n;
  a.hopping = c->p.hopping;
  a.param_stride = c->noisy_params ? c->n : 0;
  a.dt = c->p.dt;
  a.dt4 = c->dt4;
  a.inv_n = 1.0 / double(c->n);
  a.kscale = -2.0 * c->p.J;
  a.key0 = uint32_t(c->p.seed & 0xffffffffu);
  a.key1 = uint32_t(c->p.seed >> 32);
  a.px = c->px;
  a.py = c->py;
  a.twx = c->d_twx.p;
  a.twy = c->d_twy.p;
  a.kcx = c->d_kcx.p;
  a.kcy = c->d_kcy.p;
  a.csc_ptr = static_cast<const int*>(c->d_cscp.p);
  a.csc_row = static_cast<const int*>(c->d_cscr.p);
  a.csc_val = c->d_cscv.p;
  a.a_hi = c->d_ahi.p;
  a.a_lo = c->d_alo.p;
  a.np = c->np;
  a.ft_n = c->ft_n;
  a.ft_times = static_cast<const double*>(c->d_ft_times.p);
  a.ft_vals = c->d_ft_vals.p;
  a.ft_slopes = c->d_ft_slopes.p;
  a.site_uniform = site_uniform(c) ? 1 : 0;
  a.u_real = 1;
  for (int i = 0; i < c->n && a.u_real; ++i) a.u_real = c->base[1][2 * i + 1] == 0.0;
  if (c->noise[1] != TWMC_NOISE_NONE) a.u_real = 0;
  a.band_coef = static_cast<const float*>(c->d_band.p);
  a.band_w = c->band_w;
  a.band_lieb = c->band_lieb;
  return a;
}

// LDS-kernel segment with FFT hopping: the capture (if any) is the kernel's
// capture-only launch (the FFT for the k-space sums, two LDS copies of the
// state) and the reduction; the steps are the nearest-neighbour launch, one
// LDS copy, so more workgroups fit on a CU
int advance_generic_fft(Ctx* c, const StepArgs& a, hipStream_t s) {
  const int n = c->n;
  if (a.capture) {
    StepArgs ca = a;
    ca.n_steps = 0;
    hipEvent_t r0 = nullptr;
    HIPCHK(c, timed(c, c->ev_red, s, true, r0));
    HIPCHK(c, launch_generic(ca, c->precision, c->grid, c->block, c->lds, s));
    ReduceArgs r{};
    r.n = n;
    r.slots = c->grid;
    r.partials_in = static_cast<const double*>(c->d_partials.p);
    r.partials = static_cast<double*>(c->d_partials.p);
    r.obs = static_cast<double*>(c->d_obs.p) + (size_t)a.frame * kQ * n;
    HIPCHK(c, launch_reduce(r, s));
    HIPCHK(c, timed(c, c->ev_red, s, false, r0));
  }
  if (a.n_steps > 0) {
    StepArgs sa = a;
    sa.capture = 0;
    hipEvent_t e0 = nullptr;
    HIPCHK(c, timed(c, c->ev_step, s, true, e0));
    HIPCHK(c, launch_generic_nn(sa, c->precision, c->grid_nn, c->lds_nn, s));
    HIPCHK(c, timed(c, c->ev_step, s, false, e0));
    c->stats.site_steps += (uint64_t)a.n_steps * (uint64_t)n * c->n_traj;
  }
  return TWMC_OK;
}

int release_batch(Ctx* c) {
  c->batch_active = false;
  c->segs.clear();
  c->next_seg = 0;
  return TWMC_OK;
}

}  // namespace

extern "C" {

int twmc_abi_version(void) { return TWMC_ABI_VERSION; }

const char* twmc_last_error(const twmc_ctx* ctx) {
  if (ctx) return reinterpret_cast<const Ctx*>(ctx)->err.c_str();
  return g_last_error.c_str();
}

int twmc_create(const twmc_params* p, int device, twmc_ctx** ctx_out) {
  if (!p || !ctx_out) return fail(nullptr, TWMC_ERR_INVALID, "null argument");
  *ctx_out = nullptr;
  if (p->nx < 1 || p->ny < 1) return fail(nullptr, TWMC_ERR_INVALID, "nx and ny must be >= 1");
  if (p->precision != 32 && p->precision != 64)
    return fail(nullptr, TWMC_ERR_INVALID, "precision must be 32 or 64");
  if (p->frame_steps < 1) return fail(nullptr, TWMC_ERR_INVALID, "frame_steps must be >= 1");
  if (!(p->dt > 0.0)) return fail(nullptr, TWMC_ERR_INVALID, "dt must be > 0");
  if (p->hopping != TWMC_HOP_FFT && p->hopping != TWMC_HOP_MATRIX)
    return fail(nullptr, TWMC_ERR_INVALID, "unknown hopping mode");
  if (p->hopping == TWMC_HOP_MATRIX && !p->coupling)
    return fail(nullptr, TWMC_ERR_INVALID, "matrix hopping needs a coupling matrix");
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) return fail(nullptr, TWMC_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(nullptr, TWMC_ERR_INVALID, "device index out of range");

  Ctx* c = new Ctx();
  c->device = device;
  c->p = *p;
  c->n = p->nx * p->ny;
  c->precision = p->precision;
  c->es = p->precision == 64 ? 16 : 8;
  c->dt4 = std::sqrt(p->gamma * p->dt / 4.0);  // TWMCBaseSolver.cpp:210
  const twmc_matrix* mats[5] = {&p->omega, &p->U, &p->F, &p->beta_init, &p->beta_sigma};
  for (int k = 0; k < 5; ++k) {
    if (mats[k]->noise_type < 0 || mats[k]->noise_type > 2) {
      delete c;
      return fail(nullptr, TWMC_ERR_INVALID, "unknown noise type");
    }
    expand_matrix(*mats[k], c->n, c->base[k], c->amp[k], c->noise[k]);
  }
  if (p->F_n_times > 0) {
    if (!p->F_times || !p->F_table) {
      delete c;
      return fail(nullptr, TWMC_ERR_INVALID, "F_n_times > 0 needs F_times and F_table");
    }
    for (uint64_t k = 0; k < p->F_n_times; ++k)
      if (!std::isfinite(p->F_times[k]) || (k > 0 && !(p->F_times[k] > p->F_times[k - 1]))) {
        delete c;
        return fail(nullptr, TWMC_ERR_INVALID, "F_times must be finite and strictly ascending");
      }
    for (uint64_t e = 0; e < 2 * p->F_n_times * (uint64_t)c->n; ++e)
      if (!std::isfinite(p->F_table[e])) {
        delete c;
        return fail(nullptr, TWMC_ERR_INVALID, "F_table has non-finite entries (fill the gaps first)");
      }
    c->ft_n = (int)p->F_n_times;
  }
  c->p.F_times = c->p.F_table = nullptr;
  c->p.omega.base = c->p.U.base = c->p.F.base = c->p.beta_init.base = c->p.beta_sigma.base = nullptr;
  c->p.omega.amp = c->p.U.amp = c->p.F.amp = c->p.beta_init.amp = c->p.beta_sigma.amp = nullptr;
  c->noisy_params = (c->noise[0] || c->noise[1] || c->noise[2]) ? 1 : 0;

  if (p->hopping == TWMC_HOP_FFT) {
    if (!make_radix_plan(p->nx, c->px) || !make_radix_plan(p->ny, c->py)) {
      delete c;
      return fail(nullptr, TWMC_ERR_UNSUPPORTED, "FFT length needs more radix passes than supported");
    }
  } else {
    c->px.len = c->py.len = 0;
    c->px.npass = c->py.npass = 0;
  }
  if (p->hopping == TWMC_HOP_MATRIX) c->band_w = band_width(p->coupling, c->n);
  int rc = choose_kernel(c);
  if (rc != TWMC_OK) {
    std::string m = c->err;
    delete c;
    return fail(nullptr, rc, m);
  }

  e = hipSetDevice(device);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(nullptr, e, "hipSetDevice");
  }
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(nullptr, e, "hipStreamCreate");
  }
  // this device's copy of the allocation-free entry point's scratch array
  e = scratch_setup(device);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return hip_fail(nullptr, e, "hipGetSymbolAddress(scratch)");
  }

  // constant tables, computed as the reference computes them
  // (K: TWMCBaseSolver.cpp:78-88; twiddles: exp(-2 pi i m / L) in long double)
  auto tw_table = [](int L) {
    std::vector<double> t(2 * (size_t)std::max(L, 1));
    for (int m = 0; m < L; ++m) {
      long double a = -2.0L * 3.14159265358979323846264338327950288L * m / L;
      t[2 * m] = (double)cosl(a);
      t[2 * m + 1] = (double)sinl(a);
    }
    return t;
  };
  double fx = (p->nx == 1) ? 0.0 : 1.0, fy = (p->ny == 1) ? 0.0 : 1.0;
  std::vector<double> kcx(p->nx), kcy(p->ny);
  for (int i = 0; i < p->nx; ++i) kcx[i] = fx * cos(2.0 * M_PI / double(p->nx) * double(i));
  for (int j = 0; j < p->ny; ++j) kcy[j] = fy * cos(2.0 * M_PI / double(p->ny) * double(j));
  std::vector<double> twx = tw_table(p->nx), twy = tw_table(p->ny);

  hipStream_t s = c->stream;
  int fail_rc = TWMC_OK;
  auto up = [&](DevBuf& b, const std::vector<double>& v, bool is_complex_or_real_T) {
    (void)is_complex_or_real_T;
    std::vector<unsigned char> pk = c->precision == 64 ? pack_complex<double>(v) : pack_complex<float>(v);
    hipError_t ee = upload(b, pk.data(), pk.size(), s);
    if (ee != hipSuccess) fail_rc = hip_fail(c, ee, "upload constant table");
    return ee;
  };
  (void)up(c->d_twx, twx, true);  // failures are collected in fail_rc
  (void)up(c->d_twy, twy, true);
  (void)up(c->d_kcx, kcx, false);
  (void)up(c->d_kcy, kcy, false);
  for (int k = 0; k < 5 && fail_rc == TWMC_OK; ++k) {
    hipError_t ee = upload(c->d_base[k], c->base[k].data(), c->base[k].size() * 8, s);
    if (ee == hipSuccess) ee = upload(c->d_amp[k], c->amp[k].data(), c->amp[k].size() * 8, s);
    if (ee != hipSuccess) fail_rc = hip_fail(c, ee, "upload parameters");
  }
  if (fail_rc == TWMC_OK && p->hopping == TWMC_HOP_MATRIX) {
    // column-compressed nonzeros of C: h_j = sum_i beta_i C_ij
    std::vector<int> ptr(c->n + 1, 0), row;
    std::vector<double> val;
    for (int j = 0; j < c->n; ++j) {
      for (int i = 0; i < c->n; ++i) {
        double v = p->coupling[(size_t)i * c->n + j];
        if (v != 0.0) {
          row.push_back(i);
          val.push_back(v);
        }
      }
      ptr[j + 1] = (int)row.size();
    }
    c->csc_nnz = (int)row.size();
    if (row.empty()) {
      row.push_back(0);
      val.push_back(0.0);
    }
    hipError_t ee = upload(c->d_cscp, ptr.data(), ptr.size() * sizeof(int), s);
    if (ee == hipSuccess) ee = upload(c->d_cscr, row.data(), row.size() * sizeof(int), s);
    if (ee == hipSuccess) (void)up(c->d_cscv, val, false);
    if (ee != hipSuccess) fail_rc = hip_fail(c, ee, "upload coupling");
  }
  if (fail_rc == TWMC_OK && c->ft_n > 0) {
    // F(t) values and per-interval slopes (NoisyMatrix.cpp:228-239); the
    // slope after the last time is zero, so its value is held
    const size_t nt = (size_t)c->ft_n, n2 = 2 * (size_t)c->n;
    std::vector<double> vals(p->F_table, p->F_table + nt * n2), slopes(nt * n2, 0.0);
    for (size_t k = 0; k + 1 < nt; ++k)
      for (size_t e = 0; e < n2; ++e)
        slopes[k * n2 + e] = (vals[(k + 1) * n2 + e] - vals[k * n2 + e]) / (p->F_times[k + 1] - p->F_times[k]);
    hipError_t ee = upload(c->d_ft_times, p->F_times, nt * sizeof(double), s);
    if (ee == hipSuccess) ee = up(c->d_ft_vals, vals, true);
    if (ee == hipSuccess) ee = up(c->d_ft_slopes, slopes, true);
    if (ee == hipSuccess) ee = hipStreamSynchronize(s);
    if (ee != hipSuccess && fail_rc == TWMC_OK) fail_rc = hip_fail(c, ee, "upload F(t) table");
  }
  if (fail_rc == TWMC_OK && c->kernel == TWMC_KERNEL_MFMA) {
    // A = C^T, zero-padded to np x np, split into BF16 high and low parts
    c->np = mfma_padded(c->n);
    const int np = c->np;
    const int rows = mfma_rows(np);
    std::vector<double> a((size_t)rows * np, 0.0);
    for (int i = 0; i < c->n; ++i)
      for (int j = 0; j < c->n; ++j) a[(size_t)j * np + i] = p->coupling[(size_t)i * c->n + j];
    std::vector<double> as(a.size());
    mfma_swizzle(a.data(), rows, np, as.data());
    std::vector<uint16_t> hi((size_t)rows * np), lo((size_t)rows * np);
    mfma_split(as.data(), as.size(), hi.data(), lo.data());
    hipError_t ee = upload(c->d_ahi, hi.data(), hi.size() * 2, s);
    if (ee == hipSuccess) ee = upload(c->d_alo, lo.data(), lo.size() * 2, s);
    if (ee == hipSuccess) ee = hipStreamSynchronize(s);
    if (ee != hipSuccess) fail_rc = hip_fail(c, ee, "upload MFMA operand");
  }
  if (fail_rc == TWMC_OK && c->kernel == TWMC_KERNEL_BAND) {
    // dt C_(j+d) j per site j and offset d, padded to 64 S sites
    c->band_w = std::max(1, c->band_w);
    const int S = band_sites_per_lane(c->n);
    std::vector<float> coef((size_t)64 * S * (2 * c->band_w + 1));
    band_coefficients(p->coupling, c->n, c->band_w, p->dt, coef.data());
    c->band_lieb = band_lieb_pattern(coef.data(), c->n, c->band_w) ? 1 : 0;
    hipError_t ee = upload(c->d_band, coef.data(), coef.size() * sizeof(float), s);
    if (ee == hipSuccess) ee = hipStreamSynchronize(s);
    if (ee != hipSuccess) fail_rc = hip_fail(c, ee, "upload band coefficients");
  }
  if (fail_rc == TWMC_OK && c->kernel == TWMC_KERNEL_FUSED64) {
    std::vector<int> ps(4096), pk(4096);
    fused64_perms(ps.data(), pk.data());
    hipError_t ee = upload(c->d_perm_site, ps.data(), ps.size() * sizeof(int), s);
    if (ee == hipSuccess) ee = upload(c->d_perm_k, pk.data(), pk.size() * sizeof(int), s);
    if (ee != hipSuccess) fail_rc = hip_fail(c, ee, "upload permutations");
  }
  if (fail_rc == TWMC_OK) {
    hipError_t ee = hipStreamSynchronize(s);
    if (ee != hipSuccess) fail_rc = hip_fail(c, ee, "hipStreamSynchronize");
  }
  if (fail_rc != TWMC_OK) {
    std::string m = c->err;
    twmc_destroy(reinterpret_cast<twmc_ctx*>(c));
    return fail(nullptr, fail_rc, m);
  }
  *ctx_out = reinterpret_cast<twmc_ctx*>(c);
  return TWMC_OK;
}

void twmc_destroy(twmc_ctx* ctx) {
  if (!ctx) return;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  stack_release(c->stack);
  for (auto& pr : c->ev_step) {
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  for (auto& pr : c->ev_red) {
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  for (hipEvent_t e : c->seg_ev) (void)hipEventDestroy(e);
  DevBuf* bufs[] = {&c->d_twx,       &c->d_twy,      &c->d_kcx,       &c->d_kcy,  &c->d_cscp, &c->d_cscr,
                    &c->d_cscv,      &c->d_perm_site, &c->d_perm_k,   &c->d_ahi,  &c->d_alo,  &c->d_kamp,
                    &c->d_ft_times,  &c->d_ft_vals,  &c->d_ft_slopes, &c->d_band, &c->d_ids,  &c->d_beta,
                    &c->d_rsl,       &c->d_U,        &c->d_F,         &c->d_partials, &c->d_obs, &c->d_m2,
                    &c->d_m4,        &c->d_traj,     &c->d_prev};
  for (DevBuf* b : bufs) b->release();
  for (int k = 0; k < 5; ++k) {
    c->d_base[k].release();
    c->d_amp[k].release();
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int twmc_batch_begin(twmc_ctx* ctx, const twmc_batch_desc* d) {
  if (!ctx || !d) return fail(nullptr, TWMC_ERR_INVALID, "null argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (d->n_traj == 0 || !d->traj_ids) return fail(c, TWMC_ERR_INVALID, "empty trajectory batch");
  if (d->n_traj > (uint64_t)INT32_MAX) return fail(c, TWMC_ERR_INVALID, "batch too large");
  if (d->init_mode < 0 || d->init_mode > 3) return fail(c, TWMC_ERR_INVALID, "unknown init mode");
  std::vector<double> kamp;
  if (d->init_mode == TWMC_INIT_GAUSSIAN) {
    if (c->p.hopping != TWMC_HOP_FFT)
      return fail(c, TWMC_ERR_UNSUPPORTED, "TWMC_INIT_GAUSSIAN needs FFT (uniform) hopping");
    if (!d->gauss_mean || !d->gauss_nk) return fail(c, TWMC_ERR_INVALID, "TWMC_INIT_GAUSSIAN needs gauss_mean and gauss_nk");
    if (!std::isfinite(d->gauss_mean[0]) || !std::isfinite(d->gauss_mean[1]))
      return fail(c, TWMC_ERR_INVALID, "gauss_mean is not finite");
    kamp.resize(c->n);
    for (int k = 0; k < c->n; ++k) {
      const double v = d->gauss_nk[k] + 0.5;
      if (!std::isfinite(v) || v < 0.0) return fail(c, TWMC_ERR_INVALID, "gauss_nk must be finite and >= -1/2");
      kamp[k] = std::sqrt(double(c->n) * v / 2.0);
    }
  }
  if (d->init_mode == TWMC_INIT_PREVIOUS && !d->beta_prev)
    return fail(c, TWMC_ERR_INVALID, "TWMC_INIT_PREVIOUS needs beta_prev");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->desc = *d;
  c->desc.traj_ids = nullptr;
  c->desc.beta_prev = nullptr;
  c->desc.gauss_mean = c->desc.gauss_nk = nullptr;
  c->n_traj = d->n_traj;
  plan_segments(c);
  c->next_seg = 0;
  c->steps_done = 0;
  {
    std::lock_guard<std::mutex> lk(c->prog_mu);
    while (c->seg_ev.size() < c->segs.size()) {
      hipEvent_t e = nullptr;
      HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
      c->seg_ev.push_back(e);
    }
    c->seg_steps.assign(c->segs.size(), 0);
    c->seg_recorded = 0;
    c->seg_completed = 0;
  }

  const int n = c->n;
  const size_t B = d->n_traj;
  hipStream_t s = c->stream;
  // launch geometry
  int cus = 0;
  HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
  {
    Geometry geo;
    int rc = step_geometry(c, B, geo);
    if (rc != TWMC_OK) return rc;
    c->grid = geo.grid;
    c->block = geo.block;
    c->lds = geo.lds;
    c->cap_grid = geo.cap_grid;
    c->grid_nn = geo.grid_nn;
    c->lds_nn = geo.lds_nn;
  }
  if (c->kernel == TWMC_KERNEL_STACK) {
    // steps on the sweep kernel; frames stay with the LDS kernel (geometry above).
    // The site-uniform constants, formed as the init kernel forms them
    const double orr = c->base[0][0], oi = c->base[0][1];
    double rsl[2];
    if (c->p.onsite == TWMC_ONSITE_OMEGA) {
      rsl[0] = oi - c->p.gamma / 2.0;  // rsl = -i omega - gamma/2 (TWMCBaseSolver.cpp:74)
      rsl[1] = -orr;
    } else {
      const double dr = -orr + c->p.detuning, di = -oi;  // TWMCLiebSolver.cpp:44-45
      rsl[0] = di - c->p.gamma / 2.0;
      rsl[1] = -dr;
    }
    HIPCHK(c, stack_prepare(c->stack, c->p.nx, c->p.ny, (int)B, rsl, c->base[1].data(), c->base[2].data(), c->dt4,
                            -2.0 * c->p.J));
  }

  size_t sets = c->noisy_params ? B : 1;
  HIPCHK(c, upload(c->d_ids, d->traj_ids, B * sizeof(uint64_t), s));
  HIPCHK(c, c->d_beta.ensure(B * n * c->es));
  HIPCHK(c, c->d_rsl.ensure(sets * n * c->es));
  HIPCHK(c, c->d_U.ensure(sets * n * c->es));
  HIPCHK(c, c->d_F.ensure(sets * n * c->es));
  // slot rows: one per step-kernel workgroup, per capture workgroup, and per
  // trajectory chunk of site_sums_kernel (stencil and banded paths)
  int ss_chunk = 0;
  const int slots = std::max({c->grid, c->cap_grid, site_sum_chunks(B, kSiteSumChunks, &ss_chunk)});
  size_t part_bytes = (size_t)slots * kQ * n * sizeof(double);
  HIPCHK(c, c->d_partials.ensure(part_bytes));
  HIPCHK(c, hipMemsetAsync(c->d_partials.p, 0, part_bytes, s));
  size_t nf = std::max<uint64_t>(c->n_frames, 1);
  HIPCHK(c, c->d_obs.ensure(nf * kQ * n * sizeof(double)));
  HIPCHK(c, hipMemsetAsync(c->d_obs.p, 0, nf * kQ * n * sizeof(double), s));
  HIPCHK(c, c->d_m2.ensure(B * nf * sizeof(double)));
  HIPCHK(c, c->d_m4.ensure(B * nf * sizeof(double)));
  if (d->outputs & TWMC_OUT_TRAJ) HIPCHK(c, c->d_traj.ensure(B * nf * n * c->es));
  if (d->init_mode == TWMC_INIT_PREVIOUS)
    HIPCHK(c, upload(c->d_prev, d->beta_prev, B * n * 2 * sizeof(double), s));
  if (d->init_mode == TWMC_INIT_GAUSSIAN) HIPCHK(c, upload(c->d_kamp, kamp.data(), kamp.size() * 8, s));

  InitArgs ia{};
  ia.n = n;
  ia.n_traj = (int)B;
  ia.mode = d->init_mode;
  ia.per_traj_params = c->noisy_params;
  ia.onsite = c->p.onsite;
  ia.gamma = c->p.gamma;
  ia.detuning = c->p.detuning;
  ia.key0 = uint32_t(c->p.seed & 0xffffffffu);
  ia.key1 = uint32_t(c->p.seed >> 32);
  ia.traj_ids = static_cast<const uint64_t*>(c->d_ids.p);
  for (int k = 0; k < 5; ++k) {
    ia.base[k] = static_cast<const double*>(c->d_base[k].p);
    ia.amp[k] = static_cast<const double*>(c->d_amp[k].p);
    ia.noise[k] = c->noise[k];
  }
  ia.beta_prev = static_cast<const double*>(c->d_prev.p);
  ia.beta = c->d_beta.p;
  ia.rsl = c->d_rsl.p;
  ia.U = c->d_U.p;
  ia.F = c->d_F.p;
  HIPCHK(c, launch_init(ia, c->precision, s));
  if (d->init_mode == TWMC_INIT_GAUSSIAN) {
    if (gauss_lds_bytes(c->p.nx, c->p.ny, c->precision) > 160 * 1024)
      return fail(c, TWMC_ERR_UNSUPPORTED, "lattice too large for the Gaussian initial state");
    GaussArgs ga{};
    ga.n = n;
    ga.nx = c->p.nx;
    ga.ny = c->p.ny;
    ga.n_traj = (int)B;
    ga.px = c->px;
    ga.py = c->py;
    ga.twx = c->d_twx.p;
    ga.twy = c->d_twy.p;
    ga.kamp = static_cast<const double*>(c->d_kamp.p);
    ga.mean_re = d->gauss_mean[0];
    ga.mean_im = d->gauss_mean[1];
    ga.key0 = ia.key0;
    ga.key1 = ia.key1;
    ga.traj_ids = ia.traj_ids;
    ga.beta = c->d_beta.p;
    HIPCHK(c, launch_gauss_init(ga, c->precision, (int)std::min<size_t>(B, (size_t)cus * 4), s));
  }
  c->batch_active = true;
  return TWMC_OK;
}

int twmc_batch_info_get(twmc_ctx* ctx, twmc_batch_info* info) {
  if (!ctx || !info) return fail(nullptr, TWMC_ERR_INVALID, "null argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c->batch_active) return fail(c, TWMC_ERR_STATE, "no active batch");
  info->n_frames = c->n_frames;
  info->n_steps = c->n_steps;
  info->n_segments = c->segs.size();
  info->segments_done = c->next_seg;
  info->t_final = c->t_final;
  info->kernel = c->kernel;
  info->truncated = c->truncated;
  return TWMC_OK;
}

int twmc_batch_advance(twmc_ctx* ctx, uint64_t n_segments) {
  if (!ctx) return fail(nullptr, TWMC_ERR_INVALID, "null argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c->batch_active) return fail(c, TWMC_ERR_STATE, "no active batch");
  HIPCHK(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int n = c->n;
  for (uint64_t k = 0; k < n_segments && c->next_seg < c->segs.size(); ++k) {
    const Segment& sg = c->segs[c->next_seg];
    // a segment longer than 2^31 steps is split; captures stay on the first piece
    uint64_t left = sg.steps;
    bool first = true;
    do {
      int steps = (int)std::min<uint64_t>(left, 1u << 30);
      StepArgs a = shared_step_args(c);
      a.n_traj = (int)c->n_traj;
      a.step0 = c->desc.step0 + c->steps_done;
      a.n_steps = steps;
      a.capture = (first && sg.capture) ? 1 : 0;
      a.frame = sg.frame < 0 ? 0 : sg.frame;
      a.n_frames = (int)std::max<uint64_t>(c->n_frames, 1);
      a.dump = (c->desc.outputs & TWMC_OUT_TRAJ) ? 1 : 0;
      a.traj_ids = static_cast<const uint64_t*>(c->d_ids.p);
      a.beta = c->d_beta.p;
      a.rsl = c->d_rsl.p;
      a.U = c->d_U.p;
      a.F = c->d_F.p;
      a.partials = static_cast<double*>(c->d_partials.p);
      a.traj_m2 = static_cast<double*>(c->d_m2.p);
      a.traj_m4 = static_cast<double*>(c->d_m4.p);
      a.traj_out = c->d_traj.p;
      a.t0 = c->desc.t_start + double(c->steps_done) * c->p.dt;
      if (c->kernel == TWMC_KERNEL_STENCIL64) {
        int rc = advance_stencil(c, a, s);
        if (rc != TWMC_OK) return rc;
      } else if (c->kernel == TWMC_KERNEL_BAND) {
        int rc = advance_band(c, a, s);
        if (rc != TWMC_OK) return rc;
      } else if (c->kernel == TWMC_KERNEL_STACK) {
        int rc = advance_stack(c, a, s);
        if (rc != TWMC_OK) return rc;
      } else if (c->kernel == TWMC_KERNEL_GENERIC && c->p.hopping == TWMC_HOP_FFT) {
        int rc = advance_generic_fft(c, a, s);
        if (rc != TWMC_OK) return rc;
      } else if (a.n_steps > 0 || a.capture) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (c->timing) {
          HIPCHK(c, hipEventCreate(&e0));
          HIPCHK(c, hipEventCreate(&e1));
          HIPCHK(c, hipEventRecord(e0, s));
        }
        if (c->kernel == TWMC_KERNEL_FUSED64) {
          HIPCHK(c, launch_fused64(a, c->grid, s));
        } else if (c->kernel == TWMC_KERNEL_MFMA) {
          HIPCHK(c, launch_mfma(a, c->grid, s));
        } else {
          HIPCHK(c, launch_generic(a, c->precision, c->grid, c->block, c->lds, s));
        }
        if (c->timing) {
          HIPCHK(c, hipEventRecord(e1, s));
          c->ev_step.push_back({e0, e1});
        }
        c->stats.site_steps += (uint64_t)a.n_steps * (uint64_t)n * c->n_traj;
        if (a.capture) {
          ReduceArgs r{};
          r.n = n;
          r.slots = c->grid;
          r.partials_in = static_cast<const double*>(c->d_partials.p);
          r.partials = static_cast<double*>(c->d_partials.p);
          r.obs = static_cast<double*>(c->d_obs.p) + (size_t)a.frame * kQ * n;
          r.perm_site = c->kernel == TWMC_KERNEL_FUSED64 ? static_cast<const int*>(c->d_perm_site.p) : nullptr;
          r.perm_k = c->kernel == TWMC_KERNEL_FUSED64 ? static_cast<const int*>(c->d_perm_k.p) : nullptr;
          hipEvent_t r0 = nullptr, r1 = nullptr;
          if (c->timing) {
            HIPCHK(c, hipEventCreate(&r0));
            HIPCHK(c, hipEventCreate(&r1));
            HIPCHK(c, hipEventRecord(r0, s));
          }
          HIPCHK(c, launch_reduce(r, s));
          if (c->timing) {
            HIPCHK(c, hipEventRecord(r1, s));
            c->ev_red.push_back({r0, r1});
          }
        }
      }
      c->steps_done += (uint64_t)steps;
      left -= (uint64_t)steps;
      first = false;
    } while (left > 0);
    {
      std::lock_guard<std::mutex> lk(c->prog_mu);
      HIPCHK(c, hipEventRecord(c->seg_ev[c->next_seg], s));
      c->seg_steps[c->next_seg] = c->steps_done;
      c->seg_recorded = c->next_seg + 1;
    }
    c->next_seg++;
  }
  return TWMC_OK;
}

int twmc_batch_sync(twmc_ctx* ctx) {
  if (!ctx) return fail(nullptr, TWMC_ERR_INVALID, "null argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TWMC_OK;
}

int twmc_batch_read(twmc_ctx* ctx, twmc_output* out) {
  if (!ctx || !out) return fail(nullptr, TWMC_ERR_INVALID, "null argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c->batch_active) return fail(c, TWMC_ERR_STATE, "no active batch");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t n = c->n, B = c->n_traj, nf = c->n_frames;
  if (nf > 0 && (out->sum_n || out->sum_n2 || out->sum_nk || out->sum_beta)) {
    std::vector<double> obs(nf * kQ * n);
    HIPCHK(c, hipMemcpy(obs.data(), c->d_obs.p, obs.size() * 8, hipMemcpyDeviceToHost));
    for (size_t f = 0; f < nf; ++f) {
      const double* o = obs.data() + f * kQ * n;
      if (out->sum_n) std::memcpy(out->sum_n + f * n, o, n * 8);
      if (out->sum_n2) std::memcpy(out->sum_n2 + f * n, o + n, n * 8);
      if (out->sum_nk) std::memcpy(out->sum_nk + f * n, o + 2 * n, n * 8);
      if (out->sum_beta)
        for (size_t i = 0; i < n; ++i) {
          out->sum_beta[2 * (f * n + i)] = o[3 * n + i];
          out->sum_beta[2 * (f * n + i) + 1] = o[4 * n + i];
        }
    }
  }
  if (nf > 0 && out->traj_m2) HIPCHK(c, hipMemcpy(out->traj_m2, c->d_m2.p, B * nf * 8, hipMemcpyDeviceToHost));
  if (nf > 0 && out->traj_m4) HIPCHK(c, hipMemcpy(out->traj_m4, c->d_m4.p, B * nf * 8, hipMemcpyDeviceToHost));
  auto to_f64 = [&](const DevBuf& b, size_t count, double* dst) -> int {
    if (c->precision == 64) {
      HIPCHK(c, hipMemcpy(dst, b.p, count * 16, hipMemcpyDeviceToHost));
    } else {
      std::vector<float> tmp(count * 2);
      HIPCHK(c, hipMemcpy(tmp.data(), b.p, count * 8, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < count * 2; ++i) dst[i] = tmp[i];
    }
    return TWMC_OK;
  };
  if (out->traj) {
    if (!(c->desc.outputs & TWMC_OUT_TRAJ))
      return fail(c, TWMC_ERR_STATE, "trajectory frames were not requested (TWMC_OUT_TRAJ)");
    if (nf > 0) {
      int rc = to_f64(c->d_traj, B * nf * n, out->traj);
      if (rc) return rc;
    }
  }
  if (out->beta_final) {
    int rc = to_f64(c->d_beta, B * n, out->beta_final);
    if (rc) return rc;
  }
  return TWMC_OK;
}

int twmc_run(twmc_ctx* ctx, const twmc_batch_desc* d, twmc_output* out) {
  int rc = twmc_batch_begin(ctx, d);
  if (rc) return rc;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  rc = twmc_batch_advance(ctx, c->segs.size());
  if (rc) return rc;
  if (out) rc = twmc_batch_read(ctx, out);
  else rc = twmc_batch_sync(ctx);
  return rc;
}

int twmc_set_timing(twmc_ctx* ctx, int enable) {
  if (!ctx) return fail(nullptr, TWMC_ERR_INVALID, "null argument");
  reinterpret_cast<Ctx*>(ctx)->timing = enable != 0;
  return TWMC_OK;
}

int twmc_query(twmc_ctx* ctx, twmc_stats* st) {
  if (!ctx || !st) return fail(nullptr, TWMC_ERR_INVALID, "null argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto& pr : c->ev_step) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, pr.first, pr.second));
    c->stats.step_kernel_ms += ms;
    c->stats.step_launches++;
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  c->ev_step.clear();
  for (auto& pr : c->ev_red) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, pr.first, pr.second));
    c->stats.reduce_kernel_ms += ms;
    c->stats.reduce_launches++;
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  c->ev_red.clear();
  c->stats.kernel = c->kernel;
  c->stats.grid = c->grid;
  c->stats.block = c->block;
  c->stats.lds_bytes = (int)c->lds;
  *st = c->stats;
  return TWMC_OK;
}

int twmc_reset_stats(twmc_ctx* ctx) {
  if (!ctx) return fail(nullptr, TWMC_ERR_INVALID, "null argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  twmc_stats z{};
  twmc_query(ctx, &z);
  c->stats = twmc_stats{};
  return TWMC_OK;
}

float twmc_progress(const twmc_ctx* ctx) {
  if (!ctx) return 0.f;
  const Ctx* c = reinterpret_cast<const Ctx*>(ctx);
  std::lock_guard<std::mutex> lk(c->prog_mu);
  while (c->seg_completed < c->seg_recorded && hipEventQuery(c->seg_ev[c->seg_completed]) == hipSuccess)
    c->seg_completed++;
  if (c->seg_completed == 0 || c->n_steps == 0) return 0.f;
  return float(double(c->seg_steps[c->seg_completed - 1]) / double(c->n_steps));
}

int twmc_advance_noalloc(twmc_ctx* ctx, const uint64_t* d_traj_ids, uint64_t n_traj, uint64_t step0,
                         uint64_t n_steps, void* d_beta, double* d_sums, void* stream) {
  if (!ctx || !d_traj_ids || !d_beta) return fail(nullptr, TWMC_ERR_INVALID, "null argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (n_traj == 0) return fail(c, TWMC_ERR_INVALID, "empty trajectory batch");
  if (n_traj > (uint64_t)INT32_MAX) return fail(c, TWMC_ERR_INVALID, "batch too large");
  if (c->noisy_params)
    return fail(c, TWMC_ERR_UNSUPPORTED, "twmc_advance_noalloc needs site-shared parameters (no per-trajectory disorder)");
  if (c->ft_n > 0) return fail(c, TWMC_ERR_UNSUPPORTED, "twmc_advance_noalloc needs a static drive");
  if (c->kernel == TWMC_KERNEL_STACK)
    return fail(c, TWMC_ERR_UNSUPPORTED,
                "twmc_advance_noalloc does not run the stacked kernel: create the context with kernel = GENERIC");
  HIPCHK(c, hipSetDevice(c->device));
  unsigned char* base = static_cast<unsigned char*>(scratch_address(c->device));
  if (!base) return fail(c, TWMC_ERR_STATE, "no scratch address for this device");
  const int n = c->n;
  // scratch layout: rsl, U, F tables (n complex each), then [chunks][kQ][n] doubles
  const size_t table = ((size_t)n * c->es + 255) / 256 * 256;
  const size_t params = 3 * table;
  const size_t per_chunk = (size_t)kQ * n * sizeof(double);
  if (params + per_chunk > (size_t)TWMC_NOALLOC_SCRATCH_BYTES)
    return fail(c, TWMC_ERR_INVALID, "lattice too large for the allocation-free scratch array");
  const int max_chunks =
      (int)std::min<size_t>(kSiteSumChunks, ((size_t)TWMC_NOALLOC_SCRATCH_BYTES - params) / per_chunk);
  Geometry geo;
  int rc = step_geometry(c, n_traj, geo);
  if (rc != TWMC_OK) return rc;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
  void* rsl = base;
  void* U = base + table;
  void* F = base + 2 * table;
  double* partials = reinterpret_cast<double*>(base + params);

  // the site-shared parameter tables, as twmc_batch_begin forms them (the
  // init kernel with one parameter set, writing no state)
  InitArgs ia{};
  ia.n = n;
  ia.n_traj = 1;
  ia.mode = TWMC_INIT_GAUSSIAN;
  ia.per_traj_params = 0;
  ia.onsite = c->p.onsite;
  ia.gamma = c->p.gamma;
  ia.detuning = c->p.detuning;
  ia.key0 = uint32_t(c->p.seed & 0xffffffffu);
  ia.key1 = uint32_t(c->p.seed >> 32);
  ia.traj_ids = d_traj_ids;
  for (int k = 0; k < 5; ++k) {
    ia.base[k] = static_cast<const double*>(c->d_base[k].p);
    ia.amp[k] = static_cast<const double*>(c->d_amp[k].p);
    ia.noise[k] = c->noise[k];
  }
  ia.beta = d_beta;
  ia.rsl = rsl;
  ia.U = U;
  ia.F = F;
  HIPCHK(c, launch_init(ia, c->precision, s));

  uint64_t done = 0;
  while (done < n_steps) {
    const int steps = (int)std::min<uint64_t>(n_steps - done, 1u << 30);
    StepArgs a = shared_step_args(c);
    a.n_traj = (int)n_traj;
    a.step0 = step0 + done;
    a.n_steps = steps;
    a.n_frames = 1;
    a.traj_ids = d_traj_ids;
    a.beta = d_beta;
    a.rsl = rsl;
    a.U = U;
    a.F = F;
    a.partials = partials;
    a.t0 = double(step0 + done) * c->p.dt;
    if (c->kernel == TWMC_KERNEL_STENCIL64) {
      HIPCHK(c, launch_stencil64(a, geo.grid, s));
    } else if (c->kernel == TWMC_KERNEL_BAND) {
      HIPCHK(c, launch_band(a, geo.grid, s));
    } else if (c->kernel == TWMC_KERNEL_FUSED64) {
      HIPCHK(c, launch_fused64(a, geo.grid, s));
    } else if (c->kernel == TWMC_KERNEL_MFMA) {
      HIPCHK(c, launch_mfma(a, geo.grid, s));
    } else if (c->p.hopping == TWMC_HOP_FFT) {
      HIPCHK(c, launch_generic_nn(a, c->precision, geo.grid_nn, geo.lds_nn, s));
    } else {
      HIPCHK(c, launch_generic(a, c->precision, geo.grid, geo.block, geo.lds, s));
    }
    done += (uint64_t)steps;
  }
  if (d_sums) {
    // per-site sums of the final state: every row of every chunk slot is
    // written (row 2 with zeros), then folded in slot order, read-only
    SiteSumArgs ss{};
    ss.beta = d_beta;
    ss.partials = partials;
    ss.n = n;
    ss.n_traj = (int)n_traj;
    ss.zero_k = 1;
    const int chunks = site_sum_chunks(n_traj, max_chunks, &ss.chunk);
    HIPCHK(c, launch_site_sums(ss, c->precision, s));
    ReduceArgs r{};
    r.n = n;
    r.slots = chunks;
    r.keep = 1;
    r.partials_in = partials;
    r.partials = partials;
    r.obs = d_sums;
    HIPCHK(c, launch_reduce(r, s));
  }
  return TWMC_OK;
}

int twmc_noisy_matrix(const twmc_matrix* m, uint64_t n, uint64_t seed, uint64_t traj, uint32_t tag,
                      double* out) {
  if (!m || !out) return fail(nullptr, TWMC_ERR_INVALID, "null argument");
  uint32_t k0 = uint32_t(seed & 0xffffffffu), k1 = uint32_t(seed >> 32);
  for (uint64_t i = 0; i < n; ++i) {
    double re = m->base ? m->base[2 * i] : m->value[0];
    double im = m->base ? m->base[2 * i + 1] : m->value[1];
    if (m->noise_type != TWMC_NOISE_NONE) {
      u32x4 x = philox4x32_10(make_counter(traj, tag, 0u, (uint32_t)(i >> 1)), k0, k1);
      uint32_t w0 = (i & 1) ? x.z : x.x, w1 = (i & 1) ? x.w : x.y;
      double zr, zi;
      if (m->noise_type == TWMC_NOISE_GAUSSIAN) {
        box_muller_f64(w0, w1, zr, zi);
      } else {
        zr = (double(w0) + 0.5) * 2.3283064365386963e-10;
        zi = (double(w1) + 0.5) * 2.3283064365386963e-10;
      }
      double ar = m->amp ? m->amp[2 * i] : m->amp_value[0];
      double ai = m->amp ? m->amp[2 * i + 1] : m->amp_value[1];
      re += zr * ar;
      im += zi * ai;
    }
    out[2 * i] = re;
    out[2 * i + 1] = im;
  }
  return TWMC_OK;
}

}  // extern "C"
