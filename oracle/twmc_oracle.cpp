// twmc_oracle.cpp -- CPU oracle for the TWMC trajectory update.
//
// TEST INFRASTRUCTURE ONLY (see twmc_oracle.h).  FP64 restatement of the
// reference's TWMCBaseSolver / TWMCLiebSolver arithmetic with Philox
// increments, plus an own mixed-radix FFT in place of FFTW.
#include "twmc_oracle.h"

#include <atomic>
#include <chrono>
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <memory>
#include <random>
#include <thread>
#include <vector>

namespace {

using cd = std::complex<double>;
const cd ij(0.0, 1.0);

// ---------------------------------------------------------------- Philox
// Philox4x32-10 as published by Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11: round = two 32x32->64 products with
// the multipliers M0, M1, outputs permuted and xor-ed with the key; the key is
// bumped by the Weyl constants W0, W1 between rounds.
const uint32_t kM0 = 0xD2511F53u, kM1 = 0xCD9E8D57u;
const uint32_t kW0 = 0x9E3779B9u, kW1 = 0xBB67AE85u;

inline void philox10(const uint32_t c_in[4], const uint32_t k_in[2], uint32_t out[4]) {
  uint32_t c0 = c_in[0], c1 = c_in[1], c2 = c_in[2], c3 = c_in[3];
  uint32_t k0 = k_in[0], k1 = k_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += kW0; k1 += kW1; }
    uint64_t p0 = uint64_t(kM0) * c0;
    uint64_t p1 = uint64_t(kM1) * c2;
    uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
    uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

inline void make_counter(uint64_t seed, uint64_t traj, uint32_t tag, uint32_t step,
                         uint64_t site, uint32_t ctr[4], uint32_t key[2]) {
  ctr[0] = step;
  ctr[1] = uint32_t(site >> 1);
  ctr[2] = uint32_t(traj & 0xffffffffu);
  ctr[3] = (tag << 24) | uint32_t((traj >> 32) & 0xffffffu);
  key[0] = uint32_t(seed & 0xffffffffu);
  key[1] = uint32_t(seed >> 32);
}

const double kTwoPi = 6.283185307179586;
const double k2m32 = 2.3283064365386963e-10;  // 2^-32

inline void box_muller(uint32_t a, uint32_t b, double out[2]) {
  double u1 = (double(a) + 0.5) * k2m32;  // (0, 1)
  double th = double(b) * k2m32;          // [0, 1) revolutions
  double r = std::sqrt(-2.0 * std::log(u1));
  double ang = kTwoPi * th;
  out[0] = r * std::cos(ang);
  out[1] = r * std::sin(ang);
}

inline cd draw_normal(uint64_t seed, uint64_t traj, uint32_t tag, uint32_t step, uint64_t site) {
  uint32_t ctr[4], key[2], x[4];
  make_counter(seed, traj, tag, step, site, ctr, key);
  philox10(ctr, key, x);
  double z[2];
  if ((site & 1) == 0) box_muller(x[0], x[1], z); else box_muller(x[2], x[3], z);
  return cd(z[0], z[1]);
}

// Fills a whole N-site complex normal field; one Philox call per site pair.
void normal_field(uint64_t seed, uint64_t traj, uint32_t tag, uint32_t step, size_t n, cd* out) {
  uint32_t ctr[4], key[2], x[4];
  for (size_t p = 0; 2 * p < n; ++p) {
    make_counter(seed, traj, tag, step, 2 * p, ctr, key);
    philox10(ctr, key, x);
    double z[2];
    box_muller(x[0], x[1], z);
    out[2 * p] = cd(z[0], z[1]);
    if (2 * p + 1 < n) {
      box_muller(x[2], x[3], z);
      out[2 * p + 1] = cd(z[0], z[1]);
    }
  }
}

// ---------------------------------------------------------------- FFT
// Mixed-radix Stockham autosort FFT (radix 4, 2, 3, 5, generic odd), FP64.
// Pass with radix R and running span Ns: butterfly j reads in[j + r n/R],
// multiplies by exp(s 2 pi i r (j mod Ns) / (Ns R)), does a length-R DFT and
// writes out[(j / Ns) Ns R + (j mod Ns) + q Ns].
struct Plan1D {
  int n = 0;
  std::vector<int> radix;
  std::vector<cd> tw;  // exp(-2 pi i m / n), m = 0..n-1
  explicit Plan1D(int n_) : n(n_) {
    int m = n;
    while (m % 4 == 0) { radix.push_back(4); m /= 4; }
    while (m % 2 == 0) { radix.push_back(2); m /= 2; }
    for (int p = 3; m > 1; p += 2) {
      while (m % p == 0) { radix.push_back(p); m /= p; }
    }
    tw.resize(n);
    for (int k = 0; k < n; ++k) {
      long double a = -2.0L * 3.14159265358979323846264338327950288L * k / n;
      tw[k] = cd(double(std::cos(a)), double(std::sin(a)));
    }
  }
  // exp(sign 2 pi i m / n)
  inline cd w(long m, int sign) const {
    m %= n;
    if (m < 0) m += n;
    cd t = tw[m];
    return sign < 0 ? t : std::conj(t);
  }
  // strided 1D transform of src (stride st) into dst (stride st); work is n.
  void exec(const cd* src, cd* dst, long st, int sign, cd* a, cd* b) const {
    for (int i = 0; i < n; ++i) a[i] = src[i * st];
    cd* in = a;
    cd* out = b;
    long Ns = 1;
    // butterfly scratch sized by the largest radix (a prime factor can be long)
    int rmax = 1;
    for (int R : radix) rmax = R > rmax ? R : rmax;
    std::vector<cd> v(rmax), wr(rmax), o(rmax);
    for (int R : radix) {
      long nR = n / R;
      for (long j = 0; j < nR; ++j) {
        long k = j % Ns;
        for (int r = 0; r < R; ++r) v[r] = in[j + r * nR] * w(r * k * (n / (Ns * R)), sign);
        // length-R DFT of v with exp(sign 2 pi i r q / R)
        if (R == 2) {
          o[0] = v[0] + v[1];
          o[1] = v[0] - v[1];
        } else if (R == 4) {
          cd t0 = v[0] + v[2], t1 = v[0] - v[2];
          cd t2 = v[1] + v[3], t3 = (v[1] - v[3]) * cd(0.0, double(sign));
          o[0] = t0 + t2;
          o[2] = t0 - t2;
          o[1] = t1 + t3;
          o[3] = t1 - t3;
        } else {
          for (int q = 0; q < R; ++q) wr[q] = w(long(q) * (n / R), sign);
          for (int q = 0; q < R; ++q) {
            cd acc = 0.0;
            for (int r = 0; r < R; ++r) acc += v[r] * wr[(r * q) % R];
            o[q] = acc;
          }
        }
        long base = (j / Ns) * Ns * R + k;
        for (int q = 0; q < R; ++q) out[base + q * Ns] = o[q];
      }
      std::swap(in, out);
      Ns *= R;
    }
    for (int i = 0; i < n; ++i) dst[i * st] = in[i];
  }
};

const Plan1D& plan_for(int n) {
  thread_local std::map<int, std::unique_ptr<Plan1D>> cache;
  auto it = cache.find(n);
  if (it == cache.end()) it = cache.emplace(n, std::make_unique<Plan1D>(n)).first;
  return *it->second;
}

// 2D (or 1D when nx == 1) unnormalised transform, row-major nx x ny.
void fft2(int nx, int ny, const cd* in, cd* out, int sign) {
  thread_local std::vector<cd> a, b;
  int mx = nx > ny ? nx : ny;
  if ((int)a.size() < mx) { a.resize(mx); b.resize(mx); }
  const Plan1D& py = plan_for(ny);
  for (int x = 0; x < nx; ++x) py.exec(in + (long)x * ny, out + (long)x * ny, 1, sign, a.data(), b.data());
  if (nx > 1) {
    const Plan1D& px = plan_for(nx);
    for (int y = 0; y < ny; ++y) px.exec(out + y, out + y, ny, sign, a.data(), b.data());
  }
}

// ---------------------------------------------------------------- matrices
std::vector<cd> base_of(const orc_matrix& m, size_t n) {
  std::vector<cd> v(n);
  for (size_t i = 0; i < n; ++i)
    v[i] = m.base ? cd(m.base[2 * i], m.base[2 * i + 1]) : cd(m.value[0], m.value[1]);
  return v;
}

// NoisyMatrix::Generate (NoisyMatrix.cpp:118-160): base + noise, where
// Gaussian noise is N(0,1) amp.re + i N(0,1) amp.im and uniform noise is
// U(0,1) amp.re + i U(0,1) amp.im (settings-built matrices use U(0,1),
// NoisyMatrix.cpp:43-47).
std::vector<cd> generate(const orc_matrix& m, size_t n, uint64_t seed, uint64_t traj, uint32_t tag) {
  std::vector<cd> v = base_of(m, n);
  if (m.noise_type == ORC_NOISE_NONE) return v;
  for (size_t i = 0; i < n; ++i) {
    double ar = m.amp ? m.amp[2 * i] : m.amp_value[0];
    double ai = m.amp ? m.amp[2 * i + 1] : m.amp_value[1];
    double z[2];
    if (m.noise_type == ORC_NOISE_GAUSSIAN) {
      cd g = draw_normal(seed, traj, tag, 0, i);
      z[0] = g.real(); z[1] = g.imag();
    } else {
      orc_uniform(seed, traj, tag, 0, i, z);
    }
    v[i] += z[0] * ar + ij * (z[1] * ai);
  }
  return v;
}

// Lieb coupling (TWMCLiebSolver.cpp:52-85): per cell i, A(3i)-B(3i+1) J_AB,
// B(3i+1)-C(3i+2) J_BC, inter-cell C(3i+2)-B(3i+4) J_BC; PBC closes C_last-B_0.
void lieb_coupling(int ny, double jab, double jbc, int pbc, double* c) {
  int n = 3 * ny;
  std::memset(c, 0, sizeof(double) * (size_t)n * n);
  for (int i = 0; i < ny; ++i) {
    c[(3 * i) * n + 3 * i + 1] += jab;
    c[(3 * i + 1) * n + 3 * i] += jab;
    c[(3 * i + 1) * n + 3 * i + 2] += jbc;
    c[(3 * i + 2) * n + 3 * i + 1] += jbc;
    if (i == ny - 1) {
      if (pbc) {
        c[(3 * i + 2) * n + 1] += jbc;
        c[1 * n + 3 * i + 2] += jbc;
      }
      break;
    }
    c[(3 * i + 2) * n + 3 * (i + 1) + 1] += jbc;
    c[(3 * (i + 1) + 1) * n + 3 * i + 2] += jbc;
  }
}

// ---------------------------------------------------------------- solver
struct Trajectory {
  const orc_system& s;
  size_t n;
  bool fft;
  std::vector<cd> K;       // k_step_linear (TWMCBaseSolver.cpp:80-88)
  std::vector<cd> rsl;     // real_step_linear
  std::vector<cd> U, F;
  // sparse coupling rows (column list per site) for the Lieb / dense solvers
  std::vector<std::vector<std::pair<int, double>>> ccol;
  std::vector<cd> f_out, g, h;

  Trajectory(const orc_system& s_, uint64_t traj) : s(s_) {
    // a Lieb chain of ny cells has 3 ny sites (TWMCLiebSolver.cpp:29-104)
    n = (size_t)s.nx * s.ny * (s.lattice == ORC_LAT_LIEB ? 3 : 1);
    fft = (s.lattice == ORC_LAT_FFT);
    if (fft) {
      // TWMCBaseSolver::Setup: K(i,j) = -i 2 J (fx cos(2pi/nx i) + fy cos(2pi/ny j))
      K.resize(n);
      double fx = (s.nx == 1) ? 0.0 : 1.0;
      double fy = (s.ny == 1) ? 0.0 : 1.0;
      for (int i = 0; i < s.nx; ++i)
        for (int j = 0; j < s.ny; ++j)
          K[(size_t)i * s.ny + j] = -ij * 2.0 * s.J *
              (fx * std::cos(2.0 * M_PI / double(s.nx) * double(i)) +
               fy * std::cos(2.0 * M_PI / double(s.ny) * double(j)));
      std::vector<cd> om = generate(s.omega, n, s.seed, traj, ORC_TAG_OMEGA);
      rsl.resize(n);
      for (size_t i = 0; i < n; ++i) rsl[i] = (-ij * om[i] - s.gamma / 2.0);
    } else {
      // TWMCLiebSolver::Setup: delta = -E + detuning; rsl = -i delta - gamma/2
      std::vector<cd> E = generate(s.omega, n, s.seed, traj, ORC_TAG_OMEGA);
      rsl.resize(n);
      for (size_t i = 0; i < n; ++i) {
        cd delta = -E[i] + s.detuning;
        rsl[i] = (-ij * delta - s.gamma / 2.0);
      }
      std::vector<double> C((size_t)n * n);
      if (s.lattice == ORC_LAT_LIEB) lieb_coupling(s.ny, s.J_AB, s.J_BC, s.pbc, C.data());
      else std::memcpy(C.data(), s.coupling, sizeof(double) * n * n);
      ccol.resize(n);
      for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j)
          if (C[i * n + j] != 0.0) ccol[j].push_back({(int)i, C[i * n + j]});
    }
    U = generate(s.U, n, s.seed, traj, ORC_TAG_U);
    F = generate(s.F, n, s.seed, traj, ORC_TAG_F);
    f_out.resize(n); g.resize(n); h.resize(n);
    // F(t) table and per-interval slopes (NoisyMatrix.cpp:228-239); the slope
    // after the last time is zero, so the last value is held
    for (uint64_t k = 0; k < s.F_n_times; ++k) {
      ft_times.push_back(s.F_times[k]);
      std::vector<cd> v(n);
      for (size_t i = 0; i < n; ++i) v[i] = cd(s.F_table[2 * (k * n + i)], s.F_table[2 * (k * n + i) + 1]);
      ft_vals.push_back(v);
    }
    for (size_t k = 0; k < ft_times.size(); ++k) {
      std::vector<cd> d(n, cd(0.0, 0.0));
      if (k + 1 < ft_times.size())
        for (size_t i = 0; i < n; ++i) d[i] = (ft_vals[k + 1][i] - ft_vals[k][i]) / (ft_times[k + 1] - ft_times[k]);
      ft_slopes.push_back(d);
    }
  }

  std::vector<double> ft_times;
  std::vector<std::vector<cd>> ft_vals, ft_slopes;

  // F at time t (NoisyMatrix::GetAtTimeWithSuggestion, NoisyMatrix.cpp:252-264)
  void drive_at(double t) {
    if (t <= ft_times[0]) {
      F = ft_vals[0];
      return;
    }
    size_t id = size_t(std::lower_bound(ft_times.begin(), ft_times.end(), t) - ft_times.begin()) - 1;
    for (size_t i = 0; i < n; ++i) F[i] = ft_vals[id][i] + cd(t - ft_times[id], 0.0) * ft_slopes[id][i];
  }

  // hopping field h (Base: FFT^-1 (K . FFT beta) / N; Lieb: -i (beta . C))
  void hop(const cd* beta) {
    if (fft) {
      fft2(s.nx, s.ny, beta, f_out.data(), -1);
      for (size_t i = 0; i < n; ++i) g[i] = K[i] * f_out[i];
      fft2(s.nx, s.ny, g.data(), h.data(), +1);
      double norm = double(n);
      for (size_t i = 0; i < n; ++i) h[i] = h[i] / norm;
    } else {
      for (size_t j = 0; j < n; ++j) {
        cd acc = 0.0;
        for (auto& e : ccol[j]) acc += beta[e.first] * e.second;
        h[j] = -ij * acc;
      }
    }
  }

  // one Euler-Maruyama step (TWMCBaseSolver.cpp:245-253 / TWMCLiebSolver.cpp:259-267)
  void step(cd* beta, const cd* xi) {
    hop(beta);
    double dt = s.dt;
    double dt4 = std::sqrt(s.gamma * s.dt / 4.0);
    for (size_t i = 0; i < n; ++i) {
      double n2 = std::norm(beta[i]);
      cd a = ((rsl[i] + ij * U[i] * (n2 - 1.0)) * beta[i] + h[i] + ij * F[i]) * dt;
      beta[i] = beta[i] + a + dt4 * xi[i];
    }
  }
};

}  // namespace

// ================================================================ C ABI
extern "C" {

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  philox10(ctr, key, out);
}

void orc_counter(uint64_t seed, uint64_t traj, uint32_t tag, uint32_t step, uint64_t site,
                 uint32_t ctr[4], uint32_t key[2]) {
  make_counter(seed, traj, tag, step, site, ctr, key);
}

void orc_box_muller(uint32_t a, uint32_t b, double out[2]) { box_muller(a, b, out); }

void orc_normal(uint64_t seed, uint64_t traj, uint32_t tag, uint32_t step, uint64_t site,
                double out[2]) {
  cd z = draw_normal(seed, traj, tag, step, site);
  out[0] = z.real();
  out[1] = z.imag();
}

void orc_uniform(uint64_t seed, uint64_t traj, uint32_t tag, uint32_t step, uint64_t site,
                 double out[2]) {
  uint32_t ctr[4], key[2], x[4];
  make_counter(seed, traj, tag, step, site, ctr, key);
  philox10(ctr, key, x);
  uint32_t a = (site & 1) ? x[2] : x[0];
  uint32_t b = (site & 1) ? x[3] : x[1];
  out[0] = (double(a) + 0.5) * k2m32;
  out[1] = (double(b) + 0.5) * k2m32;
}

int orc_fft(int nx, int ny, const double* in, double* out, int sign) {
  if (nx < 1 || ny < 1) return -1;
  fft2(nx, ny, reinterpret_cast<const cd*>(in), reinterpret_cast<cd*>(out), sign);
  return 0;
}

void orc_dft_direct(int nx, int ny, const double* in, double* out, int sign) {
  typedef long double ld;
  const ld pi = 3.14159265358979323846264338327950288L;
  for (int kx = 0; kx < nx; ++kx)
    for (int ky = 0; ky < ny; ++ky) {
      ld re = 0, im = 0;
      for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y) {
          // exact phase reduction: (kx x mod nx)/nx + (ky y mod ny)/ny
          ld ph = 2 * pi * sign *
                  ((ld)((long)kx * x % nx) / nx + (ld)((long)ky * y % ny) / ny);
          ld c = cosl(ph), s = sinl(ph);
          ld ar = in[2 * ((long)x * ny + y)], ai = in[2 * ((long)x * ny + y) + 1];
          re += ar * c - ai * s;
          im += ar * s + ai * c;
        }
      out[2 * ((long)kx * ny + ky)] = (double)re;
      out[2 * ((long)kx * ny + ky) + 1] = (double)im;
    }
}

// TWMCSystemData ctor, schedule part (TWMCSystemData.cpp:67-115).
void orc_schedule_compute(orc_schedule* s) {
  double freq = s->frames_freq;
  double period = s->frames_period;
  double obs_n_frames = (double)s->n_frames_in;
  s->error = 0;
  if (freq == 0 && period == 0) {
    freq = obs_n_frames / (s->t_end - s->t_start);
    if (obs_n_frames == 0) { s->error = 1; return; }
  } else if (freq == 0 && period != 0) {
    freq = 1 / period;
  }
  double timestep = s->timestep;
  double t_end = s->t_end;
  size_t frame_steps = std::max(size_t(1), size_t(ceil(1 / (freq * timestep))));
  double dt_obs = timestep * frame_steps;
  size_t nFramesTot = size_t(ceil((t_end - s->t_start) / dt_obs));
  t_end = nFramesTot * dt_obs;
  size_t n_times = 0;
  double t = s->t_start;
  while (t <= t_end) { n_times++; t += timestep; }
  if (dt_obs < timestep) { dt_obs = timestep; frame_steps = 1; }
  s->frame_steps = frame_steps;
  s->dt_obs = dt_obs;
  s->n_frames_tot = nFramesTot;
  s->t_end_out = t_end;
  s->n_dt = n_times;
}

uint64_t orc_compute_n_frames(double dt_obs, double t0, double t1) {
  return uint64_t(size_t(nearbyint((t1 - t0) / dt_obs)));
}

void orc_stored_times(const orc_schedule* s, double* times) {
  size_t frames = s->n_frames_tot + 1;
  double t = s->t_start;
  size_t i_step = 0, i_frame = 1;
  times[0] = s->t_start;
  while (t <= s->t_end_out) {
    t += s->timestep;
    i_step++;
    if ((i_step % s->frame_steps == 0) && i_frame < frames) { times[i_frame] = t; i_frame++; }
  }
}

void orc_noisy_matrix(const orc_matrix* m, uint64_t n, uint64_t seed, uint64_t traj, uint32_t tag,
                      double* out) {
  std::vector<cd> v = generate(*m, n, seed, traj, tag);
  std::memcpy(out, v.data(), sizeof(cd) * n);
}

void orc_lieb_coupling(int32_t ny, double J_AB, double J_BC, int32_t pbc, double* out) {
  lieb_coupling(ny, J_AB, J_BC, pbc, out);
}

// TWMCBaseSolver::Compute / TWMCLiebSolver::Compute for one task.
void orc_gaussian_state(int nx, int ny, uint64_t seed, uint64_t traj, const double mean[2],
                        const double* nk, double* out) {
  const size_t n = (size_t)nx * ny;
  std::vector<cd> k(n), r(n);
  for (size_t i = 0; i < n; ++i)
    k[i] = std::sqrt(double(n) * (nk[i] + 0.5) / 2.0) * draw_normal(seed, traj, ORC_TAG_GAUSS, 0, i);
  k[0] += double(n) * cd(mean[0], mean[1]);
  fft2(nx, ny, k.data(), r.data(), +1);
  const double inv_n = 1.0 / double(n);
  for (size_t i = 0; i < n; ++i) {
    out[2 * i] = r[i].real() * inv_n;
    out[2 * i + 1] = r[i].imag() * inv_n;
  }
}

int64_t orc_run_trajectory(const orc_system* sp, uint64_t traj, int32_t init_mode,
                           const double* beta_prev, int32_t store_initial, uint64_t step0,
                           double t_start, double t_end, double* traj_out, uint64_t max_frames,
                           double* beta_final, double* t_final) {
  const orc_system& s = *sp;
  if (s.nx < 1 || s.ny < 1 || s.frame_steps < 1) return -1;
  Trajectory tr(s, traj);
  size_t n = tr.n;
  double dt_obs = s.dt * (double)s.frame_steps;
  size_t n_frames = size_t(nearbyint((t_end - t_start) / dt_obs));
  if (store_initial) n_frames++;

  std::vector<cd> beta(n);
  if (init_mode == ORC_INIT_SETTINGS) {
    std::vector<cd> b0 = generate(s.beta_init, n, s.seed, traj, ORC_TAG_BINIT);
    std::vector<cd> xi(n);
    normal_field(s.seed, traj, ORC_TAG_INIT, 0, n, xi.data());
    std::vector<cd> sg = generate(s.beta_sigma, n, s.seed, traj, ORC_TAG_BSIGMA);
    for (size_t i = 0; i < n; ++i) beta[i] = b0[i] + sg[i] * xi[i];
  } else if (init_mode == ORC_INIT_FIXEDPOINT) {
    beta = base_of(s.beta_init, n);
  } else {
    if (!beta_prev) return -2;
    for (size_t i = 0; i < n; ++i) beta[i] = cd(beta_prev[2 * i], beta_prev[2 * i + 1]);
  }

  size_t i_frame = 0;
  uint64_t i_step = 0;
  auto capture = [&]() {
    if (traj_out && i_frame < max_frames)
      std::memcpy(traj_out + 2 * n * i_frame, beta.data(), sizeof(cd) * n);
    i_frame++;
  };
  if (store_initial && i_frame < n_frames) capture();

  std::vector<cd> xi(n);
  double t = t_start;
  while (t <= t_end) {
    if (!tr.ft_times.empty()) tr.drive_at(t);  // TWMCLiebSolver.cpp:252-256
    normal_field(s.seed, traj, ORC_TAG_WIENER, uint32_t(step0 + i_step), n, xi.data());
    tr.step(beta.data(), xi.data());
    t += s.dt;
    i_step++;
    if ((i_step % s.frame_steps == 0) && i_frame < n_frames) {
      capture();
      if (i_frame == n_frames) break;
    }
  }
  if (beta_final) std::memcpy(beta_final, beta.data(), sizeof(cd) * n);
  if (t_final) *t_final = t;
  return (int64_t)(i_frame < max_frames ? i_frame : i_frame);
}

double orc_time_batch(const orc_system* sp, uint64_t n_traj, int32_t n_threads, uint64_t n_steps,
                      int32_t faithful) {
  const orc_system& s = *sp;
  std::atomic<uint64_t> next(0);
  auto worker = [&]() {
    for (;;) {
      uint64_t id = next.fetch_add(1);
      if (id >= n_traj) break;
      Trajectory tr(s, id);
      size_t n = tr.n;
      std::vector<cd> beta(n), xi(n);
      if (!faithful) {
        normal_field(s.seed, id, ORC_TAG_INIT, 0, n, xi.data());
        for (size_t i = 0; i < n; ++i) beta[i] = 0.5 * xi[i];
        for (uint64_t k = 0; k < n_steps; ++k) {
          normal_field(s.seed, id, ORC_TAG_WIENER, uint32_t(k), n, xi.data());
          tr.step(beta.data(), xi.data());
        }
      } else {
        // The reference's per-step structure (TWMCBaseSolver.cpp:232-266):
        // copy into the FFT input buffer, forward FFT, K product into the
        // inverse-FFT input, inverse FFT, /N, drift, randCMat, update, copy.
        std::mt19937 gen((unsigned)(s.seed + id));
        std::normal_distribution<> normal(0, 1);
        for (size_t i = 0; i < n; ++i) { double a = normal(gen); beta[i] = 0.5 * cd(a, normal(gen)); }
        std::vector<cd> f_in(n), f_out(n), i_in(n), i_out(n), a_t(n), kai(n);
        double dt = s.dt, dt4 = std::sqrt(s.gamma * s.dt / 4.0);
        for (uint64_t k = 0; k < n_steps; ++k) {
          if (tr.fft) {
            f_in = beta;
            fft2(s.nx, s.ny, f_in.data(), f_out.data(), -1);
            for (size_t i = 0; i < n; ++i) i_in[i] = tr.K[i] * f_out[i];
            fft2(s.nx, s.ny, i_in.data(), i_out.data(), +1);
            for (size_t i = 0; i < n; ++i) i_out[i] = i_out[i] / double(n);
          } else {
            tr.hop(beta.data());
            i_out = tr.h;
          }
          for (size_t i = 0; i < n; ++i)
            a_t[i] = ((tr.rsl[i] + ij * tr.U[i] * (std::norm(beta[i]) - 1.0)) * beta[i] + i_out[i] +
                      ij * tr.F[i]) * dt;
          for (size_t i = 0; i < n; ++i) { double a = normal(gen); xi[i] = cd(a, normal(gen)); }
          for (size_t i = 0; i < n; ++i) kai[i] = beta[i] + a_t[i] + dt4 * xi[i];
          beta = kai;
        }
      }
      volatile double sink = beta[0].real();
      (void)sink;
    }
  };
  auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> pool;
  int nt = n_threads < 1 ? 1 : n_threads;
  for (int i = 0; i < nt; ++i) pool.emplace_back(worker);
  for (auto& th : pool) th.join();
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
