/*
 * twmc_oracle.h -- CPU oracle for the TWMC trajectory update.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain FP64 restatement of the
 * reference solver's arithmetic, used by tests/, by __graft_entry__.smoke()
 * and by bench.py's cpu_baseline leg as the checker and the CPU timing
 * baseline.  Nothing in pincosimulator_amd/ links, loads or calls it.
 *
 * What each entry point restates (paths relative to the reference repo):
 *   orc_philox4x32_10    Philox4x32-10, Salmon et al., SC'11 (Random123); no
 *                        reference counterpart: the reference draws its
 *                        increments from std::mt19937 (TWMCBaseSolver.cpp:119-120)
 *   orc_fft              FFTW dft_1d / dft_2d, unnormalised, out of place
 *                        (TWMCBaseSolver.cpp:44-60, :236, :241)
 *   orc_schedule         TWMCSystemData ctor schedule + ComputeNFrames +
 *                        GetStoredTimes (TWMCSystemData.cpp:67-115, :117-159)
 *   orc_run_trajectory   TWMCBaseSolver::Setup + Compute (TWMCBaseSolver.cpp:24-283)
 *                        and TWMCLiebSolver::Setup + Compute
 *                        (TWMCLiebSolver.cpp:29-104, :109-299)
 *   orc_noisy_matrix     NoisyMatrix::Generate / EvaluateNoise (NoisyMatrix.cpp:118-160)
 *   orc_lieb_coupling    Lieb coupling matrix (TWMCLiebSolver.cpp:52-85)
 *   orc_gaussian_state   warm-start Gaussian sample (no reference counterpart)
 *   orc_time_batch       one trajectory per task over a std::thread pool, as
 *                        ThreadedTaskProcessor/WorkerThread do, in the
 *                        reference's own per-step structure (timing only)
 */
#ifndef TWMC_ORACLE_H
#define TWMC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Philox stream tags: which random quantity a counter addresses. */
enum {
  ORC_TAG_WIENER = 0,  /* Euler-Maruyama increment, counter step = absolute step */
  ORC_TAG_INIT = 1,    /* beta_0 = beta_init + sigma * xi                        */
  ORC_TAG_U = 2,       /* NoisyMatrix disorder of U                              */
  ORC_TAG_OMEGA = 3,   /* disorder of omega (Base) / E (Lieb)                    */
  ORC_TAG_F = 4,       /* disorder of F                                          */
  ORC_TAG_BINIT = 5,   /* disorder of beta_init                                  */
  ORC_TAG_BSIGMA = 6,  /* disorder of beta_init_sigma                            */
  ORC_TAG_BOND = 7,    /* bond (hopping) disorder                                */
  ORC_TAG_GAUSS = 8    /* Gaussian warm-start sample, counter site = k-mode      */
};

enum { ORC_NOISE_NONE = 0, ORC_NOISE_UNIFORM = 1, ORC_NOISE_GAUSSIAN = 2 };
enum { ORC_LAT_FFT = 0, ORC_LAT_LIEB = 1, ORC_LAT_DENSE = 2 };
enum { ORC_INIT_SETTINGS = 0, ORC_INIT_FIXEDPOINT = 1, ORC_INIT_PREVIOUS = 2 };

/* Gaussian Wigner state around a predicted steady state (no reference
 * counterpart: the learned warm start's initial sample).  k-space, unnormalised
 * forward DFT convention:
 *   DFT(beta)_k = N mean [k = 0] + sqrt(N (nk[k] + 1/2) / 2) (xi_re + i xi_im),
 *   xi from Philox (tag ORC_TAG_GAUSS, step 0, site = k);  beta = IDFT / N.
 * Then <|DFT(beta)_k|^2>/N - 1/2 = nk[k] (k != 0) and <beta> = mean. */
void orc_gaussian_state(int nx, int ny, uint64_t seed, uint64_t traj, const double mean[2],
                        const double* nk, double* out);

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Counter layout shared with the HIP kernels:
 *   ctr = { step, site/2, traj & 0xffffffff, (tag << 24) | ((traj >> 32) & 0xffffff) }
 *   key = { seed & 0xffffffff, seed >> 32 }
 * Site 2p uses output words (x0, x1), site 2p+1 uses (x2, x3). */
void orc_counter(uint64_t seed, uint64_t traj, uint32_t tag, uint32_t step,
                 uint64_t site, uint32_t ctr[4], uint32_t key[2]);
/* complex standard normal (Re, Im each N(0,1)) by Box-Muller on the two words */
void orc_normal(uint64_t seed, uint64_t traj, uint32_t tag, uint32_t step,
                uint64_t site, double out[2]);
/* two U(0,1) draws (x + 0.5) 2^-32 on the two words */
void orc_uniform(uint64_t seed, uint64_t traj, uint32_t tag, uint32_t step,
                 uint64_t site, double out[2]);
void orc_box_muller(uint32_t a, uint32_t b, double out[2]);

/* Unnormalised DFT of an nx x ny row-major complex array (interleaved re,im).
 * sign = -1: forward e^{-i k.r};  sign = +1: backward e^{+i k.r}.
 * nx == 1 is a 1D transform of length ny.  Returns 0 on success. */
int orc_fft(int nx, int ny, const double* in, double* out, int sign);
/* O(N^2) direct DFT accumulated in long double, for pinning orc_fft. */
void orc_dft_direct(int nx, int ny, const double* in, double* out, int sign);

typedef struct orc_schedule {
  /* inputs */
  double t_start, t_end, timestep, frames_freq, frames_period;
  uint64_t n_frames_in;
  /* outputs (TWMCSystemData.cpp:67-115) */
  uint64_t frame_steps;
  double dt_obs;
  uint64_t n_frames_tot;
  double t_end_out;
  uint64_t n_dt;
  int32_t error;
} orc_schedule;
void orc_schedule_compute(orc_schedule* s);
/* size_t(nearbyint((t1 - t0) / dt_obs)) (TWMCSystemData.cpp:117-134) */
uint64_t orc_compute_n_frames(double dt_obs, double t0, double t1);
/* GetStoredTimes (TWMCSystemData.cpp:140-159); writes n_frames_tot+1 values */
void orc_stored_times(const orc_schedule* s, double* times);

/* A site-resolved complex parameter with optional per-trajectory noise. */
typedef struct orc_matrix {
  const double* base;    /* 2N interleaved, or NULL: value everywhere */
  double value[2];
  int32_t noise_type;    /* ORC_NOISE_* */
  const double* amp;     /* 2N interleaved, or NULL: amp_value everywhere */
  double amp_value[2];
} orc_matrix;

typedef struct orc_system {
  int32_t nx, ny;        /* lattice after 1D normalisation (nx = 1 for chains);
                            ORC_LAT_LIEB: nx = 1, ny cells of 3 sites */
  int32_t lattice;       /* ORC_LAT_* */
  int32_t pbc;           /* Lieb only: periodic closing bond */
  double dt, gamma, J, J_AB, J_BC, detuning;
  orc_matrix omega;      /* Base: omega; Lieb/dense: E (delta = detuning - E) */
  orc_matrix U, F, beta_init, beta_sigma;
  const double* coupling; /* ORC_LAT_DENSE: N x N real row-major coupling C */
  uint64_t frame_steps;
  uint64_t seed;         /* Philox key */
  /* time-dependent drive (TWMCLiebSolver.cpp:252-256 with
   * NoisyMatrix::GetAtTimeWithSuggestion, NoisyMatrix.cpp:164-264): when
   * F_n_times > 0, F(t) replaces F before every step; F_table holds
   * F_n_times x N complex (gaps already filled), F_times ascending; F is
   * F_table[0] up to F_times[0], linear in between, held after the last time */
  uint64_t F_n_times;
  const double* F_times;
  const double* F_table;
} orc_system;

/* One trajectory, exactly as the reference's Compute loop runs it, with
 * Philox increments.  traj_out receives up to max_frames frames of N complex
 * (interleaved); returns the number of frames written (< 0 on error).
 * beta_final (2N) and t_final are optional (may be NULL). */
int64_t orc_run_trajectory(const orc_system* s, uint64_t traj, int32_t init_mode,
                           const double* beta_prev, int32_t store_initial,
                           uint64_t step0, double t_start, double t_end,
                           double* traj_out, uint64_t max_frames,
                           double* beta_final, double* t_final);

/* Realisation of a noisy matrix for trajectory traj (N sites, 2N doubles). */
void orc_noisy_matrix(const orc_matrix* m, uint64_t n, uint64_t seed, uint64_t traj,
                      uint32_t tag, double* out);
/* Lieb coupling, dense 3ny x 3ny real row-major (TWMCLiebSolver.cpp:52-85). */
void orc_lieb_coupling(int32_t ny, double J_AB, double J_BC, int32_t pbc, double* out);

/* Timing baseline: n_traj trajectories of n_steps steps each, one trajectory
 * per task over n_threads std::threads.  faithful = 1 reproduces the
 * reference's per-step buffer copies and std::mt19937 + normal_distribution
 * noise; faithful = 0 uses the Philox path.  Returns wall seconds. */
double orc_time_batch(const orc_system* s, uint64_t n_traj, int32_t n_threads,
                      uint64_t n_steps, int32_t faithful);

#ifdef __cplusplus
}
#endif
#endif
