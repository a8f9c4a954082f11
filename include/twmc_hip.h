/*
 * twmc_hip.h -- C ABI of libtwmc_hip.so, the MI355X (gfx950) Truncated Wigner
 * Monte Carlo trajectory engine.
 *
 * This is the drop-in boundary for the reference's trajectory solvers.  A
 * context owns one parameter set on one GPU (the reference's
 * TWMCSystemData + one Solver instance); a batch is a set of trajectories
 * that share t_start, t_end and the initial-condition mode (the reference's
 * vector of TWMCTaskData passed to one Solver::Compute call).
 *
 *   reference interface                                   replaced by
 *   Solver::Setup  (src/Base/Solver.hpp:22,
 *     TWMCBaseSolver.cpp:24-92, TWMCLiebSolver.cpp:29-104) twmc_create
 *   Solver::Compute (src/Base/Solver.hpp:23,
 *     TWMCBaseSolver.cpp:94-283, TWMCLiebSolver.cpp:109-299) twmc_run, or
 *                                                          twmc_batch_begin /
 *                                                          _advance / _sync / _read
 *   Solver::ApproximateComputationProgress
 *     (src/Base/Solver.hpp:25, TWMCBaseSolver.cpp:293-295) twmc_progress
 *   SolverGPU::SetupGPU (src/Base/SolverGPU.hpp:20)        `device` argument of twmc_create
 *   Solver::~Solver (WorkerThread.cpp:30)                 twmc_destroy
 *   std::cerr diagnostics (TWMCBaseSolver.cpp:268-274)    int return codes + twmc_last_error
 *   NoisyMatrix::Generate (src/Base/NoisyMatrix.cpp:118-160)
 *                                                         twmc_noisy_matrix (host)
 *
 * Conventions (the reference's, unchanged):
 *   - lattice row-major nx x ny; a chain is nx = 1, ny = N (TWMCSystemData.cpp:25-37)
 *   - complex arrays are interleaved (re, im) doubles
 *   - hopping FFT:  h = IDFT(K . DFT(beta)) / N,
 *                   K(i,j) = -i 2 J ([nx!=1] cos(2 pi i/nx) + [ny!=1] cos(2 pi j/ny))
 *   - hopping matrix: h = -i (beta . C), C real N x N (TWMCLiebSolver.cpp:262)
 *   - drift a = dt [ (rsl + i U (|beta|^2 - 1)) beta + h + i F ],
 *     rsl = -i omega - gamma/2 (Base) or -i (detuning - E) - gamma/2 (Lieb)
 *   - beta <- beta + a + sqrt(gamma dt / 4) xi, xi complex N(0,1) from Philox
 * No function throws; all return TWMC_OK (0) or a negative TWMC_ERR_* code.
 * A context is not thread-safe; use one per host thread.
 */
#ifndef TWMC_HIP_H
#define TWMC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TWMC_ABI_VERSION 1

/* Opaque handles, one type per kind of context, so that passing a sweep or
 * network context where an engine context is expected does not compile. */
typedef struct twmc_ctx twmc_ctx;             /* one parameter set on one GPU */
typedef struct twmc_sweep_ctx twmc_sweep_ctx; /* a batched parameter sweep */
typedef struct twmc_mlp_ctx twmc_mlp_ctx;     /* the warm-start network */

enum {
  TWMC_OK = 0,
  TWMC_ERR_INVALID = -1,     /* bad argument or unsupported combination */
  TWMC_ERR_HIP = -2,         /* HIP runtime error (no GPU, launch failure, ...) */
  TWMC_ERR_NOMEM = -3,       /* device or host allocation failed */
  TWMC_ERR_STATE = -4,       /* call out of order (e.g. advance without begin) */
  TWMC_ERR_UNSUPPORTED = -5  /* lattice or size no kernel handles */
};

enum { TWMC_NOISE_NONE = 0, TWMC_NOISE_UNIFORM = 1, TWMC_NOISE_GAUSSIAN = 2 };
enum { TWMC_HOP_FFT = 0, TWMC_HOP_MATRIX = 1 };
enum { TWMC_ONSITE_OMEGA = 0, TWMC_ONSITE_DETUNING_MINUS_E = 1 };
enum {
  TWMC_INIT_SETTINGS = 0,   /* beta_init + beta_init_sigma . xi (TWMCBaseSolver.cpp:126-134) */
  TWMC_INIT_FIXEDPOINT = 1, /* beta_init (TWMCBaseSolver.cpp:154-174) */
  TWMC_INIT_PREVIOUS = 2,   /* caller's state (TWMCBaseSolver.cpp:176-202) */
  TWMC_INIT_GAUSSIAN = 3    /* Gaussian Wigner sample around a predicted steady state (FFT hopping only):
                               DFT(beta)_k = N mean [k = 0] + sqrt(N (n_k + 1/2) / 2) (xi_re + i xi_im) */
};
enum {
  TWMC_KERNEL_AUTO = 0,     /* fastest kernel that supports the parameters */
  TWMC_KERNEL_GENERIC = 1,  /* LDS-resident mixed-radix kernel, any supported size */
  TWMC_KERNEL_FUSED64 = 2,  /* register-resident 64x64 FP32 kernel */
  TWMC_KERNEL_MFMA = 3,     /* matrix hopping as BF16x3 MFMA GEMM (FP32 state) */
  TWMC_KERNEL_STENCIL64 = 4, /* 64x64 FP32, hopping as its exact nearest-neighbour form */
  TWMC_KERNEL_BAND = 5,      /* matrix hopping of half-width <= 3, FP32, <= 512 sites, shared parameters */
  TWMC_KERNEL_STACK = 6      /* FFT hopping, FP32, site-uniform: stacked-trajectory LDS stencil (<= 4096 sites) */
};
/* output mask bits for twmc_batch_desc.outputs */
enum {
  TWMC_OUT_OBS = 1u,        /* per-frame ensemble sums (always cheap) */
  TWMC_OUT_TRAJ = 2u,       /* full per-trajectory frames (the reference's "traj") */
  TWMC_OUT_FINAL = 4u       /* final state of every trajectory */
};

/* A site-resolved complex parameter with optional per-trajectory disorder
 * (the reference's NoisyMatrix: base value plus Gaussian N(0,1) or uniform
 * U(0,1) noise scaled by amp, real and imaginary parts separately). */
typedef struct twmc_matrix {
  const double* base;     /* 2N interleaved, or NULL: value[] on every site */
  double value[2];
  int32_t noise_type;     /* TWMC_NOISE_* */
  const double* amp;      /* 2N interleaved, or NULL: amp_value[] on every site */
  double amp_value[2];
} twmc_matrix;

typedef struct twmc_params {
  int32_t nx, ny;         /* lattice after 1D normalisation */
  int32_t hopping;        /* TWMC_HOP_* */
  int32_t onsite;         /* TWMC_ONSITE_* */
  int32_t precision;      /* 32 or 64 */
  int32_t kernel;         /* TWMC_KERNEL_* */
  double dt, gamma, J, detuning;
  twmc_matrix omega;      /* omega (Base) or E (Lieb) */
  twmc_matrix U, F, beta_init, beta_sigma;
  const double* coupling; /* TWMC_HOP_MATRIX: N x N real, row-major */
  uint64_t seed;          /* Philox key */
  uint64_t frame_steps;   /* steps between stored frames */
  /* time-dependent drive F(t) (TWMCLiebSolver.cpp:252-256 with
   * NoisyMatrix::GetAtTimeWithSuggestion, NoisyMatrix.cpp:164-264): when
   * F_n_times > 0 it replaces F before every step.  F_times ascending,
   * F_table F_n_times x N complex with gaps already filled; F(t) =
   * F_table[0] for t <= F_times[0], linear between table times, held at the
   * last entry after the last time.  Generic and MFMA kernels. */
  uint64_t F_n_times;
  const double* F_times;
  const double* F_table;
} twmc_params;

typedef struct twmc_batch_desc {
  const uint64_t* traj_ids;   /* n_traj trajectory ids (Philox counter) */
  uint64_t n_traj;
  int32_t init_mode;          /* TWMC_INIT_* */
  int32_t store_initial;      /* store the state before the first step */
  const double* beta_prev;    /* TWMC_INIT_PREVIOUS: n_traj x N complex (host) */
  uint64_t step0;             /* absolute index of the first step (resume) */
  double t_start, t_end;
  uint32_t outputs;           /* TWMC_OUT_* mask */
  const double* gauss_mean;   /* TWMC_INIT_GAUSSIAN: uniform mean field (re, im) */
  const double* gauss_nk;     /* TWMC_INIT_GAUSSIAN: N populations n_k >= -1/2 of the fluctuations,
                                 row-major k in DFT order (k = 0 first) */
} twmc_batch_desc;

typedef struct twmc_batch_info {
  uint64_t n_frames;          /* frames this batch will store */
  uint64_t n_steps;           /* Euler steps it will take */
  uint64_t n_segments;        /* launches of the step kernel */
  uint64_t segments_done;
  double t_final;             /* t after the loop, as the reference reports it */
  int32_t kernel;             /* TWMC_KERNEL_* chosen */
  int32_t truncated;          /* 1 if the time loop ends before n_frames */
} twmc_batch_info;

/* Host buffers, caller-allocated; NULL entries are skipped.
 * Site and k arrays are row-major nx x ny; frames are in storing order. */
typedef struct twmc_output {
  double* traj;        /* n_traj x n_frames x N complex */
  double* beta_final;  /* n_traj x N complex */
  double* sum_n;       /* n_frames x N : sum over trajectories of |beta_i|^2 */
  double* sum_n2;      /* n_frames x N : sum of |beta_i|^4 */
  double* sum_nk;      /* n_frames x N : sum of |DFT(beta)_k|^2 (unnormalised DFT) */
  double* sum_beta;    /* n_frames x N complex : sum of beta_i */
  double* traj_m2;     /* n_traj x n_frames : site mean of |beta_i|^2 */
  double* traj_m4;     /* n_traj x n_frames : site mean of |beta_i|^4 */
} twmc_output;

typedef struct twmc_stats {
  uint64_t step_launches;    /* step-kernel launches timed */
  double step_kernel_ms;     /* summed HIP-event time of those launches */
  uint64_t site_steps;       /* trajectories x sites x steps they advanced */
  double reduce_kernel_ms;
  uint64_t reduce_launches;
  int32_t kernel;            /* kernel used by the last batch */
  int32_t grid;              /* workgroups of the step kernel */
  int32_t block;             /* threads per workgroup */
  int32_t lds_bytes;         /* dynamic LDS per workgroup */
} twmc_stats;

int twmc_abi_version(void);
const char* twmc_last_error(const twmc_ctx* ctx);  /* ctx may be NULL: last global error */

int twmc_create(const twmc_params* p, int device, twmc_ctx** ctx_out);
void twmc_destroy(twmc_ctx* ctx);

int twmc_batch_begin(twmc_ctx* ctx, const twmc_batch_desc* d);
int twmc_batch_info_get(twmc_ctx* ctx, twmc_batch_info* info);
int twmc_batch_advance(twmc_ctx* ctx, uint64_t n_segments);   /* asynchronous */
int twmc_batch_sync(twmc_ctx* ctx);
int twmc_batch_read(twmc_ctx* ctx, twmc_output* out);         /* synchronises */
int twmc_run(twmc_ctx* ctx, const twmc_batch_desc* d, twmc_output* out);

int twmc_set_timing(twmc_ctx* ctx, int enable);   /* HIP events around step launches */
int twmc_query(twmc_ctx* ctx, twmc_stats* st);
int twmc_reset_stats(twmc_ctx* ctx);
/* fraction of the current batch's steps the device has completed (an event
 * after each segment, polled without blocking); any thread may call it while
 * the owning thread advances or reads the batch */
float twmc_progress(const twmc_ctx* ctx);

/* ---- allocation-free stepping --------------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory while they run (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  A variant of twmc_batch_advance on
 * caller-owned device buffers that calls no allocating routine:
 *   d_traj_ids  n_traj trajectory ids (device)
 *   d_beta      n_traj x N complex state of the context's precision (device,
 *               interleaved, float or double), advanced in place by n_steps
 *               Euler steps whose first has absolute index step0 (the Philox
 *               counter: continuing a trajectory means step0 += n_steps)
 *   d_sums      5 x N doubles (device) or NULL: per-site sums over the batch of
 *               the state after the steps, rows sum |b|^2, sum |b|^4, 0 (no
 *               k-space sums), sum Re b, sum Im b
 *   stream      hipStream_t to enqueue on, NULL: the context's stream
 * Asynchronous; no host synchronisation.  Site-shared parameters and a static
 * drive only, on the context's kernel (the stacked kernel excepted: create
 * such a context with TWMC_KERNEL_GENERIC).  Its own device memory -- the
 * parameter tables and the partial sums -- is one fixed array of
 * TWMC_NOALLOC_SCRATCH_BYTES held by every GPU the library runs on until the
 * process ends; a lattice whose tables and one partial slot exceed it is
 * rejected.  Every call on one GPU shares that array, so calls that overlap
 * on one GPU (different streams, threads or contexts) race. */
#define TWMC_NOALLOC_SCRATCH_BYTES (32u << 20)
int twmc_advance_noalloc(twmc_ctx* ctx, const uint64_t* d_traj_ids, uint64_t n_traj, uint64_t step0,
                         uint64_t n_steps, void* d_beta, double* d_sums, void* stream);

/* Host-side NoisyMatrix realisation with the kernels' Philox layout (tag as
 * in twmc_rng.h).  Used for one shared disorder realisation. */
int twmc_noisy_matrix(const twmc_matrix* m, uint64_t n, uint64_t seed, uint64_t traj,
                      uint32_t tag, double* out);

/* ---- batched parameter sweeps -------------------------------------------
 * Many independent systems (sweep points) of uneven 1D/2D sizes, each with
 * its own uniform parameters and trajectories, advanced together by one
 * persistent launch per frame (no reference counterpart: the reference runs
 * one system per process).  Every trajectory equals the same trajectory of a
 * twmc_create/twmc_run context with the same parameters and seed. */
typedef struct twmc_sweep_point {
  int32_t nx, ny;          /* lattice; N x 1 is stored as 1 x N */
  uint64_t n_traj;         /* trajectories of this point */
  uint64_t traj_id0;       /* id of its first trajectory (Philox counter) */
  double J, gamma;
  double omega[2], U[2], F[2], beta_init[2], beta_sigma[2];  /* uniform, complex */
} twmc_sweep_point;

int twmc_sweep_create(const twmc_sweep_point* pts, uint64_t n_points, double dt, uint64_t frame_steps,
                      uint64_t seed, int device, twmc_sweep_ctx** ctx_out);
void twmc_sweep_destroy(twmc_sweep_ctx* ctx);
const char* twmc_sweep_last_error(const twmc_sweep_ctx* ctx);
/* trajectories stacked per step work item, per point (tb x sites <= 4096) */
int twmc_sweep_set_batching(twmc_sweep_ctx* ctx, const int32_t* tb);
int twmc_sweep_get_batching(twmc_sweep_ctx* ctx, int32_t* tb);
/* times the stacking candidates of every lattice shape for `steps` steps and
 * keeps the fastest; tb_out (n_points, may be NULL) receives the choice */
int twmc_sweep_autotune(twmc_sweep_ctx* ctx, int32_t steps, int32_t* tb_out);
/* vacuum-start run from t_start to t_end with the stored initial frame;
 * traj_m2 / traj_m4 (total trajectories x n_frames, points in order) receive
 * the per-trajectory site means of |beta|^2 and |beta|^4; ksum (n_points x
 * n_frames x 34) the FP64 sums over trajectories of |DFT(beta)_k|^2 per radial
 * bin (32 bins, see twmc_sweep_mode_bins) and of sum_i beta_i (Re, Im);
 * use_graph captures the frame loop into one hipGraph; kernel_ms receives the
 * HIP-event time.  Any output pointer may be NULL.  The run stores at most
 * nearbyint((t_end - t_start) / (dt frame_steps)) + 1 frames (n_frames_out
 * the actual count): size the outputs for that many. */
int twmc_sweep_run(twmc_sweep_ctx* ctx, double t_start, double t_end, int32_t use_graph, double* traj_m2, double* traj_m4,
                   double* ksum, uint64_t* n_frames_out, double* kernel_ms);
/* radial bin (0..31) of every mode of point `point` (nx * ny entries, row-major):
 * bin 0 is k = 0, others 1 + floor(|k| / (pi sqrt 2 / 31)) capped at 31 */
int twmc_sweep_mode_bins(twmc_sweep_ctx* ctx, uint64_t point, uint8_t* bins);
int twmc_sweep_query(twmc_sweep_ctx* ctx, twmc_stats* st);

/* ---- learned warm start: the network --------------------------------------
 * A fully connected MLP (n_in -> n_hidden_layers x n_hidden, exact-erf GELU ->
 * n_out, linear), FP32, trained and evaluated on the GPU by the library's own
 * kernels (no reference counterpart).  Parameters are one flat FP32 vector:
 * for each layer l, W_l [in_l][out_l] row-major, then b_l [out_l].  Rows of x,
 * t, w, y are samples (row-major).  The loss is sum_ij w_ij (y_ij - t_ij)^2 /
 * rows (w = NULL: all ones); training is minibatch Adam over a shuffled epoch
 * (the remainder rows % batch of each epoch is skipped). */
typedef struct twmc_mlp_config {
  int32_t n_in, n_hidden, n_hidden_layers, n_out;
  uint64_t seed;            /* Glorot-uniform initial weights, zero biases */
} twmc_mlp_config;

typedef struct twmc_mlp_train_opts {
  int32_t epochs, batch;
  float lr, beta1, beta2, eps;
  uint64_t shuffle_seed;
} twmc_mlp_train_opts;

int twmc_mlp_create(const twmc_mlp_config* cfg, int device, twmc_mlp_ctx** ctx_out);
void twmc_mlp_destroy(twmc_mlp_ctx* ctx);
const char* twmc_mlp_last_error(const twmc_mlp_ctx* ctx);
uint64_t twmc_mlp_n_params(const twmc_mlp_ctx* ctx);
int twmc_mlp_get_params(twmc_mlp_ctx* ctx, float* out);
int twmc_mlp_set_params(twmc_mlp_ctx* ctx, const float* in);   /* also clears the Adam moments */
int twmc_mlp_forward(twmc_mlp_ctx* ctx, const float* x, uint64_t rows, float* y);
int twmc_mlp_loss_grad(twmc_mlp_ctx* ctx, const float* x, const float* t, const float* w, uint64_t rows, double* loss,
                       float* grad);
/* epoch_loss (epochs entries, may be NULL): mean minibatch loss per epoch */
int twmc_mlp_train(twmc_mlp_ctx* ctx, const float* x, const float* t, const float* w, uint64_t rows,
                   const twmc_mlp_train_opts* opts, double* epoch_loss);

#ifdef __cplusplus
}
#endif
#endif
