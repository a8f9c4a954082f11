// twmc_system.hpp -- per-run parameters, tasks and results of a TWMC run.
//
// TWMCSystemData restates the reference's src/TWMC/TWMCSystemData.cpp:14-159
// (parameter matrices, the 1D normalisation N x 1 -> 1 x N, the frame
// schedule and the stored time grid); TWMCTaskData restates
// src/TWMC/TWMCTaskData.hpp:29-44; result dataset names follow
// src/TWMC/TWMCResults.cpp:44-49.
#ifndef PINCO_TWMC_SYSTEM_HPP
#define PINCO_TWMC_SYSTEM_HPP

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "noisy_matrix.hpp"
#include "plugin.hpp"
#include "settings.hpp"

namespace pinco {

// dataset names of a trajectory result
inline const char* kDatasetTraj = "traj";
inline const char* kDatasetU = "U_Realization";
inline const char* kDatasetF = "F_Realizatio";  // sic, as the reference writes it
inline const char* kDatasetOmega = "Omega_Realization";
inline const char* kDatasetM2 = "mean_n2";       // per-frame site mean of |beta|^2
inline const char* kDatasetM4 = "mean_n4";       // per-frame site mean of |beta|^4

// FP64 ensemble sums of a run, accumulated across batches.
struct EnsembleSums {
  std::mutex mu;
  size_t n_traj = 0, n_frames = 0, n = 0;
  std::vector<double> sum_n, sum_n2, sum_nk, sum_beta;  // frames x N (sum_beta complex)
  void Add(size_t traj, size_t frames, size_t sites, const std::vector<double>& n1, const std::vector<double>& n2,
           const std::vector<double>& nk, const std::vector<double>& b);
};

class TWMCSystemData {
 public:
  enum Dimension { D0, D1, D2 };

  explicit TWMCSystemData(const Settings& settings);
  TWMCSystemData() = default;

  size_t ComputeNFrames(double t0, double t1) const;
  std::vector<double> GetStoredTimes() const;
  std::vector<std::vector<cd>> GetStoredVariableEvolution(const NoisyMatrix& m) const;

  std::string latticeName;
  size_t nx = 1, ny = 1, nxy = 1, cellSz = 1;
  bool PBC = false;
  Dimension dimension = D0;

  std::unique_ptr<NoisyMatrix> U, F, omega, E, beta_init_val, beta_init_sigma_val;

  double J_val = 0.0, detuning = 0.0, J_AB_val = 0.0, J_BC_val = 0.0, gamma_val = 0.0;
  double t_start = 0.0, t_end = 0.0, dt = 0.0, dt_obs = 0.0;
  double frames_freq = 0.0, frames_period = 0.0;
  size_t nFramesTot = 0, n_dt = 0, frame_steps = 1;

  // engine extensions
  uint64_t seed = 0;              // Philox key (SEED)
  int precision = 64;             // 32 or 64 (key `precision`)
  std::string store = "traj";     // "traj", "observables" or "both" (key `store`)
  double J_sigma = 0.0;           // Gaussian bond disorder of the Lieb hopping (key `J_Noise_Val_0`)
  uint64_t disorder_seed = 0;     // seed of the shared bond-disorder realisation (key `disorder_seed`)
  size_t gpu_batch = 4096;        // trajectories per Compute call (key `gpu_batch`)
  std::shared_ptr<EnsembleSums> sums = std::make_shared<EnsembleSums>();
};

class TWMCTaskData : public TaskData {
 public:
  enum InitialConditions { ReadFromSettings = 0, FixedPoint = 1, ReadFromPreviousData = 2 };
  std::shared_ptr<TWMCSystemData> systemData;
  std::unique_ptr<TaskResults> prevData;
  size_t id = 0;
  double t_start = 0.0, t_end = 0.0;
  InitialConditions initialCondition = ReadFromSettings;
  bool storeInitialState = true;
  unsigned int rngSeed = 0;  // kept for the interface; increments come from (SEED, id, step)
};

}  // namespace pinco

#endif
