// hip_solver.hpp -- the GPU solvers behind the plugin interface.
//
// TWMCHipSolver replaces TWMCBaseSolver (src/TWMC/Solvers/TWMCBaseSolver.cpp,
// uniform hopping through the FFT); TWMCLiebHipSolver replaces TWMCLiebSolver
// (src/TWMC/Solvers/TWMCLiebSolver.cpp, hopping through the Lieb coupling
// matrix).  Both are SolverGPU plugins registered under the reference's
// solver names in SolverGPUFactory ("TWMCBase", "TWMCLieb") and with an _HIP
// suffix in SolverFactory, the pattern of TWMCBaseGPU.hpp:96-100.  Compute
// takes a whole batch (nTasksToRequest trajectories), groups it by
// (t_start, t_end, initial condition, storeInitialState), and runs each
// group through libtwmc_hip.so's C ABI.
#ifndef PINCO_HIP_SOLVER_HPP
#define PINCO_HIP_SOLVER_HPP

#include <atomic>
#include <mutex>
#include <vector>

#include "../../include/twmc_hip.h"
#include "plugin.hpp"
#include "twmc_system.hpp"

namespace pinco {

class TWMCHipSolver : public SolverGPU {
 public:
  TWMCHipSolver();
  ~TWMCHipSolver() override;
  void SetupGPU(size_t gpuId) override { gpu_ = (int)gpuId; }
  void SuggestBatchSize() override;
  void Setup() override;
  std::vector<std::unique_ptr<TaskResults>> Compute(const std::vector<std::unique_ptr<TaskData>>& tasks) override;
  float ApproximateComputationProgress() const override;

 protected:
  // fills the lattice-specific part of the parameters; storage keeps arrays alive
  virtual void FillLattice(const TWMCSystemData& d, twmc_params& p, std::vector<std::vector<double>>& storage);
  const TWMCSystemData* lastSharedSystemData = nullptr;
  // the context is replaced only under ctx_mu_, so that another thread's
  // ApproximateComputationProgress never reads a destroyed one
  twmc_ctx* ctx_ = nullptr;
  mutable std::mutex ctx_mu_;
  int gpu_ = 0;
  // progress of the current Compute call: its trajectories, those finished,
  // and the size of the group on the device now
  std::atomic<uint64_t> call_total_{0}, call_done_{0}, group_size_{0};
};

class TWMCLiebHipSolver : public TWMCHipSolver {
 protected:
  void FillLattice(const TWMCSystemData& d, twmc_params& p, std::vector<std::vector<double>>& storage) override;
};

// Lieb coupling matrix C (3 ny x 3 ny, row-major) of TWMCLiebSolver::Setup
// (TWMCLiebSolver.cpp:52-85), with optional Gaussian bond disorder J_b + sigma N(0,1)
// drawn from Philox (seed, stream 0, tag 7, one draw per bond).
std::vector<double> LiebCoupling(size_t ny, double J_AB, double J_BC, bool pbc, double sigma, uint64_t seed);

}  // namespace pinco

#endif
