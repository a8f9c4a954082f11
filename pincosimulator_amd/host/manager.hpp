// manager.hpp -- selects the solver for a run, creates its tasks and drives
// them through the GPU solver, saving results as they arrive.
//
// Selection and dispatch follow src/TWMC/TWMCManager.cpp:38-61 and :133-203:
// lattice == "lieb" -> TWMCLieb; PBC with a constant F -> TWMCBase; anything
// else has no eligible solver.  One task per trajectory id; ids already in the
// output archive are skipped, and stored trajectories that end before t_end
// get a continuation task (ReadFromPreviousData).  The tasks go through a
// ThreadedTaskProcessor (worker threads, one GPU solver each); results pass
// through a saver thread into the PincoFormat archive (the reference's
// ResultsSaver) and a reporter thread prints the progress.
#ifndef PINCO_MANAGER_HPP
#define PINCO_MANAGER_HPP

#include <memory>
#include <string>
#include <vector>

#include "pinco_format.hpp"
#include "plugin.hpp"
#include "settings.hpp"
#include "twmc_system.hpp"

namespace pinco {

class NoEligibleSolver : public std::runtime_error {
 public:
  NoEligibleSolver() : std::runtime_error("ERROR: No solver available.") {}
};

class TWMCManager {
 public:
  explicit TWMCManager(const Settings& s);
  void Setup();
  // runs every task; returns the number of results saved
  size_t ManagerLoop();
  std::vector<std::unique_ptr<TaskData>> DispatchTasks();
  const std::string& SolverName() const { return solverName_; }
  std::shared_ptr<TWMCSystemData> System() const { return sys_; }
  // writes the ensemble observables of this run (_observables.dat, _n_k.dat)
  void SaveObservables() const;

 private:
  const Settings& settings_;
  std::shared_ptr<TWMCSystemData> sys_;
  std::string solverName_;
  std::unique_ptr<PincoFormatDataStore> store_;
};

}  // namespace pinco

#endif
