#include "manager.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <iostream>

#include "task_processor.hpp"

namespace pinco {

TWMCManager::TWMCManager(const Settings& s) : settings_(s) {
  sys_ = std::make_shared<TWMCSystemData>(s);
  if (s.getString("Manager") == "TWMCThread") sys_->PBC = true;
  if (sys_->latticeName == "lieb") solverName_ = "TWMCLieb";
  else if (sys_->PBC) {
    if (sys_->F->HasTimeDependence()) {
      // the reference selects TWMCThermo here, a solver it does not build
      std::cerr << "ERROR: No solver available for a periodic lattice with time-dependent F." << std::endl;
      throw NoEligibleSolver();
    }
    solverName_ = "TWMCBase";
  } else {
    std::cerr << "ERROR: No solver available." << std::endl;
    throw NoEligibleSolver();
  }
}

void TWMCManager::Setup() {
  store_ = std::make_unique<PincoFormatDataStore>(settings_.GetOutputFolder(),
                                                  settings_.getDouble("chunk_size", 1.5));
  // the solver name must be registered before any worker starts
  (void)SolverGPUFactory::makeOrThrow(solverName_);
  store_->SaveFile("_t.dat", sys_->GetStoredTimes());
  if (sys_->F->HasTimeDependence()) store_->SaveFile("_F_t.dat", sys_->GetStoredVariableEvolution(*sys_->F));
}

// TWMCManager::DispatchTasks (TWMCManager.cpp:133-203)
std::vector<std::unique_ptr<TaskData>> TWMCManager::DispatchTasks() {
  std::vector<std::unique_ptr<TaskData>> tasks;
  std::set<size_t> used = store_ ? store_->UsedIds() : std::set<size_t>();
  size_t n_traj = (size_t)settings_.getInt("n_traj", 0);
  if (used.size() < n_traj) {
    size_t maxId = 0;
    for (size_t u : used) maxId = std::max(maxId, u);
    size_t todo = n_traj - used.size();
    for (size_t i = 0; i < todo; ++i) {
      auto t = std::make_unique<TWMCTaskData>();
      t->systemData = sys_;
      t->t_start = sys_->t_start;
      t->t_end = sys_->t_end;
      t->initialCondition = TWMCTaskData::ReadFromSettings;
      t->storeInitialState = true;
      t->id = ++maxId;  // ids continue after the largest stored one (1, 2, ... on a new run)
      t->rngSeed = (unsigned)t->id;
      tasks.push_back(std::move(t));
    }
  }
  double tEnd = settings_.getDouble("t_end");
  for (size_t id : used) {
    std::unique_ptr<TaskResults> old = store_->LoadEndFrame(id);
    if (!old) continue;
    if (old->extraData[1] + sys_->dt_obs < tEnd) {
      auto t = std::make_unique<TWMCTaskData>();
      t->systemData = sys_;
      t->t_start = old->extraData[1];
      t->t_end = tEnd;
      t->initialCondition = TWMCTaskData::ReadFromPreviousData;
      t->storeInitialState = false;
      t->id = id;
      t->prevData = std::move(old);
      tasks.push_back(std::move(t));
    }
  }
  return tasks;
}

// ManagerLoop (TWMCManager.cpp:113-131) with the reference's processor
// (ThreadedTaskProcessor.cpp:66-101): `gpu_workers` worker threads on GPU
// `GPU_id`, each owning one GPU solver (its own context and stream, so one
// worker's host-side work -- task grouping, result assembly, archive
// formatting -- overlaps another's kernels), a saver thread writing the
// archive and a reporter thread printing progress every
// `progress_interval` seconds.
size_t TWMCManager::ManagerLoop() {
  std::vector<std::unique_ptr<TaskData>> tasks = DispatchTasks();
  std::cout << "Enqueued " << tasks.size() << " tasks. " << std::endl;
  const size_t n_workers = (size_t)std::max<int64_t>(1, settings_.getInt("gpu_workers", 1));
  const size_t gpu = (size_t)settings_.getInt("GPU_id", 0);
  const size_t batch = std::max<size_t>(1, sys_->gpu_batch);
  const std::string name = solverName_;
  ThreadedTaskProcessor proc(
      [name, gpu, batch](size_t) {
        std::unique_ptr<SolverGPU> s = SolverGPUFactory::makeOrThrow(name);
        s->SetupGPU(gpu);
        s->nTasksToRequest = batch;
        return std::unique_ptr<Solver>(s.release());
      },
      n_workers);
  PincoFormatDataStore* store = store_.get();
  ResultsSaver saver([store](const TaskResults& r) { store->StoreData(r); });
  proc.SetConsumer(&saver);
  proc.Setup();
  proc.EnqueueTasks(std::move(tasks));
  proc.AllProducersHaveBeenTerminated();
  const double every = settings_.getDouble("progress_interval", 10.0);
  ProgressReporter reporter(proc, &saver,
                            std::chrono::milliseconds((int64_t)std::max(1.0, every * 1000.0)), std::cout);
  try {
    proc.Wait();
  } catch (...) {
    reporter.Stop();
    try {
      saver.Finish();  // keep what finished before the failure
    } catch (...) {
    }
    throw;
  }
  reporter.Stop();
  saver.Finish();
  SaveObservables();
  std::cout << "Exited the ManagerLoop()" << std::endl;
  return saver.SavedItems();
}

void TWMCManager::SaveObservables() const {
  EnsembleSums& s = *sys_->sums;
  std::lock_guard<std::mutex> lk(s.mu);
  if (s.n_traj == 0 || s.n_frames == 0) return;
  const double B = double(s.n_traj);
  const size_t n = s.n;
  std::vector<std::vector<double>> rows;
  std::vector<std::vector<double>> nk;
  for (size_t f = 0; f < s.n_frames; ++f) {
    double m2 = 0, m4 = 0;
    std::vector<double> nkrow(n);
    for (size_t i = 0; i < n; ++i) {
      m2 += s.sum_n[f * n + i] / B;
      m4 += s.sum_n2[f * n + i] / B;
      nkrow[i] = s.sum_nk[f * n + i] / B / double(n) - 0.5;
    }
    m2 /= double(n);
    m4 /= double(n);
    double g2 = (m4 - 2 * m2 + 0.5) / ((m2 - 0.5) * (m2 - 0.5));
    rows.push_back({double(f), m2 - 0.5, g2});
    nk.push_back(nkrow);
  }
  store_->SaveFile("_observables.dat", rows);
  store_->SaveFile("_n_k.dat", nk);
}

}  // namespace pinco
