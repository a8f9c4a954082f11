// task_processor_check.cpp -- the host driver's threads (worker pool, results
// saver, progress reporter) with a stub solver, no GPU.  Built with
// -fsanitize=thread by tests/test_host_driver.py; prints one line per check
// and "ALL OK" at the end, exits non-zero on a failed check.
//
// 1. 4 workers x 1000 tasks, nTasksToRequest 7: every task computed once,
//    every result saved once into a PincoFormat archive with its own bytes,
//    more than one worker did work, the reporter printed while they ran.
// 2. A solver that throws on one task: Wait() rethrows it on this thread,
//    the other workers stop, nothing hangs, and what was finished is saved.
// 3. One worker and the same tasks give the same archive contents per id.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <random>
#include <set>
#include <sstream>
#include <thread>

#include "pinco_format.hpp"
#include "task_processor.hpp"

using namespace pinco;

namespace {

int failures = 0;
void check(bool ok, const std::string& what) {
  std::printf("%s: %s\n", ok ? "ok" : "FAILED", what.c_str());
  if (!ok) ++failures;
}

struct StubTask : TaskData {
  size_t id = 0;
};

// deterministic payload of a trajectory id: 3 frames of 4 complex numbers
std::vector<double> payload(size_t id) {
  std::vector<double> v(24);
  for (size_t k = 0; k < v.size(); ++k) v[k] = double(id) * 1000.0 + double(k) * 0.25;
  return v;
}

class StubSolver : public Solver {
 public:
  StubSolver(size_t batch, size_t fail_id, unsigned seed) : fail_id_(fail_id), rng_(seed) { nTasksToRequest = batch; }
  std::vector<std::unique_ptr<TaskResults>> Compute(const std::vector<std::unique_ptr<TaskData>>& tasks) override {
    std::vector<std::unique_ptr<TaskResults>> out;
    total_ = tasks.size();
    done_ = 0;
    for (auto& t : tasks) {
      auto* st = dynamic_cast<StubTask*>(t.get());
      if (!st) throw std::runtime_error("not a stub task");
      if (st->id == fail_id_) throw std::runtime_error("stub failure on task " + std::to_string(st->id));
      std::this_thread::sleep_for(std::chrono::microseconds(20 + rng_() % 80));
      auto r = std::make_unique<TaskResults>();
      r->SetId(st->id);
      std::vector<double> p = payload(st->id);
      r->AddDataset("traj", kFormatComplex128, 3, {2, 2}, p.data(), p.size() * sizeof(double));
      r->extraData = {0.0, double(st->id)};
      out.push_back(std::move(r));
      ++done_;
    }
    return out;
  }
  float ApproximateComputationProgress() const override {
    const size_t t = total_.load();
    return t ? float(done_.load()) / float(t) : 0.f;
  }

 private:
  size_t fail_id_;
  std::minstd_rand rng_;
  std::atomic<size_t> total_{0}, done_{0};
};

std::vector<std::unique_ptr<TaskData>> make_tasks(size_t n) {
  std::vector<std::unique_ptr<TaskData>> v;
  for (size_t i = 1; i <= n; ++i) {
    auto t = std::make_unique<StubTask>();
    t->id = i;
    v.push_back(std::move(t));
  }
  return v;
}

struct RunResult {
  size_t saved = 0;
  std::vector<size_t> per_worker;
  size_t report_lines = 0;
  std::string error;
};

RunResult run(const std::string& folder, size_t workers, size_t n_tasks, size_t fail_id) {
  std::filesystem::remove_all(folder);
  RunResult rr;
  PincoFormatDataStore store(folder);
  ThreadedTaskProcessor proc(
      [fail_id](size_t w) { return std::unique_ptr<Solver>(new StubSolver(7, fail_id, 1234u + unsigned(w))); },
      workers);
  ResultsSaver saver([&store](const TaskResults& r) { store.StoreData(r); });
  proc.SetConsumer(&saver);
  proc.Setup();
  std::ostringstream log;
  ProgressReporter rep(proc, &saver, std::chrono::milliseconds(2), log);
  // tasks arrive in two pieces while the workers already run
  std::vector<std::unique_ptr<TaskData>> all = make_tasks(n_tasks);
  std::vector<std::unique_ptr<TaskData>> first, second;
  for (size_t i = 0; i < all.size(); ++i) (i < all.size() / 2 ? first : second).push_back(std::move(all[i]));
  proc.EnqueueTasks(std::move(first));
  std::this_thread::sleep_for(std::chrono::milliseconds(5));
  proc.EnqueueTasks(std::move(second));
  proc.AllProducersHaveBeenTerminated();
  try {
    proc.Wait();
  } catch (const std::exception& e) {
    rr.error = e.what();
  }
  rep.Stop();
  saver.Finish();
  rr.saved = saver.SavedItems();
  rr.per_worker = proc.CompletedPerWorker();
  rr.report_lines = rep.LinesPrinted();
  return rr;
}

bool archive_matches(const std::string& folder, size_t n) {
  PincoFormatDataStore store(folder);
  std::set<size_t> ids = store.UsedIds();
  if (ids.size() != n || *ids.begin() != 1 || *ids.rbegin() != n) return false;
  for (size_t id = 1; id <= n; ++id) {
    auto r = store.LoadResult(id, false);
    const TaskResults::Dataset* d = r ? r->GetDataset("traj") : nullptr;
    std::vector<double> p = payload(id);
    if (!d || d->frames != 3 || d->bytes.size() != p.size() * sizeof(double)) return false;
    if (std::memcmp(d->bytes.data(), p.data(), d->bytes.size()) != 0) return false;
    if (r->extraData[1] != double(id)) return false;
  }
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  const std::string root = argc > 1 ? argv[1] : "task_processor_check_out";
  const size_t n = 1000;

  RunResult a = run(root + "/pool4", 4, n, 0);
  size_t sum = 0, busy = 0;
  for (size_t c : a.per_worker) {
    sum += c;
    busy += c > 0;
  }
  check(a.error.empty(), "4 workers: no error (" + a.error + ")");
  check(a.saved == n && sum == n, "4 workers: every task computed and saved once");
  check(busy >= 2, "4 workers: " + std::to_string(busy) + " workers took tasks");
  check(a.report_lines >= 1, "reporter printed " + std::to_string(a.report_lines) + " lines");
  check(archive_matches(root + "/pool4", n), "4 workers: archive holds every id with its own bytes");

  RunResult b = run(root + "/pool1", 1, n, 0);
  check(b.error.empty() && b.saved == n, "1 worker: every task saved");
  check(archive_matches(root + "/pool1", n), "1 worker: same archive contents per id");

  RunResult c = run(root + "/failing", 3, n, 600);
  check(c.error.find("stub failure on task 600") != std::string::npos, "a worker's exception reaches Wait()");
  check(c.saved < n && c.saved > 0, "finished results before the failure were saved (" + std::to_string(c.saved) + ")");

  std::printf(failures ? "FAILURES: %d\n" : "ALL OK\n", failures);
  return failures ? 1 : 0;
}
