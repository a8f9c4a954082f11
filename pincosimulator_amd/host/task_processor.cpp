#include "task_processor.hpp"

#include <algorithm>
#include <cstdio>
#include <iomanip>
#include <sstream>

namespace pinco {

// ---------------------------------------------------------------- processor

ThreadedTaskProcessor::ThreadedTaskProcessor(SolverMaker make, size_t n_workers)
    : make_(std::move(make)), n_workers_(std::max<size_t>(1, n_workers)) {}

ThreadedTaskProcessor::~ThreadedTaskProcessor() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    closed_ = true;
    abort_ = true;
  }
  cv_.notify_all();
  for (auto& w : workers_)
    if (w->thread.joinable()) w->thread.join();
}

void ThreadedTaskProcessor::Setup() {
  for (size_t i = 0; i < n_workers_; ++i) workers_.push_back(std::make_unique<Worker>());
  for (size_t i = 0; i < n_workers_; ++i) workers_[i]->thread = std::thread(&ThreadedTaskProcessor::WorkerLoop, this, i);
}

void ThreadedTaskProcessor::EnqueueTasks(std::vector<std::unique_ptr<TaskData>>&& tasks) {
  {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& t : tasks) queue_.push_back(std::move(t));
    enqueued_ += tasks.size();
  }
  tasks.clear();
  cv_.notify_all();
}

void ThreadedTaskProcessor::AllProducersHaveBeenTerminated() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    closed_ = true;
  }
  cv_.notify_all();
}

// GetDispatchedTasks (ThreadedTaskProcessor.cpp:128-139): up to max_tasks
// tasks; an empty vector means the worker should leave.
std::vector<std::unique_ptr<TaskData>> ThreadedTaskProcessor::GetDispatchedTasks(size_t max_tasks) {
  std::vector<std::unique_ptr<TaskData>> tasks;
  std::unique_lock<std::mutex> lk(mu_);
  cv_.wait(lk, [&] { return abort_ || closed_ || !queue_.empty(); });
  if (abort_) return tasks;
  while (!queue_.empty() && tasks.size() < std::max<size_t>(1, max_tasks)) {
    tasks.push_back(std::move(queue_.front()));
    queue_.pop_front();
  }
  return tasks;
}

// WorkerLoop (WorkerThread.cpp:33-69): the solver is built, used and
// destroyed in this thread.
void ThreadedTaskProcessor::WorkerLoop(size_t id) {
  Worker& w = *workers_[id];
  try {
    std::unique_ptr<Solver> s = make_(id);
    if (!s) throw std::runtime_error("worker " + std::to_string(id) + ": no solver");
    Solver* solver = s.get();
    {
      std::lock_guard<std::mutex> lk(mu_);
      w.solver = std::move(s);
    }
    for (;;) {
      std::vector<std::unique_ptr<TaskData>> tasks = GetDispatchedTasks(solver->nTasksToRequest);
      if (tasks.empty()) break;
      w.in_flight.store(tasks.size());
      std::vector<std::unique_ptr<TaskResults>> results = solver->Compute(tasks);
      if (results.size() != tasks.size())
        throw std::runtime_error("worker " + std::to_string(id) + ": solver returned " +
                                 std::to_string(results.size()) + " results for " + std::to_string(tasks.size()) +
                                 " tasks");
      w.in_flight.store(0);
      w.completed += results.size();
      completed_ += results.size();
      if (consumer_) consumer_->EnqueueResults(std::move(results));
    }
  } catch (...) {
    std::lock_guard<std::mutex> lk(mu_);
    if (!error_) error_ = std::current_exception();
    abort_ = true;
    cv_.notify_all();
  }
  w.in_flight.store(0);
  std::unique_ptr<Solver> dead;
  {
    std::lock_guard<std::mutex> lk(mu_);
    dead = std::move(w.solver);  // destroyed outside the lock, in this thread
  }
}

void ThreadedTaskProcessor::Wait() {
  for (auto& w : workers_)
    if (w->thread.joinable()) w->thread.join();
  std::exception_ptr e;
  {
    std::lock_guard<std::mutex> lk(mu_);
    e = error_;
  }
  if (e) std::rethrow_exception(e);
}

std::vector<size_t> ThreadedTaskProcessor::CompletedPerWorker() const {
  std::vector<size_t> v;
  for (auto& w : workers_) v.push_back(w->completed.load());
  return v;
}

float ThreadedTaskProcessor::Progress() const {
  const size_t total = enqueued_.load();
  if (total == 0) return 0.f;
  double done = double(completed_.load());
  {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& w : workers_) {
      const size_t f = w->in_flight.load();
      if (f && w->solver) done += double(f) * double(std::clamp(w->solver->ApproximateComputationProgress(), 0.f, 1.f));
    }
  }
  return float(std::min(1.0, done / double(total)));
}

// ---------------------------------------------------------------- saver

ResultsSaver::ResultsSaver(StoreFn store) : store_(std::move(store)) {
  thread_ = std::thread(&ResultsSaver::Loop, this);
}

ResultsSaver::~ResultsSaver() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    finished_ = true;
  }
  cv_.notify_all();
  if (thread_.joinable()) thread_.join();
}

void ResultsSaver::EnqueueResults(std::vector<std::unique_ptr<TaskResults>>&& results) {
  {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& r : results) queue_.push_back(std::move(r));
  }
  results.clear();
  cv_.notify_all();
}

// IOThreadUpdate (ResultsSaver.cpp:34-46): one result at a time, the store
// call outside the lock so that workers never wait for the disk.
void ResultsSaver::Loop() {
  for (;;) {
    std::unique_ptr<TaskResults> r;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return finished_ || !queue_.empty(); });
      if (queue_.empty()) return;  // finished and drained
      r = std::move(queue_.front());
      queue_.pop_front();
    }
    try {
      store_(*r);
      ++saved_;
    } catch (...) {
      std::lock_guard<std::mutex> lk(mu_);
      if (!error_) error_ = std::current_exception();
    }
  }
}

void ResultsSaver::Finish() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    finished_ = true;
  }
  cv_.notify_all();
  if (thread_.joinable()) thread_.join();
  std::exception_ptr e;
  {
    std::lock_guard<std::mutex> lk(mu_);
    e = error_;
  }
  if (e) std::rethrow_exception(e);
}

// ---------------------------------------------------------------- reporter

ProgressReporter::ProgressReporter(const ThreadedTaskProcessor& p, const ResultsSaver* saver,
                                   std::chrono::milliseconds interval, std::ostream& out)
    : p_(p), saver_(saver), interval_(interval), out_(out), t0_(std::chrono::steady_clock::now()) {
  thread_ = std::thread([this] {
    std::unique_lock<std::mutex> lk(mu_);
    // a system_clock deadline waits in pthread_cond_timedwait, which thread
    // sanitizers understand (libstdc++'s steady_clock waits use
    // pthread_cond_clockwait, which older ones do not); a clock jump only
    // moves one progress line
    while (!cv_.wait_until(lk, std::chrono::system_clock::now() + interval_, [&] { return stop_; })) {
      lk.unlock();
      Report();
      lk.lock();
    }
  });
}

ProgressReporter::~ProgressReporter() { Stop(); }

void ProgressReporter::Stop() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) return;
    stop_ = true;
  }
  cv_.notify_all();
  if (thread_.joinable()) thread_.join();
  Report();
}

void ProgressReporter::Report() {
  const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
  const float f = p_.Progress();
  std::ostringstream s;
  s << "progress: " << p_.CompletedTasks() << "/" << p_.EnqueuedTasks() << " trajectories computed";
  if (saver_) s << ", " << saver_->SavedItems() << " saved";
  s << ", " << std::fixed << std::setprecision(1) << 100.0 * f << "% after " << el << " s";
  if (f > 0.f && f < 1.f) s << ", about " << el * (1.0 - f) / f << " s left";
  s << "\n";
  out_ << s.str() << std::flush;
  ++lines_;
}

}  // namespace pinco
