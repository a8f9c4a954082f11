// task_processor.hpp -- the caller of the solver boundary: a task queue feeding
// worker threads that each own one Solver, a results saver thread and a
// progress reporter thread.
//
// Same roles as the reference's ThreadedTaskProcessor / WorkerThread
// (src/Base/TaskProcessors/ThreadedTaskProcessor/ThreadedTaskProcessor.cpp:66-101,
// :128-148; WorkerThread.cpp:33-69), ResultsSaver (src/Base/Modules/ResultsSaver.cpp:34-52)
// and ProgressReporter (src/Base/Modules/ProgressReporter.cpp:44-98):
//   * every worker creates its Solver in its own thread and calls Compute only
//     from there (one Solver per thread, as WorkerThread does); a worker asks
//     the queue for up to nTasksToRequest tasks at a time;
//   * finished results go to a consumer (the saver) in the order each worker
//     produces them;
//   * the reporter reads the completed-task count and each worker's
//     ApproximateComputationProgress from its own thread.
// Differences, each deliberate: workers block on a condition variable while
// the queue is empty instead of spinning, and leave once the producer has
// declared the queue closed and it is drained (the reference's workers never
// terminate by themselves, WorkerThread.cpp:65); an exception in a worker
// stops the others after their current batch and is rethrown by Wait() on the
// calling thread (the reference has no error path between threads).
#ifndef PINCO_TASK_PROCESSOR_HPP
#define PINCO_TASK_PROCESSOR_HPP

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <ostream>
#include <string>
#include <thread>
#include <vector>

#include "plugin.hpp"

namespace pinco {

// Receives finished results from the workers (the reference's IResultConsumer).
class ResultConsumer {
 public:
  virtual ~ResultConsumer() = default;
  virtual void EnqueueResults(std::vector<std::unique_ptr<TaskResults>>&& results) = 0;
};

class ThreadedTaskProcessor {
 public:
  // make(worker index) builds that worker's solver; it runs in the worker's thread
  using SolverMaker = std::function<std::unique_ptr<Solver>(size_t worker)>;

  ThreadedTaskProcessor(SolverMaker make, size_t n_workers);
  ~ThreadedTaskProcessor();
  ThreadedTaskProcessor(const ThreadedTaskProcessor&) = delete;
  ThreadedTaskProcessor& operator=(const ThreadedTaskProcessor&) = delete;

  void SetConsumer(ResultConsumer* c) { consumer_ = c; }
  // starts the worker threads
  void Setup();
  void EnqueueTasks(std::vector<std::unique_ptr<TaskData>>&& tasks);
  // no more tasks will come: workers leave once the queue is empty
  void AllProducersHaveBeenTerminated();
  // joins the workers; rethrows the first exception a worker raised
  void Wait();

  size_t NumberOfWorkers() const { return n_workers_; }
  size_t EnqueuedTasks() const { return enqueued_.load(); }
  size_t CompletedTasks() const { return completed_.load(); }
  // tasks each worker has taken from the queue and finished
  std::vector<size_t> CompletedPerWorker() const;
  // completed tasks plus the in-flight batches weighted by their solvers'
  // ApproximateComputationProgress, over the enqueued tasks
  float Progress() const;

 private:
  struct Worker {
    std::thread thread;
    std::unique_ptr<Solver> solver;  // owned and used by the worker thread
    std::atomic<size_t> in_flight{0};
    std::atomic<size_t> completed{0};
  };
  std::vector<std::unique_ptr<TaskData>> GetDispatchedTasks(size_t max_tasks);
  void WorkerLoop(size_t id);

  SolverMaker make_;
  size_t n_workers_;
  ResultConsumer* consumer_ = nullptr;
  std::vector<std::unique_ptr<Worker>> workers_;
  mutable std::mutex mu_;            // queue, closed_, error_, the workers' solver pointers
  std::condition_variable cv_;
  std::deque<std::unique_ptr<TaskData>> queue_;
  bool closed_ = false;
  bool abort_ = false;
  std::exception_ptr error_;
  std::atomic<size_t> enqueued_{0}, completed_{0};
};

// I/O thread that drains results into a store callback (ResultsSaver.cpp:34-52).
class ResultsSaver : public ResultConsumer {
 public:
  using StoreFn = std::function<void(const TaskResults&)>;
  explicit ResultsSaver(StoreFn store);
  ~ResultsSaver() override;
  void EnqueueResults(std::vector<std::unique_ptr<TaskResults>>&& results) override;
  // stores what is queued, then stops the thread; rethrows a store error
  void Finish();
  size_t SavedItems() const { return saved_.load(); }

 private:
  void Loop();
  StoreFn store_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<std::unique_ptr<TaskResults>> queue_;
  bool finished_ = false;
  std::exception_ptr error_;
  std::atomic<size_t> saved_{0};
  std::thread thread_;
};

// Prints the run's progress every `interval` from its own thread
// (ProgressReporter.cpp:44-98): completed / enqueued tasks, the fraction
// including the batches in flight, elapsed time and an estimate of the rest.
class ProgressReporter {
 public:
  ProgressReporter(const ThreadedTaskProcessor& p, const ResultsSaver* saver, std::chrono::milliseconds interval,
                   std::ostream& out);
  ~ProgressReporter();
  void Stop();  // prints a last line and joins
  size_t LinesPrinted() const { return lines_.load(); }

 private:
  void Report();
  const ThreadedTaskProcessor& p_;
  const ResultsSaver* saver_;
  std::chrono::milliseconds interval_;
  std::ostream& out_;
  std::chrono::steady_clock::time_point t0_;
  std::mutex mu_;
  std::condition_variable cv_;
  bool stop_ = false;
  std::atomic<size_t> lines_{0};
  std::thread thread_;
};

}  // namespace pinco

#endif
