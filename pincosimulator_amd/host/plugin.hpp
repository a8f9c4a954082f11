// plugin.hpp -- the engine's plugin boundary: task/result data model, the
// Solver and SolverGPU interfaces and string-keyed factories.
//
// Same contract as the reference's interfaces (src/Base/Solver.hpp:16-33,
// src/Base/SolverGPU.hpp:18-31, src/Base/TaskData.hpp, src/Base/TaskResults.hpp:30-106,
// src/Libraries/TStaticFactory.h:30-150): a worker owns one Solver, calls
// Compute with a batch of tasks (nTasksToRequest of them) and receives one
// result per task.  Differences: results hold their datasets as named byte
// buffers, and a factory lookup of an unknown name can be made to fail loudly
// (makeOrThrow) instead of silently returning nullptr.
#ifndef PINCO_PLUGIN_HPP
#define PINCO_PLUGIN_HPP

#include <array>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace pinco {

class TaskData {
 public:
  virtual ~TaskData() = default;
};

// Storage format codes written to the register (TWMCResults.cpp:35-41).
enum DatasetFormat : unsigned char { kFormatFloat64 = 11, kFormatComplex128 = 22 };

class TaskResults : public TaskData {
 public:
  struct Dataset {
    std::string name;
    unsigned char format = kFormatComplex128;
    size_t frames = 0;            // number of stored frames ("elements")
    std::vector<size_t> dims;     // dimensions of one frame
    std::vector<unsigned char> bytes;
  };

  size_t GetId() const { return id_; }
  void SetId(size_t id) { id_ = id; }

  void AddDataset(const std::string& name, unsigned char format, size_t frames, std::vector<size_t> dims,
                  const void* data, size_t nbytes) {
    Dataset d;
    d.name = name;
    d.format = format;
    d.frames = frames;
    d.dims = std::move(dims);
    d.bytes.resize(nbytes);
    if (nbytes) std::memcpy(d.bytes.data(), data, nbytes);
    for (auto& e : datasets_)
      if (e.name == name) {
        e = std::move(d);
        return;
      }
    datasets_.push_back(std::move(d));
  }
  const Dataset* GetDataset(const std::string& name) const {
    for (auto& e : datasets_)
      if (e.name == name) return &e;
    return nullptr;
  }
  size_t NumberOfDataSets() const { return datasets_.size(); }
  const Dataset& DatasetAt(size_t i) const { return datasets_.at(i); }
  std::vector<std::string> NamesOfDatasets() const {
    std::vector<std::string> n;
    for (auto& e : datasets_) n.push_back(e.name);
    return n;
  }
  // {t_start, t_final} of the trajectory segment (TWMCBaseSolver.cpp:278-279)
  std::array<double, 2> extraData{{0.0, 0.0}};

 private:
  size_t id_ = 0;
  std::vector<Dataset> datasets_;
};

class Solver {
 public:
  virtual ~Solver() = default;
  virtual void Setup() {}
  virtual std::vector<std::unique_ptr<TaskResults>> Compute(
      const std::vector<std::unique_ptr<TaskData>>& tasks) = 0;
  virtual float ApproximateComputationProgress() const { return 0.f; }
  size_t nTasksToRequest = 1;
};

class SolverGPU : public Solver {
 public:
  virtual void SetupGPU(size_t gpuId) { (void)gpuId; }
  virtual void SuggestBatchSize() {}
};

template <class Base>
class Factory {
 public:
  using Maker = std::function<Base*()>;
  static std::map<std::string, Maker>& registry() {
    static std::map<std::string, Maker>* r = new std::map<std::string, Maker>();
    return *r;
  }
  static bool Register(const std::string& name, Maker m) {
    return registry().emplace(name, std::move(m)).second;
  }
  // nullptr for an unknown name, as the reference's factory does
  static std::unique_ptr<Base> makeUniqueNewInstance(const std::string& name) {
    auto it = registry().find(name);
    if (it == registry().end()) return nullptr;
    return std::unique_ptr<Base>(it->second());
  }
  static std::unique_ptr<Base> makeOrThrow(const std::string& name) {
    auto p = makeUniqueNewInstance(name);
    if (!p) {
      std::string known;
      for (auto& kv : registry()) known += " " + kv.first;
      throw std::runtime_error("no solver registered as '" + name + "' (known:" + known + ")");
    }
    return p;
  }
  static std::vector<std::string> getKeys() {
    std::vector<std::string> k;
    for (auto& kv : registry()) k.push_back(kv.first);
    return k;
  }
  template <class T>
  struct Registrator {
    explicit Registrator(const std::string& name) {
      Register(name, [] { return static_cast<Base*>(new T()); });
    }
  };
};

using SolverFactory = Factory<Solver>;
using SolverGPUFactory = Factory<SolverGPU>;

}  // namespace pinco

#endif
