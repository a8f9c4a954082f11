// settings.hpp -- run configuration: the reference's .ini input (TOML) plus
// command-line `key value` overrides.
//
// Same inputs as src/Base/Settings.cpp:32-108 (CLI: -i/--input, -o/--output,
// -a/--append, trailing key value pairs merged over the file) and the matrix
// forms of GetMatrix (Settings.cpp:357-497): scalar, X.real/X.imag, X.data
// (array or scalar), X.file / plain path string (column-major values), X.t or
// X_t (time table), and X.Noise_Type / X_Noise_Type with Noise_Val_0.
// The TOML reader is this project's own and covers what those files use:
// key = value lines, quoted keys, [table] headers (keys become table.key),
// numbers, booleans, strings and flat numeric arrays, # comments.
#ifndef PINCO_SETTINGS_HPP
#define PINCO_SETTINGS_HPP

#include <complex>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace pinco {

class NoisyMatrix;

struct Value {
  enum Kind { Int, Float, Bool, String, Array } kind = Int;
  int64_t i = 0;
  double f = 0.0;
  bool b = false;
  std::string s;
  std::vector<double> arr;
  double number() const { return kind == Int ? double(i) : kind == Bool ? double(b) : f; }
};

class Settings {
 public:
  Settings() = default;
  // parses argv (throws std::runtime_error on bad usage or unreadable input)
  Settings(int argc, char** argv);
  // parses TOML text; basePath resolves relative file references
  static Settings FromString(const std::string& text, const std::string& basePath = ".");
  void Merge(const std::string& text);  // applies `key = value` lines over the current values
  void Set(const std::string& key, const Value& v) { config_[key] = v; }

  bool contains(const std::string& key) const { return config_.count(key) != 0; }
  const Value* find(const std::string& key) const;
  double getDouble(const std::string& key, double def = 0.0) const;
  int64_t getInt(const std::string& key, int64_t def = 0) const;
  bool getBool(const std::string& key, bool def = false) const;
  std::string getString(const std::string& key, const std::string& def = "") const;
  // scalar, or key.real / key.imag
  std::complex<double> getComplex(const std::string& key, std::complex<double> def) const;
  std::vector<double> getArray(const std::string& key) const;
  std::vector<std::string> keys() const;

  // NoisyMatrix of nx x (ny cellSz) values from any of the matrix input forms
  std::unique_ptr<NoisyMatrix> GetMatrix(const std::string& path, size_t nx, size_t ny, size_t cellSz) const;

  // the run seed; the reference adds time(0) (Settings.cpp:349-352), which this
  // engine only does when seed_add_time = true, keeping runs reproducible
  uint64_t GlobalSeed() const;

  std::string GetInputFolder() const { return inputFolder_; }
  std::string GetOutputFolder() const { return outputFolder_; }
  std::string GetRootFolder() const;
  std::string InputPath() const { return inputPath_; }
  bool IsOutputFolderSet() const { return outputSet_; }
  bool DryRun() const { return dryRun_; }

 private:
  std::map<std::string, Value> config_;
  std::string inputPath_, inputFolder_ = ".", outputFolder_ = ".", basePath_ = ".";
  bool outputSet_ = false, dryRun_ = false, subsequentRun_ = false;
};

// parses one TOML value literal
bool ParseValue(const std::string& text, Value& out);
// whitespace-separated numbers (ReadCharFile) and time tables (ReadTemporalCharFile)
std::vector<double> ReadCharFile(const std::string& path);
std::map<double, std::vector<double>> ReadTemporalCharFile(const std::string& path);

}  // namespace pinco

#endif
