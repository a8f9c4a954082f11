// noisy_matrix.hpp -- site-resolved complex parameter with optional
// per-trajectory disorder and optional piecewise-linear time dependence.
//
// Semantics of the reference's NoisyMatrix (src/Base/NoisyMatrix.cpp):
//   Generate = base + noise (118-129); Gaussian noise N(0,1) amp.re + i N(0,1)
//   amp.im, uniform noise U(0,1) amp.re + i U(0,1) amp.im (131-160, 43-47);
//   file values are column-major, real (N) or interleaved complex (2N)
//   (src/Base/Utils/EigenUtils.cpp:61-96); time tables are filled across NaN
//   gaps by linear interpolation and evaluated piecewise-linearly (164-264).
// The random draws come from the engine's Philox stream (tag per parameter)
// instead of std::mt19937, so host and GPU realisations coincide.
#ifndef PINCO_NOISY_MATRIX_HPP
#define PINCO_NOISY_MATRIX_HPP

#include <complex>
#include <cstdint>
#include <map>
#include <string>
#include <tuple>
#include <vector>

namespace pinco {

using cd = std::complex<double>;

class NoisyMatrix {
 public:
  enum NoiseType { None = 0, Uniform = 1, Gaussian = 2, Lorentzian = 3, Poissonian = 4 };

  NoisyMatrix(size_t nx, size_t ny, size_t cellSize = 1);

  void SetValue(cd v);
  // column-major values: N real, or 2N interleaved complex
  bool SetValue(const std::vector<double>& values);
  void SetNoiseType(const std::string& type);
  NoiseType GetNoiseType() const { return noiseType_; }
  bool HasNoise() const { return noiseType_ != None; }
  size_t GetNoiseVariables() const { return noiseType_ == None ? 0 : 1; }
  void SetNoiseVal(size_t id, cd v);
  bool SetNoiseVal(size_t id, const std::vector<double>& values);
  void SetTemporalValue(const std::map<double, std::vector<double>>& data);
  bool HasTimeDependence() const { return timeDep_; }
  bool IsSet() const { return setup_; }

  size_t rows() const { return nx_; }
  size_t cols() const { return ny_; }
  size_t size() const { return nx_ * ny_; }

  const std::vector<cd>& GenerateNoNoise() const { return base_; }
  const std::vector<cd>& NoiseAmplitude() const { return amp_; }
  // base + disorder for trajectory traj (Philox tag identifies the parameter)
  std::vector<cd> Generate(uint64_t seed, uint64_t traj, uint32_t tag) const;
  // value at time t and the table interval used (GetAtTimeWithSuggestion)
  std::tuple<size_t, std::vector<cd>> GetAtTimeWithSuggestion(double t, size_t suggested = 0) const;
  std::vector<cd> GetAtTime(double t) const { return std::get<1>(GetAtTimeWithSuggestion(t)); }
  const std::vector<double>& Times() const { return times_; }
  // the filled table without the hold-after-last sentinel: times (n) and
  // values (n x size() complex, interleaved re/im)
  void TimeTable(std::vector<double>& times, std::vector<double>& values) const;
  bool IsUniform() const;

 private:
  size_t nx_, ny_;
  bool setup_ = false;
  bool timeDep_ = false;
  NoiseType noiseType_ = None;
  std::vector<cd> base_, amp_;
  std::vector<double> times_;
  std::vector<std::vector<cd>> mats_, slopes_;
};

}  // namespace pinco

#endif
