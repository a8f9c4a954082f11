#include "noisy_matrix.hpp"

#include <cmath>
#include <iostream>
#include <limits>

#include "../../include/twmc_hip.h"

namespace pinco {

NoisyMatrix::NoisyMatrix(size_t nx, size_t ny, size_t cellSize) : nx_(nx), ny_(ny * cellSize) {
  base_.assign(nx_ * ny_, cd(0.0, 0.0));
  amp_.assign(nx_ * ny_, cd(0.0, 0.0));
}

void NoisyMatrix::SetValue(cd v) {
  base_.assign(size(), v);
  setup_ = true;
}

static bool column_major(const std::vector<double>& values, size_t nx, size_t ny, std::vector<cd>& out) {
  size_t n = nx * ny;
  out.assign(n, cd(0.0, 0.0));
  if (values.size() == n) {
    for (size_t i = 0; i < nx; ++i)
      for (size_t j = 0; j < ny; ++j) out[i * ny + j] = cd(values[j * nx + i], 0.0);
    return true;
  }
  if (values.size() == 2 * n) {
    for (size_t i = 0; i < nx; ++i)
      for (size_t j = 0; j < ny; ++j)
        out[i * ny + j] = cd(values[2 * (j * nx + i)], values[2 * (j * nx + i) + 1]);
    return true;
  }
  std::cerr << "matrix input has " << values.size() << " values; the lattice needs " << n << " (real) or "
            << 2 * n << " (complex)" << std::endl;
  return false;
}

bool NoisyMatrix::SetValue(const std::vector<double>& values) {
  bool ok = column_major(values, nx_, ny_, base_);
  setup_ = ok;
  return ok;
}

void NoisyMatrix::SetNoiseType(const std::string& type) {
  if (type == "Uniform" || type == "u") noiseType_ = Uniform;
  else if (type == "Gaussian" || type == "g") noiseType_ = Gaussian;
  else if (type == "Lorentzian" || type == "l") noiseType_ = Lorentzian;
  else if (type == "Poissonian" || type == "p") noiseType_ = Poissonian;
  else {
    noiseType_ = None;
    if (type != "None")
      std::cerr << "Unrecognized Noise Type " << type << ". Defaulting to: None." << std::endl;
  }
}

void NoisyMatrix::SetNoiseVal(size_t id, cd v) {
  if (id == 0) amp_.assign(size(), v);
}

bool NoisyMatrix::SetNoiseVal(size_t id, const std::vector<double>& values) {
  if (id != 0) return false;
  return column_major(values, nx_, ny_, amp_);
}

std::vector<cd> NoisyMatrix::Generate(uint64_t seed, uint64_t traj, uint32_t tag) const {
  if (noiseType_ != Gaussian && noiseType_ != Uniform) return base_;
  twmc_matrix m{};
  std::vector<double> b(2 * size()), a(2 * size()), out(2 * size());
  for (size_t i = 0; i < size(); ++i) {
    b[2 * i] = base_[i].real();
    b[2 * i + 1] = base_[i].imag();
    a[2 * i] = amp_[i].real();
    a[2 * i + 1] = amp_[i].imag();
  }
  m.base = b.data();
  m.amp = a.data();
  m.noise_type = noiseType_ == Gaussian ? TWMC_NOISE_GAUSSIAN : TWMC_NOISE_UNIFORM;
  twmc_noisy_matrix(&m, size(), seed, traj, tag, out.data());
  std::vector<cd> v(size());
  for (size_t i = 0; i < size(); ++i) v[i] = cd(out[2 * i], out[2 * i + 1]);
  return v;
}

// Time table: NaN entries are filled from the neighbouring known times
// (linear interpolation, or the nearest value at the ends); slopes are
// stored per interval and the last value is held after the last time.
void NoisyMatrix::SetTemporalValue(const std::map<double, std::vector<double>>& data) {
  times_.clear();
  mats_.clear();
  slopes_.clear();
  for (auto& kv : data) {
    std::vector<cd> m(size(), cd(NAN, NAN));
    const auto& v = kv.second;
    if (v.size() == size()) {
      for (size_t i = 0; i < size(); ++i) m[i] = cd(v[i], std::isnan(v[i]) ? NAN : 0.0);
    } else if (v.size() == 2 * size()) {
      for (size_t i = 0; i < size(); ++i) m[i] = cd(v[2 * i], v[2 * i + 1]);
    } else {
      std::cerr << "Error in parsing a time-dependent matrix. Invalid number of elements." << std::endl;
    }
    times_.push_back(kv.first);
    mats_.push_back(m);
  }
  const int nt = (int)times_.size();
  for (size_t e = 0; e < size(); ++e) {
    for (int p = 0; p < nt; ++p) {
      if (!std::isnan(mats_[p][e].real()) && !std::isnan(mats_[p][e].imag())) continue;
      int tp = p - 1;
      while (tp >= 0 && std::isnan(mats_[tp][e].real())) --tp;
      int tn = p + 1;
      while (tn < nt && std::isnan(mats_[tn][e].real())) ++tn;
      bool hp = tp >= 0, hn = tn < nt;
      if (!hp && !hn) {
        std::cerr << "ERROR: The Time-dependent matrix element " << e << " has no known Elements!" << std::endl;
        break;
      } else if (!hn) {
        mats_[p][e] = mats_[tp][e];
      } else if (!hp) {
        mats_[p][e] = mats_[tn][e];
      } else {
        mats_[p][e] = mats_[tp][e] + (mats_[tn][e] - mats_[tp][e]) / (times_[tn] - times_[tp]) * (times_[p] - times_[tp]);
      }
    }
  }
  for (int p = 1; p < nt; ++p) {
    std::vector<cd> s(size());
    for (size_t e = 0; e < size(); ++e) s[e] = (mats_[p][e] - mats_[p - 1][e]) / (times_[p] - times_[p - 1]);
    slopes_.push_back(s);
  }
  slopes_.push_back(std::vector<cd>(size(), cd(0.0, 0.0)));
  slopes_.push_back(std::vector<cd>(size(), cd(0.0, 0.0)));
  if (nt > 0) {
    mats_.push_back(mats_[nt - 1]);
    times_.push_back(std::numeric_limits<double>::max());
    base_ = mats_[0];
    setup_ = true;
  }
  timeDep_ = nt > 0;
}

void NoisyMatrix::TimeTable(std::vector<double>& times, std::vector<double>& values) const {
  times.clear();
  values.clear();
  if (!timeDep_) return;
  const size_t nt = times_.size() - 1;  // the last entry is the hold sentinel
  for (size_t p = 0; p < nt; ++p) {
    times.push_back(times_[p]);
    for (const cd& v : mats_[p]) {
      values.push_back(v.real());
      values.push_back(v.imag());
    }
  }
}

std::tuple<size_t, std::vector<cd>> NoisyMatrix::GetAtTimeWithSuggestion(double t, size_t suggested) const {
  if (!timeDep_) return std::make_tuple(size_t(0), base_);
  if (t <= times_[0]) return std::make_tuple(size_t(0), mats_[0]);
  size_t id;
  if (suggested + 1 < times_.size() && times_[suggested + 1] > t && times_[suggested] <= t) {
    id = suggested;
  } else {
    id = 0;
    while (id + 1 < times_.size() && times_[id + 1] < t) ++id;
  }
  std::vector<cd> v(size());
  for (size_t e = 0; e < size(); ++e) v[e] = mats_[id][e] + cd(t - times_[id], 0.0) * slopes_[id][e];
  return std::make_tuple(id, v);
}

bool NoisyMatrix::IsUniform() const {
  for (size_t i = 1; i < base_.size(); ++i)
    if (base_[i] != base_[0]) return false;
  return !timeDep_;
}

}  // namespace pinco
