// pinco_format_check.cpp -- writes a small PincoFormat archive with the
// project's writer (no GPU): two trajectories with a complex128 "traj" and a
// float64 "mean_n2" dataset, then, as a second run of the same folder would,
// a continuation entry for each.  tests/test_pinco_bytes.py compares every
// file with bytes assembled from the reference's write sequence.
#include <cstdio>
#include <vector>

#include "pinco_format.hpp"

using namespace pinco;

namespace {
std::vector<double> values(size_t id, size_t count, double scale) {
  std::vector<double> v(count);
  for (size_t k = 0; k < count; ++k) v[k] = scale * double(id) + 0.125 * double(k);
  return v;
}

std::unique_ptr<TaskResults> result(size_t id, size_t frames, double t0, double t1) {
  auto r = std::make_unique<TaskResults>();
  r->SetId(id);
  std::vector<double> traj = values(id, frames * 3 * 2 * 2, 100.0);  // frames x {3, 2} complex
  std::vector<double> m2 = values(id, frames, -10.0);
  r->AddDataset("traj", kFormatComplex128, frames, {3, 2}, traj.data(), traj.size() * 8);
  r->AddDataset("mean_n2", kFormatFloat64, frames, {}, m2.data(), m2.size() * 8);
  r->extraData = {t0, t1};
  return r;
}
}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string folder = argv[1];
  {
    PincoFormatDataStore s(folder);
    s.StoreData(*result(1, 3, 0.0, 1.5));
    s.StoreData(*result(2, 3, 0.0, 1.25));
  }
  {
    PincoFormatDataStore s(folder);  // reopened: reads the register back
    s.StoreData(*result(1, 2, 1.5, 2.5));
    s.StoreData(*result(2, 2, 1.25, 2.5));
  }
  std::printf("written\n");
  return 0;
}
