#include "twmc_system.hpp"

#include <algorithm>
#include <cmath>
#include <iostream>
#include <stdexcept>

namespace pinco {

void EnsembleSums::Add(size_t traj, size_t frames, size_t sites, const std::vector<double>& n1,
                       const std::vector<double>& n2, const std::vector<double>& nk, const std::vector<double>& b) {
  std::lock_guard<std::mutex> lk(mu);
  if (n_traj == 0) {
    n_frames = frames;
    n = sites;
    sum_n.assign(frames * sites, 0.0);
    sum_n2.assign(frames * sites, 0.0);
    sum_nk.assign(frames * sites, 0.0);
    sum_beta.assign(2 * frames * sites, 0.0);
  }
  if (frames != n_frames || sites != n) return;  // continuation batches with other frame counts are not pooled
  for (size_t i = 0; i < frames * sites; ++i) {
    sum_n[i] += n1[i];
    sum_n2[i] += n2[i];
    sum_nk[i] += nk[i];
  }
  for (size_t i = 0; i < 2 * frames * sites; ++i) sum_beta[i] += b[i];
  n_traj += traj;
}

// TWMCSystemData::TWMCSystemData (src/TWMC/TWMCSystemData.cpp:14-115)
TWMCSystemData::TWMCSystemData(const Settings& s) {
  nx = (size_t)s.getInt("nx");
  ny = (size_t)s.getInt("ny");
  cellSz = (size_t)s.getInt("cellSz", 1);
  latticeName = s.getString("lattice");
  if (latticeName == "lieb") cellSz = 3;
  if (nx == 0 || ny == 0) throw std::runtime_error("nx and ny must be >= 1");
  nxy = nx * ny;
  PBC = s.getBool("PBC", false);
  if (nx == 1 && ny == 1) {
    dimension = D0;
  } else if (nx == 1 || ny == 1) {
    dimension = D1;
    if (ny == 1) {  // an N x 1 lattice is stored as 1 x N
      ny = nx;
      nx = 1;
    }
  } else {
    dimension = D2;
  }
  U = s.GetMatrix("U", nx, ny, cellSz);
  F = s.GetMatrix("F", nx, ny, cellSz);
  if (latticeName == "lieb") {
    E = s.GetMatrix("E", nx, ny, cellSz);
    omega = std::make_unique<NoisyMatrix>(nx, ny, cellSz);
  } else {
    omega = s.GetMatrix("omega", nx, ny, cellSz);
    E = s.contains("E") ? s.GetMatrix("E", nx, ny, cellSz) : std::make_unique<NoisyMatrix>(nx, ny, cellSz);
  }
  detuning = s.getDouble("detuning");
  J_AB_val = s.getDouble("J_AB");
  J_BC_val = s.getDouble("J_BC");
  J_val = s.getDouble("J");
  gamma_val = s.getDouble("gamma");
  beta_init_val = s.GetMatrix("beta_init", nx, ny, cellSz);
  beta_init_sigma_val = s.GetMatrix("beta_init_sigma", nx, ny, cellSz);

  double obs_n_frames = (double)s.getInt("n_frames", 0);
  t_start = s.getDouble("t_start");
  t_end = s.getDouble("t_end");
  frames_freq = s.getDouble("frames_freq");
  frames_period = s.getDouble("frames_period");
  if (frames_freq == 0 && frames_period == 0) {
    if (obs_n_frames == 0) throw std::runtime_error("Must have a non-zero, integer number of frames!");
    frames_freq = obs_n_frames / (t_end - t_start);
  } else if (frames_freq == 0 && frames_period != 0) {
    frames_freq = 1 / frames_period;
  }
  double timestep = s.getDouble("timestep");
  if (!(timestep > 0)) throw std::runtime_error("timestep must be > 0");
  dt = timestep;
  frame_steps = std::max(size_t(1), size_t(ceil(1 / (frames_freq * timestep))));
  dt_obs = dt * frame_steps;
  nFramesTot = size_t(ceil((t_end - t_start) / dt_obs));
  t_end = nFramesTot * dt_obs;  // without t_start, as the reference has it
  size_t n_times = 0;
  for (double t = t_start; t <= t_end; t += timestep) n_times++;
  n_dt = n_times;
  if (dt_obs < dt) {
    dt_obs = dt;
    frame_steps = 1;
  }

  seed = s.GlobalSeed();
  precision = (int)s.getInt("precision", 64);
  store = s.getString("store", "traj");
  J_sigma = s.getComplex("J_Noise_Val_0", {0.0, 0.0}).real();
  disorder_seed = (uint64_t)s.getInt("disorder_seed", (int64_t)seed);
  gpu_batch = (size_t)s.getInt("gpu_batch", 4096);
}

// ComputeNFrames (TWMCSystemData.cpp:117-134)
size_t TWMCSystemData::ComputeNFrames(double t0, double t1) const {
  return size_t(nearbyint((t1 - t0) / dt_obs));
}

// GetStoredTimes (TWMCSystemData.cpp:140-159)
std::vector<double> TWMCSystemData::GetStoredTimes() const {
  size_t frames = nFramesTot + 1;
  std::vector<double> times(frames, 0.0);
  double t = t_start;
  size_t i_step = 0, i_frame = 1;
  times[0] = t_start;
  while (t <= t_end) {
    t += dt;
    i_step++;
    if ((i_step % frame_steps == 0) && i_frame < frames) times[i_frame++] = t;
  }
  return times;
}

std::vector<std::vector<cd>> TWMCSystemData::GetStoredVariableEvolution(const NoisyMatrix& m) const {
  std::vector<double> times = GetStoredTimes();
  std::vector<std::vector<cd>> out;
  for (double t : times) out.push_back(m.GetAtTime(t));
  return out;
}

}  // namespace pinco
