#include "hip_solver.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <iostream>
#include <map>
#include <stdexcept>
#include <tuple>

namespace pinco {

namespace {

const uint32_t kTagU = 2, kTagOmega = 3, kTagF = 4, kTagBond = 7;

void to_matrix(const NoisyMatrix& m, twmc_matrix& out, std::vector<std::vector<double>>& storage) {
  std::vector<double> b(2 * m.size()), a(2 * m.size());
  for (size_t i = 0; i < m.size(); ++i) {
    b[2 * i] = m.GenerateNoNoise()[i].real();
    b[2 * i + 1] = m.GenerateNoNoise()[i].imag();
    a[2 * i] = m.NoiseAmplitude()[i].real();
    a[2 * i + 1] = m.NoiseAmplitude()[i].imag();
  }
  storage.push_back(std::move(b));
  out.base = storage.back().data();
  storage.push_back(std::move(a));
  out.amp = storage.back().data();
  out.noise_type = m.GetNoiseType() == NoisyMatrix::Gaussian  ? TWMC_NOISE_GAUSSIAN
                   : m.GetNoiseType() == NoisyMatrix::Uniform ? TWMC_NOISE_UNIFORM
                                                              : TWMC_NOISE_NONE;
  if (m.GetNoiseType() == NoisyMatrix::Lorentzian || m.GetNoiseType() == NoisyMatrix::Poissonian)
    std::cerr << "warning: Lorentzian/Poissonian disorder is not generated (as in the reference)" << std::endl;
}

void check(int rc, twmc_ctx* ctx, const char* what) {
  if (rc != TWMC_OK) {
    const char* e = twmc_last_error(ctx);
    throw std::runtime_error(std::string(what) + ": " + (e ? e : "") + " (code " + std::to_string(rc) + ")");
  }
}

}  // namespace

std::vector<double> LiebCoupling(size_t ny, double J_AB, double J_BC, bool pbc, double sigma, uint64_t seed) {
  const size_t n = 3 * ny;
  std::vector<double> C(n * n, 0.0);
  // bonds in order: per cell AB, BC, then the inter-cell C-B bond; the PBC bond last
  std::vector<std::tuple<size_t, size_t, double>> bonds;
  for (size_t i = 0; i < ny; ++i) {
    bonds.emplace_back(3 * i, 3 * i + 1, J_AB);
    bonds.emplace_back(3 * i + 1, 3 * i + 2, J_BC);
    if (i == ny - 1) {
      if (pbc) bonds.emplace_back(3 * i + 2, 1, J_BC);
      break;
    }
    bonds.emplace_back(3 * i + 2, 3 * (i + 1) + 1, J_BC);
  }
  std::vector<double> noise(2 * bonds.size(), 0.0);
  if (sigma != 0.0) {
    twmc_matrix m{};
    m.value[0] = 0.0;
    m.noise_type = TWMC_NOISE_GAUSSIAN;
    m.amp_value[0] = sigma;
    twmc_noisy_matrix(&m, bonds.size(), seed, 0, kTagBond, noise.data());
  }
  for (size_t b = 0; b < bonds.size(); ++b) {
    size_t i, j;
    double J;
    std::tie(i, j, J) = bonds[b];
    double v = J + noise[2 * b];
    C[i * n + j] += v;
    C[j * n + i] += v;
  }
  return C;
}

TWMCHipSolver::TWMCHipSolver() { nTasksToRequest = 4096; }

TWMCHipSolver::~TWMCHipSolver() {
  std::lock_guard<std::mutex> lk(ctx_mu_);
  if (ctx_) twmc_destroy(ctx_);
  ctx_ = nullptr;
}

void TWMCHipSolver::SuggestBatchSize() {
  if (lastSharedSystemData) nTasksToRequest = lastSharedSystemData->gpu_batch;
}

void TWMCHipSolver::FillLattice(const TWMCSystemData& d, twmc_params& p, std::vector<std::vector<double>>& storage) {
  if (!d.PBC) throw std::runtime_error("TWMCBase needs periodic boundaries (PBC = true)");
  p.hopping = TWMC_HOP_FFT;
  p.onsite = TWMC_ONSITE_OMEGA;
  p.J = d.J_val;
  to_matrix(*d.omega, p.omega, storage);
}

void TWMCLiebHipSolver::FillLattice(const TWMCSystemData& d, twmc_params& p,
                                    std::vector<std::vector<double>>& storage) {
  if (d.dimension == TWMCSystemData::D2) throw std::runtime_error("2D Not Supported for LIEB");
  p.hopping = TWMC_HOP_MATRIX;
  p.onsite = TWMC_ONSITE_DETUNING_MINUS_E;
  p.detuning = d.detuning;
  storage.push_back(LiebCoupling(d.ny, d.J_AB_val, d.J_BC_val, d.PBC, d.J_sigma, d.disorder_seed));
  p.coupling = storage.back().data();
  to_matrix(*d.E, p.omega, storage);
}

// TWMCBaseSolver::Setup (TWMCBaseSolver.cpp:24-92): one context per system
void TWMCHipSolver::Setup() {
  const TWMCSystemData& d = *lastSharedSystemData;
  {
    std::lock_guard<std::mutex> lk(ctx_mu_);
    if (ctx_) twmc_destroy(ctx_);
    ctx_ = nullptr;
  }
  twmc_params p{};
  std::vector<std::vector<double>> storage;
  p.nx = (int)d.nx;
  p.ny = (int)(d.ny * d.cellSz);
  p.precision = d.precision;
  p.kernel = TWMC_KERNEL_AUTO;
  p.dt = d.dt;
  p.gamma = d.gamma_val;
  p.seed = d.seed;
  p.frame_steps = d.frame_steps;
  FillLattice(d, p, storage);
  to_matrix(*d.U, p.U, storage);
  to_matrix(*d.F, p.F, storage);
  to_matrix(*d.beta_init_val, p.beta_init, storage);
  to_matrix(*d.beta_init_sigma_val, p.beta_sigma, storage);
  std::vector<double> ft_times, ft_values;
  if (d.F->HasTimeDependence()) {
    // F(t) before every step (TWMCLiebSolver.cpp:252-256)
    d.F->TimeTable(ft_times, ft_values);
    p.F_n_times = ft_times.size();
    p.F_times = ft_times.data();
    p.F_table = ft_values.data();
  }
  twmc_ctx* created = nullptr;
  check(twmc_create(&p, gpu_, &created), nullptr, "twmc_create");
  {
    std::lock_guard<std::mutex> lk(ctx_mu_);
    ctx_ = created;
  }
  nTasksToRequest = d.gpu_batch;
}

std::vector<std::unique_ptr<TaskResults>> TWMCHipSolver::Compute(
    const std::vector<std::unique_ptr<TaskData>>& tasks) {
  std::vector<std::unique_ptr<TaskResults>> all(tasks.size());
  // group by everything a batch shares
  std::map<std::tuple<const TWMCSystemData*, double, double, int, bool>, std::vector<size_t>> groups;
  for (size_t k = 0; k < tasks.size(); ++k) {
    auto* t = dynamic_cast<TWMCTaskData*>(tasks[k].get());
    if (!t) throw std::runtime_error("TWMCHipSolver received a task that is not a TWMCTaskData");
    groups[{t->systemData.get(), t->t_start, t->t_end, (int)t->initialCondition, t->storeInitialState}].push_back(k);
  }
  call_total_ = tasks.size();
  call_done_ = 0;
  for (auto& g : groups) {
    const std::vector<size_t>& idx = g.second;
    auto* first = static_cast<TWMCTaskData*>(tasks[idx[0]].get());
    if (first->systemData.get() != lastSharedSystemData) {
      lastSharedSystemData = first->systemData.get();
      Setup();
    }
    const TWMCSystemData& d = *lastSharedSystemData;
    const size_t n = d.nx * d.ny * d.cellSz;
    std::vector<uint64_t> ids(idx.size());
    std::vector<double> prev;
    for (size_t j = 0; j < idx.size(); ++j) ids[j] = static_cast<TWMCTaskData*>(tasks[idx[j]].get())->id;
    if (first->initialCondition == TWMCTaskData::ReadFromPreviousData) {
      // last stored frame of the previous segment (TWMCBaseSolver.cpp:176-183)
      prev.resize(2 * n * idx.size());
      for (size_t j = 0; j < idx.size(); ++j) {
        auto* t = static_cast<TWMCTaskData*>(tasks[idx[j]].get());
        const TaskResults::Dataset* tr = t->prevData ? t->prevData->GetDataset(kDatasetTraj) : nullptr;
        if (!tr || tr->bytes.size() < n * 16)
          throw std::runtime_error("continuation task without a stored trajectory");
        std::memcpy(&prev[2 * n * j], tr->bytes.data() + tr->bytes.size() - n * 16, n * 16);
      }
    }
    twmc_batch_desc desc{};
    desc.traj_ids = ids.data();
    desc.n_traj = ids.size();
    desc.init_mode = (int)first->initialCondition;
    desc.store_initial = first->storeInitialState ? 1 : 0;
    desc.beta_prev = prev.empty() ? nullptr : prev.data();
    // the Wiener stream continues where the previous segment stopped
    desc.step0 = (uint64_t)std::llround((first->t_start - d.t_start) / d.dt);
    if (first->initialCondition != TWMCTaskData::ReadFromPreviousData) desc.step0 = 0;
    desc.t_start = first->t_start;
    desc.t_end = first->t_end;
    // store = traj writes the reference's datasets only (TWMCResults.cpp:26-38:
    // traj and the disorder realisations); observables and both add the
    // per-trajectory site means mean_n2 / mean_n4
    const bool want_traj = d.store != "observables";
    const bool want_means = d.store != "traj";
    desc.outputs = (uint32_t)TWMC_OUT_OBS | (want_traj ? (uint32_t)TWMC_OUT_TRAJ : 0u);
    group_size_ = ids.size();
    check(twmc_batch_begin(ctx_, &desc), ctx_, "twmc_batch_begin");
    twmc_batch_info info{};
    check(twmc_batch_info_get(ctx_, &info), ctx_, "twmc_batch_info_get");
    if (info.truncated)
      std::cerr << " error: was supposed to have " << (info.n_frames + 1) << " frames but got " << info.n_frames
                << " frames; simulation ended at " << info.t_final << std::endl;
    check(twmc_batch_advance(ctx_, info.n_segments), ctx_, "twmc_batch_advance");
    const size_t nf = info.n_frames, B = ids.size();
    std::vector<double> traj(want_traj ? 2 * B * nf * n : 0), m2(B * nf), m4(B * nf);
    std::vector<double> sn(nf * n), sn2(nf * n), snk(nf * n), sb(2 * nf * n);
    twmc_output out{};
    out.traj = want_traj ? traj.data() : nullptr;
    out.traj_m2 = m2.data();
    out.traj_m4 = m4.data();
    out.sum_n = sn.data();
    out.sum_n2 = sn2.data();
    out.sum_nk = snk.data();
    out.sum_beta = sb.data();
    check(twmc_batch_read(ctx_, &out), ctx_, "twmc_batch_read");
    d.sums->Add(B, nf, n, sn, sn2, snk, sb);

    for (size_t j = 0; j < B; ++j) {
      auto* t = static_cast<TWMCTaskData*>(tasks[idx[j]].get());
      auto res = std::make_unique<TaskResults>();
      res->SetId(t->id);
      std::vector<size_t> dims = {d.nx, d.ny};
      if (d.cellSz != 1) dims.push_back(d.cellSz);
      if (want_traj) res->AddDataset(kDatasetTraj, kFormatComplex128, nf, dims, &traj[2 * n * nf * j], nf * n * 16);
      if (want_means) {
        res->AddDataset(kDatasetM2, kFormatFloat64, nf, {}, &m2[nf * j], nf * 8);
        res->AddDataset(kDatasetM4, kFormatFloat64, nf, {}, &m4[nf * j], nf * 8);
      }
      // disorder realisations, regenerated on the host from the same Philox draws
      auto add_realisation = [&](const NoisyMatrix& m, const char* name, uint32_t tag) {
        if (!m.HasNoise()) return;
        std::vector<cd> v = m.Generate(d.seed, t->id, tag);
        res->AddDataset(name, kFormatComplex128, 1, dims, v.data(), v.size() * 16);
      };
      add_realisation(*d.U, kDatasetU, kTagU);
      add_realisation(*d.F, kDatasetF, kTagF);
      add_realisation(d.latticeName == "lieb" ? *d.E : *d.omega, kDatasetOmega, kTagOmega);
      res->extraData = {t->t_start, info.t_final};
      all[idx[j]] = std::move(res);
    }
    call_done_ += B;
    group_size_ = 0;
  }
  return all;
}

// Fraction of the current Compute call done: finished groups plus the
// device's progress through the group it runs now (twmc_progress), the
// counterpart of TWMCBaseSolver.cpp:293-295 for a batch.
float TWMCHipSolver::ApproximateComputationProgress() const {
  const uint64_t tot = call_total_.load();
  if (tot == 0) return 0.f;
  double done = double(call_done_.load());
  const uint64_t g = group_size_.load();
  if (g) {
    std::lock_guard<std::mutex> lk(ctx_mu_);
    if (ctx_) done += double(g) * double(twmc_progress(ctx_));
  }
  return float(std::min(1.0, done / double(tot)));
}

namespace {
SolverGPUFactory::Registrator<TWMCHipSolver> reg_base_gpu("TWMCBase");
SolverGPUFactory::Registrator<TWMCLiebHipSolver> reg_lieb_gpu("TWMCLieb");
SolverFactory::Registrator<TWMCHipSolver> reg_base("TWMCBase_HIP");
SolverFactory::Registrator<TWMCLiebHipSolver> reg_lieb("TWMCLieb_HIP");
}  // namespace

}  // namespace pinco
