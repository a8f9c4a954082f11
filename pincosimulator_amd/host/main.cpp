// twmc_sim -- command-line driver: `twmc_sim -i run.ini [-o out] [key value ...]`,
// the reference's `sim` (src/main.cpp:21-80) for the TWMC managers, running the
// GPU solvers.  --dry-run parses the input, selects the solver and prints the
// run description as JSON without touching a GPU.
#include <cstdio>
#include <iostream>

#include "manager.hpp"

using namespace pinco;

static void print_json(const TWMCManager& m) {
  const TWMCSystemData& d = *m.System();
  std::printf("{\"solver\": \"%s\", \"lattice\": \"%s\", \"nx\": %zu, \"ny\": %zu, \"cellSz\": %zu, "
              "\"PBC\": %s, \"dimension\": %d, \"dt\": %.17g, \"dt_obs\": %.17g, \"frame_steps\": %zu, "
              "\"nFramesTot\": %zu, \"t_start\": %.17g, \"t_end\": %.17g, \"n_dt\": %zu, \"gamma\": %.17g, "
              "\"J\": %.17g, \"J_AB\": %.17g, \"J_BC\": %.17g, \"detuning\": %.17g, \"seed\": %llu, "
              "\"precision\": %d, \"F_timedep\": %s, \"U_noise\": %d, \"F_noise\": %d, \"omega_noise\": %d, "
              "\"E_noise\": %d, \"U0\": [%.17g, %.17g], \"F0\": [%.17g, %.17g], \"omega0\": [%.17g, %.17g], "
              "\"E0\": [%.17g, %.17g], \"beta_init0\": [%.17g, %.17g], \"beta_sigma0\": [%.17g, %.17g]}\n",
              m.SolverName().c_str(), d.latticeName.c_str(), d.nx, d.ny, d.cellSz, d.PBC ? "true" : "false",
              (int)d.dimension, d.dt, d.dt_obs, d.frame_steps, d.nFramesTot, d.t_start, d.t_end, d.n_dt,
              d.gamma_val, d.J_val, d.J_AB_val, d.J_BC_val, d.detuning, (unsigned long long)d.seed, d.precision,
              d.F->HasTimeDependence() ? "true" : "false", (int)d.U->GetNoiseType(), (int)d.F->GetNoiseType(),
              (int)d.omega->GetNoiseType(), (int)d.E->GetNoiseType(), d.U->GenerateNoNoise()[0].real(),
              d.U->GenerateNoNoise()[0].imag(), d.F->GenerateNoNoise()[0].real(), d.F->GenerateNoNoise()[0].imag(),
              d.omega->GenerateNoNoise()[0].real(), d.omega->GenerateNoNoise()[0].imag(),
              d.E->GenerateNoNoise()[0].real(), d.E->GenerateNoNoise()[0].imag(),
              d.beta_init_val->GenerateNoNoise()[0].real(), d.beta_init_val->GenerateNoNoise()[0].imag(),
              d.beta_init_sigma_val->GenerateNoNoise()[0].real(), d.beta_init_sigma_val->GenerateNoNoise()[0].imag());
}

int main(int argc, char** argv) {
  try {
    Settings settings(argc, argv);
    std::string mgr = settings.getString("Manager", "TWMC");
    if (mgr != "TWMC" && mgr != "TWMCThread") {
      std::cerr << "unknown Manager '" << mgr << "' (this engine provides TWMC)" << std::endl;
      return 2;
    }
    TWMCManager manager(settings);
    if (settings.DryRun()) {
      print_json(manager);
      return 0;
    }
    manager.Setup();
    size_t saved = manager.ManagerLoop();
    std::cout << "Saved " << saved << " results to " << settings.GetOutputFolder() << std::endl;
    return 0;
  } catch (const NoEligibleSolver& e) {
    std::cerr << e.what() << std::endl;
    return 3;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
}
