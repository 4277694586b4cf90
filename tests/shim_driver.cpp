// Test driver for include/pmg_parmoon.hpp (built by tests/test_shim.py with g++).
//
//   shim_driver nodevice        constructing the multigrid without a device must throw
//   shim_driver solve N OUT     Oseen N^3 hierarchy (levels 4^3..N^3), FGMRES(30) to 1e-10
//                               through TNSE_MultiGrid / TFgmresIte, one level's Defect and
//                               CellVanka through TNSE_MGLevel4; writes the solution to OUT
//                               and prints "iterations <j> residual <r>"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "pmg_parmoon.hpp"
#include "pmg_problem.h"

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  pmg_params prm;
  pmg_params_default(&prm);
  prm.sc_gmres_restart = 30;
  prm.sc_lin_red_factor_saddle = 1e-10;
  prm.sc_lin_res_norm_min_saddle = 0.0;
  prm.sc_lin_maxit_saddle = 100;
  if (std::strcmp(argv[1], "nodevice") == 0) {
    try {
      pmg_parmoon::TNSE_MultiGrid mg(prm);
      std::printf("constructed\n");
      return 1;
    } catch (const pmg_parmoon::Error &e) {
      std::printf("error %d %s\n", e.status, e.what());
      return 0;
    }
  }
  if (std::strcmp(argv[1], "solve") != 0 || argc < 4) return 2;
  const int n = std::atoi(argv[2]);
  pmgp_config cfg;
  pmgp_config_default(&cfg);
  cfg.cx = cfg.cy = cfg.cz = 4;
  int levels = 1;
  for (int m = 4; m < n; m *= 2) ++levels;
  cfg.n_levels = levels;
  cfg.nu = 1e-2;
  cfg.conv_kind = PMGP_CONV_FOURIER;
  cfg.seed_conv = 1001;
  cfg.rhs_kind = PMGP_RHS_FOURIER;
  cfg.seed_rhs = 1002;
  pmgp_problem *prob = nullptr;
  if (pmgp_create(&cfg, &prob) != 0) return 3;
  try {
    pmg_parmoon::TNSE_MultiGrid mg(prm);
    for (int l = 0; l < levels; ++l) {
      pmg_level_desc d;
      pmgp_level_desc(prob, l, &d);
      mg.AddLevel(d);
    }
    mg.Setup();
    const int nd = pmgp_n_dof(prob, levels - 1);
    std::vector<double> rhs(pmgp_rhs(prob, levels - 1), pmgp_rhs(prob, levels - 1) + nd), sol(nd, 0.0);
    // level operations through the TNSE_MGLevel4 face
    pmg_parmoon::TNSE_MGLevel4 *fine = mg.GetLevel(levels - 1);
    std::vector<double> u(nd, 0.0), d(nd), aux(nd);
    double res0 = 0, res1 = 0;
    fine->Defect(u.data(), rhs.data(), d.data(), res0);
    fine->CellVanka(u.data(), rhs.data(), aux.data(), 0, nullptr, 2, levels);
    fine->Defect(u.data(), rhs.data(), d.data(), res1);
    std::printf("defect %.6e after_sweep %.6e\n", res0, res1);
    pmg_parmoon::TFgmresIte gm(mg);
    const int it = gm.Iterate(sol.data(), rhs.data());
    std::printf("iterations %d residual %.6e converged %d\n", it, gm.FinalResidual(), int(gm.Converged()));
    FILE *f = std::fopen(argv[3], "wb");
    if (!f) return 4;
    std::fwrite(sol.data(), sizeof(double), sol.size(), f);
    std::fclose(f);
  } catch (const pmg_parmoon::Error &e) {
    std::printf("error %d %s\n", e.status, e.what());
    pmgp_destroy(prob);
    return 5;
  }
  pmgp_destroy(prob);
  return 0;
}
