// Test driver for include/pmg_parmoon.hpp (built by tests/test_shim.py with g++).
//
//   shim_driver nodevice        constructing the multigrid without a device must throw
//   shim_driver solve N OUT     Oseen N^3 hierarchy (levels 4^3..N^3), FGMRES(30) to 1e-10
//                               through TNSE_MultiGrid / TFgmresIte, one level's Defect and
//                               CellVanka through TNSE_MGLevel4; writes the solution to OUT
//                               and prints "iterations <j> residual <r>"; the same
//                               operations through the reference's signatures (matrix and
//                               space pointers registered with AddLevel) must agree bit for
//                               bit: prints "dropin <max difference>"
#include <algorithm>
#include <cmath>
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
  // stand-ins for ParMooN's matrix and space objects: only their addresses are used
  std::vector<char> objs(size_t(levels) * 4);
  std::vector<TSquareMatrix *> sq(size_t(levels) * 9);
  std::vector<TMatrix *> mt(size_t(levels) * 6);
  auto sqm = [&](int l) { return sq.data() + 9 * l; };
  auto mat = [&](int l) { return mt.data() + 6 * l; };
  auto uspace = [&](int l) { return reinterpret_cast<TFESpace3D *>(&objs[4 * l + 1]); };
  auto pspace = [&](int l) { return reinterpret_cast<TFESpace3D *>(&objs[4 * l + 2]); };
  for (int l = 0; l < levels; ++l) {
    for (int b = 0; b < 9; ++b) sqm(l)[b] = reinterpret_cast<TSquareMatrix *>(&objs[4 * l]);
    for (int b = 0; b < 6; ++b) mat(l)[b] = reinterpret_cast<TMatrix *>(&objs[4 * l + 3]);
  }
  try {
    pmg_parmoon::TNSE_MultiGrid mg(prm);
    for (int l = 0; l < levels; ++l) {
      pmg_level_desc d;
      pmgp_level_desc(prob, l, &d);
      mg.AddLevel(d, sqm(l), mat(l), uspace(l), pspace(l));
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
    {
      // the reference's signatures against the level-index forms
      const int lf = levels - 1, lc = levels - 2;
      const int ndc = pmgp_n_dof(prob, lc);
      double diff = 0;
      auto cmp = [&](const std::vector<double> &a, const std::vector<double> &b) {
        for (size_t i = 0; i < a.size(); ++i) diff = std::max(diff, std::fabs(a[i] - b[i]));
      };
      std::vector<double> x(nd), y1(nd), y2(nd);
      for (int i = 0; i < nd; ++i) x[i] = std::sin(0.37 * i);
      pmg_parmoon::MatVect_NSE4(mg, lf, x.data(), y1.data());
      pmg_parmoon::MatVect_NSE4(sqm(lf), mat(lf), x.data(), y2.data());
      cmp(y1, y2);
      pmg_parmoon::Defect_NSE4(mg, lf, x.data(), rhs.data(), y1.data());
      pmg_parmoon::Defect_NSE4(sqm(lf), mat(lf), x.data(), rhs.data(), y2.data());
      cmp(y1, y2);
      std::vector<double> c1(ndc), c2(ndc), a(nd);
      pmg_parmoon::DefectRestriction(mg, lf, x.data(), c1.data());
      const int nuf = 3 * mg.N_UDOF(lf), nuc = 3 * mg.N_UDOF(lc);
      pmg_parmoon::DefectRestriction(uspace(lc), uspace(lf), 3, c2.data(), x.data(), a.data());
      pmg_parmoon::DefectRestriction(pspace(lc), pspace(lf), 1, c2.data() + nuc, x.data() + nuf, a.data());
      cmp(c1, c2);
      std::vector<double> xc(ndc);
      for (int i = 0; i < ndc; ++i) xc[i] = std::cos(0.21 * i);
      pmg_parmoon::Prolongate(mg, lf, xc.data(), y1.data());
      pmg_parmoon::Prolongate(uspace(lc), uspace(lf), 3, xc.data(), y2.data(), a.data());
      pmg_parmoon::Prolongate(pspace(lc), pspace(lf), 1, xc.data() + nuc, y2.data() + nuf, a.data());
      cmp(y1, y2);
      std::printf("dropin %.3e\n", diff);
    }
    pmg_parmoon::TFgmresIte gm(mg);
    const int it = gm.Iterate(sqm(levels - 1), mat(levels - 1), sol.data(), rhs.data());
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
