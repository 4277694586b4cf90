// pmg_parmoon.hpp - C++ host classes with the method names, argument meaning and call
// order of ParMooN's NSE multigrid (include/FE/NSE_MGLevel.h:57-338,
// include/FE/NSE_MGLevel4.h:48-469, include/FE/NSE_MultiGrid.h:56-302,
// include/FE/MultiGridIte.h:45-69, include/FE/FgmresIte.h:52-125), implemented over the C
// ABI of pmg.h.  Header only; link with libpmg.so.
//
// Differences a caller sees:
//  * levels are described by pmg_level_desc (the reference's CSR arrays, dof maps and
//    prolongation tables) instead of TSquareMatrix3D / TFESpace3D objects;
//  * parameters come from a pmg_params (the TDatabase::ParamDB fields of this path)
//    instead of the global ParamDB;
//  * where the reference calls exit(4711) or prints an error, these classes throw
//    pmg_parmoon::Error carrying the pmg status code.
//
// The free functions MatVect_NSE4 / Defect_NSE4 (the MatVecProc / DefectProc signatures of
// include/General/Constants.h:159-160), DefectRestriction / Prolongate (the TFESpace3D
// signatures of MGComponents3D.C:184-1240) and TFgmresIte::Iterate(A, B, sol, rhs)
// (ItMethod.h:105) also exist with the reference's arguments: the caller's matrix and space
// pointers are keys of the levels registered with them (AddLevel(desc, A, B, U, P)); they
// are only compared, never dereferenced, so ParMooN's classes stay declared, not defined.
#pragma once
#include <algorithm>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "pmg.h"

class TSquareMatrix;
class TMatrix;
class TFESpace3D;

namespace pmg_parmoon {

struct Error : std::runtime_error {
  int status;
  Error(int st, const std::string &what) : std::runtime_error(what), status(st) {}
};

inline void check(pmg_ctx *ctx, int st, const char *where) {
  if (st == PMG_OK) return;
  std::string msg = std::string(where) + ": " + pmg_status_string(st);
  if (ctx) {
    const char *e = pmg_last_error_string(ctx);
    if (e && *e) msg += std::string(" (") + e + ")";
  }
  throw Error(st, msg);
}

class TNSE_MultiGrid;

// level lookup of the reference-signature entry points: matrix / space pointer -> level
struct LevelKey {
  TNSE_MultiGrid *mg;
  int level;
  int kind;   // 0: the level's first square matrix, 1: velocity space, 2: pressure space
};
inline std::map<const void *, LevelKey> &level_registry() {
  static std::map<const void *, LevelKey> r;
  return r;
}
inline LevelKey lookup(const void *key, int kind, const char *where) {
  auto it = level_registry().find(key);
  if (it == level_registry().end() || it->second.kind != kind)
    throw Error(PMG_ERR_INVALID_ARGUMENT, std::string(where) + ": matrix or space not registered with AddLevel");
  return it->second;
}

// One level: the TNSE_MGLevel4 operations on that level's vectors.  Vectors are laid out as
// in the reference ([u1 | u2 | u3 | p], n_dof = 3 N_UDOF + N_PDOF) and live in host memory
// unless the multigrid was created with device vectors.
class TNSE_MGLevel4 {
 public:
  TNSE_MGLevel4(TNSE_MultiGrid &mg, int level) : mg_(&mg), level_(level) {}
  // TNSE_MGLevel4::Defect (NSE_MGLevel4.C:465-539)
  void Defect(double *u1, double *f1, double *d1, double &res);
  // TNSE_MGLevel4::CorrectNodes (NSE_MGLevel4.C:542-601)
  void CorrectNodes(double *u1);
  // TNSE_MGLevel4::CellVanka (NSE_MGLevel4.C:603-1649) / NodalVanka (:4098-4549): one
  // sweep of the smoother the multigrid was created with (sc_smoother_saddle 1-4); aux,
  // the parameter array, smoother and N_Levels are accepted for signature compatibility
  // (the sweep keeps its old iterate internally)
  void CellVanka(double *u1, double *rhs1, double *aux, int N_Parameters, double *Parameters, int smoother,
                 int N_Levels);
  // TNSE_MGLevel4::SolveExact (NSE_MGLevel4.C:5728-5846), coarsest level only
  void SolveExact(double *u1, double *rhs1);
  // TNSE_MGLevel::Update (NSE_MGLevel.C:121-129): u1 += beta v1
  void Update(double *u1, double *v1);
  // TNSE_MGLevel::CorrectDefect (NSE_MGLevel.C:202-278)
  void CorrectDefect(double *v1);
  int GetN_DOF() const;
  int GetLevel() const { return level_; }

 private:
  TNSE_MultiGrid *mg_;
  int level_;
};

// TNSE_MultiGrid (NSE_MultiGrid.C): levels added coarsest first; Setup() does the
// value-dependent factorizations the reference repeats inside every sweep and coarse visit.
class TNSE_MultiGrid {
 public:
  explicit TNSE_MultiGrid(const pmg_params &params, int device = 0, bool device_vectors = false)
      : mem_(device_vectors ? PMG_MEM_DEVICE : PMG_MEM_HOST) {
    check(nullptr, pmg_ctx_create(device, &params, &ctx_), "pmg_ctx_create");
  }
  ~TNSE_MultiGrid() {
    auto &r = level_registry();
    for (auto it = r.begin(); it != r.end();) it = it->second.mg == this ? r.erase(it) : std::next(it);
    if (ctx_) pmg_ctx_destroy(ctx_);
  }
  TNSE_MultiGrid(const TNSE_MultiGrid &) = delete;
  TNSE_MultiGrid &operator=(const TNSE_MultiGrid &) = delete;

  // TNSE_MultiGrid::AddLevel (NSE_MultiGrid.C:94-102) + the TNSE_MGLevel4 constructor
  int AddLevel(const pmg_level_desc &desc) {
    int lvl = -1;
    check(ctx_, pmg_level_add(ctx_, &desc, &lvl), "AddLevel");
    levels_.emplace_back(*this, lvl);
    n_udof_.push_back(desc.n_udof);
    n_pdof_.push_back(desc.n_pdof);
    return lvl;
  }
  // the same, registering the level's sqmatrices / matrices arrays (first entries) and FE
  // spaces for the reference-signature entry points
  int AddLevel(const pmg_level_desc &desc, TSquareMatrix **A, TMatrix **B, TFESpace3D *U, TFESpace3D *P) {
    const int lvl = AddLevel(desc);
    auto &r = level_registry();
    if (A && A[0]) r[A[0]] = LevelKey{this, lvl, 0};
    if (U) r[U] = LevelKey{this, lvl, 1};
    if (P) r[P] = LevelKey{this, lvl, 2};
    (void)B;
    return lvl;
  }
  int N_UDOF(int level) const { return n_udof_.at(level); }
  int N_PDOF(int level) const { return n_pdof_.at(level); }
  TNSE_MGLevel4 *GetLevel(int i) { return &levels_.at(i); }
  int GetN_Levels() const { return int(levels_.size()); }
  void Setup() { check(ctx_, pmg_setup(ctx_), "Setup"); }
  // new matrix values on the same pattern, then Setup()
  void UpdateValues(int level, const pmg_level_desc &desc) {
    check(ctx_, pmg_level_update_values(ctx_, level, &desc), "UpdateValues");
  }
  // TNSE_MultiGrid::SetDirichletNodes (NSE_MultiGrid.C:1781-1802) on level i
  void SetDirichletNodes(int i, double *u, double *f) {
    check(ctx_, pmg_set_dirichlet(ctx_, i, u, f, mem_), "SetDirichletNodes");
  }
  // TMultiGridIte::Iterate with zero start (MultiGridIte.C:76-130): one V-cycle,
  // TNSE_MultiGrid::Cycle on the finest level, z = M^-1 v
  void Cycle(const double *v, double *z) { check(ctx_, pmg_vcycle(ctx_, v, z, mem_), "Cycle"); }
  void SetParam(const pmg_params &params) { check(ctx_, pmg_set_params(ctx_, &params), "SetParam"); }

  pmg_ctx *ctx() const { return ctx_; }
  int mem() const { return mem_; }

 private:
  pmg_ctx *ctx_ = nullptr;
  int mem_;
  std::vector<TNSE_MGLevel4> levels_;
  std::vector<int> n_udof_, n_pdof_;
};

// TFgmresIte (FgmresIte.C:75-677), flexible branch, preconditioned by the multigrid.
class TFgmresIte {
 public:
  explicit TFgmresIte(TNSE_MultiGrid &mg) : mg_(&mg) {}
  // TItMethod::Iterate (ItMethod.h:105): returns the iteration count j; sol is updated in
  // place.  Divergence (the reference exits) throws; non-convergence within maxit returns
  // the count like the reference (which only prints) and records it in Converged().
  int Iterate(double *sol, double *rhs) {
    int it = 0;
    double res = 0;
    const int st = pmg_fgmres_solve(mg_->ctx(), sol, rhs, &it, &res, mg_->mem());
    final_res_ = res;
    converged_ = st == PMG_OK;
    if (st != PMG_OK && st != PMG_ERR_NOT_CONVERGED) check(mg_->ctx(), st, "TFgmresIte::Iterate");
    return it;
  }
  // the reference's signature (ItMethod.h:105): A must be the finest level's sqmatrices
  int Iterate(TSquareMatrix **A, TMatrix **B, double *sol, double *rhs) {
    const LevelKey k = lookup(A ? A[0] : nullptr, 0, "TFgmresIte::Iterate");
    if (k.mg != mg_ || k.level != mg_->GetN_Levels() - 1)
      throw Error(PMG_ERR_INVALID_ARGUMENT, "TFgmresIte::Iterate: not the finest level of this multigrid");
    (void)B;
    return Iterate(sol, rhs);
  }
  double FinalResidual() const { return final_res_; }
  bool Converged() const { return converged_; }
  std::vector<double> History() const {
    std::vector<double> h(100000);
    h.resize(pmg_fgmres_history(mg_->ctx(), h.data(), int(h.size())));
    return h;
  }

 private:
  TNSE_MultiGrid *mg_;
  double final_res_ = 0;
  bool converged_ = false;
};

// MatVect_NSE4 / Defect_NSE4 (MGComponents3D.C:2872-3231) on one level
inline void MatVect_NSE4(TNSE_MultiGrid &mg, int level, double *x, double *y) {
  check(mg.ctx(), pmg_matvec(mg.ctx(), level, x, y, mg.mem()), "MatVect_NSE4");
}
inline void Defect_NSE4(TNSE_MultiGrid &mg, int level, double *x, double *b, double *r) {
  check(mg.ctx(), pmg_defect_nse4(mg.ctx(), level, x, b, r, mg.mem()), "Defect_NSE4");
}
// DefectRestriction / Prolongate (MGComponents3D.C:184-1240); 'fine' is the fine level
inline void DefectRestriction(TNSE_MultiGrid &mg, int fine, double *d_fine, double *f_coarse) {
  check(mg.ctx(), pmg_restrict_defect(mg.ctx(), fine, d_fine, f_coarse, mg.mem()), "DefectRestriction");
}
inline void Prolongate(TNSE_MultiGrid &mg, int fine, double *u_coarse, double *d_fine) {
  check(mg.ctx(), pmg_prolongate(mg.ctx(), fine, u_coarse, d_fine, mg.mem()), "Prolongate");
}

// The reference's signatures.  MatVect_NSE4 / Defect_NSE4: A = the level's sqmatrices
// (A11..A33), B = its matrices (B1T..B3T, B1..B3), vectors [u1 | u2 | u3 | p] of that level.
inline void MatVect_NSE4(TSquareMatrix **A, TMatrix **B, double *x, double *y) {
  const LevelKey k = lookup(A ? A[0] : nullptr, 0, "MatVect_NSE4");
  (void)B;
  MatVect_NSE4(*k.mg, k.level, x, y);
}
inline void Defect_NSE4(TSquareMatrix **A, TMatrix **B, double *x, double *b, double *r) {
  const LevelKey k = lookup(A ? A[0] : nullptr, 0, "Defect_NSE4");
  (void)B;
  Defect_NSE4(*k.mg, k.level, x, b, r);
}
namespace detail {
// one component block (velocity: N_Functions = 3 blocks of N_UDOF; pressure: 1 of N_PDOF)
// of a transfer, through the whole-vector entry point; host vectors only
inline void transfer_part(TFESpace3D *coarse, TFESpace3D *fine, int n_functions, double *coarse_f, double *fine_f,
                          bool restrict_, const char *where) {
  const auto it = level_registry().find(fine);
  if (it == level_registry().end() || it->second.kind == 0)
    throw Error(PMG_ERR_INVALID_ARGUMENT, std::string(where) + ": fine space not registered with AddLevel");
  const LevelKey f = it->second;
  const auto ic = level_registry().find(coarse);
  if (ic == level_registry().end() || ic->second.mg != f.mg || ic->second.kind != f.kind ||
      ic->second.level != f.level - 1)
    throw Error(PMG_ERR_INVALID_ARGUMENT, std::string(where) + ": coarse space is not the next coarser level");
  TNSE_MultiGrid &mg = *f.mg;
  if (mg.mem() != PMG_MEM_HOST)
    throw Error(PMG_ERR_UNSUPPORTED, std::string(where) + ": the space form takes host vectors");
  const bool vel = f.kind == 1;
  if (n_functions != (vel ? 3 : 1))
    throw Error(PMG_ERR_INVALID_ARGUMENT, std::string(where) + ": N_Functions must be 3 (velocity) or 1 (pressure)");
  const int lf = f.level, lc = f.level - 1;
  const size_t nf = 3 * size_t(mg.N_UDOF(lf)) + mg.N_PDOF(lf), nc = 3 * size_t(mg.N_UDOF(lc)) + mg.N_PDOF(lc);
  const size_t of = vel ? 0 : 3 * size_t(mg.N_UDOF(lf)), oc = vel ? 0 : 3 * size_t(mg.N_UDOF(lc));
  const size_t lenf = vel ? 3 * size_t(mg.N_UDOF(lf)) : size_t(mg.N_PDOF(lf));
  const size_t lenc = vel ? 3 * size_t(mg.N_UDOF(lc)) : size_t(mg.N_PDOF(lc));
  std::vector<double> F(nf, 0.0), Cv(nc, 0.0);
  if (restrict_) {
    std::copy(fine_f, fine_f + lenf, F.begin() + of);
    check(mg.ctx(), pmg_restrict_defect(mg.ctx(), lf, F.data(), Cv.data(), PMG_MEM_HOST), where);
    std::copy(Cv.begin() + oc, Cv.begin() + oc + lenc, coarse_f);
  } else {
    std::copy(coarse_f, coarse_f + lenc, Cv.begin() + oc);
    check(mg.ctx(), pmg_prolongate(mg.ctx(), lf, Cv.data(), F.data(), PMG_MEM_HOST), where);
    std::copy(F.begin() + of, F.begin() + of + lenf, fine_f);
  }
}
}  // namespace detail
// DefectRestriction (MGComponents3D.C:711-1240): CoarseFunction := R FineFunction for the
// velocity space (N_Functions = 3) or the pressure space (N_Functions = 1); aux unused
inline void DefectRestriction(TFESpace3D *CoarseSpace, TFESpace3D *FineSpace, int N_Functions, double *CoarseFunction,
                              double *FineFunction, double *aux) {
  (void)aux;
  detail::transfer_part(CoarseSpace, FineSpace, N_Functions, CoarseFunction, FineFunction, true, "DefectRestriction");
}
// Prolongate (MGComponents3D.C:184-708): FineFunction := P CoarseFunction; aux unused
inline void Prolongate(TFESpace3D *CoarseSpace, TFESpace3D *FineSpace, int N_Functions, double *CoarseFunction,
                       double *FineFunction, double *aux) {
  (void)aux;
  detail::transfer_part(CoarseSpace, FineSpace, N_Functions, CoarseFunction, FineFunction, false, "Prolongate");
}

// ------------------------------------------------------------------ TNSE_MGLevel4 bodies
inline void TNSE_MGLevel4::Defect(double *u1, double *f1, double *d1, double &res) {
  check(mg_->ctx(), pmg_defect(mg_->ctx(), level_, u1, f1, d1, &res, mg_->mem()), "Defect");
}
inline void TNSE_MGLevel4::CorrectNodes(double *u1) {
  check(mg_->ctx(), pmg_correct_nodes(mg_->ctx(), level_, u1, mg_->mem()), "CorrectNodes");
}
inline void TNSE_MGLevel4::CellVanka(double *u1, double *rhs1, double *, int, double *, int smoother, int) {
  if (smoother < 1 || smoother > 4) throw Error(PMG_ERR_UNSUPPORTED, "CellVanka: smoother must be 1 .. 4");
  check(mg_->ctx(), pmg_vanka_smooth(mg_->ctx(), level_, u1, rhs1, 1, mg_->mem()), "CellVanka");
}
inline void TNSE_MGLevel4::SolveExact(double *u1, double *rhs1) {
  if (level_ != 0) throw Error(PMG_ERR_UNSUPPORTED, "SolveExact: coarsest level only");
  check(mg_->ctx(), pmg_coarse_solve(mg_->ctx(), u1, rhs1, mg_->mem()), "SolveExact");
}
inline void TNSE_MGLevel4::Update(double *u1, double *v1) {
  check(mg_->ctx(), pmg_update(mg_->ctx(), level_, u1, v1, mg_->mem()), "Update");
}
inline void TNSE_MGLevel4::CorrectDefect(double *v1) {
  check(mg_->ctx(), pmg_correct_defect(mg_->ctx(), level_, v1, mg_->mem()), "CorrectDefect");
}
inline int TNSE_MGLevel4::GetN_DOF() const {
  int n = 0;
  check(mg_->ctx(), pmg_level_info(mg_->ctx(), level_, &n, nullptr, nullptr), "GetN_DOF");
  return n;
}

}  // namespace pmg_parmoon
