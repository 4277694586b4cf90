// pmg_oracle.cpp - FP64 CPU restatement of ParMooN's NSE type-4 GMG-Vanka + FGMRES path.
//
// TEST INFRASTRUCTURE ONLY (correctness oracle and timed CPU baseline); the product
// library never calls it.  Each function follows the reference lines cited above it,
// statement by statement, with ParMooN's variable names where that helps the reader.
// Compiled with -ffp-contract=off; the dense LU and triangular sweeps use explicit fma()
// so that a device implementation performing the same operations in the same order
// reproduces them bit for bit.
#include "pmg_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int GEO_DIM = 3;

struct OLevel {
  pmg_level_desc d;
  int N_UDOF, N_PDOF, N_DOF, N_Active, HangingNodeBound, N_Dirichlet;
  int N_U, N_P, N_LocalDOF;
  // colorCell output (ptrCellColors / CellReorder / maxCellsPerColor)
  int N_CIntCell = 0;
  std::vector<int> ptrCellColors, CellReorder;
  int maxCellsPerColor = 0;
  // MG vectors: Sol, Rhs, Aux[0] (defect), Aux[1] (aux), as TNSE_MGLevel4 owns them
  std::vector<double> sol, rhs, aux0, aux1;
  // cached factors (refactor == 0)
  std::vector<double> lu;
  std::vector<int> piv;
  std::vector<char> lu_ready;
};

}  // namespace

struct oracle_mg {
  pmg_params prm;
  std::vector<OLevel> L;
  int refactor = 1;
  int variant = 0;
  std::vector<double> coarse_lu;
  std::vector<int> coarse_piv;
  bool coarse_ready = false;
  std::vector<double> hist;
  std::vector<double> cyclelog;
};

namespace {

inline double Ddot(int n, const double *a, const double *b) {
  double s = 0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}
inline void Daxpy(int n, double alpha, const double *x, double *y) {
  for (int i = 0; i < n; ++i) y[i] += alpha * x[i];
}
inline void Dscal(int n, double alpha, double *x) {
  for (int i = 0; i < n; ++i) x[i] *= alpha;
}

// IntoL20Vector3D, case -11 (P1disc on hexahedra): MGComponents3D.C:1756-1790
void IntoL20Vector3D(double *v, int Length) {
  double s = 0;
  for (int i = 0; i < Length; i += 4) s += v[i];
  s /= Length;
  s *= 4;
  for (int i = 0; i < Length; i += 4) v[i] -= s;
}

// CoupledDefect, NSTYPE 4: MGComponents3D.C:3038-3209.  r := b - K x
void CoupledDefect(const OLevel &L, const double *x, const double *b, double *r) {
  const pmg_level_desc &d = L.d;
  const int N_UDOF = L.N_UDOF, N_PDOF = L.N_PDOF, N_Active = L.N_Active;
  const int *ARowPtr = d.a_rowptr, *AKCol = d.a_kcol;
  const double *A11 = d.a[0], *A12 = d.a[1], *A13 = d.a[2], *A21 = d.a[3], *A22 = d.a[4],
               *A23 = d.a[5], *A31 = d.a[6], *A32 = d.a[7], *A33 = d.a[8];
  const double *u1 = x, *u2 = u1 + N_UDOF, *u3 = u2 + N_UDOF, *p = u3 + N_UDOF;
  const double *v1 = b, *v2 = v1 + N_UDOF, *v3 = v2 + N_UDOF, *q = v3 + N_UDOF;
  double *r1 = r, *r2 = r1 + N_UDOF, *r3 = r2 + N_UDOF, *r4 = r3 + N_UDOF;
#pragma omp parallel
  {
#pragma omp for
    for (int i = 0; i < N_Active; i++) {
      double s = v1[i], t = v2[i], u = v3[i];
      int k = ARowPtr[i + 1];
      for (int j = ARowPtr[i]; j < k; j++) {
        int index = AKCol[j];
        double value = A11[j], value1 = A12[j], value2 = A13[j], value3 = A21[j], value4 = A22[j],
               value5 = A23[j], value6 = A31[j], value7 = A32[j], value8 = A33[j];
        double value9 = u1[index], value10 = u2[index], value11 = u3[index];
        s -= value * value9 + value1 * value10 + value2 * value11;
        t -= value3 * value9 + value4 * value10 + value5 * value11;
        u -= value6 * value9 + value7 * value10 + value8 * value11;
      }
      r1[i] = s; r2[i] = t; r3[i] = u;
    }
#pragma omp for
    for (int i = N_Active; i < N_UDOF; i++) {
      double s = v1[i], t = v2[i], u = v3[i];
      int k = ARowPtr[i + 1];
      for (int j = ARowPtr[i]; j < k; j++) {
        int index = AKCol[j];
        s -= A11[j] * u1[index];
        t -= A22[j] * u2[index];
        u -= A33[j] * u3[index];
      }
      r1[i] = s; r2[i] = t; r3[i] = u;
    }
#pragma omp for
    for (int i = 0; i < N_PDOF; i++) {
      double s = q[i];
      int k = d.b_rowptr[i + 1];
      for (int j = d.b_rowptr[i]; j < k; j++) {
        int index = d.b_kcol[j];
        s -= d.b[0][j] * u1[index] + d.b[1][j] * u2[index] + d.b[2][j] * u3[index];
      }
      r4[i] = s;
    }
#pragma omp for
    for (int i = 0; i < N_Active; i++) {
      double s = 0, t = 0, u = 0;
      int k = d.bt_rowptr[i + 1];
      for (int j = d.bt_rowptr[i]; j < k; j++) {
        int index = d.bt_kcol[j];
        double value = p[index];
        s += d.bt[0][j] * value;
        t += d.bt[1][j] * value;
        u += d.bt[2][j] * value;
      }
      r1[i] -= s; r2[i] -= t; r3[i] -= u;
    }
  }
}

// CoupledMatVect, NSTYPE 4: MGComponents3D.C:2872-3035.  y := K x
void CoupledMatVect(const OLevel &L, const double *x, double *y) {
  const pmg_level_desc &d = L.d;
  const int N_UDOF = L.N_UDOF, N_PDOF = L.N_PDOF, N_Active = L.N_Active;
  const double *u1 = x, *u2 = u1 + N_UDOF, *u3 = u2 + N_UDOF, *p = u3 + N_UDOF;
  double *v1 = y, *v2 = v1 + N_UDOF, *v3 = v2 + N_UDOF, *q = v3 + N_UDOF;
#pragma omp parallel
  {
#pragma omp for
    for (int i = 0; i < N_Active; i++) {
      double s = 0, t = 0, u = 0;
      for (int j = d.a_rowptr[i]; j < d.a_rowptr[i + 1]; j++) {
        int index = d.a_kcol[j];
        double value9 = u1[index], value10 = u2[index], value11 = u3[index];
        s += d.a[0][j] * value9 + d.a[1][j] * value10 + d.a[2][j] * value11;
        t += d.a[3][j] * value9 + d.a[4][j] * value10 + d.a[5][j] * value11;
        u += d.a[6][j] * value9 + d.a[7][j] * value10 + d.a[8][j] * value11;
      }
      v1[i] = s; v2[i] = t; v3[i] = u;
    }
#pragma omp for
    for (int i = N_Active; i < N_UDOF; i++) {
      double s = 0, t = 0, u = 0;
      for (int j = d.a_rowptr[i]; j < d.a_rowptr[i + 1]; j++) {
        int index = d.a_kcol[j];
        s += d.a[0][j] * u1[index];
        t += d.a[4][j] * u2[index];
        u += d.a[8][j] * u3[index];
      }
      v1[i] = s; v2[i] = t; v3[i] = u;
    }
#pragma omp for
    for (int i = 0; i < N_PDOF; i++) {
      double s = 0;
      for (int j = d.b_rowptr[i]; j < d.b_rowptr[i + 1]; j++) {
        int index = d.b_kcol[j];
        s += d.b[0][j] * u1[index] + d.b[1][j] * u2[index] + d.b[2][j] * u3[index];
      }
      q[i] = s;
    }
#pragma omp for
    for (int i = 0; i < N_Active; i++) {
      double s = 0, t = 0, u = 0;
      for (int j = d.bt_rowptr[i]; j < d.bt_rowptr[i + 1]; j++) {
        double value = p[d.bt_kcol[j]];
        s += d.bt[0][j] * value;
        t += d.bt[1][j] * value;
        u += d.bt[2][j] * value;
      }
      v1[i] += s; v2[i] += t; v3[i] += u;
    }
  }
}

// TNSE_MGLevel4::Defect: NSE_MGLevel4.C:465-539
void Defect(const oracle_mg &mg, const OLevel &L, const double *u1, const double *f1, double *d1,
            double &res) {
  CoupledDefect(L, u1, f1, d1);
  if (mg.prm.internal_project_pressure) IntoL20Vector3D(d1 + GEO_DIM * L.N_UDOF, L.N_PDOF);
  const int n = L.N_UDOF - L.N_Active;
  std::memset(d1 + L.N_Active, 0, sizeof(double) * n);
  std::memset(d1 + L.N_UDOF + L.N_Active, 0, sizeof(double) * n);
  std::memset(d1 + 2 * L.N_UDOF + L.N_Active, 0, sizeof(double) * n);
  res = std::sqrt(Ddot(L.N_DOF, d1, d1));
}

// TNSE_MGLevel4::CorrectNodes: NSE_MGLevel4.C:542-601 (no hanging nodes on these meshes)
void CorrectNodes(const OLevel &L, double *u1) {
  for (int c = 0; c < GEO_DIM; ++c)
    std::memset(u1 + c * L.N_UDOF + L.HangingNodeBound, 0, L.N_Dirichlet * sizeof(double));
}

// TNSE_MultiGrid::SetDirichletNodes: NSE_MultiGrid.C:1781-1802
void SetDirichletNodes(const OLevel &L, double *u, const double *f) {
  for (int j = 0; j < GEO_DIM; j++)
    std::memcpy(u + j * L.N_UDOF + L.HangingNodeBound, f + j * L.N_UDOF + L.HangingNodeBound,
                sizeof(double) * L.N_Dirichlet);
}

// TNSE_MGLevel::CorrectDefect: NSE_MGLevel.C:202-278
void CorrectDefect(const oracle_mg &mg, const OLevel &L, double *v1) {
  for (int c = 0; c < GEO_DIM; ++c)
    std::memset(v1 + c * L.N_UDOF + L.N_Active, 0, sizeof(double) * (L.N_UDOF - L.N_Active));
  if (mg.prm.internal_project_pressure) IntoL20Vector3D(v1 + GEO_DIM * L.N_UDOF, L.N_PDOF);
}

// TNSE_MGLevel::Update: NSE_MGLevel.C:121-129
void Update(const oracle_mg &mg, const OLevel &L, double *u1, const double *v1) {
  double omega = mg.prm.sc_gmg_damp_factor_saddle;
  for (int i = 0; i < L.N_DOF; i++) u1[i] += omega * v1[i];
}

// ---------------------------------------------------------------- dense LU (LinAlg.C:106-130)
// dgetf2: pivot = first index of max |a(i,j)|, i >= j; reciprocal scaling; rank-1 update.
int lu_factor(int n, double *a, int *ipiv) {
  int info = 0;
  for (int j = 0; j < n; ++j) {
    double *colj = a + size_t(j) * n;
    int p = j;
    double mx = std::fabs(colj[j]);
    for (int i = j + 1; i < n; ++i) {
      double v = std::fabs(colj[i]);
      if (v > mx) { mx = v; p = i; }
    }
    ipiv[j] = p;
    if (colj[p] != 0.0) {
      if (p != j)
        for (int c = 0; c < n; ++c) std::swap(a[size_t(c) * n + j], a[size_t(c) * n + p]);
      double r = 1.0 / colj[j];
      for (int i = j + 1; i < n; ++i) colj[i] *= r;
    } else if (info == 0) {
      info = j + 1;
    }
    // trailing update, column by column (each column independent -> deterministic)
#pragma omp parallel for schedule(static) if (n > 512)
    for (int c = j + 1; c < n; ++c) {
      double *colc = a + size_t(c) * n;
      double u = colc[j];
      for (int i = j + 1; i < n; ++i) colc[i] = std::fma(-colj[i], u, colc[i]);
    }
  }
  return info;
}

// dgetrs('N', nrhs = 1): dlaswp, dtrsm(L, unit), dtrsm(U) in column sweeps
void lu_solve(int n, const double *a, const int *ipiv, double *b) {
  for (int k = 0; k < n; ++k)
    if (ipiv[k] != k) std::swap(b[k], b[ipiv[k]]);
  for (int k = 0; k < n; ++k) {
    double bk = b[k];
    const double *col = a + size_t(k) * n;
    for (int i = k + 1; i < n; ++i) b[i] = std::fma(-bk, col[i], b[i]);
  }
  for (int k = n - 1; k >= 0; --k) {
    const double *col = a + size_t(k) * n;
    b[k] = b[k] / col[k];
    double bk = b[k];
    for (int i = 0; i < k; ++i) b[i] = std::fma(-bk, col[i], b[i]);
  }
}

// ---------------------------------------------------------------- Vanka patch pieces
// local residual of one cell, NSE_MGLevel4.C:902-1061 (colored) / the same lines of the
// _SEQ variant: velocity rows j, j+N_U, j+2N_U, pressure rows 3N_U+j
void PatchResidual(const OLevel &L, int ii, const double *u1, const double *rhs1, double *Rhs) {
  const pmg_level_desc &d = L.d;
  const int N_U = L.N_U, N_P = L.N_P, N_UDOF = L.N_UDOF, ActiveBound = L.N_Active;
  const double *u2 = u1 + N_UDOF, *u3 = u2 + N_UDOF, *p = u1 + GEO_DIM * N_UDOF;
  const double *rhs2 = rhs1 + N_UDOF, *rhs3 = rhs2 + N_UDOF, *rhsp = rhs1 + GEO_DIM * N_UDOF;
  const int *UDOFs = d.u_global + d.u_begin[ii];
  const int *PDOFs = d.p_global + d.p_begin[ii];
  for (int j = 0; j < N_U; j++) {
    int j1 = j, j2 = j + N_U, j3 = j2 + N_U;
    int UDOF = UDOFs[j];
    int begin = d.a_rowptr[UDOF], end = d.a_rowptr[UDOF + 1];
    Rhs[j1] = rhs1[UDOF];
    Rhs[j2] = rhs2[UDOF];
    Rhs[j3] = rhs3[UDOF];
    for (int k = begin; k < end; k++) {
      int l = d.a_kcol[k];
      double value11 = d.a[0][k], value12 = d.a[1][k], value13 = d.a[2][k];
      double value21 = d.a[3][k], value22 = d.a[4][k], value23 = d.a[5][k];
      double value31 = d.a[6][k], value32 = d.a[7][k], value33 = d.a[8][k];
      if (UDOF >= ActiveBound)  // Dirichlet node
        value12 = value13 = value21 = value23 = value31 = value32 = 0;
      Rhs[j1] -= value11 * u1[l] + value12 * u2[l] + value13 * u3[l];
      Rhs[j2] -= value21 * u1[l] + value22 * u2[l] + value23 * u3[l];
      Rhs[j3] -= value31 * u1[l] + value32 * u2[l] + value33 * u3[l];
    }
    if (UDOF < ActiveBound) {
      for (int k = d.bt_rowptr[UDOF]; k < d.bt_rowptr[UDOF + 1]; k++) {
        int l = d.bt_kcol[k];
        double value = p[l];
        Rhs[j1] -= d.bt[0][k] * value;
        Rhs[j2] -= d.bt[1][k] * value;
        Rhs[j3] -= d.bt[2][k] * value;
      }
    }
  }
  for (int j = 0; j < N_P; j++) {
    int j1 = j + GEO_DIM * N_U;
    int PDOF = PDOFs[j];
    Rhs[j1] = rhsp[PDOF];
    for (int k = d.b_rowptr[PDOF]; k < d.b_rowptr[PDOF + 1]; k++) {
      int l = d.b_kcol[k];
      Rhs[j1] -= d.b[0][k] * u1[l];
      Rhs[j1] -= d.b[1][k] * u2[l];
      Rhs[j1] -= d.b[2][k] * u3[l];
    }
  }
}

// local 85x85 matrix, column-major System[col*85+row], by the reference's linear search:
// NSE_MGLevel4.C:1066-1290
void PatchMatrix(const OLevel &L, int ii, double *System) {
  const pmg_level_desc &d = L.d;
  const int N_U = L.N_U, N_P = L.N_P, N_LocalDOF = L.N_LocalDOF, ActiveBound = L.N_Active;
  std::memset(System, 0, sizeof(double) * N_LocalDOF * N_LocalDOF);
  const int *UDOFs = d.u_global + d.u_begin[ii];
  const int *PDOFs = d.p_global + d.p_begin[ii];
  for (int j = 0; j < N_U; j++) {
    int j1 = j, j2 = j + N_U, j3 = j2 + N_U;
    int UDOF = UDOFs[j];
    for (int k = d.a_rowptr[UDOF]; k < d.a_rowptr[UDOF + 1]; k++) {
      int l = d.a_kcol[k];
      double value11 = d.a[0][k], value12 = d.a[1][k], value13 = d.a[2][k];
      double value21 = d.a[3][k], value22 = d.a[4][k], value23 = d.a[5][k];
      double value31 = d.a[6][k], value32 = d.a[7][k], value33 = d.a[8][k];
      if (UDOF >= ActiveBound) value12 = value13 = value21 = value23 = value31 = value32 = 0;
      for (int m = 0; m < N_U; m++)
        if (UDOFs[m] == l) {
          int k1 = m * N_LocalDOF, k2 = (m + N_U) * N_LocalDOF, k3 = (m + 2 * N_U) * N_LocalDOF;
          System[k1 + j1] = value11;
          System[k2 + j1] = value12;
          System[k1 + j2] = value21;
          System[k2 + j2] = value22;
          System[k3 + j1] = value13;
          System[k3 + j2] = value23;
          System[k1 + j3] = value31;
          System[k2 + j3] = value32;
          System[k3 + j3] = value33;
          break;
        }
    }
    if (UDOF < ActiveBound) {
      for (int k = d.bt_rowptr[UDOF]; k < d.bt_rowptr[UDOF + 1]; k++) {
        int l = d.bt_kcol[k];
        for (int m = 0; m < N_P; m++)
          if (PDOFs[m] == l) {
            int k1 = (m + GEO_DIM * N_U) * N_LocalDOF;
            System[k1 + j1] = d.bt[0][k];
            System[k1 + j2] = d.bt[1][k];
            System[k1 + j3] = d.bt[2][k];
            break;
          }
      }
    }
  }
  for (int j = 0; j < N_P; j++) {
    int j1 = j + GEO_DIM * N_U;
    int PDOF = PDOFs[j];
    for (int k = d.b_rowptr[PDOF]; k < d.b_rowptr[PDOF + 1]; k++) {
      int l = d.b_kcol[k];
      for (int m = 0; m < N_U; m++)
        if (UDOFs[m] == l) {
          System[m * N_LocalDOF + j1] = d.b[0][k];
          System[(m + N_U) * N_LocalDOF + j1] = d.b[1][k];
          System[(m + 2 * N_U) * N_LocalDOF + j1] = d.b[2][k];
          break;
        }
    }
  }
}

// SolveLinearSystemLapack on the patch (LinAlg.C:106-130); with refactor == 0 the factors of
// the identical matrix are computed once and reused
int PatchSolve(oracle_mg &mg, OLevel &L, int ii, double *System, double *Rhs) {
  const int n = L.N_LocalDOF;
  if (mg.refactor) {
    std::vector<int> ipiv(n);
    PatchMatrix(L, ii, System);
    int info = lu_factor(n, System, ipiv.data());
    lu_solve(n, System, ipiv.data(), Rhs);
    return info;
  }
  double *F = &L.lu[size_t(ii) * n * n];
  int *pv = &L.piv[size_t(ii) * n];
  int info = 0;
  if (!L.lu_ready[ii]) {
    PatchMatrix(L, ii, F);
    info = lu_factor(n, F, pv);
    L.lu_ready[ii] = 1;
  }
  lu_solve(n, F, pv, Rhs);
  return info;
}

// TNSE_MGLevel4::CellVanka, colored (_HYBRID) variant: NSE_MGLevel4.C:603-1649
int CellVankaColored(oracle_mg &mg, OLevel &L, double *u1, const double *rhs1, double *aux) {
  const int N_U = L.N_U, N_P = L.N_P, N_LocalDOF = L.N_LocalDOF, N_UDOF = L.N_UDOF;
  const double damp = mg.prm.sc_smooth_damp_factor_coarse_saddle;
  const double alpha = mg.prm.sc_smooth_damp_factor_saddle;
  double *u2 = u1 + N_UDOF, *u3 = u2 + N_UDOF, *p = u1 + GEO_DIM * N_UDOF;
  SetDirichletNodes(L, u1, rhs1);  // :683-695 (memcpy u_D := rhs_D)
  double *uold = aux;
  std::memcpy(uold, u1, L.N_DOF * sizeof(double));
  const int mc = L.maxCellsPerColor;
  std::vector<double> Rhs(size_t(mc) * N_LocalDOF);
  std::vector<double> System(mg.refactor ? size_t(mc) * N_LocalDOF * N_LocalDOF : 0);
  int info_all = 0;
  for (int i = 0; i < L.N_CIntCell; i++) {
    const int c0 = L.ptrCellColors[i], c1 = L.ptrCellColors[i + 1];
#pragma omp parallel for schedule(dynamic, 16)
    for (int jj = c0; jj < c1; jj++)
      PatchResidual(L, L.CellReorder[jj], u1, rhs1, &Rhs[size_t(jj - c0) * N_LocalDOF]);
#pragma omp parallel for schedule(dynamic, 16) reduction(max : info_all)
    for (int jj = c0; jj < c1; jj++) {
      int ii = L.CellReorder[jj];
      double *R = &Rhs[size_t(jj - c0) * N_LocalDOF];
      double *S = mg.refactor ? &System[size_t(jj - c0) * N_LocalDOF * N_LocalDOF] : nullptr;
      int info = PatchSolve(mg, L, ii, S, R);
      if (info) info_all = std::max(info_all, ii + 1);
      const int *UDOFs = L.d.u_global + L.d.u_begin[ii];
      const int *PDOFs = L.d.p_global + L.d.p_begin[ii];
      for (int j = 0; j < N_U; j++) {
        int l = UDOFs[j];
        u1[l] += damp * R[j];
        u2[l] += damp * R[j + N_U];
        u3[l] += damp * R[j + 2 * N_U];
      }
      for (int j = 0; j < N_P; j++) p[PDOFs[j]] += damp * R[j + GEO_DIM * N_U];
    }
  }
  if (std::fabs(1 - alpha) > 1e-3)  // :1615-1617
    for (int j = 0; j < L.N_DOF; j++) u1[j] = uold[j] + alpha * (u1[j] - uold[j]);
  SetDirichletNodes(L, u1, rhs1);  // :1620-1630
  return info_all;
}

// TNSE_MGLevel4::CellVanka, Gauss-Seidel (_SEQ) variant: NSE_MGLevel4.C:1653-2085
int CellVankaGS(oracle_mg &mg, OLevel &L, double *u1, const double *rhs1, double *aux) {
  const int N_U = L.N_U, N_P = L.N_P, N_LocalDOF = L.N_LocalDOF, N_UDOF = L.N_UDOF;
  const double damp = mg.prm.sc_smooth_damp_factor_coarse_saddle;
  const double alpha = mg.prm.sc_smooth_damp_factor_saddle;
  double *u2 = u1 + N_UDOF, *u3 = u2 + N_UDOF, *p = u1 + GEO_DIM * N_UDOF;
  SetDirichletNodes(L, u1, rhs1);
  double *uold = aux;
  std::memcpy(uold, u1, L.N_DOF * sizeof(double));
  std::vector<double> System(size_t(N_LocalDOF) * N_LocalDOF), Rhs(N_LocalDOF);
  int info_all = 0;
  for (int ii = 0; ii < L.d.n_cells; ii++) {
    PatchResidual(L, ii, u1, rhs1, Rhs.data());
    int info = PatchSolve(mg, L, ii, System.data(), Rhs.data());
    if (info) info_all = std::max(info_all, ii + 1);
    const int *UDOFs = L.d.u_global + L.d.u_begin[ii];
    const int *PDOFs = L.d.p_global + L.d.p_begin[ii];
    for (int j = 0; j < N_U; j++) {
      int l = UDOFs[j];
      u1[l] += damp * Rhs[j];
      u2[l] += damp * Rhs[j + N_U];
      u3[l] += damp * Rhs[j + 2 * N_U];
    }
    for (int j = 0; j < N_P; j++) p[PDOFs[j]] += damp * Rhs[j + GEO_DIM * N_U];
  }
  if (std::fabs(1 - alpha) > 1e-3)
    for (int j = 0; j < L.N_DOF; j++) u1[j] = uold[j] + alpha * (u1[j] - uold[j]);
  SetDirichletNodes(L, u1, rhs1);
  return info_all;
}

// SolveDiagonalVanka3D with N_P = 1 (MGComponents3D.C:32-85): the velocity block replaced
// by its diagonal, taken from the first component's rows for all three (Ai[i] = 1/A(i,i),
// i < N_U); a = System column-major with leading dimension LDA, b: rhs in, solution out.
void SolveDiagonalVanka3D(const double *a, double *b, int N_U, int LDA) {
  auto AT = [&](int i, int j) { return a[size_t(j) * LDA + i]; };
  const int row = 3 * N_U;
  std::vector<double> Ai(N_U);
  for (int i = 0; i < N_U; i++) Ai[i] = 1 / AT(i, i);
  double dp = 0, ffp = b[row];
  for (int i = 0, j = N_U, k = 2 * N_U; i < N_U; i++, j++, k++) {
    const double tmp = Ai[i];
    dp -= tmp * (AT(row, i) * AT(i, row) + AT(row, j) * AT(j, row) + AT(row, k) * AT(k, row));
    ffp -= tmp * (AT(row, i) * b[i] + AT(row, j) * b[j] + AT(row, k) * b[k]);
  }
  const double pp = ffp / dp;
  b[row] = pp;
  for (int i = 0, j = N_U, k = 2 * N_U; i < N_U; i++, j++, k++) {
    const double tmp = Ai[i];
    b[i] = tmp * (b[i] - AT(i, row) * pp);
    b[j] = tmp * (b[j] - AT(j, row) * pp);
    b[k] = tmp * (b[k] - AT(k, row) * pp);
  }
}

// TNSE_MGLevel4::NodalVanka, smoothers 3 (diagonal) and 4 (full LU): NSE_MGLevel4.C:4098-4549
// (one process, DOF_Reorder = 0): pressure dofs in order, each with the velocity dofs of
// its B row, local system assembled (columns found by linear search in UDOFs), solved,
// updates applied at once; then the global relaxation (without the |1 - alpha| test of
// CellVanka) and the Dirichlet copy.
int NodalVanka(oracle_mg &mg, OLevel &L, double *u1, const double *rhs1, double *aux, int smoother) {
  const pmg_level_desc &d = L.d;
  const int N_UDOF = L.N_UDOF, HangingBound = L.HangingNodeBound;
  const double damp = mg.prm.sc_smooth_damp_factor_coarse_saddle;
  const double alpha = mg.prm.sc_smooth_damp_factor_saddle;
  double *u2 = u1 + N_UDOF, *u3 = u2 + N_UDOF, *p = u1 + 3 * N_UDOF;
  const double *rhs2 = rhs1 + N_UDOF, *rhs3 = rhs2 + N_UDOF, *rhsp = rhs1 + 3 * N_UDOF;
  SetDirichletNodes(L, u1, rhs1);                      // :4182-4188
  double *uold = aux;
  std::memcpy(uold, u1, sizeof(double) * L.N_DOF);
  int info_all = 0;
  std::vector<int> UDOFs;
  std::vector<double> System, Rhs;
  std::vector<int> piv;
  for (int i = 0; i < L.N_PDOF; i++) {
    const int begin = d.b_rowptr[i], end = d.b_rowptr[i + 1];
    double value = rhsp[i];
    UDOFs.clear();
    std::vector<double> bl;
    for (int k = begin; k < end; k++) {
      const int l = d.b_kcol[k];
      const double v1 = d.b[0][k], v2 = d.b[1][k], v3 = d.b[2][k];
      value -= v1 * u1[l] + v2 * u2[l] + v3 * u3[l];
      UDOFs.push_back(l);
      bl.push_back(v1); bl.push_back(v2); bl.push_back(v3);
    }
    const int N_U = int(UDOFs.size()), N_U2 = 2 * N_U, N_UGEO = 3 * N_U, N_LocalDOF = N_UGEO + 1;
    System.assign(size_t(N_LocalDOF) * N_LocalDOF, 0.0);
    Rhs.assign(N_LocalDOF, 0.0);
    Rhs[N_LocalDOF - 1] = value;
    for (int k = 0; k < N_U; k++) {                    // local B in the pressure row
      System[size_t(k) * N_LocalDOF + N_UGEO] = bl[3 * k];
      System[size_t(k + N_U) * N_LocalDOF + N_UGEO] = bl[3 * k + 1];
      System[size_t(k + N_U2) * N_LocalDOF + N_UGEO] = bl[3 * k + 2];
    }
    for (int j = 0; j < N_U; j++) {
      const int j1 = j, j2 = j + N_U, j3 = j2 + N_U;
      const int UDOF = UDOFs[j];
      Rhs[j1] = rhs1[UDOF];
      Rhs[j2] = rhs2[UDOF];
      Rhs[j3] = rhs3[UDOF];
      for (int k = d.a_rowptr[UDOF]; k < d.a_rowptr[UDOF + 1]; k++) {
        const int l = d.a_kcol[k];
        double v11 = d.a[0][k], v12 = d.a[1][k], v13 = d.a[2][k], v21 = d.a[3][k], v22 = d.a[4][k], v23 = d.a[5][k],
               v31 = d.a[6][k], v32 = d.a[7][k], v33 = d.a[8][k];
        if (UDOF >= HangingBound) v12 = v13 = v21 = v23 = v31 = v32 = 0;   // Dirichlet node
        Rhs[j1] -= v11 * u1[l] + v12 * u2[l] + v13 * u3[l];
        Rhs[j2] -= v21 * u1[l] + v22 * u2[l] + v23 * u3[l];
        Rhs[j3] -= v31 * u1[l] + v32 * u2[l] + v33 * u3[l];
        for (int m = 0; m < N_U; m++)
          if (UDOFs[m] == l) {
            const size_t k1 = size_t(m) * N_LocalDOF, k2 = size_t(m + N_U) * N_LocalDOF,
                         k3 = size_t(m + 2 * N_U) * N_LocalDOF;
            System[k1 + j1] = v11; System[k2 + j1] = v12; System[k3 + j1] = v13;
            System[k1 + j2] = v21; System[k2 + j2] = v22; System[k3 + j2] = v23;
            System[k1 + j3] = v31; System[k2 + j3] = v32; System[k3 + j3] = v33;
            break;
          }
      }
      if (UDOF < HangingBound) {                        // active dof: BT row
        for (int k = d.bt_rowptr[UDOF]; k < d.bt_rowptr[UDOF + 1]; k++) {
          const int l = d.bt_kcol[k];
          const double v1 = d.bt[0][k], v2 = d.bt[1][k], v3 = d.bt[2][k], pv = p[l];
          Rhs[j1] -= v1 * pv;
          Rhs[j2] -= v2 * pv;
          Rhs[j3] -= v3 * pv;
          if (i == l) {
            const size_t j4 = size_t(N_UGEO) * N_LocalDOF;
            System[j4 + j1] = v1; System[j4 + j2] = v2; System[j4 + j3] = v3;
          }
        }
      }
    }
    if (smoother == 3) {
      SolveDiagonalVanka3D(System.data(), Rhs.data(), N_U, N_LocalDOF);
    } else {
      piv.resize(N_LocalDOF);
      const int info = lu_factor(N_LocalDOF, System.data(), piv.data());
      if (info) info_all = std::max(info_all, i + 1);
      else lu_solve(N_LocalDOF, System.data(), piv.data(), Rhs.data());
    }
    for (int j = 0; j < N_U; j++) {
      const int l = UDOFs[j];
      u1[l] += damp * Rhs[j];
      u2[l] += damp * Rhs[j + N_U];
      u3[l] += damp * Rhs[j + N_U2];
    }
    p[i] += damp * Rhs[N_UGEO];
  }
  for (int j = 0; j < L.N_DOF; j++) u1[j] = uold[j] + alpha * (u1[j] - uold[j]);
  SetDirichletNodes(L, u1, rhs1);
  return info_all;
}

// children of every coarse cell in child order, from the fine level's parent/child maps
// (cell->GetParent / parent->GetChild(j))
void ChildrenOf(const OLevel &F, int n_coarse, std::vector<int> &kids) {
  const int nch = F.d.n_children;
  kids.assign(size_t(n_coarse) * nch, -1);
  for (int c = 0; c < F.d.n_cells; ++c) kids[size_t(F.d.parent[c]) * nch + F.d.child[c]] = c;
}

// Prolongate, N_Functions components: MGComponents3D.C:468-708 (U, 3 functions) and
// :184-466 (P, 1 function).  Cells visited in fine collection order, each parent's
// children processed together in child order the first time one of them is met.
void ProlongateSpace(const OLevel &C, const OLevel &F, int N_Functions, bool velocity,
                     const double *CoarseFunction, double *FineFunction, double *aux) {
  const pmg_level_desc &cd = C.d, &fd = F.d;
  const int nloc_c = velocity ? cd.n_ulocal : cd.n_plocal;
  const int nloc_f = velocity ? fd.n_ulocal : fd.n_plocal;
  const int N_CoarseDOFs = velocity ? C.N_UDOF : C.N_PDOF;
  const int N_FineDOFs = velocity ? F.N_UDOF : F.N_PDOF;
  const double *QQall = velocity ? fd.qq_u : fd.qq_p;
  const int *cbeg = velocity ? cd.u_begin : cd.p_begin, *cglob = velocity ? cd.u_global : cd.p_global;
  const int *fbeg = velocity ? fd.u_begin : fd.p_begin, *fglob = velocity ? fd.u_global : fd.p_global;
  std::memset(aux, 0, sizeof(double) * N_FineDOFs);
  std::memset(FineFunction, 0, sizeof(double) * N_Functions * N_FineDOFs);
  std::vector<int> kids;
  ChildrenOf(F, cd.n_cells, kids);
  std::vector<char> done(fd.n_cells, 0);
  double Val[64], Val2[64];
  for (int i = 0; i < fd.n_cells; i++) {
    if (done[i]) continue;
    int CoarseNumber = fd.parent[i];
    for (int j = 0; j < fd.n_children; j++) {
      int FineNumber = kids[size_t(CoarseNumber) * fd.n_children + j];
      done[FineNumber] = 1;
      const double *QQ = QQall + size_t(j) * nloc_f * nloc_c;
      const int *FineDOF = fglob + fbeg[FineNumber];
      const int *CoarseDOF = cglob + cbeg[CoarseNumber];
      for (int IFunct = 0; IFunct < N_Functions; IFunct++) {
        int CoarseOffset = IFunct * N_CoarseDOFs, FineOffset = IFunct * N_FineDOFs;
        for (int l = 0; l < nloc_c; l++) Val[l] = CoarseFunction[CoarseOffset + CoarseDOF[l]];
        for (int k = 0; k < nloc_f; k++) {
          double s = 0;
          const double *entry = QQ + k * nloc_c;
          for (int l = 0; l < nloc_c; l++) s += entry[l] * Val[l];
          Val2[k] = s;
        }
        for (int k = 0; k < nloc_f; k++) {
          int Index = FineDOF[k];
          FineFunction[FineOffset + Index] += Val2[k];
          aux[Index] += 1;
        }
      }
    }
  }
  Dscal(N_FineDOFs, 1.0 / N_Functions, aux);
  for (int IFunct = 0; IFunct < N_Functions; IFunct++) {
    int FineOffset = IFunct * N_FineDOFs;
    for (int i = 0; i < N_FineDOFs; i++)
      if (aux[i]) FineFunction[FineOffset + i] /= aux[i];
  }
}

// TFEDatabase3D::GetRestrictionMatrix3D for Q2 hexahedra and the regular refinement
// (FEDatabase3D.C:1791-1970): G = coarse basis mass on the reference cube, R_j[l2][l1] =
// sum over the quadrature points of child j of w |det| phi_coarse_l1(X) phi_fine_l2(xi),
// QQ_j[l1][l2] = (G^{-1} R_j[l2])[l1].  Quadrature: 3 x 3 x 3 Gauss points (exact for
// these degree-4 products).  Child j covers the octant of the reference cube with centre
// ((j & 1) ? 0.5 : -0.5, (j & 2) ? 0.5 : -0.5, (j & 4) ? 0.5 : -0.5), as the children of
// the level descriptors are numbered (the prolongation tables fix the same octants).
void Q2Values(double x, double y, double z, double *v) {
  auto l = [](double t, double *o) { o[0] = 0.5 * t * (t - 1); o[1] = 1 - t * t; o[2] = 0.5 * t * (t + 1); };
  double a[3], b[3], c[3];
  l(x, a); l(y, b); l(z, c);
  for (int k = 0; k < 27; ++k) v[k] = a[k % 3] * b[(k / 3) % 3] * c[k / 9];
}
const std::vector<double> &Q2RestrictionMatrices() {
  static std::vector<double> qq;
  if (!qq.empty()) return qq;
  const double gp[3] = {-std::sqrt(0.6), 0.0, std::sqrt(0.6)}, gw[3] = {5.0 / 9, 8.0 / 9, 5.0 / 9};
  std::vector<double> G(27 * 27, 0.0);
  double v[27], vc[27];
  for (int q = 0; q < 27; ++q) {
    const double w = gw[q % 3] * gw[(q / 3) % 3] * gw[q / 9];
    Q2Values(gp[q % 3], gp[(q / 3) % 3], gp[q / 9], v);
    for (int i = 0; i < 27; ++i)
      for (int j = 0; j < 27; ++j) G[i * 27 + j] += w * v[i] * v[j];
  }
  std::vector<int> piv(27);
  std::vector<double> Gf(G);   // symmetric: row- and column-major agree
  lu_factor(27, Gf.data(), piv.data());
  qq.assign(8 * 27 * 27, 0.0);
  for (int ch = 0; ch < 8; ++ch) {
    const double s[3] = {(ch & 1) ? 0.5 : -0.5, (ch & 2) ? 0.5 : -0.5, (ch & 4) ? 0.5 : -0.5};
    std::vector<double> R(27 * 27, 0.0);   // R[l2][l1]
    for (int q = 0; q < 27; ++q) {
      const double w = gw[q % 3] * gw[(q / 3) % 3] * gw[q / 9] / 8.0;   // |det| of the child map
      const double xi = gp[q % 3], eta = gp[(q / 3) % 3], zeta = gp[q / 9];
      Q2Values(xi, eta, zeta, v);                                                // fine basis
      Q2Values(s[0] + 0.5 * xi, s[1] + 0.5 * eta, s[2] + 0.5 * zeta, vc);        // coarse basis
      for (int l1 = 0; l1 < 27; ++l1)
        for (int l2 = 0; l2 < 27; ++l2) R[l2 * 27 + l1] += w * vc[l1] * v[l2];
    }
    for (int l2 = 0; l2 < 27; ++l2) {
      lu_solve(27, Gf.data(), piv.data(), &R[l2 * 27]);
      for (int l1 = 0; l1 < 27; ++l1) qq[size_t(ch) * 729 + l1 * 27 + l2] = R[l2 * 27 + l1];
    }
  }
  return qq;
}

// RestrictFunction, N_Functions = 3 (velocity): MGComponents3D.C:1509-1755.  Parents are
// visited when the first of their children is met in fine collection order; a coarse
// dof's value is the sum over its cells divided by the number of cells (aux).
void RestrictFunctionU(const OLevel &C, const OLevel &F, const double *fine, double *coarse, double *aux) {
  const std::vector<double> &QQ = Q2RestrictionMatrices();
  const int Nc = C.N_UDOF, Nf = F.N_UDOF;
  std::memset(aux, 0, sizeof(double) * Nc);
  std::memset(coarse, 0, sizeof(double) * 3 * Nc);
  std::vector<int> kids;
  ChildrenOf(F, C.d.n_cells, kids);
  std::vector<char> done(C.d.n_cells, 0);
  double Val[27], Val2[3 * 27];
  for (int i = 0; i < F.d.n_cells; i++) {
    const int pc = F.d.parent[i];
    if (done[pc]) continue;
    done[pc] = 1;
    std::memset(Val2, 0, sizeof(Val2));
    for (int j = 0; j < F.d.n_children; j++) {
      const int fc = kids[size_t(pc) * F.d.n_children + j];
      const int *FineDOF = F.d.u_global + F.d.u_begin[fc];
      for (int IFunct = 0; IFunct < 3; IFunct++) {
        for (int l = 0; l < 27; l++) Val[l] = fine[IFunct * Nf + FineDOF[l]];
        for (int k = 0; k < 27; k++) {
          double s = 0;
          for (int l = 0; l < 27; l++) s += QQ[size_t(j) * 729 + k * 27 + l] * Val[l];
          Val2[IFunct * 27 + k] += s;
        }
      }
    }
    const int *CoarseDOF = C.d.u_global + C.d.u_begin[pc];
    for (int IFunct = 0; IFunct < 3; IFunct++)
      for (int k = 0; k < 27; k++) {
        const int l = CoarseDOF[k];
        aux[l] += 1;
        coarse[IFunct * Nc + l] += Val2[IFunct * 27 + k];
      }
  }
  Dscal(Nc, 1.0 / 3, aux);
  for (int IFunct = 0; IFunct < 3; IFunct++)
    for (int i = 0; i < Nc; i++) coarse[IFunct * Nc + i] /= aux[i];
}

// DefectRestriction: MGComponents3D.C:996-1240 (U, 3 functions) and :711-993 (P)
void RestrictSpace(const OLevel &C, const OLevel &F, int N_Functions, bool velocity,
                   double *CoarseFunction, double *FineFunction, double *aux) {
  const pmg_level_desc &cd = C.d, &fd = F.d;
  const int nloc_c = velocity ? cd.n_ulocal : cd.n_plocal;
  const int nloc_f = velocity ? fd.n_ulocal : fd.n_plocal;
  const int N_CoarseDOFs = velocity ? C.N_UDOF : C.N_PDOF;
  const int N_FineDOFs = velocity ? F.N_UDOF : F.N_PDOF;
  const double *QQall = velocity ? fd.qq_u : fd.qq_p;
  const int *cbeg = velocity ? cd.u_begin : cd.p_begin, *cglob = velocity ? cd.u_global : cd.p_global;
  const int *fbeg = velocity ? fd.u_begin : fd.p_begin, *fglob = velocity ? fd.u_global : fd.p_global;
  std::memset(aux, 0, sizeof(double) * N_FineDOFs);
  std::memset(CoarseFunction, 0, sizeof(double) * N_CoarseDOFs * N_Functions);
  for (int i = 0; i < fd.n_cells; i++) {
    const int *DOF = fglob + fbeg[i];
    for (int j = 0; j < nloc_f; j++) aux[DOF[j]] += 1;
  }
  for (int IFunct = 0; IFunct < N_Functions; IFunct++) {
    int FineOffset = IFunct * N_FineDOFs;
    for (int i = 0; i < N_FineDOFs; i++) FineFunction[FineOffset + i] /= aux[i];
  }
  std::vector<int> kids;
  ChildrenOf(F, cd.n_cells, kids);
  std::vector<char> done(fd.n_cells, 0);
  double Val[64], Val2[64];
  for (int i = 0; i < fd.n_cells; i++) {
    if (done[i]) continue;
    int CoarseNumber = fd.parent[i];
    for (int j = 0; j < fd.n_children; j++) {
      int FineNumber = kids[size_t(CoarseNumber) * fd.n_children + j];
      done[FineNumber] = 1;
      const double *QQ = QQall + size_t(j) * nloc_f * nloc_c;
      const int *FineDOF = fglob + fbeg[FineNumber];
      const int *CoarseDOF = cglob + cbeg[CoarseNumber];
      for (int IFunct = 0; IFunct < N_Functions; IFunct++) {
        int FineOffset = IFunct * N_FineDOFs, CoarseOffset = IFunct * N_CoarseDOFs;
        for (int l = 0; l < nloc_f; l++) Val[l] = FineFunction[FineOffset + FineDOF[l]];
        for (int k = 0; k < nloc_c; k++) {
          double s = 0;
          for (int l = 0; l < nloc_f; l++) s += QQ[l * nloc_c + k] * Val[l];
          Val2[k] = s;
        }
        for (int k = 0; k < nloc_c; k++) CoarseFunction[CoarseOffset + CoarseDOF[k]] += Val2[k];
      }
    }
  }
  for (int IFunct = 0; IFunct < N_Functions; IFunct++) {
    int FineOffset = IFunct * N_FineDOFs;
    for (int i = 0; i < N_FineDOFs; i++) FineFunction[FineOffset + i] *= aux[i];
  }
}

// TNSE_MGLevel4::SolveExact: NSE_MGLevel4.C:5728-5846 (dense column-major a[col*N+row])
void BuildCoarseDense(const oracle_mg &mg, const OLevel &L, std::vector<double> &a) {
  const pmg_level_desc &d = L.d;
  const size_t N_DOF = L.N_DOF;
  const int N_UDOF = L.N_UDOF, N_UDOF2 = 2 * N_UDOF, N_UDOFGEO = GEO_DIM * N_UDOF;
  a.assign(N_DOF * N_DOF, 0.0);
  int j = d.a_rowptr[0];
  for (int i = 0; i < L.N_Active; i++) {
    int k = d.a_rowptr[i + 1];
    for (; j < k; j++) {
      size_t index = d.a_kcol[j];
      a[index * N_DOF + i] = d.a[0][j];
      a[(index + N_UDOF) * N_DOF + i] = d.a[1][j];
      a[index * N_DOF + i + N_UDOF] = d.a[3][j];
      a[(index + N_UDOF) * N_DOF + i + N_UDOF] = d.a[4][j];
      a[(index + N_UDOF2) * N_DOF + i] = d.a[2][j];
      a[(index + N_UDOF2) * N_DOF + i + N_UDOF] = d.a[5][j];
      a[index * N_DOF + i + N_UDOF2] = d.a[6][j];
      a[(index + N_UDOF) * N_DOF + i + N_UDOF2] = d.a[7][j];
      a[(index + N_UDOF2) * N_DOF + i + N_UDOF2] = d.a[8][j];
    }
  }
  for (int i = L.N_Active; i < N_UDOF; i++) {
    int k = d.a_rowptr[i + 1];
    for (; j < k; j++) {
      size_t index = d.a_kcol[j];
      a[i + N_DOF * index] = d.a[0][j];
      a[(index + N_UDOF) * N_DOF + i + N_UDOF] = d.a[4][j];
      a[(index + N_UDOF2) * N_DOF + i + N_UDOF2] = d.a[8][j];
    }
  }
  j = d.bt_rowptr[0];
  for (int i = 0; i < L.N_Active; i++) {
    int k = d.bt_rowptr[i + 1];
    for (; j < k; j++) {
      size_t l = (size_t(d.bt_kcol[j]) + N_UDOFGEO) * N_DOF;
      a[l + i] = d.bt[0][j];
      a[l + i + N_UDOF] = d.bt[1][j];
      a[l + i + 2 * N_UDOF] = d.bt[2][j];
    }
  }
  j = d.b_rowptr[0];
  for (int i = 0; i < L.N_PDOF; i++) {
    int k = d.b_rowptr[i + 1];
    for (; j < k; j++) {
      size_t index = d.b_kcol[j];
      a[index * N_DOF + i + N_UDOFGEO] = d.b[0][j];
      a[(index + N_UDOF) * N_DOF + i + N_UDOFGEO] = d.b[1][j];
      a[(index + N_UDOF2) * N_DOF + i + N_UDOFGEO] = d.b[2][j];
    }
  }
  if (mg.prm.internal_project_pressure) {
    for (size_t i = 0; i < N_DOF; i++) a[i * N_DOF + N_UDOFGEO] = 0;
    a[size_t(N_UDOFGEO) * N_DOF + N_UDOFGEO] = 1;
  }
}

int SolveExact(oracle_mg &mg, OLevel &L, double *u1, const double *rhs1) {
  const int N = L.N_DOF;
  std::vector<double> b(rhs1, rhs1 + N);
  if (mg.prm.internal_project_pressure) b[GEO_DIM * L.N_UDOF] = 0;
  int info = 0;
  if (mg.refactor) {
    std::vector<double> a;
    std::vector<int> ipiv(N);
    BuildCoarseDense(mg, L, a);
    info = lu_factor(N, a.data(), ipiv.data());
    lu_solve(N, a.data(), ipiv.data(), b.data());
  } else {
    if (!mg.coarse_ready) {
      BuildCoarseDense(mg, L, mg.coarse_lu);
      mg.coarse_piv.resize(N);
      info = lu_factor(N, mg.coarse_lu.data(), mg.coarse_piv.data());
      mg.coarse_ready = true;
    }
    lu_solve(N, mg.coarse_lu.data(), mg.coarse_piv.data(), b.data());
  }
  std::memcpy(u1, b.data(), N * sizeof(double));
  return info;
}

int CellVanka(oracle_mg &mg, OLevel &L, double *u, const double *f, double *aux) {
  // smoother dispatch of Cycle (NSE_MultiGrid.C:1193-1329): 3/4 nodal Vanka
  if (mg.prm.sc_smoother_saddle == 3 || mg.prm.sc_smoother_saddle == 4)
    return NodalVanka(mg, L, u, f, aux, mg.prm.sc_smoother_saddle);
  return mg.variant == 0 ? CellVankaColored(mg, L, u, f, aux) : CellVankaGS(mg, L, u, f, aux);
}

// TNSE_MultiGrid::Cycle: NSE_MultiGrid.C:434-1779 (smoother 1/2, coarse smoother 17,
// no step length control; the recursion loop runs once, :1459-1460)
int Cycle(oracle_mg &mg, int i, double &res) {
  OLevel &Cur = mg.L[i];
  double *CurrentU = Cur.sol.data(), *CurrentRhsU = Cur.rhs.data();
  double *CurrentDefectU = Cur.aux0.data(), *CurrentAux = Cur.aux1.data();
  int status = 0;
  CorrectNodes(Cur, CurrentU);                                  // :501
  SetDirichletNodes(Cur, CurrentU, CurrentRhsU);                // :502
  Defect(mg, Cur, CurrentU, CurrentRhsU, CurrentAux, res);      // :503
  mg.cyclelog.push_back(res);
  if (i == 0) {
    status = SolveExact(mg, Cur, CurrentU, CurrentRhsU);        // :546-557
    Defect(mg, Cur, CurrentU, CurrentRhsU, CurrentDefectU, res);
    mg.cyclelog.push_back(res);
    return status ? PMG_ERR_SINGULAR_COARSE : 0;
  }
  OLevel &Coa = mg.L[i - 1];
  for (int j = 0; j < mg.prm.sc_pre_smooth_saddle; j++)        // :1195-1203
    if (CellVanka(mg, Cur, CurrentU, CurrentRhsU, CurrentAux)) status = PMG_ERR_SINGULAR_PATCH;
  CorrectNodes(Cur, CurrentU);                                  // :1340
  Defect(mg, Cur, CurrentU, CurrentRhsU, CurrentDefectU, res);  // :1341
  mg.cyclelog.push_back(res);
  // :1407-1408
  RestrictSpace(Coa, Cur, GEO_DIM, true, Coa.rhs.data(), CurrentDefectU, CurrentAux);
  RestrictSpace(Coa, Cur, 1, false, Coa.rhs.data() + GEO_DIM * Coa.N_UDOF,
                CurrentDefectU + GEO_DIM * Cur.N_UDOF, CurrentAux);
  CorrectDefect(mg, Coa, Coa.rhs.data());                       // :1449
  std::memset(Coa.sol.data(), 0, sizeof(double) * Coa.N_DOF);   // :1450 Reset
  int st = Cycle(mg, i - 1, res);                               // :1459-1460
  if (st) status = st;
  // :1510-1514
  ProlongateSpace(Coa, Cur, GEO_DIM, true, Coa.sol.data(), CurrentDefectU, CurrentAux);
  ProlongateSpace(Coa, Cur, 1, false, Coa.sol.data() + GEO_DIM * Coa.N_UDOF,
                  CurrentDefectU + GEO_DIM * Cur.N_UDOF, CurrentAux);
  CorrectNodes(Cur, CurrentDefectU);                            // :1548
  Update(mg, Cur, CurrentU, CurrentDefectU);                    // :1549
  Defect(mg, Cur, CurrentU, CurrentRhsU, CurrentDefectU, res);  // :1550
  mg.cyclelog.push_back(res);
  CorrectNodes(Cur, CurrentU);                                  // :1580
  for (int j = 0; j < mg.prm.sc_post_smooth_saddle; j++)       // :1610-1618
    if (CellVanka(mg, Cur, CurrentU, CurrentRhsU, CurrentAux)) status = PMG_ERR_SINGULAR_PATCH;
  Defect(mg, Cur, CurrentU, CurrentRhsU, CurrentDefectU, res);  // :1753
  mg.cyclelog.push_back(res);
  return status;
}

// TMultiGridIte::Iterate with Zero_Start: MultiGridIte.C:76-130
int MGIterate(oracle_mg &mg, double *sol, const double *rhs) {
  OLevel &F = mg.L.back();
  std::memset(F.sol.data(), 0, sizeof(double) * F.N_DOF);
  std::memcpy(F.rhs.data(), rhs, sizeof(double) * F.N_DOF);
  SetDirichletNodes(F, F.sol.data(), F.rhs.data());
  double res;
  mg.cyclelog.clear();
  int st = Cycle(mg, int(mg.L.size()) - 1, res);
  std::memcpy(sol, F.sol.data(), sizeof(double) * F.N_DOF);
  return st;
}

// GeneratePlaneRotation / ApplyPlaneRotation / UpdateGmresIterate: FgmresIte.C:199-247
void GeneratePlaneRotation(double dx, double dy, double *cs, double *sn) {
  if (dy == 0.0) {
    cs[0] = 1.0; sn[0] = 0.0;
  } else if (std::fabs(dy) > std::fabs(dx)) {
    double temp = dx / dy;
    sn[0] = 1.0 / std::sqrt(1.0 + temp * temp);
    cs[0] = temp * sn[0];
  } else {
    double temp = dy / dx;
    cs[0] = 1.0 / std::sqrt(1.0 + temp * temp);
    sn[0] = temp * cs[0];
  }
}
void ApplyPlaneRotation(double *dx, double *dy, double cs, double sn) {
  double temp = cs * dx[0] + sn * dy[0];
  dy[0] = -sn * dx[0] + cs * dy[0];
  dx[0] = temp;
}
void UpdateGmresIterate(double *x, int Len_x, int k, std::vector<std::vector<double>> &h, double *s,
                        std::vector<std::vector<double>> &v) {
  for (int i = k; i >= 0; i--) {
    s[i] /= h[i][i];
    for (int j = i - 1; j >= 0; j--) s[j] -= h[j][i] * s[i];
  }
  for (int j = 0; j <= k; j++)
    for (int i = 0; i < Len_x; i++) x[i] += v[j][i] * s[j];
}

// TFgmresIte::Iterate, flexible branch: FgmresIte.C:250-677
int Fgmres(oracle_mg &mg, double *sol, const double *rhs, int *iters, double *final_res) {
  OLevel &F = mg.L.back();
  const int N_DOF = F.N_DOF;
  const pmg_params &P = mg.prm;
  int maxit = P.sc_lin_maxit_saddle;
  int restart = P.sc_gmres_restart;
  if (restart > maxit) restart = maxit;
  if (restart <= 0) restart = 10;
  const double res_norm_min = P.sc_lin_res_norm_min_saddle, red_factor = P.sc_lin_red_factor_saddle;
  const double div_factor = P.sc_div_factor;
  const int minit = P.sc_minit, prec_maxit = std::max(1, P.sc_amg_prec_it);
  std::vector<double> s(restart + 1), cosi(restart + 1), sn(restart + 1);
  std::vector<std::vector<double>> H(restart + 1, std::vector<double>(restart, 0.0));
  std::vector<std::vector<double>> v(restart + 1, std::vector<double>(N_DOF, 0.0));
  std::vector<std::vector<double>> zv(restart + 1, std::vector<double>(N_DOF, 0.0));
  std::vector<double> defect(N_DOF);
  mg.hist.clear();
  auto matvecdefect = [&](const double *x, const double *b, double *r) {  // Defect_NSE4
    CoupledDefect(F, x, b, r);
    if (P.internal_project_pressure) IntoL20Vector3D(r + GEO_DIM * F.N_UDOF, F.N_PDOF);
  };
  int status = 0;
  matvecdefect(sol, rhs, v[0].data());
  double res = std::sqrt(Ddot(N_DOF, v[0].data(), v[0].data()));
  double res0 = res, beta = res;
  int maxite = maxit;
  if (!std::isfinite(res)) { *iters = 0; *final_res = res; return PMG_ERR_NONFINITE; }
  if (beta <= res_norm_min) {
    if (minit == 0 || beta < 1e-20) {
      *iters = 0; *final_res = res; return PMG_OK;
    } else {
      maxite = minit;
    }
  }
  mg.hist.push_back(beta);
  int j = 1, i = 0;
  while (j <= maxite) {
    Dscal(N_DOF, 1.0 / beta, v[0].data());
    std::fill(s.begin(), s.end(), 0.0);
    s[0] = beta;
    for (i = 0; i < restart && j <= maxite; i++, j++) {
      std::fill(zv[i].begin(), zv[i].end(), 0.0);
      for (int k = 0; k < prec_maxit; k++) {
        int st = MGIterate(mg, zv[i].data(), v[i].data());
        if (st) status = st;
        if (k < prec_maxit - 1) {
          matvecdefect(zv[i].data(), v[i].data(), defect.data());
          res = res0 = std::sqrt(Ddot(N_DOF, defect.data(), defect.data()));
        }
      }
      CoupledMatVect(F, zv[i].data(), defect.data());
      for (int k = 0; k <= i; k++) {
        H[k][i] = Ddot(N_DOF, defect.data(), v[k].data());
        Daxpy(N_DOF, -H[k][i], v[k].data(), defect.data());
      }
      H[i + 1][i] = std::sqrt(Ddot(N_DOF, defect.data(), defect.data()));
      std::memcpy(v[i + 1].data(), defect.data(), sizeof(double) * N_DOF);
      if (H[i + 1][i] != 0) Dscal(N_DOF, 1 / H[i + 1][i], v[i + 1].data());
      for (int k = 0; k < i; k++) ApplyPlaneRotation(&H[k][i], &H[k + 1][i], cosi[k], sn[k]);
      GeneratePlaneRotation(H[i][i], H[i + 1][i], &cosi[i], &sn[i]);
      ApplyPlaneRotation(&H[i][i], &H[i + 1][i], cosi[i], sn[i]);
      ApplyPlaneRotation(&s[i], &s[i + 1], cosi[i], sn[i]);
      res = std::fabs(s[i + 1]);
      mg.hist.push_back(res);
      if (!std::isfinite(res)) { *iters = j; *final_res = res; return PMG_ERR_NONFINITE; }
      if ((j >= minit) && ((res < res_norm_min) || (res < red_factor * res0))) {
        std::fill(defect.begin(), defect.end(), 0.0);
        UpdateGmresIterate(defect.data(), N_DOF, i, H, s.data(), zv);
        Daxpy(N_DOF, 1.0, defect.data(), sol);
        *iters = j; *final_res = res;
        return status;
      }
      if (res > div_factor * res0) { *iters = j; *final_res = res; return PMG_ERR_DIVERGED; }
    }
    std::fill(v[0].begin(), v[0].end(), 0.0);
    UpdateGmresIterate(v[0].data(), N_DOF, restart - 1, H, s.data(), zv);
    Daxpy(N_DOF, 1.0, v[0].data(), sol);
    matvecdefect(sol, rhs, v[0].data());
    res = res0 = beta = std::sqrt(Ddot(N_DOF, v[0].data(), v[0].data()));
    mg.hist.push_back(res);
  }
  if (j > maxite) j = maxite;
  *iters = j; *final_res = res;
  if (res > res_norm_min && res > res0 * red_factor) return status ? status : PMG_ERR_NOT_CONVERGED;
  return status;
}

void BuildColoringBrute(OLevel &L) {
  std::vector<int> ptr(L.d.n_cells + 1), reo(L.d.n_cells);
  int nc = 0;
  oracle_color_cells(L.d.n_cells, L.d.n_ulocal, L.d.u_begin, L.d.u_global, &nc, ptr.data(), reo.data());
  L.N_CIntCell = nc;
  L.ptrCellColors.assign(ptr.begin(), ptr.begin() + nc + 1);
  L.CellReorder = reo;
}

}  // namespace

extern "C" {

void oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int oracle_create(const pmg_params *params, const pmg_level_desc *levels, int n_levels, int refactor,
                  oracle_mg **out) {
  if (!params || !levels || n_levels < 1 || !out) return PMG_ERR_INVALID_ARGUMENT;
  auto *mg = new oracle_mg();
  mg->prm = *params;
  mg->refactor = refactor;
  mg->L.resize(n_levels);
  for (int l = 0; l < n_levels; ++l) {
    OLevel &L = mg->L[l];
    L.d = levels[l];
    L.N_UDOF = L.d.n_udof; L.N_PDOF = L.d.n_pdof; L.N_Active = L.d.n_active;
    L.HangingNodeBound = L.d.hanging_bound; L.N_Dirichlet = L.d.n_dirichlet;
    L.N_DOF = GEO_DIM * L.N_UDOF + L.N_PDOF;
    L.N_U = L.d.n_ulocal; L.N_P = L.d.n_plocal; L.N_LocalDOF = GEO_DIM * L.N_U + L.N_P;
    if (L.d.ptr_colors && L.d.cell_reorder && L.d.n_colors > 0) {
      L.N_CIntCell = L.d.n_colors;
      L.ptrCellColors.assign(L.d.ptr_colors, L.d.ptr_colors + L.d.n_colors + 1);
      L.CellReorder.assign(L.d.cell_reorder, L.d.cell_reorder + L.d.n_cells);
    } else {
      BuildColoringBrute(L);
    }
    L.maxCellsPerColor = 0;
    for (int c = 0; c < L.N_CIntCell; ++c)
      L.maxCellsPerColor = std::max(L.maxCellsPerColor, L.ptrCellColors[c + 1] - L.ptrCellColors[c]);
    L.sol.assign(L.N_DOF, 0.0); L.rhs.assign(L.N_DOF, 0.0);
    L.aux0.assign(L.N_DOF, 0.0); L.aux1.assign(L.N_DOF, 0.0);
    if (!refactor) {  // every level can be smoothed (the coarsest one too)
      L.lu.assign(size_t(L.d.n_cells) * L.N_LocalDOF * L.N_LocalDOF, 0.0);
      L.piv.assign(size_t(L.d.n_cells) * L.N_LocalDOF, 0);
      L.lu_ready.assign(L.d.n_cells, 0);
    }
  }
  *out = mg;
  return PMG_OK;
}

void oracle_destroy(oracle_mg *mg) { delete mg; }
int oracle_n_dof(const oracle_mg *mg, int level) { return mg->L[level].N_DOF; }
void oracle_set_vanka_variant(oracle_mg *mg, int variant) { mg->variant = variant; }

void oracle_defect(oracle_mg *mg, int level, const double *u, const double *f, double *d, double *res) {
  double r;
  Defect(*mg, mg->L[level], u, f, d, r);
  if (res) *res = r;
}
void oracle_matvec(oracle_mg *mg, int level, const double *x, double *y) { CoupledMatVect(mg->L[level], x, y); }
void oracle_defect_nse4(oracle_mg *mg, int level, const double *x, const double *b, double *r) {
  const OLevel &L = mg->L[level];
  CoupledDefect(L, x, b, r);
  if (mg->prm.internal_project_pressure) IntoL20Vector3D(r + GEO_DIM * L.N_UDOF, L.N_PDOF);
}
int oracle_cell_vanka(oracle_mg *mg, int level, double *u, const double *f, double *aux) {
  return CellVanka(*mg, mg->L[level], u, f, aux);
}
void oracle_correct_nodes(oracle_mg *mg, int level, double *u) { CorrectNodes(mg->L[level], u); }
void oracle_set_dirichlet(oracle_mg *mg, int level, double *u, const double *f) {
  SetDirichletNodes(mg->L[level], u, f);
}
void oracle_correct_defect(oracle_mg *mg, int level, double *d) { CorrectDefect(*mg, mg->L[level], d); }
void oracle_update(oracle_mg *mg, int level, double *u, const double *d) { Update(*mg, mg->L[level], u, d); }
void oracle_restrict_defect(oracle_mg *mg, int fine_level, double *d_fine, double *f_coarse, double *aux) {
  OLevel &F = mg->L[fine_level], &C = mg->L[fine_level - 1];
  RestrictSpace(C, F, GEO_DIM, true, f_coarse, d_fine, aux);
  RestrictSpace(C, F, 1, false, f_coarse + GEO_DIM * C.N_UDOF, d_fine + GEO_DIM * F.N_UDOF, aux);
}
void oracle_prolongate(oracle_mg *mg, int fine_level, const double *u_coarse, double *d_fine, double *aux) {
  OLevel &F = mg->L[fine_level], &C = mg->L[fine_level - 1];
  ProlongateSpace(C, F, GEO_DIM, true, u_coarse, d_fine, aux);
  ProlongateSpace(C, F, 1, false, u_coarse + GEO_DIM * C.N_UDOF, d_fine + GEO_DIM * F.N_UDOF, aux);
}
int oracle_solve_exact(oracle_mg *mg, double *u, const double *f) { return SolveExact(*mg, mg->L[0], u, f); }
void oracle_restrict_function(oracle_mg *mg, int fine_level, const double *u_fine, double *u_coarse, double *aux) {
  RestrictFunctionU(mg->L[fine_level - 1], mg->L[fine_level], u_fine, u_coarse, aux);
}
int oracle_vcycle(oracle_mg *mg, const double *v, double *z) { return MGIterate(*mg, z, v); }
int oracle_fgmres(oracle_mg *mg, double *sol, const double *rhs, int *iters, double *final_res) {
  return Fgmres(*mg, sol, rhs, iters, final_res);
}
int oracle_fgmres_history(const oracle_mg *mg, double *out, int cap) {
  int n = std::min<int>(cap, int(mg->hist.size()));
  for (int i = 0; i < n; ++i) out[i] = mg->hist[i];
  return n;
}
int oracle_cycle_log(const oracle_mg *mg, double *out, int cap) {
  int n = std::min<int>(cap, int(mg->cyclelog.size()));
  for (int i = 0; i < n; ++i) out[i] = mg->cyclelog[i];
  return n;
}
int oracle_lu_factor(int n, double *a, int *ipiv) { return lu_factor(n, a, ipiv); }
void oracle_lu_solve(int n, const double *a, const int *ipiv, double *b) { lu_solve(n, a, ipiv, b); }

// colorCell: ParFEMapper3D.C:750-1084 with the pairwise match() adjacency (:725-748)
int oracle_color_cells(int n_cells, int n_ulocal, const int *u_begin, const int *u_global, int *n_colors,
                       int *ptr_colors, int *cell_reorder) {
  const int N_OwnCells = n_cells;
  std::vector<std::vector<int>> adjlist(N_OwnCells);
#pragma omp parallel for schedule(dynamic, 8)
  for (int pp = 0; pp < N_OwnCells; pp++) {
    const int *l1 = u_global + u_begin[pp];
    for (int qq = 0; qq < N_OwnCells; qq++) {
      const int *l2 = u_global + u_begin[qq];
      bool m = false;
      for (int a = 0; a < n_ulocal && !m; a++)
        for (int b = 0; b < n_ulocal; b++)
          if (l1[a] == l2[b]) { m = true; break; }
      if (m) adjlist[pp].push_back(qq);
    }
  }
  std::vector<int> allocatedColor(n_cells, -1);
  std::vector<char> avail(N_OwnCells + 1, 0);
  int numColors = 0;
  for (int ii = 0; ii < N_OwnCells; ii++) {
    for (int nb : adjlist[ii])
      if (allocatedColor[nb] != -1) avail[allocatedColor[nb]] = 1;
    int cr;
    for (cr = 0; cr < N_OwnCells; cr++)
      if (!avail[cr]) break;
    allocatedColor[ii] = cr;
    for (int nb : adjlist[ii])
      if (allocatedColor[nb] != -1) avail[allocatedColor[nb]] = 0;
    if (numColors < allocatedColor[ii]) numColors = allocatedColor[ii];
  }
  numColors++;
  int k = 0;
  for (int i = 0; i < numColors; i++) {
    ptr_colors[i] = k;
    for (int j = 0; j < n_cells; j++)
      if (allocatedColor[j] == i) { cell_reorder[k++] = j; allocatedColor[j] = -1; }
  }
  ptr_colors[numColors] = k;
  *n_colors = numColors;
  return PMG_OK;
}

int oracle_patch_matrix(oracle_mg *mg, int level, int cell, double *sys) {
  PatchMatrix(mg->L[level], cell, sys);
  return PMG_OK;
}

}  // extern "C"
