// pmg_oracle2d.cpp - FP64 CPU restatement of the 2D twins of the NSE type-4 GMG-Vanka /
// FGMRES path and the generator of its test problems (see pmg_oracle2d.h).  Test
// infrastructure only.
#include "pmg_oracle2d.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

extern "C" int oracle_lu_factor(int n, double *a, int *ipiv);
extern "C" void oracle_lu_solve(int n, const double *a, const int *ipiv, double *b);
extern "C" int oracle_color_cells(int n_cells, int n_ulocal, const int *u_begin, const int *u_global, int *n_colors,
                                  int *ptr_colors, int *cell_reorder);

namespace {

constexpr int GEO_DIM = 2;
constexpr int N_U = 9;                          // Q2 quadrilateral
constexpr int N_P = 3;                          // P1disc quadrilateral: 1, 3 xi, 3 eta
constexpr int N_LocalDOF = GEO_DIM * N_U + N_P; // 21
constexpr int N_CHILD = 4;

struct Level2 {
  int n = 0;                       // cells per direction
  int N_Cells = 0, N_UDOF = 0, N_Active = 0, N_PDOF = 0, N_DOF = 0;
  std::vector<int> a_rowptr, a_kcol, b_rowptr, b_kcol, bt_rowptr, bt_kcol;
  std::vector<double> a[4], b[2], bt[2];   // A11 A12 A21 A22; B1 B2; B1T B2T
  std::vector<int> ug, pg;         // 9 / 3 dofs per cell
  std::vector<int> parent, child;
  std::vector<int> cell_i, cell_j; // cell position (column, row) on this level
  std::vector<double> xy;          // velocity dof coordinates
  int N_CIntCell = 0, maxCellsPerColor = 0;
  std::vector<int> ptrCellColors, CellReorder;
  std::vector<double> rhs, exact;
  std::vector<double> sol, frhs, aux0, aux1;   // multigrid vectors
};

// ---------------------------------------------------------------- reference element
// 1D Q2 Lagrange basis on [-1, 1] with nodes -1, 0, 1
inline void lag(double t, double v[3], double d[3]) {
  v[0] = 0.5 * t * (t - 1.0); v[1] = 1.0 - t * t; v[2] = 0.5 * t * (t + 1.0);
  d[0] = t - 0.5; d[1] = -2.0 * t; d[2] = t + 0.5;
}
// Q2 quadrilateral, node a + 3b (xi fastest)
void q2(double xi, double eta, double v[9], double dx[9], double dy[9]) {
  double vx[3], dxv[3], vy[3], dyv[3];
  lag(xi, vx, dxv);
  lag(eta, vy, dyv);
  for (int b = 0; b < 3; ++b)
    for (int a = 0; a < 3; ++a) {
      v[a + 3 * b] = vx[a] * vy[b];
      dx[a + 3 * b] = dxv[a] * vy[b];
      dy[a + 3 * b] = vx[a] * dyv[b];
    }
}
// P1disc: include/FE2D/BF_D_Q_P1_2D.h
void p1(double xi, double eta, double v[3]) {
  v[0] = 1.0; v[1] = 3.0 * xi; v[2] = 3.0 * eta;
}
const double GP[3] = {-0.774596669241483377, 0.0, 0.774596669241483377};
const double GW[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};

// ---------------------------------------------------------------- manufactured solutions
struct Exact {
  int kind;
  double nu;
  // u, p and f = -nu lap u + grad p
  void eval(double x, double y, double u[2], double &p, double f[2]) const {
    const double pi = 3.14159265358979323846;
    if (kind == O2D_RHS_STOKES_CURL) {
      // psi = sin^2(pi x) sin^2(pi y); u = (psi_y, -psi_x)
      const double sx = std::sin(pi * x), sy = std::sin(pi * y);
      const double s2x = std::sin(2 * pi * x), s2y = std::sin(2 * pi * y);
      const double c2x = std::cos(2 * pi * x), c2y = std::cos(2 * pi * y);
      u[0] = pi * sx * sx * s2y;
      u[1] = -pi * s2x * sy * sy;
      p = x * x * x + y * y * y - 0.5;
      const double lap0 = 2 * pi * pi * pi * s2y * (2 * c2x - 1);
      const double lap1 = -2 * pi * pi * pi * s2x * (2 * c2y - 1);
      f[0] = -nu * lap0 + 3 * x * x;
      f[1] = -nu * lap1 + 3 * y * y;
    } else {
      // u in Q2 (divergence free), p linear
      u[0] = x * x + 2 * x * y;
      u[1] = -y * y - 2 * x * y + x;
      p = 1.0 + x - 2.0 * y;
      f[0] = -nu * 2.0 + 1.0;
      f[1] = -nu * (-2.0) - 2.0;
    }
  }
};

// ---------------------------------------------------------------- mesh and assembly
// cells of level l: children of the cells of level l-1 in parent order, child c of a
// parent at (i, j) sits at (2i + (c & 1), 2j + (c >> 1)); level 0 is the coarse grid in
// row order
void build_cells(std::vector<Level2> &L, int coarse) {
  for (size_t l = 0; l < L.size(); ++l) {
    Level2 &V = L[l];
    V.n = coarse << l;
    V.N_Cells = V.n * V.n;
    V.cell_i.resize(V.N_Cells);
    V.cell_j.resize(V.N_Cells);
    if (l == 0) {
      for (int c = 0; c < V.N_Cells; ++c) { V.cell_i[c] = c % V.n; V.cell_j[c] = c / V.n; }
    } else {
      const Level2 &C = L[l - 1];
      V.parent.resize(V.N_Cells);
      V.child.resize(V.N_Cells);
      for (int pc = 0; pc < C.N_Cells; ++pc)
        for (int ch = 0; ch < N_CHILD; ++ch) {
          const int c = N_CHILD * pc + ch;
          V.cell_i[c] = 2 * C.cell_i[pc] + (ch & 1);
          V.cell_j[c] = 2 * C.cell_j[pc] + (ch >> 1);
          V.parent[c] = pc;
          V.child[c] = ch;
        }
    }
  }
}

void assemble(Level2 &V, const Exact &ex) {
  const int n = V.n, m = 2 * n + 1;
  const double h = 1.0 / n;
  // velocity dofs: lattice nodes, interior (active) first, then boundary, lattice order
  std::vector<int> num(size_t(m) * m);
  int na = 0;
  for (int J = 0; J < m; ++J)
    for (int I = 0; I < m; ++I) {
      const bool bd = I == 0 || J == 0 || I == m - 1 || J == m - 1;
      if (!bd) num[size_t(J) * m + I] = na++;
    }
  int nb = na;
  for (int J = 0; J < m; ++J)
    for (int I = 0; I < m; ++I) {
      const bool bd = I == 0 || J == 0 || I == m - 1 || J == m - 1;
      if (bd) num[size_t(J) * m + I] = nb++;
    }
  V.N_UDOF = nb; V.N_Active = na; V.N_PDOF = N_P * V.N_Cells; V.N_DOF = GEO_DIM * V.N_UDOF + V.N_PDOF;
  V.xy.assign(2 * size_t(V.N_UDOF), 0.0);
  for (int J = 0; J < m; ++J)
    for (int I = 0; I < m; ++I) {
      const int g = num[size_t(J) * m + I];
      V.xy[2 * g] = I * 0.5 * h;
      V.xy[2 * g + 1] = J * 0.5 * h;
    }
  V.ug.resize(size_t(V.N_Cells) * N_U);
  V.pg.resize(size_t(V.N_Cells) * N_P);
  for (int c = 0; c < V.N_Cells; ++c) {
    for (int b = 0; b < 3; ++b)
      for (int a = 0; a < 3; ++a)
        V.ug[size_t(c) * N_U + a + 3 * b] = num[size_t(2 * V.cell_j[c] + b) * m + 2 * V.cell_i[c] + a];
    for (int k = 0; k < N_P; ++k) V.pg[size_t(c) * N_P + k] = N_P * c + k;
  }
  // element matrices (affine square cells: x = x0 + h (xi + 1) / 2)
  double Ael[9][9] = {}, Bx[3][9] = {}, By[3][9] = {};
  for (int qb = 0; qb < 3; ++qb)
    for (int qa = 0; qa < 3; ++qa) {
      double v[9], dxi[9], deta[9], pv[3];
      q2(GP[qa], GP[qb], v, dxi, deta);
      p1(GP[qa], GP[qb], pv);
      const double w = GW[qa] * GW[qb] * (h * h / 4.0), s = 2.0 / h;
      for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j) Ael[i][j] += w * ex.nu * (dxi[j] * dxi[i] + deta[j] * deta[i]) * s * s;
      for (int mm = 0; mm < 3; ++mm)
        for (int j = 0; j < 9; ++j) {
          Bx[mm][j] += -w * pv[mm] * dxi[j] * s;
          By[mm][j] += -w * pv[mm] * deta[j] * s;
        }
    }
  // CSR structure: A rows over the velocity dofs sharing a cell (sorted), Dirichlet rows
  // only the diagonal; B rows (pressure) list the cell's 9 velocity dofs in local order;
  // BT rows over the pressure dofs of the cells containing the velocity dof
  std::vector<std::vector<int>> cellsof(V.N_UDOF);
  for (int c = 0; c < V.N_Cells; ++c)
    for (int k = 0; k < N_U; ++k) {
      auto &cl = cellsof[V.ug[size_t(c) * N_U + k]];
      if (std::find(cl.begin(), cl.end(), c) == cl.end()) cl.push_back(c);
    }
  V.a_rowptr.assign(1, 0);
  V.a_kcol.clear();
  V.bt_rowptr.assign(1, 0);
  V.bt_kcol.clear();
  for (int r = 0; r < V.N_UDOF; ++r) {
    std::vector<int> cols;
    if (r < V.N_Active) {
      for (int c : cellsof[r])
        for (int k = 0; k < N_U; ++k) cols.push_back(V.ug[size_t(c) * N_U + k]);
      std::sort(cols.begin(), cols.end());
      cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    } else {
      cols.push_back(r);
    }
    V.a_kcol.insert(V.a_kcol.end(), cols.begin(), cols.end());
    V.a_rowptr.push_back(int(V.a_kcol.size()));
    std::vector<int> pc;
    for (int c : cellsof[r])
      for (int k = 0; k < N_P; ++k) pc.push_back(V.pg[size_t(c) * N_P + k]);
    std::sort(pc.begin(), pc.end());
    V.bt_kcol.insert(V.bt_kcol.end(), pc.begin(), pc.end());
    V.bt_rowptr.push_back(int(V.bt_kcol.size()));
  }
  for (int b = 0; b < 4; ++b) V.a[b].assign(V.a_kcol.size(), 0.0);
  for (int b = 0; b < 2; ++b) V.bt[b].assign(V.bt_kcol.size(), 0.0);
  V.b_rowptr.assign(1, 0);
  V.b_kcol.clear();
  for (int c = 0; c < V.N_Cells; ++c)
    for (int k = 0; k < N_P; ++k) {
      for (int j = 0; j < N_U; ++j) V.b_kcol.push_back(V.ug[size_t(c) * N_U + j]);
      V.b_rowptr.push_back(int(V.b_kcol.size()));
    }
  for (int b = 0; b < 2; ++b) V.b[b].assign(V.b_kcol.size(), 0.0);
  auto apos = [&](int r, int col) {
    for (int k = V.a_rowptr[r]; k < V.a_rowptr[r + 1]; ++k)
      if (V.a_kcol[k] == col) return k;
    return -1;
  };
  auto btpos = [&](int r, int col) {
    for (int k = V.bt_rowptr[r]; k < V.bt_rowptr[r + 1]; ++k)
      if (V.bt_kcol[k] == col) return k;
    return -1;
  };
  V.rhs.assign(V.N_DOF, 0.0);
  for (int c = 0; c < V.N_Cells; ++c) {
    const int *U = &V.ug[size_t(c) * N_U];
    const int *P = &V.pg[size_t(c) * N_P];
    for (int i = 0; i < N_U; ++i) {
      const int r = U[i];
      if (r >= V.N_Active) continue;
      for (int j = 0; j < N_U; ++j) {
        const int k = apos(r, U[j]);
        V.a[0][k] += Ael[i][j];
        V.a[3][k] += Ael[i][j];
      }
      for (int mm = 0; mm < N_P; ++mm) {
        const int k = btpos(r, P[mm]);
        V.bt[0][k] += Bx[mm][i];
        V.bt[1][k] += By[mm][i];
      }
    }
    for (int mm = 0; mm < N_P; ++mm) {
      const int row = P[mm];
      for (int j = 0; j < N_U; ++j) {
        const int k = V.b_rowptr[row] + j;
        V.b[0][k] += Bx[mm][j];
        V.b[1][k] += By[mm][j];
      }
    }
    // f_i = (f, phi_i)
    const double x0 = V.cell_i[c] * h, y0 = V.cell_j[c] * h;
    for (int qb = 0; qb < 3; ++qb)
      for (int qa = 0; qa < 3; ++qa) {
        double v[9], dxi[9], deta[9];
        q2(GP[qa], GP[qb], v, dxi, deta);
        const double x = x0 + h * (GP[qa] + 1) / 2, y = y0 + h * (GP[qb] + 1) / 2;
        double u[2], p, f[2];
        ex.eval(x, y, u, p, f);
        const double w = GW[qa] * GW[qb] * (h * h / 4.0);
        for (int i = 0; i < N_U; ++i)
          if (U[i] < V.N_Active) {
            V.rhs[U[i]] += w * f[0] * v[i];
            V.rhs[V.N_UDOF + U[i]] += w * f[1] * v[i];
          }
      }
  }
  // Dirichlet rows: identity in A11 and A22, right-hand side = boundary value
  V.exact.assign(V.N_DOF, 0.0);
  for (int r = 0; r < V.N_UDOF; ++r) {
    double u[2], p, f[2];
    ex.eval(V.xy[2 * r], V.xy[2 * r + 1], u, p, f);
    V.exact[r] = u[0];
    V.exact[V.N_UDOF + r] = u[1];
    if (r >= V.N_Active) {
      const int k = V.a_rowptr[r];
      V.a[0][k] = 1.0;
      V.a[3][k] = 1.0;
      V.rhs[r] = u[0];
      V.rhs[V.N_UDOF + r] = u[1];
    }
  }
  // pressure: the moment functionals of include/FE2D/NF_D_Q_P1_2D.h (1/4 of the integral
  // of p, p xi, p eta over the reference cell)
  for (int c = 0; c < V.N_Cells; ++c) {
    const double x0 = V.cell_i[c] * h, y0 = V.cell_j[c] * h;
    double mom[3] = {0, 0, 0};
    for (int qb = 0; qb < 3; ++qb)
      for (int qa = 0; qa < 3; ++qa) {
        const double x = x0 + h * (GP[qa] + 1) / 2, y = y0 + h * (GP[qb] + 1) / 2;
        double u[2], p, f[2];
        ex.eval(x, y, u, p, f);
        const double w = GW[qa] * GW[qb];
        mom[0] += 0.25 * w * p;
        mom[1] += 0.25 * w * p * GP[qa];
        mom[2] += 0.25 * w * p * GP[qb];
      }
    for (int k = 0; k < N_P; ++k) V.exact[GEO_DIM * V.N_UDOF + N_P * c + k] = mom[k];
  }
}

// per-child prolongation matrices QQ[child][k * nloc + l]: coarse basis l at the fine nodal
// functional k (Q2: Lagrange values at the child's nodes; P1disc: the child's moments of the
// coarse linear function), as TFEDatabase2D::GetProlongationMatrix2D builds them
void build_qq(std::vector<double> &qu, std::vector<double> &qp) {
  qu.assign(size_t(N_CHILD) * N_U * N_U, 0.0);
  qp.assign(size_t(N_CHILD) * N_P * N_P, 0.0);
  for (int ch = 0; ch < N_CHILD; ++ch) {
    const double ox = (ch & 1) ? 0.0 : -1.0, oy = (ch >> 1) ? 0.0 : -1.0;   // child lower corner
    for (int kb = 0; kb < 3; ++kb)
      for (int ka = 0; ka < 3; ++ka) {
        const double xi = ox + 0.5 * ka, eta = oy + 0.5 * kb;   // child node in parent coordinates
        double v[9], dx[9], dy[9];
        q2(xi, eta, v, dx, dy);
        for (int l = 0; l < N_U; ++l) qu[size_t(ch) * N_U * N_U + (ka + 3 * kb) * N_U + l] = v[l];
      }
    // coarse basis 1, 3 xi_p, 3 eta_p with xi_p = ox + (xi_c + 1) / 2: moments in the child
    for (int k = 0; k < N_P; ++k)
      for (int l = 0; l < N_P; ++l) {
        double s = 0;
        for (int qb = 0; qb < 3; ++qb)
          for (int qa = 0; qa < 3; ++qa) {
            const double xc = GP[qa], yc = GP[qb];
            double pv[3];
            p1(ox + (xc + 1) / 2, oy + (yc + 1) / 2, pv);
            const double fk = k == 0 ? 1.0 : (k == 1 ? xc : yc);
            s += 0.25 * GW[qa] * GW[qb] * pv[l] * fk;
          }
        qp[size_t(ch) * N_P * N_P + k * N_P + l] = s;
      }
  }
}

}  // namespace

struct oracle2d {
  pmg_params prm;
  std::vector<Level2> L;
  std::vector<double> qu, qp;
  std::vector<double> coarse_lu;
  std::vector<int> coarse_piv;
  bool coarse_ready = false;
};

namespace {

inline double Ddot(int n, const double *a, const double *b) {
  double s = 0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

// IntoL20Vector2D case -11: MGComponents2D.C:1491-1508
void IntoL20Vector2D(double *v, int Length) {
  double s = 0;
  for (int i = 0; i < Length; i += 3) s += v[i];
  s /= Length;
  s *= 3;
  for (int i = 0; i < Length; i += 3) v[i] -= s;
}

// CoupledDefect NSE4 2D: MGComponents2D.C:2708-2833
void CoupledDefect(const Level2 &V, const double *x, const double *b, double *r) {
  const int N = V.N_UDOF;
  const double *u1 = x, *u2 = x + N, *p = x + 2 * N;
  const double *v1 = b, *v2 = b + N, *q = b + 2 * N;
  double *r1 = r, *r2 = r + N, *r3 = r + 2 * N;
  for (int i = 0; i < V.N_Active; i++) {
    double s = v1[i], t = v2[i];
    for (int j = V.a_rowptr[i]; j < V.a_rowptr[i + 1]; j++) {
      const int index = V.a_kcol[j];
      s -= V.a[0][j] * u1[index] + V.a[1][j] * u2[index];
      t -= V.a[2][j] * u1[index] + V.a[3][j] * u2[index];
    }
    r1[i] = s;
    r2[i] = t;
  }
  for (int i = V.N_Active; i < N; i++) {
    double s = v1[i], t = v2[i];
    for (int j = V.a_rowptr[i]; j < V.a_rowptr[i + 1]; j++) {
      const int index = V.a_kcol[j];
      s -= V.a[0][j] * u1[index];
      t -= V.a[3][j] * u2[index];
    }
    r1[i] = s;
    r2[i] = t;
  }
  for (int i = 0; i < V.N_PDOF; i++) {
    double s = q[i];
    for (int j = V.b_rowptr[i]; j < V.b_rowptr[i + 1]; j++) {
      const int index = V.b_kcol[j];
      s -= V.b[0][j] * u1[index] + V.b[1][j] * u2[index];
    }
    r3[i] = s;
  }
  for (int i = 0; i < V.N_Active; i++) {
    double s = 0, t = 0;
    for (int j = V.bt_rowptr[i]; j < V.bt_rowptr[i + 1]; j++) {
      const double value = p[V.bt_kcol[j]];
      s += V.bt[0][j] * value;
      t += V.bt[1][j] * value;
    }
    r1[i] -= s;
    r2[i] -= t;
  }
}

// CoupledMatVect NSE4 2D: MGComponents2D.C:2578-2700 (y = K x)
void CoupledMatVect(const Level2 &V, const double *x, double *y) {
  std::vector<double> zero(V.N_DOF, 0.0);
  CoupledDefect(V, x, zero.data(), y);
  for (int i = 0; i < V.N_DOF; ++i) y[i] = -y[i];
}

void SetDirichletNodes(const Level2 &V, double *u, const double *f) {
  for (int c = 0; c < GEO_DIM; ++c)
    std::memcpy(u + c * V.N_UDOF + V.N_Active, f + c * V.N_UDOF + V.N_Active,
                sizeof(double) * (V.N_UDOF - V.N_Active));
}
void CorrectNodes(const Level2 &V, double *u) {
  for (int c = 0; c < GEO_DIM; ++c)
    std::memset(u + c * V.N_UDOF + V.N_Active, 0, sizeof(double) * (V.N_UDOF - V.N_Active));
}
void Defect(const oracle2d &o, const Level2 &V, const double *u, const double *f, double *d, double &res) {
  CoupledDefect(V, u, f, d);
  if (o.prm.internal_project_pressure) IntoL20Vector2D(d + GEO_DIM * V.N_UDOF, V.N_PDOF);
  CorrectNodes(V, d);
  res = std::sqrt(Ddot(V.N_DOF, d, d));
}
void CorrectDefect(const oracle2d &o, const Level2 &V, double *d) {
  CorrectNodes(V, d);
  if (o.prm.internal_project_pressure) IntoL20Vector2D(d + GEO_DIM * V.N_UDOF, V.N_PDOF);
}

// local residual and matrix of a cell (the 2D branches of NSE_MGLevel4.C:902-1290):
// rows j (u1), 9 + j (u2), 18 + m (p), System[col * 21 + row]
void PatchResidual(const Level2 &V, int ii, const double *u1, const double *rhs1, double *Rhs) {
  const int N = V.N_UDOF;
  const double *u2 = u1 + N, *p = u1 + 2 * N;
  const int *UDOFs = &V.ug[size_t(ii) * N_U], *PDOFs = &V.pg[size_t(ii) * N_P];
  for (int j = 0; j < N_U; j++) {
    const int UDOF = UDOFs[j];
    Rhs[j] = rhs1[UDOF];
    Rhs[j + N_U] = rhs1[N + UDOF];
    for (int k = V.a_rowptr[UDOF]; k < V.a_rowptr[UDOF + 1]; k++) {
      const int l = V.a_kcol[k];
      double v11 = V.a[0][k], v12 = V.a[1][k], v21 = V.a[2][k], v22 = V.a[3][k];
      if (UDOF >= V.N_Active) v12 = v21 = 0;
      Rhs[j] -= v11 * u1[l] + v12 * u2[l];
      Rhs[j + N_U] -= v21 * u1[l] + v22 * u2[l];
    }
    if (UDOF < V.N_Active)
      for (int k = V.bt_rowptr[UDOF]; k < V.bt_rowptr[UDOF + 1]; k++) {
        const double value = p[V.bt_kcol[k]];
        Rhs[j] -= V.bt[0][k] * value;
        Rhs[j + N_U] -= V.bt[1][k] * value;
      }
  }
  for (int j = 0; j < N_P; j++) {
    const int PDOF = PDOFs[j];
    double s = rhs1[2 * N + PDOF];
    for (int k = V.b_rowptr[PDOF]; k < V.b_rowptr[PDOF + 1]; k++) {
      const int l = V.b_kcol[k];
      s -= V.b[0][k] * u1[l];
      s -= V.b[1][k] * u2[l];
    }
    Rhs[j + GEO_DIM * N_U] = s;
  }
}
void PatchMatrix(const Level2 &V, int ii, double *System) {
  std::memset(System, 0, sizeof(double) * N_LocalDOF * N_LocalDOF);
  const int *UDOFs = &V.ug[size_t(ii) * N_U], *PDOFs = &V.pg[size_t(ii) * N_P];
  for (int j = 0; j < N_U; j++) {
    const int UDOF = UDOFs[j];
    for (int k = V.a_rowptr[UDOF]; k < V.a_rowptr[UDOF + 1]; k++) {
      const int l = V.a_kcol[k];
      double v11 = V.a[0][k], v12 = V.a[1][k], v21 = V.a[2][k], v22 = V.a[3][k];
      if (UDOF >= V.N_Active) v12 = v21 = 0;
      for (int m = 0; m < N_U; m++)
        if (UDOFs[m] == l) {
          System[m * N_LocalDOF + j] = v11;
          System[(m + N_U) * N_LocalDOF + j] = v12;
          System[m * N_LocalDOF + j + N_U] = v21;
          System[(m + N_U) * N_LocalDOF + j + N_U] = v22;
          break;
        }
    }
    if (UDOF < V.N_Active)
      for (int k = V.bt_rowptr[UDOF]; k < V.bt_rowptr[UDOF + 1]; k++) {
        const int l = V.bt_kcol[k];
        for (int m = 0; m < N_P; m++)
          if (PDOFs[m] == l) {
            System[(m + GEO_DIM * N_U) * N_LocalDOF + j] = V.bt[0][k];
            System[(m + GEO_DIM * N_U) * N_LocalDOF + j + N_U] = V.bt[1][k];
            break;
          }
      }
  }
  for (int j = 0; j < N_P; j++) {
    const int PDOF = PDOFs[j];
    for (int k = V.b_rowptr[PDOF]; k < V.b_rowptr[PDOF + 1]; k++) {
      const int l = V.b_kcol[k];
      for (int m = 0; m < N_U; m++)
        if (UDOFs[m] == l) {
          System[m * N_LocalDOF + j + GEO_DIM * N_U] = V.b[0][k];
          System[(m + N_U) * N_LocalDOF + j + GEO_DIM * N_U] = V.b[1][k];
          break;
        }
    }
  }
}

// colored cell Vanka (NSE_MGLevel4.C:603-1649 in its 2D form): per color the local
// residuals of all its cells from the state before the color, then the 21x21 solves
// (LinAlg.C:106-130) and u += damp * delta; global relaxation alpha; Dirichlet copy
int CellVanka(oracle2d &o, Level2 &V, double *u1, const double *rhs1, double *aux) {
  const int N = V.N_UDOF;
  const double damp = o.prm.sc_smooth_damp_factor_coarse_saddle, alpha = o.prm.sc_smooth_damp_factor_saddle;
  double *u2 = u1 + N, *p = u1 + 2 * N;
  SetDirichletNodes(V, u1, rhs1);
  std::memcpy(aux, u1, sizeof(double) * V.N_DOF);
  std::vector<double> Rhs(size_t(V.maxCellsPerColor) * N_LocalDOF), System(N_LocalDOF * N_LocalDOF);
  std::vector<int> ipiv(N_LocalDOF);
  int info_all = 0;
  for (int i = 0; i < V.N_CIntCell; i++) {
    const int c0 = V.ptrCellColors[i], c1 = V.ptrCellColors[i + 1];
    for (int jj = c0; jj < c1; jj++) PatchResidual(V, V.CellReorder[jj], u1, rhs1, &Rhs[size_t(jj - c0) * N_LocalDOF]);
    for (int jj = c0; jj < c1; jj++) {
      const int ii = V.CellReorder[jj];
      double *R = &Rhs[size_t(jj - c0) * N_LocalDOF];
      PatchMatrix(V, ii, System.data());
      if (oracle_lu_factor(N_LocalDOF, System.data(), ipiv.data())) info_all = ii + 1;
      oracle_lu_solve(N_LocalDOF, System.data(), ipiv.data(), R);
      const int *UDOFs = &V.ug[size_t(ii) * N_U], *PDOFs = &V.pg[size_t(ii) * N_P];
      for (int j = 0; j < N_U; j++) {
        u1[UDOFs[j]] += damp * R[j];
        u2[UDOFs[j]] += damp * R[j + N_U];
      }
      for (int j = 0; j < N_P; j++) p[PDOFs[j]] += damp * R[j + GEO_DIM * N_U];
    }
  }
  if (std::fabs(1 - alpha) > 1e-3)
    for (int j = 0; j < V.N_DOF; j++) u1[j] = aux[j] + alpha * (u1[j] - aux[j]);
  SetDirichletNodes(V, u1, rhs1);
  return info_all;
}

// Prolongate / DefectRestriction 2D (MGComponents2D.C:118-1022): as in 3D, values of
// shared fine dofs averaged over their cells; the restriction is the transpose
void ProlongateSpace(const oracle2d &o, const Level2 &C, const Level2 &F, bool velocity, const double *cf,
                     double *ff) {
  const int nloc = velocity ? N_U : N_P;
  const int ncomp = velocity ? GEO_DIM : 1;
  const int nC = velocity ? C.N_UDOF : C.N_PDOF, nF = velocity ? F.N_UDOF : F.N_PDOF;
  const std::vector<int> &cg = velocity ? C.ug : C.pg, &fg = velocity ? F.ug : F.pg;
  const double *QQall = velocity ? o.qu.data() : o.qp.data();
  std::vector<double> mult(nF, 0.0);
  std::memset(ff, 0, sizeof(double) * ncomp * nF);
  for (int fc = 0; fc < F.N_Cells; ++fc) {
    const int pc = F.parent[fc];
    const double *QQ = QQall + size_t(F.child[fc]) * nloc * nloc;
    for (int k = 0; k < nloc; ++k) {
      const int fd = fg[size_t(fc) * nloc + k];
      mult[fd] += 1;
      for (int comp = 0; comp < ncomp; ++comp) {
        double s = 0;
        for (int l = 0; l < nloc; ++l) s += QQ[k * nloc + l] * cf[comp * nC + cg[size_t(pc) * nloc + l]];
        ff[comp * nF + fd] += s;
      }
    }
  }
  for (int comp = 0; comp < ncomp; ++comp)
    for (int i = 0; i < nF; ++i)
      if (mult[i]) ff[comp * nF + i] /= mult[i];
}
void RestrictSpace(const oracle2d &o, const Level2 &C, const Level2 &F, bool velocity, double *cf,
                   const double *ff) {
  const int nloc = velocity ? N_U : N_P;
  const int ncomp = velocity ? GEO_DIM : 1;
  const int nC = velocity ? C.N_UDOF : C.N_PDOF, nF = velocity ? F.N_UDOF : F.N_PDOF;
  const std::vector<int> &cg = velocity ? C.ug : C.pg, &fg = velocity ? F.ug : F.pg;
  const double *QQall = velocity ? o.qu.data() : o.qp.data();
  std::vector<double> mult(nF, 0.0);
  for (int fc = 0; fc < F.N_Cells; ++fc)
    for (int k = 0; k < nloc; ++k) mult[fg[size_t(fc) * nloc + k]] += 1;
  std::memset(cf, 0, sizeof(double) * ncomp * nC);
  for (int fc = 0; fc < F.N_Cells; ++fc) {
    const int pc = F.parent[fc];
    const double *QQ = QQall + size_t(F.child[fc]) * nloc * nloc;
    for (int comp = 0; comp < ncomp; ++comp)
      for (int l = 0; l < nloc; ++l) {
        double s = 0;
        for (int k = 0; k < nloc; ++k) {
          const int fd = fg[size_t(fc) * nloc + k];
          s += QQ[k * nloc + l] * ff[comp * nF + fd] / mult[fd];
        }
        cf[comp * nC + cg[size_t(pc) * nloc + l]] += s;
      }
  }
}

// SolveExact 2D: dense system of all blocks, pressure dof 0 pinned when projecting
int SolveExact(oracle2d &o, Level2 &V, double *u, const double *f) {
  const int Nd = V.N_DOF, N = V.N_UDOF;
  if (!o.coarse_ready) {
    std::vector<double> &a = o.coarse_lu;
    a.assign(size_t(Nd) * Nd, 0.0);
    for (int i = 0; i < N; ++i)
      for (int j = V.a_rowptr[i]; j < V.a_rowptr[i + 1]; ++j) {
        const size_t col = V.a_kcol[j];
        if (i < V.N_Active) {
          a[col * Nd + i] = V.a[0][j];
          a[(col + N) * Nd + i] = V.a[1][j];
          a[col * Nd + i + N] = V.a[2][j];
          a[(col + N) * Nd + i + N] = V.a[3][j];
        } else {
          a[col * Nd + i] = V.a[0][j];
          a[(col + N) * Nd + i + N] = V.a[3][j];
        }
      }
    for (int i = 0; i < V.N_Active; ++i)
      for (int j = V.bt_rowptr[i]; j < V.bt_rowptr[i + 1]; ++j) {
        const size_t col = size_t(V.bt_kcol[j]) + 2 * N;
        a[col * Nd + i] = V.bt[0][j];
        a[col * Nd + i + N] = V.bt[1][j];
      }
    for (int i = 0; i < V.N_PDOF; ++i)
      for (int j = V.b_rowptr[i]; j < V.b_rowptr[i + 1]; ++j) {
        const size_t col = V.b_kcol[j];
        a[col * Nd + i + 2 * N] = V.b[0][j];
        a[(col + N) * Nd + i + 2 * N] = V.b[1][j];
      }
    if (o.prm.internal_project_pressure) {
      for (int i = 0; i < Nd; ++i) a[size_t(i) * Nd + 2 * N] = 0;
      a[size_t(2 * N) * Nd + 2 * N] = 1;
    }
    o.coarse_piv.resize(Nd);
    if (oracle_lu_factor(Nd, a.data(), o.coarse_piv.data())) return 1;
    o.coarse_ready = true;
  }
  std::vector<double> b(f, f + Nd);
  if (o.prm.internal_project_pressure) b[2 * N] = 0;
  oracle_lu_solve(Nd, o.coarse_lu.data(), o.coarse_piv.data(), b.data());
  std::memcpy(u, b.data(), sizeof(double) * Nd);
  return 0;
}

// TNSE_MultiGrid::Cycle (NSE_MultiGrid.C:434-1779), the 2D level calls of the 3D sequence
int Cycle(oracle2d &o, int i) {
  Level2 &Cur = o.L[i];
  double *u = Cur.sol.data(), *f = Cur.frhs.data(), *d = Cur.aux0.data(), *aux = Cur.aux1.data();
  double res;
  CorrectNodes(Cur, u);
  SetDirichletNodes(Cur, u, f);
  if (i == 0) return SolveExact(o, Cur, u, f) ? PMG_ERR_SINGULAR_COARSE : 0;
  Level2 &Coa = o.L[i - 1];
  int status = 0;
  for (int j = 0; j < o.prm.sc_pre_smooth_saddle; j++)
    if (CellVanka(o, Cur, u, f, aux)) status = PMG_ERR_SINGULAR_PATCH;
  CorrectNodes(Cur, u);
  Defect(o, Cur, u, f, d, res);
  RestrictSpace(o, Coa, Cur, true, Coa.frhs.data(), d);
  RestrictSpace(o, Coa, Cur, false, Coa.frhs.data() + 2 * Coa.N_UDOF, d + 2 * Cur.N_UDOF);
  CorrectDefect(o, Coa, Coa.frhs.data());
  std::fill(Coa.sol.begin(), Coa.sol.end(), 0.0);
  if (int st = Cycle(o, i - 1)) status = st;
  ProlongateSpace(o, Coa, Cur, true, Coa.sol.data(), d);
  ProlongateSpace(o, Coa, Cur, false, Coa.sol.data() + 2 * Coa.N_UDOF, d + 2 * Cur.N_UDOF);
  CorrectNodes(Cur, d);
  for (int k = 0; k < Cur.N_DOF; ++k) u[k] += o.prm.sc_gmg_damp_factor_saddle * d[k];
  CorrectNodes(Cur, u);
  for (int j = 0; j < o.prm.sc_post_smooth_saddle; j++)
    if (CellVanka(o, Cur, u, f, aux)) status = PMG_ERR_SINGULAR_PATCH;
  return status;
}

int MGIterate(oracle2d &o, double *z, const double *v) {
  Level2 &F = o.L.back();
  std::fill(F.sol.begin(), F.sol.end(), 0.0);
  std::memcpy(F.frhs.data(), v, sizeof(double) * F.N_DOF);
  SetDirichletNodes(F, F.sol.data(), F.frhs.data());
  const int st = Cycle(o, int(o.L.size()) - 1);
  std::memcpy(z, F.sol.data(), sizeof(double) * F.N_DOF);
  return st;
}

void rot_gen(double dx, double dy, double &cs, double &sn) {
  if (dy == 0.0) {
    cs = 1.0; sn = 0.0;
  } else if (std::fabs(dy) > std::fabs(dx)) {
    const double t = dx / dy;
    sn = 1.0 / std::sqrt(1.0 + t * t);
    cs = t * sn;
  } else {
    const double t = dy / dx;
    cs = 1.0 / std::sqrt(1.0 + t * t);
    sn = t * cs;
  }
}
void rot_apply(double &dx, double &dy, double cs, double sn) {
  const double t = cs * dx + sn * dy;
  dy = -sn * dx + cs * dy;
  dx = t;
}

// TFgmresIte::Iterate, flexible branch (FgmresIte.C:250-677), sc_minit = 0
int Fgmres(oracle2d &o, double *sol, const double *rhs, int *iters, double *final_res) {
  Level2 &F = o.L.back();
  const int n = F.N_DOF;
  const pmg_params &P = o.prm;
  int maxit = P.sc_lin_maxit_saddle, restart = P.sc_gmres_restart;
  if (restart > maxit) restart = maxit;
  if (restart <= 0) restart = 10;
  std::vector<double> s(restart + 1), cs(restart + 1), sn(restart + 1), w(n);
  std::vector<std::vector<double>> H(restart + 1, std::vector<double>(restart, 0.0));
  std::vector<std::vector<double>> v(restart + 1, std::vector<double>(n)), z(restart + 1, std::vector<double>(n));
  auto defect = [&](const double *x, double *r) {   // Defect_NSE4 (MGComponents2D.C:2834-2845)
    CoupledDefect(F, x, rhs, r);
    if (P.internal_project_pressure) IntoL20Vector2D(r + 2 * F.N_UDOF, F.N_PDOF);
  };
  auto update = [&](int k) {
    for (int i = k; i >= 0; i--) {
      s[i] /= H[i][i];
      for (int j = i - 1; j >= 0; j--) s[j] -= H[j][i] * s[i];
    }
    for (int j = 0; j <= k; j++)
      for (int i = 0; i < n; i++) sol[i] += z[j][i] * s[j];
  };
  defect(sol, v[0].data());
  double res = std::sqrt(Ddot(n, v[0].data(), v[0].data())), res0 = res, beta = res;
  if (beta <= P.sc_lin_res_norm_min_saddle) { *iters = 0; *final_res = res; return 0; }
  int status = 0, j = 1;
  while (j <= maxit) {
    for (int i = 0; i < n; ++i) v[0][i] /= beta;
    std::fill(s.begin(), s.end(), 0.0);
    s[0] = beta;
    for (int i = 0; i < restart && j <= maxit; i++, j++) {
      if (int st = MGIterate(o, z[i].data(), v[i].data())) status = st;
      CoupledMatVect(F, z[i].data(), w.data());
      for (int k = 0; k <= i; k++) {
        H[k][i] = Ddot(n, w.data(), v[k].data());
        for (int q = 0; q < n; ++q) w[q] -= H[k][i] * v[k][q];
      }
      H[i + 1][i] = std::sqrt(Ddot(n, w.data(), w.data()));
      for (int q = 0; q < n; ++q) v[i + 1][q] = H[i + 1][i] != 0 ? w[q] / H[i + 1][i] : w[q];
      for (int k = 0; k < i; k++) rot_apply(H[k][i], H[k + 1][i], cs[k], sn[k]);
      rot_gen(H[i][i], H[i + 1][i], cs[i], sn[i]);
      rot_apply(H[i][i], H[i + 1][i], cs[i], sn[i]);
      rot_apply(s[i], s[i + 1], cs[i], sn[i]);
      res = std::fabs(s[i + 1]);
      if (!std::isfinite(res)) { *iters = j; *final_res = res; return PMG_ERR_NONFINITE; }
      if (res < P.sc_lin_res_norm_min_saddle || res < P.sc_lin_red_factor_saddle * res0) {
        update(i);
        *iters = j; *final_res = res;
        return status;
      }
      if (res > P.sc_div_factor * res0) { *iters = j; *final_res = res; return PMG_ERR_DIVERGED; }
    }
    update(restart - 1);
    defect(sol, v[0].data());
    res = res0 = beta = std::sqrt(Ddot(n, v[0].data(), v[0].data()));
  }
  *iters = std::min(j, maxit);
  *final_res = res;
  if (res > P.sc_lin_res_norm_min_saddle && res > res0 * P.sc_lin_red_factor_saddle)
    return status ? status : PMG_ERR_NOT_CONVERGED;
  return status;
}

}  // namespace

extern "C" {

int oracle2d_create(const pmg_params *params, int coarse, int n_levels, double nu, int rhs_kind, oracle2d **out) {
  if (!params || !out || coarse <= 0 || n_levels <= 0 || n_levels > 10) return PMG_ERR_INVALID_ARGUMENT;
  auto *o = new oracle2d;
  o->prm = *params;
  o->L.resize(n_levels);
  build_cells(o->L, coarse);
  Exact ex{rhs_kind, nu};
  for (auto &V : o->L) {
    assemble(V, ex);
    std::vector<int> ub(V.N_Cells);
    for (int c = 0; c < V.N_Cells; ++c) ub[c] = c * N_U;
    V.ptrCellColors.resize(V.N_Cells + 1);
    V.CellReorder.resize(V.N_Cells);
    oracle_color_cells(V.N_Cells, N_U, ub.data(), V.ug.data(), &V.N_CIntCell, V.ptrCellColors.data(),
                       V.CellReorder.data());
    V.ptrCellColors.resize(V.N_CIntCell + 1);
    V.maxCellsPerColor = 0;
    for (int k = 0; k < V.N_CIntCell; ++k)
      V.maxCellsPerColor = std::max(V.maxCellsPerColor, V.ptrCellColors[k + 1] - V.ptrCellColors[k]);
    V.sol.assign(V.N_DOF, 0.0);
    V.frhs.assign(V.N_DOF, 0.0);
    V.aux0.assign(V.N_DOF, 0.0);
    V.aux1.assign(V.N_DOF, 0.0);
  }
  build_qq(o->qu, o->qp);
  *out = o;
  return PMG_OK;
}
void oracle2d_destroy(oracle2d *o) { delete o; }
void oracle2d_sizes(const oracle2d *o, int level, int *n_udof, int *n_active, int *n_pdof, int *n_cells,
                    int *n_colors) {
  const Level2 &V = o->L[level];
  if (n_udof) *n_udof = V.N_UDOF;
  if (n_active) *n_active = V.N_Active;
  if (n_pdof) *n_pdof = V.N_PDOF;
  if (n_cells) *n_cells = V.N_Cells;
  if (n_colors) *n_colors = V.N_CIntCell;
}
int oracle2d_n_dof(const oracle2d *o, int level) { return o->L[level].N_DOF; }
void oracle2d_rhs(const oracle2d *o, int level, double *out) {
  std::copy(o->L[level].rhs.begin(), o->L[level].rhs.end(), out);
}
void oracle2d_exact(const oracle2d *o, int level, double *out) {
  std::copy(o->L[level].exact.begin(), o->L[level].exact.end(), out);
}
void oracle2d_dof_coords(const oracle2d *o, int level, double *xy) {
  std::copy(o->L[level].xy.begin(), o->L[level].xy.end(), xy);
}
int oracle2d_nnz(const oracle2d *o, int level, int *a_nnz, int *b_nnz, int *bt_nnz) {
  const Level2 &V = o->L[level];
  if (a_nnz) *a_nnz = int(V.a_kcol.size());
  if (b_nnz) *b_nnz = int(V.b_kcol.size());
  if (bt_nnz) *bt_nnz = int(V.bt_kcol.size());
  return 0;
}
void oracle2d_csr(const oracle2d *o, int level, int *a_rowptr, int *a_kcol, double *a4, int *b_rowptr, int *b_kcol,
                  double *b2, int *bt_rowptr, int *bt_kcol, double *bt2) {
  const Level2 &V = o->L[level];
  std::copy(V.a_rowptr.begin(), V.a_rowptr.end(), a_rowptr);
  std::copy(V.a_kcol.begin(), V.a_kcol.end(), a_kcol);
  for (int b = 0; b < 4; ++b) std::copy(V.a[b].begin(), V.a[b].end(), a4 + b * V.a_kcol.size());
  std::copy(V.b_rowptr.begin(), V.b_rowptr.end(), b_rowptr);
  std::copy(V.b_kcol.begin(), V.b_kcol.end(), b_kcol);
  for (int b = 0; b < 2; ++b) std::copy(V.b[b].begin(), V.b[b].end(), b2 + b * V.b_kcol.size());
  std::copy(V.bt_rowptr.begin(), V.bt_rowptr.end(), bt_rowptr);
  std::copy(V.bt_kcol.begin(), V.bt_kcol.end(), bt_kcol);
  for (int b = 0; b < 2; ++b) std::copy(V.bt[b].begin(), V.bt[b].end(), bt2 + b * V.bt_kcol.size());
}
void oracle2d_cell_dofs(const oracle2d *o, int level, int *u_global9, int *p_global3) {
  std::copy(o->L[level].ug.begin(), o->L[level].ug.end(), u_global9);
  std::copy(o->L[level].pg.begin(), o->L[level].pg.end(), p_global3);
}
void oracle2d_matvec(oracle2d *o, int level, const double *x, double *y) { CoupledMatVect(o->L[level], x, y); }
void oracle2d_defect(oracle2d *o, int level, const double *u, const double *f, double *d, double *res) {
  double r;
  Defect(*o, o->L[level], u, f, d, r);
  if (res) *res = r;
}
int oracle2d_cell_vanka(oracle2d *o, int level, double *u, const double *f) {
  return CellVanka(*o, o->L[level], u, f, o->L[level].aux1.data());
}
void oracle2d_prolongate(oracle2d *o, int fine_level, const double *u_coarse, double *d_fine) {
  const Level2 &F = o->L[fine_level], &C = o->L[fine_level - 1];
  ProlongateSpace(*o, C, F, true, u_coarse, d_fine);
  ProlongateSpace(*o, C, F, false, u_coarse + 2 * C.N_UDOF, d_fine + 2 * F.N_UDOF);
}
void oracle2d_restrict_defect(oracle2d *o, int fine_level, double *d_fine, double *f_coarse) {
  const Level2 &F = o->L[fine_level], &C = o->L[fine_level - 1];
  RestrictSpace(*o, C, F, true, f_coarse, d_fine);
  RestrictSpace(*o, C, F, false, f_coarse + 2 * C.N_UDOF, d_fine + 2 * F.N_UDOF);
}
int oracle2d_solve_exact(oracle2d *o, double *u, const double *f) { return SolveExact(*o, o->L[0], u, f); }
int oracle2d_vcycle(oracle2d *o, const double *v, double *z) { return MGIterate(*o, z, v); }
int oracle2d_fgmres(oracle2d *o, double *sol, const double *rhs, int *iters, double *final_res) {
  return Fgmres(*o, sol, rhs, iters, final_res);
}
int oracle2d_patch_matrix(oracle2d *o, int level, int cell, double *sys) {
  PatchMatrix(o->L[level], cell, sys);
  return 0;
}

}  // extern "C"
