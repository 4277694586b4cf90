// pmg_problem.cpp - structured Q2/P1disc Oseen hierarchies (see pmg_problem.h).
//
// Element matrices follow NSType4Galerkin3D (src/FE/NSE3D_FixPo.C:1919-2071):
//   A11 = A22 = A33 = nu (grad phi_j, grad phi_i) + (b . grad phi_j, phi_i), A_rc = 0 (r != c)
//   BkT(i,m) = -(psi_m, d_k phi_i),  Bk(m,j) = -(psi_m, d_k phi_j),  f_i = (f, phi_i)
// with a 3x3x3 Gauss rule and the trilinear cell map.  Dirichlet rows of the square
// blocks hold only the diagonal (src/FE/SquareStructure3D.C:119-126), active rows have
// sorted columns (TSquareStructure::Sort, src/FE/SquareStructure.C:65-84); B rows list
// the cell's 27 velocity dofs in local order (TStructure3D, src/FE/Structure3D.C:24-200).
#include "pmg_problem.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace {

constexpr int NU_LOC = 27;
constexpr int NP_LOC = 4;
constexpr double kGaussP1 = 0.5773502691896257645091489;  // NF_D_H_P1_3D point coordinate

struct Level {
  int nx = 0, ny = 0, nz = 0, n_cells = 0;
  std::vector<int> cell_ijk;           // 3 per cell
  std::vector<double> vert;            // (nx+1)(ny+1)(nz+1)*3
  int n_udof = 0, n_active = 0, n_dirichlet = 0, n_pdof = 0;
  std::vector<int> u_begin, u_global, p_begin, p_global;
  std::vector<int> a_rowptr, a_kcol;
  std::vector<double> a[9];
  std::vector<int> b_rowptr, b_kcol;
  std::vector<double> b[3];
  std::vector<int> bt_rowptr, bt_kcol;
  std::vector<double> bt[3];
  std::vector<double> rhs, exact, dof_xyz, cell_xyz;
  std::vector<int> parent, child;
};

struct Mode { double k[3]; double amp[3]; double phase; };

}  // namespace

struct pmgp_problem {
  pmgp_config cfg;
  std::vector<Level> levels;
  std::vector<Mode> conv_modes, rhs_modes;
  double conv_scale = 1.0;
  std::vector<double> qq_u, qq_p;      // 8 children
};

namespace {

inline double L0(double t) { return 0.5 * t * (t - 1.0); }
inline double L1(double t) { return 1.0 - t * t; }
inline double L2(double t) { return 0.5 * t * (t + 1.0); }
inline double dL0(double t) { return t - 0.5; }
inline double dL1(double t) { return -2.0 * t; }
inline double dL2(double t) { return t + 0.5; }

inline void lag(double t, double v[3], double d[3]) {
  v[0] = L0(t); v[1] = L1(t); v[2] = L2(t);
  d[0] = dL0(t); d[1] = dL1(t); d[2] = dL2(t);
}

// Q2 hex basis at reference point (local index a + 3b + 9c, xi fastest, as
// include/FE3D/NF_C_H_Q2_3D.h lists its 27 nodes)
void q2_basis(double xi, double eta, double zeta, double val[27], double dxi[27],
              double deta[27], double dzeta[27]) {
  double vx[3], dx[3], vy[3], dy[3], vz[3], dz[3];
  lag(xi, vx, dx); lag(eta, vy, dy); lag(zeta, vz, dz);
  for (int c = 0; c < 3; ++c)
    for (int b = 0; b < 3; ++b)
      for (int a = 0; a < 3; ++a) {
        int k = a + 3 * b + 9 * c;
        val[k] = vx[a] * vy[b] * vz[c];
        if (dxi) {
          dxi[k] = dx[a] * vy[b] * vz[c];
          deta[k] = vx[a] * dy[b] * vz[c];
          dzeta[k] = vx[a] * vy[b] * dz[c];
        }
      }
}

// trilinear map of a hex with corner vertices X[8][3] (corner index a + 2b + 4c)
void trilinear(const double X[8][3], double xi, double eta, double zeta, double x[3],
               double J[3][3]) {
  for (int r = 0; r < 3; ++r) { x[r] = 0; J[r][0] = J[r][1] = J[r][2] = 0; }
  for (int c = 0; c < 2; ++c)
    for (int b = 0; b < 2; ++b)
      for (int a = 0; a < 2; ++a) {
        double sx = a ? 1 : -1, sy = b ? 1 : -1, sz = c ? 1 : -1;
        double nx = 0.5 * (1 + sx * xi), ny = 0.5 * (1 + sy * eta), nz = 0.5 * (1 + sz * zeta);
        double N = nx * ny * nz;
        double dNx = 0.5 * sx * ny * nz, dNy = 0.5 * sy * nx * nz, dNz = 0.5 * sz * nx * ny;
        const double *P = X[a + 2 * b + 4 * c];
        for (int r = 0; r < 3; ++r) {
          x[r] += N * P[r];
          J[r][0] += dNx * P[r];
          J[r][1] += dNy * P[r];
          J[r][2] += dNz * P[r];
        }
      }
}

double inv3(const double J[3][3], double Ji[3][3]) {
  double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
               J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  double id = 1.0 / det;
  Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  return det;
}

// ---------------------------------------------------------------- exact solution (quad/lin)
// Examples/NSE_3D/AnsatzQuadLin.h: u in Q2, p in P1, div u = 0
void exact_quadlin(const double x[3], double u[3], double gu[3][3], double lap[3], double &p,
                   double gp[3]) {
  double X = x[0], Y = x[1], Z = x[2];
  u[0] = X * X + Y * Y + Z * Z;
  u[1] = X * X + 2 * X * Z + 13;
  u[2] = -2 * X * Z + 5 * Y * Y;
  gu[0][0] = 2 * X; gu[0][1] = 2 * Y; gu[0][2] = 2 * Z;
  gu[1][0] = 2 * X + 2 * Z; gu[1][1] = 0; gu[1][2] = 2 * X;
  gu[2][0] = -2 * Z; gu[2][1] = 10 * Y; gu[2][2] = -2 * X;
  lap[0] = 6; lap[1] = 2; lap[2] = 10;
  p = 3 * X - 2 * Y + 7 * Z - 4;
  gp[0] = 3; gp[1] = -2; gp[2] = 7;
}

void conv_at(const pmgp_problem &P, const double x[3], double bvec[3]) {
  const pmgp_config &c = P.cfg;
  bvec[0] = bvec[1] = bvec[2] = 0;
  if (c.conv_kind == PMGP_CONV_CONST) {
    bvec[0] = c.conv_const[0]; bvec[1] = c.conv_const[1]; bvec[2] = c.conv_const[2];
  } else if (c.conv_kind == PMGP_CONV_FOURIER) {
    // b = curl psi, psi = sum_m a_m sin(pi k_m.x + phi_m)  ->  b = sum pi cos(.) (k x a)
    for (const Mode &m : P.conv_modes) {
      double th = M_PI * (m.k[0] * x[0] + m.k[1] * x[1] + m.k[2] * x[2]) + m.phase;
      double cs = M_PI * std::cos(th);
      double kxa[3] = {m.k[1] * m.amp[2] - m.k[2] * m.amp[1], m.k[2] * m.amp[0] - m.k[0] * m.amp[2],
                       m.k[0] * m.amp[1] - m.k[1] * m.amp[0]};
      for (int r = 0; r < 3; ++r) bvec[r] += cs * kxa[r];
    }
    for (int r = 0; r < 3; ++r) bvec[r] *= P.conv_scale;
  } else if (c.conv_kind == PMGP_CONV_CHANNEL) {
    // parabolic profile in y,z (channel along x), max = conv_max
    double y = x[1] / c.ly, z = x[2] / c.lz;
    bvec[0] = c.conv_max * 16.0 * y * (1 - y) * z * (1 - z);
  }
}

void rhs_at(const pmgp_problem &P, const double x[3], double f[3]) {
  const pmgp_config &c = P.cfg;
  f[0] = f[1] = f[2] = 0;
  if (c.rhs_kind == PMGP_RHS_FOURIER) {
    for (const Mode &m : P.rhs_modes) {
      double th = M_PI * (m.k[0] * x[0] + m.k[1] * x[1] + m.k[2] * x[2]) + m.phase;
      double s = std::sin(th);
      for (int r = 0; r < 3; ++r) f[r] += s * m.amp[r];
    }
  } else if (c.rhs_kind == PMGP_RHS_EXACT && c.exact_kind == PMGP_EXACT_QUADLIN) {
    double u[3], gu[3][3], lap[3], p, gp[3], b[3];
    exact_quadlin(x, u, gu, lap, p, gp);
    conv_at(P, x, b);
    for (int r = 0; r < 3; ++r)
      f[r] = -c.nu * lap[r] + b[0] * gu[r][0] + b[1] * gu[r][1] + b[2] * gu[r][2] + gp[r];
  }
}

std::vector<Mode> make_modes(unsigned long long seed, int n) {
  std::mt19937_64 rng(seed);
  std::vector<std::array<int, 3>> ks;
  for (int a = -3; a <= 3; ++a)
    for (int b = -3; b <= 3; ++b)
      for (int c = -3; c <= 3; ++c) {
        int r2 = a * a + b * b + c * c;
        if (r2 > 0 && r2 <= 9) ks.push_back({a, b, c});
      }
  std::uniform_int_distribution<size_t> pick(0, ks.size() - 1);
  std::normal_distribution<double> nrm(0.0, 1.0);
  std::uniform_real_distribution<double> ph(0.0, 2 * M_PI);
  std::vector<Mode> out(n);
  for (int m = 0; m < n; ++m) {
    auto k = ks[pick(rng)];
    for (int r = 0; r < 3; ++r) out[m].k[r] = k[r];
    for (int r = 0; r < 3; ++r) out[m].amp[r] = nrm(rng);
    out[m].phase = ph(rng);
  }
  return out;
}

inline int vid(const Level &L, int i, int j, int k) { return i + (L.nx + 1) * (j + (L.ny + 1) * k); }

void cell_corners(const Level &L, int c, double X[8][3]) {
  int i = L.cell_ijk[3 * c], j = L.cell_ijk[3 * c + 1], k = L.cell_ijk[3 * c + 2];
  for (int cc = 0; cc < 2; ++cc)
    for (int bb = 0; bb < 2; ++bb)
      for (int aa = 0; aa < 2; ++aa) {
        const double *v = &L.vert[3 * vid(L, i + aa, j + bb, k + cc)];
        double *o = X[aa + 2 * bb + 4 * cc];
        o[0] = v[0]; o[1] = v[1]; o[2] = v[2];
      }
}

void build_geometry(pmgp_problem &P) {
  const pmgp_config &c = P.cfg;
  P.levels.resize(c.n_levels);
  // coarsest level: lexicographic cells, uniform vertices + interior perturbation
  {
    Level &L = P.levels[0];
    L.nx = c.cx; L.ny = c.cy; L.nz = c.cz;
    L.n_cells = L.nx * L.ny * L.nz;
    L.cell_ijk.resize(3 * L.n_cells);
    int n = 0;
    for (int k = 0; k < L.nz; ++k)
      for (int j = 0; j < L.ny; ++j)
        for (int i = 0; i < L.nx; ++i) {
          L.cell_ijk[3 * n] = i; L.cell_ijk[3 * n + 1] = j; L.cell_ijk[3 * n + 2] = k; ++n;
        }
    L.vert.resize(3 * (L.nx + 1) * (L.ny + 1) * (L.nz + 1));
    double hx = c.lx / L.nx, hy = c.ly / L.ny, hz = c.lz / L.nz;
    std::mt19937_64 rng(c.seed_perturb);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (int k = 0; k <= L.nz; ++k)
      for (int j = 0; j <= L.ny; ++j)
        for (int i = 0; i <= L.nx; ++i) {
          double *v = &L.vert[3 * vid(L, i, j, k)];
          v[0] = i * hx; v[1] = j * hy; v[2] = k * hz;
          if (i == L.nx) v[0] = c.lx;
          if (j == L.ny) v[1] = c.ly;
          if (k == L.nz) v[2] = c.lz;
          bool interior = i > 0 && i < L.nx && j > 0 && j < L.ny && k > 0 && k < L.nz;
          if (c.perturb != 0.0 && interior) {
            v[0] += c.perturb * hx * U(rng);
            v[1] += c.perturb * hy * U(rng);
            v[2] += c.perturb * hz * U(rng);
          }
        }
  }
  // refined levels: children contiguous (cell = 8*parent + child, child = a + 2b + 4c),
  // vertices by trilinear subdivision (edge/face/cell midpoints = averages)
  for (int l = 1; l < c.n_levels; ++l) {
    const Level &C = P.levels[l - 1];
    Level &L = P.levels[l];
    L.nx = 2 * C.nx; L.ny = 2 * C.ny; L.nz = 2 * C.nz;
    L.n_cells = 8 * C.n_cells;
    L.cell_ijk.resize(3 * L.n_cells);
    L.parent.resize(L.n_cells);
    L.child.resize(L.n_cells);
    for (int pc = 0; pc < C.n_cells; ++pc)
      for (int ch = 0; ch < 8; ++ch) {
        int cc = 8 * pc + ch;
        L.cell_ijk[3 * cc] = 2 * C.cell_ijk[3 * pc] + (ch & 1);
        L.cell_ijk[3 * cc + 1] = 2 * C.cell_ijk[3 * pc + 1] + ((ch >> 1) & 1);
        L.cell_ijk[3 * cc + 2] = 2 * C.cell_ijk[3 * pc + 2] + ((ch >> 2) & 1);
        L.parent[cc] = pc;
        L.child[cc] = ch;
      }
    L.vert.resize(3 * (L.nx + 1) * (L.ny + 1) * (L.nz + 1));
    for (int k = 0; k <= L.nz; ++k)
      for (int j = 0; j <= L.ny; ++j)
        for (int i = 0; i <= L.nx; ++i) {
          double s[3] = {0, 0, 0};
          int cnt = 0;
          int i0 = i / 2, i1 = (i + 1) / 2, j0 = j / 2, j1 = (j + 1) / 2, k0 = k / 2, k1 = (k + 1) / 2;
          for (int kk = k0; kk <= k1; ++kk)
            for (int jj = j0; jj <= j1; ++jj)
              for (int ii = i0; ii <= i1; ++ii) {
                const double *v = &C.vert[3 * vid(C, ii, jj, kk)];
                s[0] += v[0]; s[1] += v[1]; s[2] += v[2]; ++cnt;
              }
          double *o = &L.vert[3 * vid(L, i, j, k)];
          o[0] = s[0] / cnt; o[1] = s[1] / cnt; o[2] = s[2] / cnt;
        }
  }
}

void number_dofs(const pmgp_problem &P, Level &L) {
  const pmgp_config &c = P.cfg;
  const int NX = 2 * L.nx + 1, NY = 2 * L.ny + 1, NZ = 2 * L.nz + 1;
  std::vector<int> node(size_t(NX) * NY * NZ, -1);
  std::vector<char> isdir(size_t(NX) * NY * NZ, 0);
  // first-encounter numbering over cells in collection order, active and Dirichlet
  // counted separately; Dirichlet numbers are shifted behind the active block
  int n_act = 0, n_dir = 0;
  L.u_begin.resize(L.n_cells);
  L.u_global.resize(size_t(L.n_cells) * NU_LOC);
  for (int cell = 0; cell < L.n_cells; ++cell) {
    L.u_begin[cell] = cell * NU_LOC;
    int ci = L.cell_ijk[3 * cell], cj = L.cell_ijk[3 * cell + 1], ck = L.cell_ijk[3 * cell + 2];
    for (int cc = 0; cc < 3; ++cc)
      for (int bb = 0; bb < 3; ++bb)
        for (int aa = 0; aa < 3; ++aa) {
          int I = 2 * ci + aa, J = 2 * cj + bb, K = 2 * ck + cc;
          size_t id = I + size_t(NX) * (J + size_t(NY) * K);
          if (node[id] < 0) {
            bool dir = ((c.dirichlet_faces & 1) && I == 0) || ((c.dirichlet_faces & 2) && I == NX - 1) ||
                       ((c.dirichlet_faces & 4) && J == 0) || ((c.dirichlet_faces & 8) && J == NY - 1) ||
                       ((c.dirichlet_faces & 16) && K == 0) || ((c.dirichlet_faces & 32) && K == NZ - 1);
            isdir[id] = dir;
            node[id] = dir ? n_dir++ : n_act++;
          }
          L.u_global[size_t(cell) * NU_LOC + aa + 3 * bb + 9 * cc] = int(id);
        }
  }
  for (auto &g : L.u_global) {
    size_t id = size_t(g);
    g = isdir[id] ? n_act + node[id] : node[id];
  }
  L.n_active = n_act;
  L.n_dirichlet = n_dir;
  L.n_udof = n_act + n_dir;
  L.n_pdof = NP_LOC * L.n_cells;
  L.p_begin.resize(L.n_cells);
  L.p_global.resize(size_t(L.n_cells) * NP_LOC);
  for (int cell = 0; cell < L.n_cells; ++cell) {
    L.p_begin[cell] = cell * NP_LOC;
    for (int m = 0; m < NP_LOC; ++m) L.p_global[size_t(cell) * NP_LOC + m] = NP_LOC * cell + m;
  }
  // coordinates of the dofs (nodal points F_K(xi_k)) and of the cell centres
  L.dof_xyz.assign(size_t(3) * L.n_udof, 0.0);
  L.cell_xyz.assign(size_t(3) * L.n_cells, 0.0);
  for (int cell = 0; cell < L.n_cells; ++cell) {
    double X[8][3], J[3][3], x[3];
    cell_corners(L, cell, X);
    for (int k = 0; k < NU_LOC; ++k) {
      double xi = (k % 3) - 1.0, eta = ((k / 3) % 3) - 1.0, zeta = (k / 9) - 1.0;
      trilinear(X, xi, eta, zeta, x, J);
      int g = L.u_global[size_t(cell) * NU_LOC + k];
      for (int r = 0; r < 3; ++r) L.dof_xyz[3 * size_t(g) + r] = x[r];
    }
    trilinear(X, 0, 0, 0, x, J);
    for (int r = 0; r < 3; ++r) L.cell_xyz[3 * size_t(cell) + r] = x[r];
  }
}

void build_structure(Level &L) {
  const int N = L.n_udof;
  // dof -> cells
  std::vector<int> cnt(N + 1, 0);
  for (int cell = 0; cell < L.n_cells; ++cell)
    for (int k = 0; k < NU_LOC; ++k) cnt[L.u_global[size_t(cell) * NU_LOC + k] + 1]++;
  for (int i = 0; i < N; ++i) cnt[i + 1] += cnt[i];
  std::vector<int> d2c(cnt[N]);
  std::vector<int> pos(cnt.begin(), cnt.end() - 1);
  for (int cell = 0; cell < L.n_cells; ++cell)
    for (int k = 0; k < NU_LOC; ++k) d2c[pos[L.u_global[size_t(cell) * NU_LOC + k]]++] = cell;
  // A: active rows = sorted union of the dofs of the cells around the row dof
  L.a_rowptr.assign(N + 1, 0);
  std::vector<std::vector<int>> rows(N);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < N; ++i) {
    std::vector<int> &r = rows[i];
    if (i < L.n_active) {
      for (int q = cnt[i]; q < cnt[i + 1]; ++q) {
        const int *ud = &L.u_global[size_t(d2c[q]) * NU_LOC];
        r.insert(r.end(), ud, ud + NU_LOC);
      }
      std::sort(r.begin(), r.end());
      r.erase(std::unique(r.begin(), r.end()), r.end());
    } else {
      r.push_back(i);
    }
  }
  for (int i = 0; i < N; ++i) L.a_rowptr[i + 1] = L.a_rowptr[i] + int(rows[i].size());
  L.a_kcol.resize(L.a_rowptr[N]);
  for (int i = 0; i < N; ++i) std::copy(rows[i].begin(), rows[i].end(), L.a_kcol.begin() + L.a_rowptr[i]);
  rows.clear();
  rows.shrink_to_fit();
  for (int b = 0; b < 9; ++b) L.a[b].assign(L.a_kcol.size(), 0.0);
  // B: pressure row -> the cell's velocity dofs in local order
  L.b_rowptr.resize(L.n_pdof + 1);
  L.b_kcol.resize(size_t(L.n_pdof) * NU_LOC);
  for (int pr = 0; pr <= L.n_pdof; ++pr) L.b_rowptr[pr] = pr * NU_LOC;
  for (int cell = 0; cell < L.n_cells; ++cell)
    for (int m = 0; m < NP_LOC; ++m)
      for (int k = 0; k < NU_LOC; ++k)
        L.b_kcol[(size_t(cell) * NP_LOC + m) * NU_LOC + k] = L.u_global[size_t(cell) * NU_LOC + k];
  for (int b = 0; b < 3; ++b) L.b[b].assign(L.b_kcol.size(), 0.0);
  // BT: velocity row -> pressure dofs of the surrounding cells (cell order, mode order)
  L.bt_rowptr.assign(N + 1, 0);
  for (int i = 0; i < N; ++i) L.bt_rowptr[i + 1] = L.bt_rowptr[i] + NP_LOC * (cnt[i + 1] - cnt[i]);
  L.bt_kcol.resize(L.bt_rowptr[N]);
  for (int i = 0; i < N; ++i) {
    int o = L.bt_rowptr[i];
    for (int q = cnt[i]; q < cnt[i + 1]; ++q)
      for (int m = 0; m < NP_LOC; ++m) L.bt_kcol[o++] = NP_LOC * d2c[q] + m;
  }
  for (int b = 0; b < 3; ++b) L.bt[b].assign(L.bt_kcol.size(), 0.0);
}

inline int find_col(const int *kcol, int lo, int hi, int col) {
  const int *p = std::lower_bound(kcol + lo, kcol + hi, col);
  return int(p - kcol);
}

void assemble(const pmgp_problem &P, Level &L) {
  const pmgp_config &c = P.cfg;
  static const double gp[3] = {-0.7745966692414833770358531, 0.0, 0.7745966692414833770358531};
  static const double gw[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
  // reference basis at the 27 quadrature points
  double Bv[27][27], Bx[27][27], By[27][27], Bz[27][27];
  for (int q = 0; q < 27; ++q) q2_basis(gp[q % 3], gp[(q / 3) % 3], gp[q / 9], Bv[q], Bx[q], By[q], Bz[q]);
  const size_t NC = L.n_cells;
  const int LA = NU_LOC * NU_LOC, LB = NP_LOC * NU_LOC;
  std::vector<double> eA(NC * LA), eB(NC * 3 * LB), eF(NC * 3 * NU_LOC);
#pragma omp parallel for schedule(static)
  for (long cell = 0; cell < long(NC); ++cell) {
    double X[8][3];
    cell_corners(L, int(cell), X);
    double *A = &eA[cell * LA], *Bm = &eB[cell * 3 * LB], *F = &eF[cell * 3 * NU_LOC];
    std::fill(A, A + LA, 0.0);
    std::fill(Bm, Bm + 3 * LB, 0.0);
    std::fill(F, F + 3 * NU_LOC, 0.0);
    for (int q = 0; q < 27; ++q) {
      double xi = gp[q % 3], eta = gp[(q / 3) % 3], zeta = gp[q / 9];
      double w = gw[q % 3] * gw[(q / 3) % 3] * gw[q / 9];
      double x[3], J[3][3], Ji[3][3];
      trilinear(X, xi, eta, zeta, x, J);
      double det = inv3(J, Ji);
      double mult = w * std::fabs(det);
      double gx[27], gy[27], gz[27];
      for (int k = 0; k < 27; ++k) {
        // physical gradient = J^{-T} grad_ref
        gx[k] = Ji[0][0] * Bx[q][k] + Ji[1][0] * By[q][k] + Ji[2][0] * Bz[q][k];
        gy[k] = Ji[0][1] * Bx[q][k] + Ji[1][1] * By[q][k] + Ji[2][1] * Bz[q][k];
        gz[k] = Ji[0][2] * Bx[q][k] + Ji[1][2] * By[q][k] + Ji[2][2] * Bz[q][k];
      }
      double b[3], f[3];
      conv_at(P, x, b);
      rhs_at(P, x, f);
      double psi[4] = {1.0, xi, eta, zeta};
      for (int i = 0; i < 27; ++i) {
        double t0 = Bv[q][i];
        double val1 = mult * t0;
        F[i] += val1 * f[0];
        F[27 + i] += val1 * f[1];
        F[54 + i] += val1 * f[2];
        for (int j = 0; j < 27; ++j) {
          double val = c.nu * (gx[i] * gx[j] + gy[i] * gy[j] + gz[i] * gz[j]);
          val += (b[0] * gx[j] + b[1] * gy[j] + b[2] * gz[j]) * t0;
          if (c.mass != 0.0) val += c.mass * Bv[q][j] * t0;
          A[i * 27 + j] += mult * val;
        }
      }
      for (int m = 0; m < 4; ++m) {
        double v1 = mult * psi[m];
        for (int j = 0; j < 27; ++j) {
          Bm[0 * LB + m * 27 + j] += -v1 * gx[j];
          Bm[1 * LB + m * 27 + j] += -v1 * gy[j];
          Bm[2 * LB + m * 27 + j] += -v1 * gz[j];
        }
      }
    }
  }
  // scatter
  L.rhs.assign(size_t(3) * L.n_udof + L.n_pdof, 0.0);
  for (size_t cell = 0; cell < NC; ++cell) {
    const int *ud = &L.u_global[cell * NU_LOC];
    const double *A = &eA[cell * LA], *Bm = &eB[cell * 3 * LB], *F = &eF[cell * 3 * NU_LOC];
    for (int i = 0; i < 27; ++i) {
      int row = ud[i];
      if (row >= L.n_active) continue;
      int lo = L.a_rowptr[row], hi = L.a_rowptr[row + 1];
      for (int j = 0; j < 27; ++j) {
        int p = find_col(L.a_kcol.data(), lo, hi, ud[j]);
        double v = A[i * 27 + j];
        L.a[0][p] += v;
        L.a[4][p] += v;
        L.a[8][p] += v;
      }
      for (int r = 0; r < 3; ++r) L.rhs[size_t(r) * L.n_udof + row] += F[r * 27 + i];
    }
    for (int m = 0; m < 4; ++m) {
      size_t pr = cell * NP_LOC + m;
      for (int j = 0; j < 27; ++j)
        for (int r = 0; r < 3; ++r) L.b[r][pr * NU_LOC + j] = Bm[r * LB + m * 27 + j];
    }
  }
  // BT(i, m) = B(m, i): walk the BT rows
  {
    // position of (cell, m, local j) in B is (cell*4+m)*27 + j; BT row i lists the cells
    // around dof i in increasing order, so fill from the cell side
    std::vector<int> fillpos(L.n_udof);
    for (int i = 0; i < L.n_udof; ++i) fillpos[i] = L.bt_rowptr[i];
    for (size_t cell = 0; cell < NC; ++cell) {
      const int *ud = &L.u_global[cell * NU_LOC];
      for (int j = 0; j < 27; ++j) {
        int row = ud[j];
        int o = fillpos[row];
        for (int m = 0; m < 4; ++m)
          for (int r = 0; r < 3; ++r) L.bt[r][o + m] = L.b[r][(cell * NP_LOC + m) * NU_LOC + j];
        fillpos[row] = o + 4;
      }
    }
  }
  // Dirichlet rows: identity in the diagonal blocks, rhs = boundary value
  for (int i = L.n_active; i < L.n_udof; ++i) {
    int p = L.a_rowptr[i];
    L.a[0][p] = L.a[4][p] = L.a[8][p] = 1.0;
    double u[3] = {0, 0, 0};
    if (c.exact_kind == PMGP_EXACT_QUADLIN) {
      double gu[3][3], lap[3], pp, gpv[3];
      exact_quadlin(&L.dof_xyz[3 * size_t(i)], u, gu, lap, pp, gpv);
    } else if (c.conv_kind == PMGP_CONV_CHANNEL) {
      conv_at(P, &L.dof_xyz[3 * size_t(i)], u);   // inflow profile, zero on the walls
    }
    for (int r = 0; r < 3; ++r) L.rhs[size_t(r) * L.n_udof + i] = u[r];
  }
  // nodal interpolant of the exact solution
  if (c.exact_kind == PMGP_EXACT_QUADLIN) {
    L.exact.assign(L.rhs.size(), 0.0);
    for (int i = 0; i < L.n_udof; ++i) {
      double u[3], gu[3][3], lap[3], pp, gpv[3];
      exact_quadlin(&L.dof_xyz[3 * size_t(i)], u, gu, lap, pp, gpv);
      for (int r = 0; r < 3; ++r) L.exact[size_t(r) * L.n_udof + i] = u[r];
    }
    // P1disc functionals (NF_D_H_P1_3D): Gauss-point moments of p o F_K
    for (int cell = 0; cell < L.n_cells; ++cell) {
      double X[8][3];
      cell_corners(L, cell, X);
      double pv[8], gx[8], gy[8], gz[8];
      for (int q = 0; q < 8; ++q) {
        gx[q] = (q & 1) ? kGaussP1 : -kGaussP1;
        gy[q] = (q & 2) ? kGaussP1 : -kGaussP1;
        gz[q] = (q & 4) ? kGaussP1 : -kGaussP1;
        double x[3], J[3][3], u[3], gu[3][3], lap[3], gpv[3];
        trilinear(X, gx[q], gy[q], gz[q], x, J);
        exact_quadlin(x, u, gu, lap, pv[q], gpv);
      }
      double f0 = 0, f1 = 0, f2 = 0, f3 = 0;
      for (int q = 0; q < 8; ++q) { f0 += pv[q]; f1 += pv[q] * gx[q]; f2 += pv[q] * gy[q]; f3 += pv[q] * gz[q]; }
      size_t o = size_t(3) * L.n_udof + size_t(NP_LOC) * cell;
      L.exact[o] = f0 * 0.125; L.exact[o + 1] = f1 * 0.375;
      L.exact[o + 2] = f2 * 0.375; L.exact[o + 3] = f3 * 0.375;
    }
  }
}

// QQ[child][k*n+l] = fine functional k applied to coarse basis function l
void build_qq(pmgp_problem &P) {
  P.qq_u.assign(8 * 27 * 27, 0.0);
  P.qq_p.assign(8 * 4 * 4, 0.0);
  for (int ch = 0; ch < 8; ++ch) {
    double s[3] = {(ch & 1) ? 0.5 : -0.5, (ch & 2) ? 0.5 : -0.5, (ch & 4) ? 0.5 : -0.5};
    // Q2: values of the parent basis at the child's nodes (NF_C_H_Q2_3D point values)
    for (int k = 0; k < 27; ++k) {
      double xi = s[0] + 0.5 * ((k % 3) - 1.0), eta = s[1] + 0.5 * (((k / 3) % 3) - 1.0),
             zeta = s[2] + 0.5 * ((k / 9) - 1.0);
      double v[27];
      q2_basis(xi, eta, zeta, v, nullptr, nullptr, nullptr);
      for (int l = 0; l < 27; ++l) P.qq_u[ch * 729 + k * 27 + l] = v[l];
    }
    // P1disc: NF_D_H_P1_3D moments at the child's 8 Gauss points of the parent's {1,xi,eta,zeta}
    double pts[8][3];
    for (int q = 0; q < 8; ++q) {
      pts[q][0] = (q & 1) ? kGaussP1 : -kGaussP1;
      pts[q][1] = (q & 2) ? kGaussP1 : -kGaussP1;
      pts[q][2] = (q & 4) ? kGaussP1 : -kGaussP1;
    }
    for (int l = 0; l < 4; ++l) {
      double pv[8];
      for (int q = 0; q < 8; ++q) {
        double xp[3] = {s[0] + 0.5 * pts[q][0], s[1] + 0.5 * pts[q][1], s[2] + 0.5 * pts[q][2]};
        pv[q] = (l == 0) ? 1.0 : xp[l - 1];
      }
      double f[4] = {0, 0, 0, 0};
      for (int q = 0; q < 8; ++q) {
        f[0] += pv[q];
        f[1] += pv[q] * pts[q][0];
        f[2] += pv[q] * pts[q][1];
        f[3] += pv[q] * pts[q][2];
      }
      f[0] *= 0.125; f[1] *= 0.375; f[2] *= 0.375; f[3] *= 0.375;
      for (int k = 0; k < 4; ++k) P.qq_p[ch * 16 + k * 4 + l] = f[k];
    }
  }
}

}  // namespace

extern "C" {

void pmgp_config_default(pmgp_config *c) {
  std::memset(c, 0, sizeof(*c));
  c->cx = c->cy = c->cz = 1;
  c->lx = c->ly = c->lz = 1.0;
  c->n_levels = 3;
  c->nu = 1.0;
  c->dirichlet_faces = 63;
  c->conv_kind = PMGP_CONV_NONE;
  c->conv_max = 1.0;
  c->rhs_kind = PMGP_RHS_FOURIER;
  c->seed_conv = 1001;
  c->seed_rhs = 1002;
  c->seed_perturb = 4000;
  c->n_modes = 8;
}

int pmgp_create(const pmgp_config *c, pmgp_problem **out) {
  if (!c || !out || c->n_levels < 1 || c->cx < 1 || c->cy < 1 || c->cz < 1) return PMG_ERR_INVALID_ARGUMENT;
  auto *P = new pmgp_problem();
  P->cfg = *c;
  if (P->cfg.n_modes <= 0) P->cfg.n_modes = 8;
  P->conv_modes = make_modes(c->seed_conv, P->cfg.n_modes);
  P->rhs_modes = make_modes(c->seed_rhs, P->cfg.n_modes);
  build_geometry(*P);
  build_qq(*P);
  // scale of the Fourier convection: max over the finest level's quadrature points
  if (c->conv_kind == PMGP_CONV_FOURIER) {
    P->conv_scale = 1.0;
    const Level &L = P->levels.back();
    static const double gp[3] = {-0.7745966692414833770358531, 0.0, 0.7745966692414833770358531};
    double mx = 0;
#pragma omp parallel for reduction(max : mx)
    for (int cell = 0; cell < L.n_cells; ++cell) {
      double X[8][3];
      cell_corners(L, cell, X);
      for (int q = 0; q < 27; ++q) {
        double x[3], J[3][3], b[3];
        trilinear(X, gp[q % 3], gp[(q / 3) % 3], gp[q / 9], x, J);
        conv_at(*P, x, b);
        mx = std::max(mx, std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]));
      }
    }
    P->conv_scale = (mx > 0) ? c->conv_max / mx : 1.0;
  }
  for (auto &L : P->levels) {
    number_dofs(*P, L);
    build_structure(L);
    assemble(*P, L);
  }
  *out = P;
  return PMG_OK;
}

void pmgp_destroy(pmgp_problem *p) { delete p; }

int pmgp_n_levels(const pmgp_problem *p) { return p ? int(p->levels.size()) : 0; }

int pmgp_level_desc(const pmgp_problem *p, int level, pmg_level_desc *o) {
  if (!p || !o || level < 0 || level >= int(p->levels.size())) return PMG_ERR_INVALID_ARGUMENT;
  const Level &L = p->levels[level];
  std::memset(o, 0, sizeof(*o));
  o->n_udof = L.n_udof; o->n_active = L.n_active; o->hanging_bound = L.n_active;
  o->n_dirichlet = L.n_dirichlet; o->n_pdof = L.n_pdof;
  o->a_rowptr = L.a_rowptr.data(); o->a_kcol = L.a_kcol.data(); o->a_nnz = int(L.a_kcol.size());
  for (int b = 0; b < 9; ++b) o->a[b] = L.a[b].data();
  o->b_rowptr = L.b_rowptr.data(); o->b_kcol = L.b_kcol.data(); o->b_nnz = int(L.b_kcol.size());
  o->bt_rowptr = L.bt_rowptr.data(); o->bt_kcol = L.bt_kcol.data(); o->bt_nnz = int(L.bt_kcol.size());
  for (int b = 0; b < 3; ++b) { o->b[b] = L.b[b].data(); o->bt[b] = L.bt[b].data(); }
  o->n_cells = L.n_cells; o->n_ulocal = NU_LOC; o->n_plocal = NP_LOC;
  o->u_begin = L.u_begin.data(); o->u_global = L.u_global.data();
  o->p_begin = L.p_begin.data(); o->p_global = L.p_global.data();
  o->n_colors = 0; o->ptr_colors = nullptr; o->cell_reorder = nullptr;
  if (level > 0) {
    o->parent = L.parent.data(); o->child = L.child.data(); o->n_children = 8;
    o->qq_u = p->qq_u.data(); o->qq_p = p->qq_p.data();
  }
  return PMG_OK;
}

int pmgp_n_dof(const pmgp_problem *p, int level) {
  const Level &L = p->levels[level];
  return 3 * L.n_udof + L.n_pdof;
}
const double *pmgp_rhs(const pmgp_problem *p, int level) { return p->levels[level].rhs.data(); }
const double *pmgp_exact(const pmgp_problem *p, int level) {
  return p->levels[level].exact.empty() ? nullptr : p->levels[level].exact.data();
}
const double *pmgp_dof_coords(const pmgp_problem *p, int level) { return p->levels[level].dof_xyz.data(); }
const double *pmgp_cell_centers(const pmgp_problem *p, int level) { return p->levels[level].cell_xyz.data(); }
double pmgp_conv_scale(const pmgp_problem *p) { return p->conv_scale; }
int pmgp_level_dims(const pmgp_problem *p, int level, int *nx, int *ny, int *nz) {
  const Level &L = p->levels[level];
  *nx = L.nx; *ny = L.ny; *nz = L.nz;
  return PMG_OK;
}

int pmgp_cell_vertices(const pmgp_problem *p, int level, double *out) {
  const Level &L = p->levels[level];
  for (int c = 0; c < L.n_cells; ++c) {
    double X[8][3];
    cell_corners(L, c, X);
    for (int k = 0; k < 8; ++k)
      for (int r = 0; r < 3; ++r) out[(size_t(c) * 8 + k) * 3 + r] = X[k][r];
  }
  return PMG_OK;
}

int pmgp_conv_quad(const pmgp_problem *p, int level, double *out) {
  static const double gp[3] = {-0.7745966692414833770358531, 0.0, 0.7745966692414833770358531};
  const Level &L = p->levels[level];
  for (int c = 0; c < L.n_cells; ++c) {
    double X[8][3];
    cell_corners(L, c, X);
    for (int q = 0; q < 27; ++q) {
      double x[3], J[3][3], b[3];
      trilinear(X, gp[q % 3], gp[(q / 3) % 3], gp[q / 9], x, J);
      conv_at(*p, x, b);
      for (int r = 0; r < 3; ++r) out[(size_t(c) * 27 + q) * 3 + r] = b[r];
    }
  }
  return PMG_OK;
}

}  // extern "C"
