// k_assemble.hip - device-side re-assembly of the Oseen velocity blocks (gfx950): the
// element matrices of NSType4Galerkin3D (src/FE/NSE3D_FixPo.C:1919-2071) with the
// convection field of the fixed-point linearisation, and the evaluation of a Q2 velocity
// field at the quadrature points (the 'param' values the reference's assembly framework
// hands to it, src/FE/Assemble3D.C).
//
// Q2 hexahedra (27 nodes, xi fastest, include/FE3D/NF_C_H_Q2_3D.h), trilinear cell map from
// the 8 corners (corner index a + 2b + 4c), 3x3x3 Gauss rule (point q: xi = g[q % 3],
// eta = g[q / 3 % 3], zeta = g[q / 9]).  Element entry (test i, ansatz j):
//   sum_q w_q |det J| ( nu grad phi_i . grad phi_j + (b_q . grad phi_j) phi_i + sigma phi_j phi_i )
// added to A11, A22 and A33 of the active rows; A12..A32 are zero for this Galerkin form.
// Cells of one color share no velocity dof, so one launch per color scatters without
// atomics.
#include "pmg_internal.hpp"

namespace pmg {

namespace {

__device__ const double kGP[3] = {-0.7745966692414833770358531, 0.0, 0.7745966692414833770358531};
__device__ const double kGW[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};

__device__ inline void lag1(double t, double v[3], double d[3]) {
  v[0] = 0.5 * t * (t - 1.0); v[1] = 1.0 - t * t; v[2] = 0.5 * t * (t + 1.0);
  d[0] = t - 0.5; d[1] = -2.0 * t; d[2] = t + 0.5;
}

// value and reference gradient of Q2 basis i at quadrature point q
__device__ inline void q2_at(int q, int i, double &v, double &dx, double &dy, double &dz) {
  double vx[3], ax[3], vy[3], ay[3], vz[3], az[3];
  lag1(kGP[q % 3], vx, ax);
  lag1(kGP[(q / 3) % 3], vy, ay);
  lag1(kGP[q / 9], vz, az);
  const int a = i % 3, b = (i / 3) % 3, c = i / 9;
  v = vx[a] * vy[b] * vz[c];
  dx = ax[a] * vy[b] * vz[c];
  dy = vx[a] * ay[b] * vz[c];
  dz = vx[a] * vy[b] * az[c];
}

// zero the nine A blocks of the active rows (entries [0, a_rowptr[n_active]))
__global__ void k_zero_active(double *__restrict__ a, long nnz, long n_act_entries) {
  for (long k = blockIdx.x * long(blockDim.x) + threadIdx.x; k < 9 * n_act_entries;
       k += long(gridDim.x) * blockDim.x) {
    const long b = k / n_act_entries, e = k - b * n_act_entries;
    a[b * nnz + e] = 0.0;
  }
}

// Geometry and basis data of one cell in LDS: weights w_q |det J| (W), values Vq[q][i] and
// physical gradients G*[q][i] of the 27 Q2 basis functions at the 27 Gauss points, and the
// convection at the points (Bc).  Called by all 256 threads; ends with a barrier.
struct CellData {
  double X[8][3], Bc[27][3], Ji[27][9], W[27];
  double Vq[27][27], Gx[27][27], Gy[27][27], Gz[27][27];
};

__device__ void cell_setup(CellData &D, int cell, const double *__restrict__ geom, const double *__restrict__ bq) {
  const int t = threadIdx.x;
  if (t < 24) D.X[t / 3][t % 3] = geom[long(cell) * 24 + t];
  if (t < 81) D.Bc[t / 3][t % 3] = bq[long(cell) * 81 + t];
  __syncthreads();
  if (t < 27) {
    const double xi = kGP[t % 3], eta = kGP[(t / 3) % 3], zeta = kGP[t / 9];
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int c = 0; c < 2; ++c)
      for (int b = 0; b < 2; ++b)
        for (int aa = 0; aa < 2; ++aa) {
          const double sx = aa ? 1 : -1, sy = b ? 1 : -1, sz = c ? 1 : -1;
          const double nx = 0.5 * (1 + sx * xi), ny = 0.5 * (1 + sy * eta), nz = 0.5 * (1 + sz * zeta);
          const double dNx = 0.5 * sx * ny * nz, dNy = 0.5 * sy * nx * nz, dNz = 0.5 * sz * nx * ny;
          const double *P = D.X[aa + 2 * b + 4 * c];
          for (int r = 0; r < 3; ++r) {
            J[r][0] += dNx * P[r];
            J[r][1] += dNy * P[r];
            J[r][2] += dNz * P[r];
          }
        }
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    const double id = 1.0 / det;
    double *I = D.Ji[t];
    I[0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
    I[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    I[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    I[3] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
    I[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    I[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    I[6] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
    I[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    I[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    D.W[t] = kGW[t % 3] * kGW[(t / 3) % 3] * kGW[t / 9] * fabs(det);
  }
  __syncthreads();
  // physical gradients: J^{-T} times the reference gradient
  for (int idx = t; idx < 27 * 27; idx += 256) {
    const int q = idx / 27, i = idx - 27 * q;
    double v, dx, dy, dz;
    q2_at(q, i, v, dx, dy, dz);
    const double *I = D.Ji[q];
    D.Vq[q][i] = v;
    D.Gx[q][i] = I[0] * dx + I[3] * dy + I[6] * dz;
    D.Gy[q][i] = I[1] * dx + I[4] * dy + I[7] * dz;
    D.Gz[q][i] = I[2] * dx + I[5] * dy + I[8] * dz;
  }
  __syncthreads();
}

// element entry (test i, ansatz j) of NSType4Galerkin3D with the mass term sigma
__device__ inline double element_entry(const CellData &D, int i, int j, double nu, double sigma) {
  double s = 0.0;
  for (int q = 0; q < 27; ++q) {
    double val = nu * (D.Gx[q][i] * D.Gx[q][j] + D.Gy[q][i] * D.Gy[q][j] + D.Gz[q][i] * D.Gz[q][j]);
    val += (D.Bc[q][0] * D.Gx[q][j] + D.Bc[q][1] * D.Gy[q][j] + D.Bc[q][2] * D.Gz[q][j]) * D.Vq[q][i];
    if (sigma != 0.0) val += sigma * D.Vq[q][j] * D.Vq[q][i];
    s += D.W[q] * val;
  }
  return s;
}

__global__ __launch_bounds__(256) void k_assemble_oseen(const int *__restrict__ cells, int ncells,
                                                        const double *__restrict__ geom,
                                                        const double *__restrict__ bq,
                                                        const int *__restrict__ u_global,
                                                        const int *__restrict__ pmap,
                                                        const int *__restrict__ a_rowptr, int n_active, long nnz,
                                                        double *__restrict__ a, double nu, double sigma) {
  __shared__ CellData D;
  if (int(blockIdx.x) >= ncells) return;
  const int cell = cells[blockIdx.x];
  cell_setup(D, cell, geom, bq);
  const int *ug = u_global + long(cell) * NU_LOC;
  const int *mp = pmap + long(cell) * PM_STRIDE;
  for (int idx = threadIdx.x; idx < 27 * 27; idx += 256) {
    const int i = idx / 27, j = idx - 27 * i;
    const int row = ug[i];
    if (row >= n_active) continue;   // Dirichlet rows keep their identity
    const double s = element_entry(D, i, j, nu, sigma);
    // entry (row, col j) in the CSR: the gather map holds its row-blocked offset
    // 9 k0 + (k - k0) of block 0
    const int off = mp[PM_A + i * NU_LOC + j];
    if (off < 0) continue;
    const long k = long(off) - 8L * a_rowptr[row];
    a[k] += s;
    a[4 * nnz + k] += s;
    a[8 * nnz + k] += s;
  }
}

// rhs_c(row i) -= sum_j E_ij(nu, sigma') u_c(j) for the active rows of one color's cells:
// with sigma' = -sigma, sigma M u - A u of the theta scheme's right-hand side
__global__ __launch_bounds__(256) void k_theta_rhs(const int *__restrict__ cells, int ncells,
                                                   const double *__restrict__ geom, const double *__restrict__ bq,
                                                   const int *__restrict__ u_global, int n_udof, int n_active,
                                                   const double *__restrict__ u, double *__restrict__ rhs, double nu,
                                                   double sigma_neg) {
  __shared__ CellData D;
  __shared__ double E[27][28];
  if (int(blockIdx.x) >= ncells) return;
  const int cell = cells[blockIdx.x];
  cell_setup(D, cell, geom, bq);
  for (int idx = threadIdx.x; idx < 27 * 27; idx += 256) {
    const int i = idx / 27, j = idx - 27 * i;
    E[i][j] = element_entry(D, i, j, nu, sigma_neg);
  }
  __syncthreads();
  const int *ug = u_global + long(cell) * NU_LOC;
  const int t = threadIdx.x;
  if (t < 81) {
    const int i = t / 3, c = t - 3 * i;
    const int row = ug[i];
    if (row < n_active) {
      double s = 0.0;
      for (int j = 0; j < 27; ++j) s += E[i][j] * u[long(c) * n_udof + ug[j]];
      rhs[long(c) * n_udof + row] -= s;
    }
  }
}

// theta-scheme right-hand side outside the active velocity rows: Dirichlet values from g
// (velocity rows >= n_active), zero pressure rows; the active velocity rows start at 0
__global__ void k_theta_rhs_init(double *__restrict__ rhs, const double *__restrict__ g, int n_udof, int n_active,
                                 int n_pdof) {
  const long i = blockIdx.x * long(blockDim.x) + threadIdx.x;
  const long n = 3L * n_udof + n_pdof;
  if (i >= n) return;
  if (i < 3L * n_udof) {
    const int r = int(i % n_udof);
    rhs[i] = r < n_active ? 0.0 : g[i];
  } else {
    rhs[i] = 0.0;
  }
}

// b[cell][q][c] = sum_i u_c(UDOF_i) phi_i(q)
__global__ void k_eval_velocity(int ncells, const double *__restrict__ u, int N, const int *__restrict__ u_global,
                                double *__restrict__ bq) {
  const long idx = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (idx >= long(ncells) * 27) return;
  const int cell = int(idx / 27), q = int(idx - 27L * cell);
  const int *ug = u_global + long(cell) * NU_LOC;
  double s0 = 0, s1 = 0, s2 = 0;
  for (int i = 0; i < 27; ++i) {
    double v, dx, dy, dz;
    q2_at(q, i, v, dx, dy, dz);
    const int g = ug[i];
    s0 += v * u[g];
    s1 += v * u[N + g];
    s2 += v * u[2L * N + g];
  }
  bq[idx * 3 + 0] = s0;
  bq[idx * 3 + 1] = s1;
  bq[idx * 3 + 2] = s2;
}

}  // namespace

void launch_assemble_oseen(const DevLevel &L, const double *geom, const double *bq, double nu, double sigma,
                           hipStream_t s) {
  std::vector<int> rp(2);
  PMG_HIP_CHECK(hipMemcpyAsync(rp.data(), L.a_rowptr + L.n_active, sizeof(int), hipMemcpyDeviceToHost, s));
  PMG_HIP_CHECK(hipStreamSynchronize(s));
  const long n_act = rp[0];
  k_zero_active<<<2048, 256, 0, s>>>(L.a, L.a_nnz, n_act);
  PMG_HIP_CHECK(hipGetLastError());
  for (int col = 0; col + 1 < int(L.ptr_colors.size()); ++col) {
    const int p0 = L.ptr_colors[col], nc = L.ptr_colors[col + 1] - p0;
    if (nc <= 0) continue;
    k_assemble_oseen<<<nc, 256, 0, s>>>(L.cell_reorder + p0, nc, geom, bq, L.u_global, L.pmap, L.a_rowptr, L.n_active,
                                        L.a_nnz, L.a, nu, sigma);
    PMG_HIP_CHECK(hipGetLastError());
  }
}

void launch_theta_rhs(const DevLevel &L, const double *geom, const double *bq, double nu, double sigma,
                      const double *u, const double *g, double *rhs, hipStream_t s) {
  const long n = 3L * L.n_udof + L.n_pdof;
  k_theta_rhs_init<<<int((n + 255) / 256), 256, 0, s>>>(rhs, g, L.n_udof, L.n_active, L.n_pdof);
  PMG_HIP_CHECK(hipGetLastError());
  for (int col = 0; col + 1 < int(L.ptr_colors.size()); ++col) {
    const int p0 = L.ptr_colors[col], nc = L.ptr_colors[col + 1] - p0;
    if (nc <= 0) continue;
    k_theta_rhs<<<nc, 256, 0, s>>>(L.cell_reorder + p0, nc, geom, bq, L.u_global, L.n_udof, L.n_active, u, rhs, nu,
                                   -sigma);
    PMG_HIP_CHECK(hipGetLastError());
  }
}

void launch_eval_velocity(const DevLevel &L, const double *u, double *bq, hipStream_t s) {
  const long n = long(L.n_cells) * 27;
  k_eval_velocity<<<int((n + 255) / 256), 256, 0, s>>>(L.n_cells, u, L.n_udof, L.u_global, bq);
  PMG_HIP_CHECK(hipGetLastError());
}

}  // namespace pmg
