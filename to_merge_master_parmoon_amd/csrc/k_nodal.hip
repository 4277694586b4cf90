// k_nodal.hip - nodal Vanka smoother (TNSE_MGLevel4::NodalVanka, smoothers 3 and 4,
// src/FE/NSE_MGLevel4.C:4098-4549) for gfx950.
//
// Every pressure dof i is a patch with the velocity dofs of its B row (N_U of them, the
// 27 nodes of its cell for Q2/P1disc) and itself: 3 N_U + 1 local unknowns.  The reference
// visits the pressure dofs in order with immediate updates; the GPU runs them in the level
// sets of that order (no two coupled patches in one set, every coupled pair in order), one
// launch per set, one workgroup per patch doing residual, local solve and update at once
// (the patches of a set read none of each other's dofs, so no barrier between them is
// needed).  Smoother 3 solves with the velocity block replaced by its diagonal (the first
// component's diagonal for all three, SolveDiagonalVanka3D, MGComponents3D.C:32-85);
// smoother 4 assembles the local matrix and factors it with partial pivoting (dgetf2 order)
// in LDS, as the reference factors it at every visit.
#include "pmg_internal.hpp"
#include "pmg_rows.hpp"

namespace pmg {

namespace {

constexpr int NV_T = 128;                    // threads per patch
constexpr int NV_LD = 3 * NODAL_MAXU + 1;    // leading dimension of the local matrix

template <bool FULL>
__global__ __launch_bounds__(NV_T) void k_nodal_vanka(OpView<double> op, const int *__restrict__ pdofs, int n,
                                                      const double *__restrict__ f, double *__restrict__ u,
                                                      double damp, int *__restrict__ info) {
  // products and sums as written (no contraction), explicit fma where the reference's
  // factorisation uses one: the sweep reproduces the reference's arithmetic
#pragma clang fp contract(off)
  __shared__ int ud[NODAL_MAXU];
  __shared__ double bl[3][NODAL_MAXU], btl[3][NODAL_MAXU], dg[NODAL_MAXU];
  __shared__ double r[NV_LD];
  __shared__ double S[FULL ? NV_LD * NV_LD : 1];   // column-major S[col * NV_LD + row]
  __shared__ int pv_s;
  if (int(blockIdx.x) >= n) return;
  const int i = pdofs[blockIdx.x];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int N = op.n_udof, NA = op.n_active;
  const int k0 = op.b_rowptr[i], nU = op.b_rowptr[i + 1] - k0;
  const int NG = 3 * nU, NL = NG + 1;
  if (t < nU) {
    ud[t] = op.b_kcol[k0 + t];
    for (int c = 0; c < 3; ++c) {
      bl[c][t] = op.b[c * op.b_nnz + k0 + t];
      btl[c][t] = 0.0;
    }
    dg[t] = 0.0;
  }
  if (FULL)
    for (int idx = t; idx < NL * NL; idx += NV_T) S[(idx / NL) * NV_LD + idx % NL] = 0.0;
  __syncthreads();
  // local residual of the velocity rows (one thread per row and component), the
  // first component's diagonal, the BT entries of column i and (FULL) the local matrix
  if (t < NG) {
    const int c = t / nU, j = t - c * nU;
    const int row = ud[j];
    const bool dir = row >= NA;
    double s = f[long(c) * N + row];
    for (int k = op.a_rowptr[row]; k < op.a_rowptr[row + 1]; ++k) {
      const int l = op.a_kcol[k];
      double v[3];
      for (int cc = 0; cc < 3; ++cc) v[cc] = (dir && cc != c) ? 0.0 : op.a[(3 * c + cc) * op.a_nnz + k];
      s -= v[0] * u[l] + v[1] * u[long(N) + l] + v[2] * u[2L * N + l];
      if (c == 0 && l == row) dg[j] = v[0];
      if (FULL) {
        int m = 0;
        while (m < nU && ud[m] != l) ++m;
        if (m < nU)
          for (int cc = 0; cc < 3; ++cc) S[(m + cc * nU) * NV_LD + t] = v[cc];
      }
    }
    if (!dir)
      for (int k = op.bt_rowptr[row]; k < op.bt_rowptr[row + 1]; ++k) {
        const int q = op.bt_kcol[k];
        const double v = op.bt[c * op.bt_nnz + k];
        s -= v * u[3L * N + q];
        if (q == i) {
          btl[c][j] = v;
          if (FULL) S[NG * NV_LD + t] = v;
        }
      }
    r[t] = s;
  }
  // pressure row: g_i - B_i u, in the reference's order (one lane)
  if (t == 0) {
    double s = f[3L * N + i];
    for (int j = 0; j < nU; ++j) {
      const int l = ud[j];
      s -= bl[0][j] * u[l] + bl[1][j] * u[long(N) + l] + bl[2][j] * u[2L * N + l];
    }
    r[NG] = s;
  }
  if (FULL && t < nU)
    for (int c = 0; c < 3; ++c) S[(t + c * nU) * NV_LD + NG] = bl[c][t];
  __syncthreads();
  if (!FULL) {
    // SolveDiagonalVanka3D: Schur complement of the diagonal block, summed in the
    // reference's order by one lane, then the velocities
    if (t == 0) {
      double dp = 0.0, ffp = r[NG];
      for (int j = 0; j < nU; ++j) {
        const double tmp = 1.0 / dg[j];
        dp -= tmp * (bl[0][j] * btl[0][j] + bl[1][j] * btl[1][j] + bl[2][j] * btl[2][j]);
        ffp -= tmp * (bl[0][j] * r[j] + bl[1][j] * r[nU + j] + bl[2][j] * r[2 * nU + j]);
      }
      r[NG] = ffp / dp;
    }
    __syncthreads();
    if (t < NG) {
      const int c = t / nU, j = t - c * nU;
      r[t] = (1.0 / dg[j]) * (r[t] - btl[c][j] * r[NG]);
    }
    __syncthreads();
  } else {
    // dgetf2 with the right-hand side carried along: pivot = first row of largest |S(r, k)|
    bool singular = false;
    for (int k = 0; k < NL; ++k) {
      if (wave == 0) {
        double bv = -1.0;
        int bi = NL;
        for (int rr = k + lane; rr < NL; rr += 64) {
          const double x = fabs(S[k * NV_LD + rr]);
          if (x > bv) {
            bv = x;
            bi = rr;
          }
        }
        for (int o = 32; o > 0; o >>= 1) {
          const double ov = __shfl_xor(bv, o, 64);
          const int oi = __shfl_xor(bi, o, 64);
          if (ov > bv || (ov == bv && oi < bi)) {
            bv = ov;
            bi = oi;
          }
        }
        if (lane == 0) pv_s = bv > 0.0 ? bi : -1;
      }
      __syncthreads();
      const int p = pv_s;
      if (p < 0) {
        singular = true;
        break;
      }
      if (p != k) {
        for (int c = t; c < NL; c += NV_T) {
          const double x = S[c * NV_LD + k];
          S[c * NV_LD + k] = S[c * NV_LD + p];
          S[c * NV_LD + p] = x;
        }
        if (t == 0) {
          const double x = r[k];
          r[k] = r[p];
          r[p] = x;
        }
      }
      __syncthreads();
      const double rp = 1.0 / S[k * NV_LD + k];
      for (int rr = k + 1 + t; rr < NL; rr += NV_T) S[k * NV_LD + rr] *= rp;
      __syncthreads();
      const int m = NL - k - 1;
      for (int idx = t; idx < m * m; idx += NV_T) {
        const int c = k + 1 + idx / m, rr = k + 1 + idx % m;
        S[c * NV_LD + rr] = fma(-S[k * NV_LD + rr], S[c * NV_LD + k], S[c * NV_LD + rr]);
      }
      __syncthreads();
    }
    if (singular) {
      if (t == 0) atomicMax(info, i + 1);
      return;   // no update for a singular patch
    }
    // L y = P r (unit lower), U x = y: one wavefront
    if (wave == 0) {
      for (int k = 0; k < NL; ++k) {
        const double y = r[k];
        for (int rr = k + 1 + lane; rr < NL; rr += 64) r[rr] = fma(-S[k * NV_LD + rr], y, r[rr]);
        __builtin_amdgcn_wave_barrier();
      }
      for (int k = NL - 1; k >= 0; --k) {
        const double x = r[k] / S[k * NV_LD + k];
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) r[k] = x;
        for (int rr = lane; rr < k; rr += 64) r[rr] = fma(-S[k * NV_LD + rr], x, r[rr]);
        __builtin_amdgcn_wave_barrier();
      }
    }
    __syncthreads();
  }
  // damped updates of the patch's dofs
  if (t < NG) {
    const int c = t / nU, j = t - c * nU;
    u[long(c) * N + ud[j]] += damp * r[t];
  }
  if (t == 0) u[3L * N + i] += damp * r[NG];
}

}  // namespace

void launch_nodal_vanka(const OpView<double> &op, bool full, const int *pdofs, int n, const double *f, double *u,
                        double damp, int *info, hipStream_t s) {
  if (n <= 0) return;
  if (full)
    k_nodal_vanka<true><<<n, NV_T, 0, s>>>(op, pdofs, n, f, u, damp, info);
  else
    k_nodal_vanka<false><<<n, NV_T, 0, s>>>(op, pdofs, n, f, u, damp, info);
  PMG_HIP_CHECK(hipGetLastError());
}

}  // namespace pmg
