// coarse_rows.hpp - per-row device code of one visit of the coarse solve (SolveExact,
// NSE_MGLevel4.C:5728-5846), shared by the per-problem kernels and the batched kernels that
// visit the coarse systems of several independent problems in one launch, so both run the
// same arithmetic.  Vector entries are read and written through accessors: fa(j) returns
// entry j of the problem's own right-hand side as a double, put(j, v) stores entry j of its
// solution.
#pragma once
#include "pmg_internal.hpp"
#include "pmg_rows.hpp"

namespace pmg {

// x_D = f_D / diag(A_cc) on the Dirichlet velocity dof c of component b
template <typename FA>
__device__ inline double coarse_dirichlet(const OpView<double> &op, const FA &fa, int b, int c) {
  const int k = op.a_rowptr[c];   // Dirichlet rows hold only their diagonal
  return fa(long(b) * op.n_udof + c) / op.a[(4 * b) * op.a_nnz + k];
}

// entry k of the reduced right-hand side (f_I - K_ID x_D): 16 lanes, every lane returns it
template <typename FA>
__device__ inline double coarse_rhs_row(const OpView<double> &op, const FA &fa, int k, int pin, int lane) {
  const int N = op.n_udof, NA = op.n_active;
  double v = 0.0;
  double f0;
  if (k < 3 * NA) {
    const int br = k / NA, r = k - br * NA;
    f0 = fa(long(br) * N + r);
    for (int q = op.a_rowptr[r] + lane; q < op.a_rowptr[r + 1]; q += 16) {
      const int c = op.a_kcol[q];
      if (c < NA) continue;
      for (int bc = 0; bc < 3; ++bc) v += op.a[(3 * br + bc) * op.a_nnz + q] * coarse_dirichlet(op, fa, bc, c);
    }
  } else {
    const int pi = k - 3 * NA + pin;
    f0 = fa(3L * N + pi);
    for (int q = op.b_rowptr[pi] + lane; q < op.b_rowptr[pi + 1]; q += 16) {
      const int c = op.b_kcol[q];
      if (c < NA) continue;
      for (int b = 0; b < 3; ++b) v += op.b[b * op.b_nnz + q] * coarse_dirichlet(op, fa, b, c);
    }
  }
  v = red16(v);
  return f0 - v;
}

// block form, phase 1, one wavefront per row: y_b = Ainv r_b (rows < m, b = 0..2) and
// p = [-Q | S^{-1}] r (rows m .. m + q - 1)
__device__ inline void cb_phase1_row(const double *__restrict__ Ainv, int lda, const double *__restrict__ P1,
                                     const double *__restrict__ rr, double *__restrict__ y, double *__restrict__ p,
                                     int m, int q, int row, int lane) {
  if (row < m) {
    const double *ar = Ainv + long(row) * lda;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int c = lane; c < m; c += 64) {
      const double v = ar[c];
      s0 = fma(v, rr[c], s0);
      s1 = fma(v, rr[m + c], s1);
      s2 = fma(v, rr[2 * m + c], s2);
    }
    s0 = red64(s0);
    s1 = red64(s1);
    s2 = red64(s2);
    if (lane == 0) {
      y[row] = s0;
      y[m + row] = s1;
      y[2 * m + row] = s2;
    }
  } else if (row < m + q) {
    const int i = row - m;
    const long nr = 3L * m + q;
    const double *pr = P1 + long(i) * nr;
    double s = 0.0;
    for (long k = lane; k < nr; k += 64) s = fma(pr[k], rr[k], s);
    s = red64(s);
    if (lane == 0) p[i] = s;
  }
}

// block form, phase 2, active velocity row r < m (one wavefront): x_b = y_b - W_b p
template <typename PUT>
__device__ inline void cb_phase2_vel(const OpView<double> &op, const double *__restrict__ W, int ldw,
                                     const double *__restrict__ y, const double *__restrict__ p, const PUT &put,
                                     int m, int q, int r, int lane) {
  const long N = op.n_udof;
  const double *wr = W + long(r) * ldw;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int i = lane; i < q; i += 64) {
    const double pi = p[i];
    s0 = fma(wr[i], pi, s0);
    s1 = fma(wr[q + i], pi, s1);
    s2 = fma(wr[2 * q + i], pi, s2);
  }
  s0 = red64(s0);
  s1 = red64(s1);
  s2 = red64(s2);
  if (lane == 0) {
    put(r, y[r] - s0);
    put(N + r, y[m + r] - s1);
    put(2 * N + r, y[2 * m + r] - s2);
  }
}

// block form, phase 2, entry idx of the rest: the Dirichlet velocity dofs (f_D / diagonal),
// then the pressure (0 at the pinned dof)
template <typename FA, typename PUT>
__device__ inline void cb_phase2_rest(const OpView<double> &op, const double *__restrict__ p, const FA &fa,
                                      const PUT &put, int pin, long idx) {
  const int N = op.n_udof, NA = op.n_active, NP = op.n_pdof;
  const long nd = 3L * (N - NA);
  if (idx < nd) {
    const int b = int(idx / (N - NA)), c = NA + int(idx - long(b) * (N - NA));
    put(long(b) * N + c, coarse_dirichlet(op, fa, b, c));
  } else if (idx < nd + NP) {
    const int i = int(idx - nd);
    put(3L * N + i, (pin && i == 0) ? 0.0 : p[i - pin]);
  }
}

}  // namespace pmg
