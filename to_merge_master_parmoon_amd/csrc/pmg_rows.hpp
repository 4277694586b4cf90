// pmg_rows.hpp - one CSR row of the type-4 operator against a vector, by 16 lanes.
//
// Row-blocked values (see OpView): the n entries of a row are read in batches of
// U*16: every lane first loads the column indices and the values of its U entries of the
// batch (all independent loads), then the U gathers of x, then accumulates, so a whole
// row of up to 16*U entries costs one index -> gather round trip.  Entries past the end of
// the row are clamped to the last entry and their x values zeroed (no branches).
#pragma once
#include <hip/hip_runtime.h>

namespace pmg {

// sum over the 16 lanes of a row group (every lane gets the total)
template <typename T>
__device__ inline T red16(T v) {
  v += __shfl_xor(v, 8, 16);
  v += __shfl_xor(v, 4, 16);
  v += __shfl_xor(v, 2, 16);
  v += __shfl_xor(v, 1, 16);
  return v;
}

// A row with all nine blocks: s_r += sum_c A_rc(row, col) x_c(col)
template <typename T, int U>
__device__ inline void row_a9(const int *__restrict__ kc, const T *__restrict__ v, int n, int lane,
                              const T *__restrict__ x1, const T *__restrict__ x2, const T *__restrict__ x3,
                              T &s1, T &s2, T &s3) {
  for (int j0 = 0; j0 < n; j0 += 16 * U) {
    int jj[U];
    bool ok[U];
    int c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = j0 + 16 * u + lane;
      ok[u] = j < n;
      jj[u] = ok[u] ? j : n - 1;
      c[u] = kc[jj[u]];
    }
    T a[9][U];
#pragma unroll
    for (int b = 0; b < 9; ++b)
#pragma unroll
      for (int u = 0; u < U; ++u) a[b][u] = v[b * n + jj[u]];
    T y1[U], y2[U], y3[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      y1[u] = x1[c[u]];
      y2[u] = x2[c[u]];
      y3[u] = x3[c[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const T z1 = ok[u] ? y1[u] : T(0), z2 = ok[u] ? y2[u] : T(0), z3 = ok[u] ? y3[u] : T(0);
      s1 += a[0][u] * z1 + a[1][u] * z2 + a[2][u] * z3;
      s2 += a[3][u] * z1 + a[4][u] * z2 + a[5][u] * z3;
      s3 += a[6][u] * z1 + a[7][u] * z2 + a[8][u] * z3;
    }
  }
}

// A row of a Dirichlet dof: only the diagonal blocks A11, A22, A33
template <typename T, int U>
__device__ inline void row_a_diag(const int *__restrict__ kc, const T *__restrict__ v, int n, int lane,
                                  const T *__restrict__ x1, const T *__restrict__ x2, const T *__restrict__ x3,
                                  T &s1, T &s2, T &s3) {
  for (int j0 = 0; j0 < n; j0 += 16 * U) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = j0 + 16 * u + lane;
      if (j < n) {
        const int c = kc[j];
        s1 += v[j] * x1[c];
        s2 += v[4 * n + j] * x2[c];
        s3 += v[8 * n + j] * x3[c];
      }
    }
  }
}

// BT row (velocity row, pressure columns): s_r += BT_r(row, q) p(q)
template <typename T, int U>
__device__ inline void row_bt(const int *__restrict__ kc, const T *__restrict__ w, int m, int lane,
                              const T *__restrict__ p, T &s1, T &s2, T &s3) {
  for (int j0 = 0; j0 < m; j0 += 16 * U) {
    int jj[U];
    bool ok[U];
    T pv[U], b[3][U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = j0 + 16 * u + lane;
      ok[u] = j < m;
      jj[u] = ok[u] ? j : m - 1;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      pv[u] = p[kc[jj[u]]];
#pragma unroll
      for (int r = 0; r < 3; ++r) b[r][u] = w[r * m + jj[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const T z = ok[u] ? pv[u] : T(0);
      s1 += b[0][u] * z;
      s2 += b[1][u] * z;
      s3 += b[2][u] * z;
    }
  }
}

// B row (pressure row, velocity columns): s += sum_c B_c(row, col) x_c(col)
template <typename T, int U>
__device__ inline void row_b(const int *__restrict__ kc, const T *__restrict__ v, int n, int lane,
                             const T *__restrict__ x1, const T *__restrict__ x2, const T *__restrict__ x3, T &s) {
  for (int j0 = 0; j0 < n; j0 += 16 * U) {
    int jj[U];
    bool ok[U];
    int c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = j0 + 16 * u + lane;
      ok[u] = j < n;
      jj[u] = ok[u] ? j : n - 1;
      c[u] = kc[jj[u]];
    }
    T b[3][U];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int u = 0; u < U; ++u) b[r][u] = v[r * n + jj[u]];
    T y1[U], y2[U], y3[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      y1[u] = x1[c[u]];
      y2[u] = x2[c[u]];
      y3[u] = x3[c[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const T z1 = ok[u] ? y1[u] : T(0), z2 = ok[u] ? y2[u] : T(0), z3 = ok[u] ? y3[u] : T(0);
      s += b[0][u] * z1 + b[1][u] * z2 + b[2][u] * z3;
    }
  }
}

}  // namespace pmg
