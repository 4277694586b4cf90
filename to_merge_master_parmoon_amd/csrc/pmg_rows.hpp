// pmg_rows.hpp - one CSR row of the type-4 operator against a vector, by 16 lanes.
//
// Row-blocked values (see OpView): the n entries of a row are read in batches of 64:
// every lane first loads the column indices and the values of its 4 consecutive entries
// of the batch as 16-byte vectors (all independent loads; a 16-lane group reads 256 B
// (FP32) or 512 B (FP64) of each block per batch), then the 4 gathers of x, then
// accumulates, so a row of up to 64 entries costs one index -> gather round trip.
// Entries a lane does not own are masked (zero x values, no branches).
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

// sum over the 64 lanes of a wavefront (every lane gets the total)
template <typename T>
__device__ inline T red64(T v) {
  v += __shfl_xor(v, 32, 64);
  v += __shfl_xor(v, 16, 64);
  return red16(v);
}

// 16-byte vector loads of 4 consecutive entries (the addresses are 4-byte aligned; the
// hardware runs in unaligned mode, so one dwordx4 per float4/int4 and two per double4)
typedef float pmg_vf4 __attribute__((ext_vector_type(4)));
typedef double pmg_vd2 __attribute__((ext_vector_type(2)));
typedef int pmg_vi4 __attribute__((ext_vector_type(4)));
__device__ inline void ld4(const float *p, float (&o)[4]) {
  pmg_vf4 a;
  __builtin_memcpy(&a, p, 16);
  o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
}
__device__ inline void ld4(const double *p, double (&o)[4]) {
  pmg_vd2 a, b;
  __builtin_memcpy(&a, p, 16);
  __builtin_memcpy(&b, p + 2, 16);
  o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
}
__device__ inline void ld4(const int *p, int (&o)[4]) {
  pmg_vi4 a;
  __builtin_memcpy(&a, p, 16);
  o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
}

// Batches of 64 entries: lane l holds entries j0 + 4l .. j0 + 4l + 3.  The window is
// clamped into the row (start min(j0 + 4l, n - 4)), so the loads of a row with n >= 4
// entries never leave it; ok[u] marks the entries this lane owns.  A row shorter than 4
// reads up to 3 entries past its end (the next row, or the zeroed tail padding of
// ROW_PAD entries every streamed operator array carries).
__device__ inline int row_window(int j0, int n, int lane, bool (&ok)[4]) {
  const int jb = j0 + 4 * lane;
  const int ld = n >= 4 ? (jb < n - 4 ? jb : n - 4) : 0;
#pragma unroll
  for (int u = 0; u < 4; ++u) ok[u] = (ld + u >= jb) && (ld + u < n);
  return ld;
}

// A row with all nine blocks: s_r += sum_c A_rc(row, col) x_c(col)
template <typename T>
__device__ inline void row_a9_v(const int *__restrict__ kc, const T *__restrict__ v, int n, int lane,
                              const T *__restrict__ x1, const T *__restrict__ x2, const T *__restrict__ x3,
                              T &s1, T &s2, T &s3) {
  for (int j0 = 0; j0 < n; j0 += 64) {
    bool ok[4];
    const int ld = row_window(j0, n, lane, ok);
    int c[4];
    ld4(kc + ld, c);
#pragma unroll
    for (int u = 0; u < 4; ++u) c[u] = ok[u] ? c[u] : 0;
    T a[9][4];
#pragma unroll
    for (int b = 0; b < 9; ++b) ld4(v + b * n + ld, a[b]);
    T y1[4], y2[4], y3[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      y1[u] = x1[c[u]];
      y2[u] = x2[c[u]];
      y3[u] = x3[c[u]];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const T z1 = ok[u] ? y1[u] : T(0), z2 = ok[u] ? y2[u] : T(0), z3 = ok[u] ? y3[u] : T(0);
      s1 += a[0][u] * z1 + a[1][u] * z2 + a[2][u] * z3;
      s2 += a[3][u] * z1 + a[4][u] * z2 + a[5][u] * z3;
      s3 += a[6][u] * z1 + a[7][u] * z2 + a[8][u] * z3;
    }
  }
}

// A row whose velocity coupling is one shared block (A11 = A22 = A33, A_rc = 0 for r != c):
// s_r += A(row, col) x_r(col), one value stream for the three components
template <typename T>
__device__ inline void row_a1_v(const int *__restrict__ kc, const T *__restrict__ v, int n, int lane,
                                const T *__restrict__ x1, const T *__restrict__ x2, const T *__restrict__ x3,
                                T &s1, T &s2, T &s3) {
  for (int j0 = 0; j0 < n; j0 += 64) {
    bool ok[4];
    const int ld = row_window(j0, n, lane, ok);
    int c[4];
    ld4(kc + ld, c);
#pragma unroll
    for (int u = 0; u < 4; ++u) c[u] = ok[u] ? c[u] : 0;
    T a[4];
    ld4(v + ld, a);
    T y1[4], y2[4], y3[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      y1[u] = x1[c[u]];
      y2[u] = x2[c[u]];
      y3[u] = x3[c[u]];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const T w = ok[u] ? a[u] : T(0);
      s1 += w * y1[u];
      s2 += w * y2[u];
      s3 += w * y3[u];
    }
  }
}

// row_a1_v with x gathered from an interleaved copy (x1, x2, x3, 0 per dof): two 16-byte
// loads per entry instead of three separate ones; the same entries per lane in the same order
__device__ inline void row_a1_v_aos(const int *__restrict__ kc, const double *__restrict__ v, int n, int lane,
                                    const double *__restrict__ xa, double &s1, double &s2, double &s3) {
  for (int j0 = 0; j0 < n; j0 += 64) {
    bool ok[4];
    const int ld = row_window(j0, n, lane, ok);
    int c[4];
    ld4(kc + ld, c);
#pragma unroll
    for (int u = 0; u < 4; ++u) c[u] = ok[u] ? c[u] : 0;
    double a[4];
    ld4(v + ld, a);
    double2 p[4], q[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      p[u] = *reinterpret_cast<const double2 *>(xa + 4L * c[u]);
      q[u] = *reinterpret_cast<const double2 *>(xa + 4L * c[u] + 2);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double w = ok[u] ? a[u] : 0.0;
      s1 += w * p[u].x;
      s2 += w * p[u].y;
      s3 += w * q[u].x;
    }
  }
}
// row_b_v with x gathered from the interleaved copy
__device__ inline void row_b_v_aos(const int *__restrict__ kc, const double *__restrict__ v, int n, int lane,
                                   const double *__restrict__ xa, double &s) {
  for (int j0 = 0; j0 < n; j0 += 64) {
    bool ok[4];
    const int ld = row_window(j0, n, lane, ok);
    int c[4];
    ld4(kc + ld, c);
#pragma unroll
    for (int u = 0; u < 4; ++u) c[u] = ok[u] ? c[u] : 0;
    double b[3][4];
#pragma unroll
    for (int r = 0; r < 3; ++r) ld4(v + r * n + ld, b[r]);
    double2 p[4], q[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      p[u] = *reinterpret_cast<const double2 *>(xa + 4L * c[u]);
      q[u] = *reinterpret_cast<const double2 *>(xa + 4L * c[u] + 2);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double z1 = ok[u] ? p[u].x : 0.0, z2 = ok[u] ? p[u].y : 0.0, z3 = ok[u] ? q[u].x : 0.0;
      s += b[0][u] * z1 + b[1][u] * z2 + b[2][u] * z3;
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
template <typename T>
__device__ inline void row_bt_v(const int *__restrict__ kc, const T *__restrict__ w, int m, int lane,
                              const T *__restrict__ p, T &s1, T &s2, T &s3) {
  for (int j0 = 0; j0 < m; j0 += 64) {
    bool ok[4];
    const int ld = row_window(j0, m, lane, ok);
    int c[4];
    ld4(kc + ld, c);
    T b[3][4];
#pragma unroll
    for (int r = 0; r < 3; ++r) ld4(w + r * m + ld, b[r]);
    T pv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) pv[u] = p[ok[u] ? c[u] : 0];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const T z = ok[u] ? pv[u] : T(0);
      s1 += b[0][u] * z;
      s2 += b[1][u] * z;
      s3 += b[2][u] * z;
    }
  }
}

// B row (pressure row, velocity columns): s += sum_c B_c(row, col) x_c(col)
template <typename T>
__device__ inline void row_b_v(const int *__restrict__ kc, const T *__restrict__ v, int n, int lane,
                             const T *__restrict__ x1, const T *__restrict__ x2, const T *__restrict__ x3, T &s) {
  for (int j0 = 0; j0 < n; j0 += 64) {
    bool ok[4];
    const int ld = row_window(j0, n, lane, ok);
    int c[4];
    ld4(kc + ld, c);
#pragma unroll
    for (int u = 0; u < 4; ++u) c[u] = ok[u] ? c[u] : 0;
    T b[3][4];
#pragma unroll
    for (int r = 0; r < 3; ++r) ld4(v + r * n + ld, b[r]);
    T y1[4], y2[4], y3[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      y1[u] = x1[c[u]];
      y2[u] = x2[c[u]];
      y3[u] = x3[c[u]];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const T z1 = ok[u] ? y1[u] : T(0), z2 = ok[u] ? y2[u] : T(0), z3 = ok[u] ? y3[u] : T(0);
      s += b[0][u] * z1 + b[1][u] * z2 + b[2][u] * z3;
    }
  }
}

// ---- one entry per lane per step (FP32)
// A row with all nine blocks: s_r += sum_c A_rc(row, col) x_c(col)
template <typename T, int U>
__device__ inline void row_a9_s(const int *__restrict__ kc, const T *__restrict__ v, int n, int lane,
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

// A row with one shared velocity block (see row_a1_v)
template <typename T, int U>
__device__ inline void row_a1_s(const int *__restrict__ kc, const T *__restrict__ v, int n, int lane,
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
    T a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] = v[jj[u]];
    T y1[U], y2[U], y3[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      y1[u] = x1[c[u]];
      y2[u] = x2[c[u]];
      y3[u] = x3[c[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const T w = ok[u] ? a[u] : T(0);
      s1 += w * y1[u];
      s2 += w * y2[u];
      s3 += w * y3[u];
    }
  }
}

// row_a1_s with the velocity gathered from the interleaved copy xa (x_c(col) at 4 col + c):
// one 16-byte gather per entry instead of three
template <int U>
__device__ inline void row_a1_aos(const int *__restrict__ kc, const float *__restrict__ v, int n, int lane,
                                  const float *__restrict__ xa, float &s1, float &s2, float &s3) {
  typedef float vf4 __attribute__((ext_vector_type(4)));
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
    float a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] = v[jj[u]];
    vf4 y[U];
#pragma unroll
    for (int u = 0; u < U; ++u) __builtin_memcpy(&y[u], xa + 4L * c[u], 16);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const float w = ok[u] ? a[u] : 0.0f;
      s1 += w * y[u].x;
      s2 += w * y[u].y;
      s3 += w * y[u].z;
    }
  }
}

// row_b_s with the velocity gathered from the interleaved copy xa (x_c(col) at 4 col + c)
template <int U>
__device__ inline void row_b_aos(const int *__restrict__ kc, const float *__restrict__ v, int n, int lane,
                                 const float *__restrict__ xa, float &s) {
  typedef float vf4 __attribute__((ext_vector_type(4)));
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
    float b[3][U];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int u = 0; u < U; ++u) b[r][u] = v[r * n + jj[u]];
    vf4 y[U];
#pragma unroll
    for (int u = 0; u < U; ++u) __builtin_memcpy(&y[u], xa + 4L * c[u], 16);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const float z1 = ok[u] ? y[u].x : 0.0f, z2 = ok[u] ? y[u].y : 0.0f, z3 = ok[u] ? y[u].z : 0.0f;
      s += b[0][u] * z1 + b[1][u] * z2 + b[2][u] * z3;
    }
  }
}

// BT row (velocity row, pressure columns): s_r += BT_r(row, q) p(q)
template <typename T, int U>
__device__ inline void row_bt_s(const int *__restrict__ kc, const T *__restrict__ w, int m, int lane,
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
__device__ inline void row_b_s(const int *__restrict__ kc, const T *__restrict__ v, int n, int lane,
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

// FP64 rows read 4-entry vectors per lane (fewer, wider loads: measured faster); FP32
// rows read one entry per lane per 16-entry step, which keeps the 16 lanes' gathers of x
// on consecutive columns of the row (the vector form spreads them over 4x as many cache
// lines, and the FP32 kernels are bound by those gathers: measured slower)
template <typename T>
__device__ inline void row_a9(const int *__restrict__ kc, const T *__restrict__ v, int n, int lane,
                              const T *__restrict__ x1, const T *__restrict__ x2, const T *__restrict__ x3,
                              T &s1, T &s2, T &s3) {
  if constexpr (sizeof(T) == 8)
    row_a9_v<T>(kc, v, n, lane, x1, x2, x3, s1, s2, s3);
  else
    row_a9_s<T, 4>(kc, v, n, lane, x1, x2, x3, s1, s2, s3);
}
template <typename T>
__device__ inline void row_a1(const int *__restrict__ kc, const T *__restrict__ v, int n, int lane,
                              const T *__restrict__ x1, const T *__restrict__ x2, const T *__restrict__ x3,
                              T &s1, T &s2, T &s3) {
  if constexpr (sizeof(T) == 8)
    row_a1_v<T>(kc, v, n, lane, x1, x2, x3, s1, s2, s3);
  else
    row_a1_s<T, 4>(kc, v, n, lane, x1, x2, x3, s1, s2, s3);
}
template <typename T>
__device__ inline void row_bt(const int *__restrict__ kc, const T *__restrict__ w, int m, int lane,
                              const T *__restrict__ p, T &s1, T &s2, T &s3) {
  if constexpr (sizeof(T) == 8)
    row_bt_v<T>(kc, w, m, lane, p, s1, s2, s3);
  else
    row_bt_s<T, 2>(kc, w, m, lane, p, s1, s2, s3);
}
template <typename T>
__device__ inline void row_b(const int *__restrict__ kc, const T *__restrict__ v, int n, int lane,
                             const T *__restrict__ x1, const T *__restrict__ x2, const T *__restrict__ x3, T &s) {
  if constexpr (sizeof(T) == 8)
    row_b_v<T>(kc, v, n, lane, x1, x2, x3, s);
  else
    row_b_s<T, 2>(kc, v, n, lane, x1, x2, x3, s);
}

}  // namespace pmg
