// k_operator.hip - type-4 block operator (CoupledDefect / CoupledMatVect / Defect_NSE4 /
// TNSE_MGLevel4::Defect) and the vector kernels around it, for gfx950.
//
// Mapping: a group of 16 lanes owns one CSR row; lanes stride over the row's entries so
// that every value stream (9 A blocks, 3 B or BT blocks, column indices) is read as
// contiguous 16-entry segments; the three velocity components of a row are produced
// together so x is gathered once per entry.  Pressure rows follow the velocity rows in
// the same launch.  Reductions that the reference performs after the product (IntoL20
// mean of the P1disc constants, the Euclidean norm) are folded into the same launch with a
// deterministic last-block finish: every block writes its partial sums, the block that
// arrives last sums them in a fixed order and applies the projection.
#include "pmg_internal.hpp"
#include "pmg_rows.hpp"

namespace pmg {

namespace {


// block-wide sum of two doubles (blockDim.x == 256), fixed order: wave shuffles, then the
// four wave sums in wave order; the result is valid in every thread
__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline void block_sum2(double &a, double &b, double *sh) {
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) {
    sh[w] = a;
    sh[4 + w] = b;
  }
  __syncthreads();
  a = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  b = (sh[4] + sh[5]) + (sh[6] + sh[7]);
}

// Last block of the grid sums partials[0..nblk) (two values each) in a fixed order.
// Returns true in every thread of the last block.
__device__ inline bool last_block_sum2(double pa, double pb, double *partials, unsigned *counter,
                                       double &ta, double &tb, double *sh) {
  __shared__ bool am_last;
  block_sum2(pa, pb, sh);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = pa;
    partials[2 * blockIdx.x + 1] = pb;
    __threadfence();
    unsigned prev = atomicAdd(counter, 1u);
    am_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!am_last) return false;
  __threadfence();
  double a = 0, b = 0;
  for (int i = threadIdx.x; i < int(gridDim.x); i += 256) {
    a += static_cast<volatile double *>(partials)[2 * i];
    b += static_cast<volatile double *>(partials)[2 * i + 1];
  }
  __syncthreads();
  block_sum2(a, b, sh);
  ta = a;
  tb = b;
  if (threadIdx.x == 0) *counter = 0u;
  return true;
}

// SH: the velocity coupling is the shared block op.a1 (one value stream instead of nine)
template <int MODE, typename T, bool SH>
__global__ __launch_bounds__(256) void k_type4(OpView<T> op, const T *__restrict__ x,
                                               const T *__restrict__ bv, T *__restrict__ y,
                                               int project, int want_norm, double *partials,
                                               unsigned *counter, double *res, double *res_slots) {
  __shared__ double sh[8];
  const int lane = threadIdx.x & 15, grp = threadIdx.x >> 4;
  const int N = op.n_udof, NA = op.n_active, NP = op.n_pdof;
  const T *__restrict__ x1 = x;
  const T *__restrict__ x2 = x + N;
  const T *__restrict__ x3 = x + 2 * N;
  const T *__restrict__ xp = x + 3 * N;
  double pc = 0.0, sq = 0.0;
  // this block's chunk of rows (of the listed rows, when the level runs a subset)
  int r0, r1;
  chunk_of_block(op.rows ? long(op.n_rows) : long(N) + NP, r0, r1);
  for (int ri = r0 + grp; ri < r1; ri += 16) {
    const int r = op.rows ? op.rows[ri] : ri;
    if (r < N) {
      T s1 = 0, s2 = 0, s3 = 0;
      const int k0 = op.a_rowptr[r], n = op.a_rowptr[r + 1] - k0;
      const int *__restrict__ kc = op.a_kcol + k0;
      const T *__restrict__ v = op.av + 9L * k0;
      if constexpr (SH && sizeof(T) == 8) {
        if (op.xa)
          row_a1_v_aos(kc, op.a1 + k0, n, lane, op.xa, s1, s2, s3);
        else
          row_a1<T>(kc, op.a1 + k0, n, lane, x1, x2, x3, s1, s2, s3);
      } else if constexpr (SH) {
        row_a1<T>(kc, op.a1 + k0, n, lane, x1, x2, x3, s1, s2, s3);
      } else if (r < NA) {
        row_a9<T>(kc, v, n, lane, x1, x2, x3, s1, s2, s3);
      } else {
        // rows N_Active..N_UDOF: diagonal blocks only (MGComponents3D.C:3136-3156)
        row_a_diag<T, 1>(kc, v, n, lane, x1, x2, x3, s1, s2, s3);
      }
      if (r < NA) {
        const int q0 = op.bt_rowptr[r], m = op.bt_rowptr[r + 1] - q0;
        row_bt<T>(op.bt_kcol + q0, op.btv + 3L * q0, m, lane, xp, s1, s2, s3);
      }
      s1 = red16(s1);
      s2 = red16(s2);
      s3 = red16(s3);
      if (lane == 0) {
        T y1, y2, y3;
        if (MODE == 0) {
          y1 = s1; y2 = s2; y3 = s3;
        } else {
          y1 = bv[r] - s1; y2 = bv[N + r] - s2; y3 = bv[2 * N + r] - s3;
        }
        if (MODE == 2 && r >= NA) y1 = y2 = y3 = 0;
        y[r] = y1; y[N + r] = y2; y[2 * N + r] = y3;
        if (want_norm) sq += double(y1) * y1 + double(y2) * y2 + double(y3) * y3;
      }
    } else {
      const int pr = r - N;
      T s = 0;
      const int k0 = op.b_rowptr[pr], n = op.b_rowptr[pr + 1] - k0;
      if constexpr (SH && sizeof(T) == 8) {
        if (op.xa)
          row_b_v_aos(op.b_kcol + k0, op.bv + 3L * k0, n, lane, op.xa, s);
        else
          row_b<T>(op.b_kcol + k0, op.bv + 3L * k0, n, lane, x1, x2, x3, s);
      } else {
        row_b<T>(op.b_kcol + k0, op.bv + 3L * k0, n, lane, x1, x2, x3, s);
      }
      s = red16(s);
      if (lane == 0) {
        const T v = (MODE == 0) ? s : T(bv[3 * N + pr] - s);
        y[3 * N + pr] = v;
        if (project && (pr & 3) == 0)
          pc += double(v);
        else if (want_norm)
          sq += double(v) * v;
      }
    }
  }
  if (!(project || want_norm)) return;
  double tpc, tsq;
  if (!last_block_sum2(pc, sq, partials, counter, tpc, tsq, sh)) return;
  if (threadIdx.x == 0) {
    if (project) {
      // IntoL20Vector3D case -11 (MGComponents3D.C:1766-1786): s = sum/Length*4, applied by
      // k_l20_apply; the squares of the constants are added there
      res_slots[0] = tpc / double(NP) * 4.0;
      res_slots[1] = tsq;
    } else {
      *res = sqrt(tsq);
    }
  }
}

// v[4i] -= s for the P1disc constants; with res: ||.||^2 = base + sum of the new squares
template <typename T>
__global__ __launch_bounds__(256) void k_l20_apply(T *__restrict__ p, int NP, const double *slots, double *res,
                                                   double *partials, unsigned *counter) {
  __shared__ double sh[8];
  const double m = slots[0];
  double add = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < NP / 4; i += gridDim.x * 256) {
    const T v = T(p[4 * i] - T(m));
    p[4 * i] = v;
    add += double(v) * v;
  }
  if (!res) return;
  double t, u;
  if (last_block_sum2(add, 0.0, partials, counter, t, u, sh) && threadIdx.x == 0) *res = sqrt(slots[1] + t);
}

template <typename T>
__global__ void k_fill(T *x, T v, long n) {
  for (long i = blockIdx.x * long(blockDim.x) + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x) x[i] = v;
}
template <typename T>
__global__ void k_copy(T *__restrict__ d, const T *__restrict__ s, long n) {
  for (long i = blockIdx.x * long(blockDim.x) + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x) d[i] = s[i];
}
template <typename T>
__global__ void k_axpy(T *__restrict__ y, T a, const T *__restrict__ x, long n) {
  for (long i = blockIdx.x * long(blockDim.x) + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x) y[i] += a * x[i];
}
template <typename T>
__global__ void k_relax(T *__restrict__ u, const T *__restrict__ uo, T a, long n) {
  for (long i = blockIdx.x * long(blockDim.x) + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x)
    u[i] = uo[i] + a * (u[i] - uo[i]);
}
__global__ void k_d2f(float *__restrict__ d, const double *__restrict__ s, long n) {
  for (long i = blockIdx.x * long(blockDim.x) + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x) d[i] = float(s[i]);
}
__global__ void k_f2d(double *__restrict__ d, const float *__restrict__ s, long n) {
  for (long i = blockIdx.x * long(blockDim.x) + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x) d[i] = double(s[i]);
}
// u[c*N + start + i] = f ? f[c*N + start + i] : 0, c = 0..2
template <typename T>
__global__ void k_dir(T *__restrict__ u, const T *__restrict__ f, int N, int start, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * count) return;
  const int c = i / count, j = i - c * count;
  const long o = long(c) * N + start + j;
  u[o] = f ? f[o] : T(0);
}
// Start of an FP32 sweep with interleaved velocity gathers: the Dirichlet slices of u take
// f (as k_dir does), and ua[i] = (u1, u2, u3, 0) of every velocity dof i
__global__ void k_sweep_start(float *__restrict__ u, const float *__restrict__ f, int N, int start, int count,
                              float4 *__restrict__ ua) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  float v1 = u[i], v2 = u[N + i], v3 = u[2 * N + i];
  if (i >= start && i < start + count) {
    v1 = f[i];
    v2 = f[N + i];
    v3 = f[2 * N + i];
    u[i] = v1;
    u[N + i] = v2;
    u[2 * N + i] = v3;
  }
  ua[i] = make_float4(v1, v2, v3, 0.0f);
}

// A zero start: u = 0 with the velocity Dirichlet slices from f, and (ua non-null) the
// interleaved copy of that velocity
template <typename T>
__global__ void k_start_zero(T *__restrict__ u, const T *__restrict__ f, int N, int NP, int start, int count,
                             float4 *__restrict__ ua) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) {
    const bool dir = i >= start && i < start + count;
    const T v1 = dir ? f[i] : T(0), v2 = dir ? f[N + i] : T(0), v3 = dir ? f[2 * N + i] : T(0);
    u[i] = v1;
    u[N + i] = v2;
    u[2 * N + i] = v3;
    if (ua) ua[i] = make_float4(float(v1), float(v2), float(v3), 0.0f);
  } else if (i < N + NP) {
    u[2 * N + i] = T(0);
  }
}
__global__ void k_ua_to_u(const float4 *__restrict__ ua, float *__restrict__ u, int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float4 v = ua[i];
  u[i] = v.x;
  u[N + i] = v.y;
  u[2 * N + i] = v.z;
}

// IntoL20 of the P1disc constants in two passes without atomics: every block sums its
// chunk of constants (k_l20_part), then every block adds all partials in the same order
// and subtracts the mean from its chunk (k_l20_sub)
template <typename T>
__global__ __launch_bounds__(256) void k_l20_part(const T *__restrict__ p, int NP, double *__restrict__ part) {
  __shared__ double sh[8];
  const int nc = NP / 4;
  const int c0 = int(long(nc) * blockIdx.x / gridDim.x), c1 = int(long(nc) * (blockIdx.x + 1) / gridDim.x);
  double s = 0, d = 0;
  for (int i = c0 + threadIdx.x; i < c1; i += 256) s += double(p[4 * i]);
  block_sum2(s, d, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
template <typename T>
__global__ __launch_bounds__(256) void k_l20_sub(T *__restrict__ p, int NP, const double *__restrict__ part) {
  __shared__ double sh[8];
  double s = 0, d = 0;
  for (int i = threadIdx.x; i < int(gridDim.x); i += 256) s += part[i];
  block_sum2(s, d, sh);
  const double m = s / double(NP) * 4.0;
  const int nc = NP / 4;
  const int c0 = int(long(nc) * blockIdx.x / gridDim.x), c1 = int(long(nc) * (blockIdx.x + 1) / gridDim.x);
  for (int i = c0 + threadIdx.x; i < c1; i += 256) p[4 * i] = T(p[4 * i] - T(m));
}

// standalone IntoL20 of the P1disc constants (single block)
template <typename T>
__global__ __launch_bounds__(256) void k_l20(T *p, int NP) {
  __shared__ double sh[8];
  double s = 0, d = 0;
  for (int i = threadIdx.x; i < NP / 4; i += 256) s += double(p[4 * i]);
  block_sum2(s, d, sh);
  const double m = s / double(NP) * 4.0;
  for (int i = threadIdx.x; i < NP / 4; i += 256) p[4 * i] = T(p[4 * i] - T(m));
}

// ------------------------------------------------------------------ BLAS-1 for FGMRES
// <x, y> and w -= h v ; <w, vnext>: each block writes its partial sum, k_sum_partials adds
// them in block order (no per-block fence and counter atomic as a last-block finish needs)
__global__ __launch_bounds__(256) void k_dot_part(const double *__restrict__ x, const double *__restrict__ y,
                                                  long n, double *__restrict__ part) {
  __shared__ double sh[8];
  double s = 0, d = 0;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += long(gridDim.x) * 256) s += x[i] * y[i];
  block_sum2(s, d, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_axpy_dot_part(double *__restrict__ w, const double *h,
                                                       const double *__restrict__ v, const double *__restrict__ vn,
                                                       long n, double *__restrict__ part) {
  __shared__ double sh[8];
  const double a = -h[0];
  double s = 0, d = 0;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += long(gridDim.x) * 256) {
    const double wi = w[i] + a * v[i];
    w[i] = wi;
    if (vn) s += wi * vn[i];
  }
  if (!vn) return;
  block_sum2(s, d, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_sum_partials(const double *__restrict__ part, int np, double *out) {
  __shared__ double sh[8];
  double s = 0, d = 0;
  for (int i = threadIdx.x; i < np; i += 256) s += part[i];
  block_sum2(s, d, sh);
  if (threadIdx.x == 0) *out = s;
}

__global__ void k_scale_dev(double *__restrict__ y, const double *__restrict__ x, const double *den, long n) {
  const double d = 1.0 / den[0];
  for (long i = blockIdx.x * long(blockDim.x) + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x)
    y[i] = x[i] * d;
}
__global__ void k_lincomb(double *__restrict__ x, const double *__restrict__ z, long stride,
                          const double *__restrict__ coef, int m, long n) {
  for (long i = blockIdx.x * long(blockDim.x) + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x) {
    double t = 0.0;
    for (int j = 0; j < m; ++j) t += z[j * stride + i] * coef[j];
    x[i] += t;
  }
}
__global__ void k_scal(double *__restrict__ y, const double *__restrict__ x, double a, long n) {
  for (long i = blockIdx.x * long(blockDim.x) + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x)
    y[i] = x[i] * a;
}

inline int grid_for(long n, int bs = 256, int cap = 8192) {
  long g = (n + bs - 1) / bs;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return int(g);
}

}  // namespace

namespace {
__global__ void k_interleave3(double *__restrict__ xa, const double *__restrict__ x, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 p = {x[i], x[n + i]}, q = {x[2L * n + i], 0.0};
  reinterpret_cast<double2 *>(xa)[2L * i] = p;
  reinterpret_cast<double2 *>(xa)[2L * i + 1] = q;
}
}  // namespace
void launch_interleave3(double *xa, const double *x, int n, hipStream_t s) {
  if (n <= 0) return;
  k_interleave3<<<(n + 255) / 256, 256, 0, s>>>(xa, x, n);
  PMG_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_type4(const OpView<T> &op, int mode, const T *x, const T *b, T *y, bool project, double *res,
                  RedScratch &red, hipStream_t s) {
  const long n_rows = op.rows ? long(op.n_rows) : long(op.n_udof) + op.n_pdof;
  const int want_norm = res != nullptr;
  const int proj = (mode != 0 && project) ? 1 : 0;
  if (op.rows && (want_norm || proj))
    throw DeviceError("launch_type4: a row subset is supported without the norm and the projection only");
  if (n_rows <= 0) return;
  const int grid = chunk_grid(n_rows, op.rpb > 0 ? op.rpb : 64, RED_BLOCKS);
  auto go = [&](auto mode_c, int pr, int wn, double *rs, double *sl) {
    constexpr int M = decltype(mode_c)::value;
    if (op.a1)
      k_type4<M, T, true><<<grid, 256, 0, s>>>(op, x, b, y, pr, wn, red.partials, red.counter, rs, sl);
    else
      k_type4<M, T, false><<<grid, 256, 0, s>>>(op, x, b, y, pr, wn, red.partials, red.counter, rs, sl);
  };
  using M0 = std::integral_constant<int, 0>;
  using M1 = std::integral_constant<int, 1>;
  using M2 = std::integral_constant<int, 2>;
  if (proj && !want_norm) {
    // the product without a grid-wide reduction (no per-block fence and counter atomic),
    // then the projection in two small passes
    if (mode == 1)
      go(M1{}, 0, 0, nullptr, nullptr);
    else
      go(M2{}, 0, 0, nullptr, nullptr);
    PMG_HIP_CHECK(hipGetLastError());
    const int g = std::max(1, std::min(64, op.n_pdof / 4 / 2048));
    k_l20_part<T><<<g, 256, 0, s>>>(y + 3L * op.n_udof, op.n_pdof, red.partials);
    k_l20_sub<T><<<g, 256, 0, s>>>(y + 3L * op.n_udof, op.n_pdof, red.partials);
    PMG_HIP_CHECK(hipGetLastError());
    return;
  }
  double *slots = red.result + 8;
  if (mode == 0)
    go(M0{}, 0, want_norm, res, slots);
  else if (mode == 1)
    go(M1{}, proj, want_norm, res, slots);
  else
    go(M2{}, proj, want_norm, res, slots);
  PMG_HIP_CHECK(hipGetLastError());
  if (proj) {
    const int nc = op.n_pdof / 4;
    const int g = std::max(1, std::min(512, (nc + 255) / 256));
    k_l20_apply<T><<<g, 256, 0, s>>>(y + 3L * op.n_udof, op.n_pdof, slots, want_norm ? res : nullptr,
                                     red.partials, red.counter);
    PMG_HIP_CHECK(hipGetLastError());
  }
}
template void launch_type4<double>(const OpView<double> &, int, const double *, const double *, double *, bool,
                                   double *, RedScratch &, hipStream_t);
template void launch_type4<float>(const OpView<float> &, int, const float *, const float *, float *, bool,
                                  double *, RedScratch &, hipStream_t);

namespace {
// one 16-lane group per row: row r of an nb-block CSR matrix, SoA values src[b*nnz + k],
// goes to dst[nb*k0 + b*n + (k-k0)]
template <typename T, int NB>
__global__ void k_row_blocked(const int *__restrict__ rowptr, int rows, const double *__restrict__ src, long nnz,
                              T *__restrict__ dst) {
  const int lane = threadIdx.x & 15;
  const long grp = (long(blockIdx.x) * blockDim.x + threadIdx.x) >> 4;
  const long ngrp = (long(gridDim.x) * blockDim.x) >> 4;
  for (long r = grp; r < rows; r += ngrp) {
    const int k0 = rowptr[r], n = rowptr[r + 1] - k0;
    T *out = dst + long(NB) * k0;
    for (int j = lane; j < n; j += 16)
#pragma unroll
      for (int b = 0; b < NB; ++b) out[long(b) * n + j] = T(src[long(b) * nnz + k0 + j]);
  }
}
// 1 in part[block] when the block's entries break the shared velocity block form: on every
// row A22 = A11 and A33 = A11, on the active rows (entries [0, n_act)) also A_rc = 0 for
// r != c.  NaN compares unequal, so such values keep the nine-block form.
__global__ __launch_bounds__(256) void k_shared_check(const double *__restrict__ a, long nnz, long n_act,
                                                      double *__restrict__ part) {
  int bad = 0;
  for (long k = blockIdx.x * 256L + threadIdx.x; k < nnz; k += long(gridDim.x) * 256) {
    const double v = a[k];
    bad |= (a[4 * nnz + k] != v) | (a[8 * nnz + k] != v);
    if (k < n_act)
      bad |= (a[nnz + k] != 0.0) | (a[2 * nnz + k] != 0.0) | (a[3 * nnz + k] != 0.0) |
             (a[5 * nnz + k] != 0.0) | (a[6 * nnz + k] != 0.0) | (a[7 * nnz + k] != 0.0);
  }
  bad = __syncthreads_or(bad);
  if (threadIdx.x == 0) part[blockIdx.x] = bad ? 1.0 : 0.0;
}
}  // namespace

bool shared_velocity_block(const DevLevel &L, double *part, hipStream_t s) {
  constexpr int G = 256;
  k_shared_check<<<G, 256, 0, s>>>(L.a, L.a_nnz, L.a_act_entries, part);
  PMG_HIP_CHECK(hipGetLastError());
  double h[G];
  PMG_HIP_CHECK(hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, s));
  PMG_HIP_CHECK(hipStreamSynchronize(s));
  for (int g = 0; g < G; ++g)
    if (h[g] != 0.0) return false;
  return true;
}

template <typename T>
void launch_row_blocked(const DevLevel &L, T *av, T *bv, T *btv, hipStream_t s) {
  k_row_blocked<T, 9><<<grid_for(16L * L.n_udof), 256, 0, s>>>(L.a_rowptr, L.n_udof, L.a, L.a_nnz, av);
  PMG_HIP_CHECK(hipGetLastError());
  k_row_blocked<T, 3><<<grid_for(16L * L.n_pdof), 256, 0, s>>>(L.b_rowptr, L.n_pdof, L.b, L.b_nnz, bv);
  PMG_HIP_CHECK(hipGetLastError());
  k_row_blocked<T, 3><<<grid_for(16L * L.n_udof), 256, 0, s>>>(L.bt_rowptr, L.n_udof, L.bt, L.bt_nnz, btv);
  PMG_HIP_CHECK(hipGetLastError());
}
template void launch_row_blocked<double>(const DevLevel &, double *, double *, double *, hipStream_t);
template void launch_row_blocked<float>(const DevLevel &, float *, float *, float *, hipStream_t);

template <typename T>
void launch_fill(T *x, T v, long n, hipStream_t s) {
  if (n <= 0) return;
  k_fill<T><<<grid_for(n), 256, 0, s>>>(x, v, n);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_copy(T *d, const T *src, long n, hipStream_t s) {
  if (n <= 0) return;
  k_copy<T><<<grid_for(n), 256, 0, s>>>(d, src, n);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_axpy(T *y, T a, const T *x, long n, hipStream_t s) {
  if (n <= 0) return;
  k_axpy<T><<<grid_for(n), 256, 0, s>>>(y, a, x, n);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_relax(T *u, const T *uo, T a, long n, hipStream_t s) {
  if (n <= 0) return;
  k_relax<T><<<grid_for(n), 256, 0, s>>>(u, uo, a, n);
  PMG_HIP_CHECK(hipGetLastError());
}
void launch_d2f(float *d, const double *src, long n, hipStream_t s) {
  if (n <= 0) return;
  k_d2f<<<grid_for(n), 256, 0, s>>>(d, src, n);
  PMG_HIP_CHECK(hipGetLastError());
}
void launch_f2d(double *d, const float *src, long n, hipStream_t s) {
  if (n <= 0) return;
  k_f2d<<<grid_for(n), 256, 0, s>>>(d, src, n);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_dirichlet_slices(T *u, const T *f, int N, int start, int count, hipStream_t s) {
  if (count <= 0) return;
  k_dir<T><<<(3 * count + 255) / 256, 256, 0, s>>>(u, f, N, start, count);
  PMG_HIP_CHECK(hipGetLastError());
}
void launch_ua_to_u(const float *ua, float *u, int N, hipStream_t s) {
  if (N <= 0) return;
  k_ua_to_u<<<(N + 255) / 256, 256, 0, s>>>(reinterpret_cast<const float4 *>(ua), u, N);
  PMG_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_start_zero(T *u, const T *f, int N, int NP, int start, int count, float *ua, hipStream_t s) {
  if (N + NP <= 0) return;
  k_start_zero<T><<<(N + NP + 255) / 256, 256, 0, s>>>(u, f, N, NP, start, count, reinterpret_cast<float4 *>(ua));
  PMG_HIP_CHECK(hipGetLastError());
}
template void launch_start_zero<float>(float *, const float *, int, int, int, int, float *, hipStream_t);
template void launch_start_zero<double>(double *, const double *, int, int, int, int, float *, hipStream_t);

void launch_sweep_start(float *u, const float *f, int N, int start, int count, float *ua, hipStream_t s) {
  if (N <= 0) return;
  k_sweep_start<<<(N + 255) / 256, 256, 0, s>>>(u, f, N, start, count, reinterpret_cast<float4 *>(ua));
  PMG_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_l20(T *p, int NP, RedScratch &, hipStream_t s) {
  if (NP <= 0) return;
  k_l20<T><<<1, 256, 0, s>>>(p, NP);
  PMG_HIP_CHECK(hipGetLastError());
}

#define PMG_INST_VEC(T)                                                        \
  template void launch_fill<T>(T *, T, long, hipStream_t);                    \
  template void launch_copy<T>(T *, const T *, long, hipStream_t);            \
  template void launch_axpy<T>(T *, T, const T *, long, hipStream_t);         \
  template void launch_relax<T>(T *, const T *, T, long, hipStream_t);        \
  template void launch_dirichlet_slices<T>(T *, const T *, int, int, int, hipStream_t); \
  template void launch_l20<T>(T *, int, RedScratch &, hipStream_t);
PMG_INST_VEC(double)
PMG_INST_VEC(float)
template void launch_fill<int>(int *, int, long, hipStream_t);
template void launch_fill<uint16_t>(uint16_t *, uint16_t, long, hipStream_t);

void launch_dot(const double *x, const double *y, long n, double *out, RedScratch &red, hipStream_t s) {
  const int g = grid_for(n, 256, 512);
  k_dot_part<<<g, 256, 0, s>>>(x, y, n, red.partials);
  k_sum_partials<<<1, 256, 0, s>>>(red.partials, g, out);
  PMG_HIP_CHECK(hipGetLastError());
}
void launch_axpy_dev_dot(double *w, const double *h, const double *v, const double *vnext, long n, double *out,
                         RedScratch &red, hipStream_t s) {
  const int g = grid_for(n, 256, 512);
  k_axpy_dot_part<<<g, 256, 0, s>>>(w, h, v, vnext, n, red.partials);
  if (vnext) k_sum_partials<<<1, 256, 0, s>>>(red.partials, g, out);
  PMG_HIP_CHECK(hipGetLastError());
}
void launch_scale_dev(double *y, const double *x, const double *den, long n, hipStream_t s) {
  k_scale_dev<<<grid_for(n), 256, 0, s>>>(y, x, den, n);
  PMG_HIP_CHECK(hipGetLastError());
}
void launch_lincomb(double *x, const double *z, long stride, const double *coef, int m, long n, hipStream_t s) {
  if (m <= 0) return;
  k_lincomb<<<grid_for(n), 256, 0, s>>>(x, z, stride, coef, m, n);
  PMG_HIP_CHECK(hipGetLastError());
}
void launch_scal(double *y, const double *x, double a, long n, hipStream_t s) {
  k_scal<<<grid_for(n), 256, 0, s>>>(y, x, a, n);
  PMG_HIP_CHECK(hipGetLastError());
}

}  // namespace pmg
