// k_dense.hip - coarse-level dense direct solve (TNSE_MGLevel4::SolveExact,
// src/FE/NSE_MGLevel4.C:5728-5846) and the grid-transfer SpMVs, for gfx950.
//
// SolveExact builds the dense column-major matrix of the coarsest level with the row of
// pressure dof 0 replaced by the identity and calls LAPACK dgetrf/dgetrs at every visit.
// Here the matrix is built and factored on the device once per operator, with a blocked
// right-looking LU that applies, to every entry, the same fma sequence as the unblocked
// dgetf2 of the oracle (panel factorization, row swaps, triangular solve of the U block
// row, and a trailing update whose per-entry fma chain runs over the panel in order), so
// the factors agree bit for bit.  Each visit then runs blocked column sweeps that keep
// the oracle's per-entry update order.
#include "pmg_internal.hpp"

namespace pmg {

namespace {

constexpr int NB = 32;     // LU panel width
constexpr int TB = 64;     // triangular-solve block (one wavefront)

__device__ inline double bcast(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// ------------------------------------------------------------------ dense assembly
// :5739-5833: A blocks (active rows all nine, other rows the diagonal blocks), BT, B
__global__ void k_coarse_scatter(OpView<double> op, double *__restrict__ a, long n) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = op.n_udof, NA = op.n_active, NP = op.n_pdof;
  const long G = 3L * N;
  if (r < N) {
    for (int k = op.a_rowptr[r]; k < op.a_rowptr[r + 1]; ++k) {
      const long c = op.a_kcol[k];
      if (r < NA) {
        for (int br = 0; br < 3; ++br)
          for (int bc = 0; bc < 3; ++bc)
            a[(c + bc * long(N)) * n + r + br * long(N)] = op.a[(3 * br + bc) * op.a_nnz + k];
      } else {
        for (int b = 0; b < 3; ++b) a[(c + b * long(N)) * n + r + b * long(N)] = op.a[(4 * b) * op.a_nnz + k];
      }
    }
    if (r < NA)
      for (int k = op.bt_rowptr[r]; k < op.bt_rowptr[r + 1]; ++k) {
        const long l = (op.bt_kcol[k] + G) * n;
        for (int b = 0; b < 3; ++b) a[l + r + b * long(N)] = op.bt[b * op.bt_nnz + k];
      }
  } else if (r < N + NP) {
    const int i = r - N;
    for (int k = op.b_rowptr[i]; k < op.b_rowptr[i + 1]; ++k) {
      const long c = op.b_kcol[k];
      for (int b = 0; b < 3; ++b) a[(c + b * long(N)) * n + i + G] = op.b[b * op.b_nnz + k];
    }
  }
}
// :5824-5830: row of pressure dof 0 := e_{3N}
__global__ void k_coarse_pin(double *a, long n, long pin) {
  const long c = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (c < n) a[c * n + pin] = (c == pin) ? 1.0 : 0.0;
}

__global__ void k_fill_d(double *x, double v, long n) {
  for (long i = blockIdx.x * long(blockDim.x) + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x) x[i] = v;
}

// ------------------------------------------------------------------ blocked LU
// panel [j0, j1): dgetf2 restricted to the panel columns (single workgroup)
__global__ __launch_bounds__(1024) void k_lu_panel(double *__restrict__ a, int n, int j0, int j1,
                                                   int *__restrict__ ipiv, int *__restrict__ info) {
  __shared__ double sv[1024];
  __shared__ int si[1024];
  const int t = threadIdx.x;
  for (int j = j0; j < j1; ++j) {
    double *colj = a + long(j) * n;
    double bv = -1.0;
    int bi = n;
    for (int i = j + t; i < n; i += 1024) {
      const double v = fabs(colj[i]);
      if (v > bv) { bv = v; bi = i; }   // rows visited in increasing order per thread
    }
    sv[t] = bv;
    si[t] = bi;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
      if (t < o) {
        const double ov = sv[t + o];
        const int oi = si[t + o];
        if (ov > sv[t] || (ov == sv[t] && oi < si[t])) { sv[t] = ov; si[t] = oi; }
      }
      __syncthreads();
    }
    const int p = si[0];
    const bool nz = colj[p] != 0.0;
    __syncthreads();
    if (t == 0) ipiv[j] = p;
    if (nz) {
      if (p != j && t < j1 - j0) {
        const long c = j0 + t;
        const double x = a[c * n + j];
        a[c * n + j] = a[c * n + p];
        a[c * n + p] = x;
      }
      __syncthreads();
      const double r = 1.0 / colj[j];
      for (int i = j + 1 + t; i < n; i += 1024) colj[i] *= r;
    } else if (t == 0 && info[0] == 0) {
      info[0] = j + 1;
    }
    __syncthreads();
    for (int c = j + 1; c < j1; ++c) {
      double *colc = a + long(c) * n;
      const double u = colc[j];
      for (int i = j + 1 + t; i < n; i += 1024) colc[i] = fma(-colj[i], u, colc[i]);
    }
    __syncthreads();
  }
}
// apply the panel's row interchanges to the columns outside [j0, j1)
__global__ void k_lu_swap(double *__restrict__ a, int n, int j0, int j1, const int *__restrict__ ipiv) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int nout = n - (j1 - j0);
  if (c >= nout) return;
  if (c >= j0) c += j1 - j0;
  double *col = a + long(c) * n;
  for (int j = j0; j < j1; ++j) {
    const int p = ipiv[j];
    if (p != j) {
      const double x = col[j];
      col[j] = col[p];
      col[p] = x;
    }
  }
}
// U block row: rows [j0, j1) of the columns >= j1, unit-lower forward sweep with L11
__global__ __launch_bounds__(256) void k_lu_trsm(double *__restrict__ a, int n, int j0, int j1) {
  __shared__ double L[NB * NB];
  const int w = j1 - j0;
  for (int idx = threadIdx.x; idx < w * w; idx += 256) {
    const int cc = idx / w, rr = idx % w;
    L[cc * NB + rr] = a[long(j0 + cc) * n + j0 + rr];
  }
  __syncthreads();
  const int c = j1 + blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  double *col = a + long(c) * n + j0;
  double v[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) v[i] = i < w ? col[i] : 0.0;
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    if (k >= w) break;
    const double uk = v[k];
#pragma unroll
    for (int i = k + 1; i < NB; ++i)
      if (i < w) v[i] = fma(-L[k * NB + i], uk, v[i]);
  }
#pragma unroll
  for (int i = 0; i < NB; ++i)
    if (i < w) col[i] = v[i];
}
// trailing update a[i][c] -= sum_k l[i][k] u[k][c], per-entry fma chain over k ascending
__global__ __launch_bounds__(256) void k_lu_gemm(double *__restrict__ a, int n, int j0, int j1) {
  __shared__ double Ls[NB][64 + 1];
  __shared__ double Us[NB][64 + 1];
  const int w = j1 - j0;
  const int i0 = j1 + blockIdx.y * 64, c0 = j1 + blockIdx.x * 64;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  for (int idx = threadIdx.x; idx < NB * 64; idx += 256) {
    const int k = idx / 64, r = idx % 64;
    Ls[k][r] = (k < w && i0 + r < n) ? a[long(j0 + k) * n + i0 + r] : 0.0;
    Us[k][r] = (k < w && c0 + r < n) ? a[long(c0 + r) * n + j0 + k] : 0.0;
  }
  __syncthreads();
  double acc[4][4];
#pragma unroll
  for (int ii = 0; ii < 4; ++ii)
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
      const int i = i0 + ty * 4 + ii, c = c0 + tx * 4 + cc;
      acc[ii][cc] = (i < n && c < n) ? a[long(c) * n + i] : 0.0;
    }
  for (int k = 0; k < w; ++k) {
    double l[4], u[4];
#pragma unroll
    for (int ii = 0; ii < 4; ++ii) l[ii] = Ls[k][ty * 4 + ii];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) u[cc] = Us[k][tx * 4 + cc];
#pragma unroll
    for (int ii = 0; ii < 4; ++ii)
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) acc[ii][cc] = fma(-l[ii], u[cc], acc[ii][cc]);
  }
#pragma unroll
  for (int ii = 0; ii < 4; ++ii)
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
      const int i = i0 + ty * 4 + ii, c = c0 + tx * 4 + cc;
      if (i < n && c < n) a[long(c) * n + i] = acc[ii][cc];
    }
}
// b'[i] = b[sigma[i]] of the sequential interchanges (permutation kept in LDS)
__global__ void k_sigma(const int *ipiv, int *sigma, int n) {
  extern __shared__ int sg[];
  if (threadIdx.x != 0) return;
  for (int i = 0; i < n; ++i) sg[i] = i;
  for (int k = 0; k < n; ++k) {
    const int p = ipiv[k];
    if (p != k) {
      const int x = sg[k];
      sg[k] = sg[p];
      sg[p] = x;
    }
  }
  for (int i = 0; i < n; ++i) sigma[i] = sg[i];
}

// ------------------------------------------------------------------ per-visit solve
__global__ void k_coarse_rhs(const double *__restrict__ f, const int *__restrict__ sigma, double *__restrict__ b,
                             int n, int pin) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const int s = sigma[i];
    b[i] = (s == pin) ? 0.0 : f[s];
  }
}
// forward sweep, block [K0, K0+TB): every workgroup solves the diagonal block in wave 0
// (rows of the block in b are final), workgroup 0 stores it in y, all workgroups update
// the rows >= K0+TB of b
__global__ __launch_bounds__(256) void k_trsv_fwd(const double *__restrict__ a, double *__restrict__ b,
                                                  double *__restrict__ y, int n, int K0) {
  __shared__ double xs[TB];
  const int t = threadIdx.x;
  const int kb = min(TB, n - K0);
  if (t < 64) {
    double x = t < kb ? b[K0 + t] : 0.0;
    for (int k = 0; k < kb; ++k) {
      const double xk = bcast(x, k);
      if (t > k && t < kb) x = fma(-xk, a[long(K0 + k) * n + K0 + t], x);
    }
    xs[t] = x;
    if (blockIdx.x == 0 && t < kb) y[K0 + t] = x;
  }
  __syncthreads();
  const int i = K0 + kb + blockIdx.x * 256 + t;
  if (i >= n) return;
  double v = b[i];
  for (int k = 0; k < kb; ++k) v = fma(-xs[k], a[long(K0 + k) * n + i], v);
  b[i] = v;
}
// backward sweep, block [K0, K0+kb): diagonal block of y (division) into x, then the
// rows < K0 of y
__global__ __launch_bounds__(256) void k_trsv_bwd(const double *__restrict__ a, double *__restrict__ y,
                                                  double *__restrict__ xo, int n, int K0) {
  __shared__ double xs[TB];
  const int t = threadIdx.x;
  const int kb = min(TB, n - K0);
  if (t < 64) {
    double x = t < kb ? y[K0 + t] : 0.0;
    for (int k = kb - 1; k >= 0; --k) {
      const double xk = bcast(x, k) / a[long(K0 + k) * n + K0 + k];
      if (t == k) x = xk;
      if (t < k) x = fma(-xk, a[long(K0 + k) * n + K0 + t], x);
    }
    xs[t] = x;
    if (blockIdx.x == 0 && t < kb) xo[K0 + t] = x;
  }
  __syncthreads();
  const int i = blockIdx.x * 256 + t;
  if (i >= K0) return;
  double v = y[i];
  for (int k = kb - 1; k >= 0; --k) v = fma(-xs[k], a[long(K0 + k) * n + i], v);
  y[i] = v;
}

// ------------------------------------------------------------------ transfers
template <typename T>
__global__ void k_csr(int n_rows, const int *__restrict__ rp, const int *__restrict__ col,
                      const double *__restrict__ val, int ncomp, const T *__restrict__ x, long xs,
                      T *__restrict__ y, long ys) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rows) return;
  const int k0 = rp[r], k1 = rp[r + 1];
  for (int c = 0; c < ncomp; ++c) {
    T s = 0;
    const T *xc = x + c * xs;
    for (int k = k0; k < k1; ++k) s += T(val[k]) * xc[col[k]];
    y[c * ys + r] = s;
  }
}

}  // namespace

void launch_coarse_build(const DevLevel &L, double *a, bool pin, hipStream_t s) {
  const long n = L.n_dof;
  k_fill_d<<<4096, 256, 0, s>>>(a, 0.0, n * n);
  PMG_HIP_CHECK(hipGetLastError());
  const int rows = L.n_udof + L.n_pdof;
  k_coarse_scatter<<<(rows + 127) / 128, 128, 0, s>>>(L.op64(), a, n);
  PMG_HIP_CHECK(hipGetLastError());
  if (pin) {
    k_coarse_pin<<<int((n + 255) / 256), 256, 0, s>>>(a, n, 3L * L.n_udof);
    PMG_HIP_CHECK(hipGetLastError());
  }
}

int coarse_lu_factor(Coarse &C, hipStream_t s) {
  const int n = C.n;
  PMG_HIP_CHECK(hipMemsetAsync(C.info, 0, sizeof(int), s));
  for (int j0 = 0; j0 < n; j0 += NB) {
    const int j1 = std::min(n, j0 + NB);
    k_lu_panel<<<1, 1024, 0, s>>>(C.a, n, j0, j1, C.ipiv, C.info);
    PMG_HIP_CHECK(hipGetLastError());
    const int nout = n - (j1 - j0);
    if (nout > 0) {
      k_lu_swap<<<(nout + 255) / 256, 256, 0, s>>>(C.a, n, j0, j1, C.ipiv);
      PMG_HIP_CHECK(hipGetLastError());
    }
    if (j1 < n) {
      k_lu_trsm<<<(n - j1 + 255) / 256, 256, 0, s>>>(C.a, n, j0, j1);
      PMG_HIP_CHECK(hipGetLastError());
      dim3 g((n - j1 + 63) / 64, (n - j1 + 63) / 64);
      k_lu_gemm<<<g, 256, 0, s>>>(C.a, n, j0, j1);
      PMG_HIP_CHECK(hipGetLastError());
    }
  }
  k_sigma<<<1, 64, size_t(n) * sizeof(int), s>>>(C.ipiv, C.sigma, n);
  PMG_HIP_CHECK(hipGetLastError());
  int info = 0;
  PMG_HIP_CHECK(hipMemcpyAsync(&info, C.info, sizeof(int), hipMemcpyDeviceToHost, s));
  PMG_HIP_CHECK(hipStreamSynchronize(s));
  return info;
}

void coarse_lu_solve(Coarse &C, const double *f, double *u, int pin_index, hipStream_t s) {
  const int n = C.n;
  k_coarse_rhs<<<(n + 255) / 256, 256, 0, s>>>(f, C.sigma, C.b, n, pin_index);
  PMG_HIP_CHECK(hipGetLastError());
  for (int K0 = 0; K0 < n; K0 += TB) {
    const int below = n - std::min(n, K0 + TB);
    const int g = std::max(1, (below + 255) / 256);
    k_trsv_fwd<<<g, 256, 0, s>>>(C.a, C.b, C.y, n, K0);
    PMG_HIP_CHECK(hipGetLastError());
  }
  const int last = ((n - 1) / TB) * TB;
  for (int K0 = last; K0 >= 0; K0 -= TB) {
    const int g = std::max(1, (K0 + 255) / 256);
    k_trsv_bwd<<<g, 256, 0, s>>>(C.a, C.y, u, n, K0);
    PMG_HIP_CHECK(hipGetLastError());
  }
}

template <typename T>
void launch_csr_spmv(const Csr &m, int ncomp, const T *x, long xs, T *y, long ys, hipStream_t s) {
  if (m.n_rows <= 0) return;
  k_csr<T><<<(m.n_rows + 255) / 256, 256, 0, s>>>(m.n_rows, m.rowptr, m.col, m.val, ncomp, x, xs, y, ys);
  PMG_HIP_CHECK(hipGetLastError());
}
template void launch_csr_spmv<double>(const Csr &, int, const double *, long, double *, long, hipStream_t);
template void launch_csr_spmv<float>(const Csr &, int, const float *, long, float *, long, hipStream_t);

}  // namespace pmg
