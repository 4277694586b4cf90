// k_dense.hip - coarse-level dense direct solve (TNSE_MGLevel4::SolveExact,
// src/FE/NSE_MGLevel4.C:5728-5846) and the grid-transfer SpMVs, for gfx950.
//
// SolveExact builds the dense column-major matrix of the coarsest level with the row of
// pressure dof 0 replaced by the identity and calls LAPACK dgetrf/dgetrs at every visit.
// Its Dirichlet rows are identity rows too (only the diagonal entry of the diagonal
// blocks), so x_D = f_D / diag and x_pin = 0 are known, and the remaining unknowns solve
// the dense system of the active velocity and free pressure dofs with the right-hand side
// f_I - K_ID x_D: the same solution with 3*N_Dirichlet + 1 fewer unknowns (for the 4^3
// level 1284 instead of 2443, 7x fewer LU flops).  That system is built and factored on the
// device once per operator by a blocked right-looking LU with partial pivoting whose
// per-entry update order is that of the unblocked dgetf2 (panel factorization, row swaps,
// triangular solve of the U block row, trailing update as an fma chain over the panel in
// order); each visit runs blocked column sweeps in the dgetrs order.
#include "pmg_internal.hpp"

namespace pmg {

namespace {

constexpr int NB = 32;     // max LU panel width (trsm / gemm tiles)
constexpr int NBR = 16;    // panel width of the register-resident panel factorization
constexpr int PNT = 512;   // threads of the register-resident panel factorization
constexpr int TB = 64;     // triangular-solve block (one wavefront)

__device__ inline double bcast(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// ------------------------------------------------------------------ dense assembly
// reduced numbering: velocity (component b, active dof r) -> b*na + r, pressure q -> 3na+q-pin
__device__ inline long cvel(int b, int r, int na) { return long(b) * na + r; }
__device__ inline long cpre(int q, int na, int pin) { return 3L * na + q - pin; }

// :5739-5833 restricted to the kept rows and columns: A blocks of active rows/columns,
// BT columns except the pinned one, B rows except the pinned one
__global__ void k_coarse_scatter(OpView<double> op, double *__restrict__ a, long n, int pin) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = op.n_udof, NA = op.n_active, NP = op.n_pdof;
  if (r < NA) {
    for (int k = op.a_rowptr[r]; k < op.a_rowptr[r + 1]; ++k) {
      const int c = op.a_kcol[k];
      if (c >= NA) continue;
      for (int br = 0; br < 3; ++br)
        for (int bc = 0; bc < 3; ++bc) a[cvel(bc, c, NA) * n + cvel(br, r, NA)] = op.a[(3 * br + bc) * op.a_nnz + k];
    }
    for (int k = op.bt_rowptr[r]; k < op.bt_rowptr[r + 1]; ++k) {
      const int q = op.bt_kcol[k];
      if (pin && q == 0) continue;
      for (int b = 0; b < 3; ++b) a[cpre(q, NA, pin) * n + cvel(b, r, NA)] = op.bt[b * op.bt_nnz + k];
    }
  } else if (r >= N && r < N + NP) {
    const int i = r - N;
    if (pin && i == 0) return;
    for (int k = op.b_rowptr[i]; k < op.b_rowptr[i + 1]; ++k) {
      const int c = op.b_kcol[k];
      if (c >= NA) continue;
      for (int b = 0; b < 3; ++b) a[cvel(b, c, NA) * n + cpre(i, NA, pin)] = op.b[b * op.b_nnz + k];
    }
  }
}

__global__ void k_fill_d(double *x, double v, long n) {
  for (long i = blockIdx.x * long(blockDim.x) + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x) x[i] = v;
}

// ------------------------------------------------------------------ blocked LU
// panel [j0, j1): dgetf2 restricted to the panel columns (single workgroup)
__global__ __launch_bounds__(1024) void k_lu_panel(double *__restrict__ a, int n, int j0, int j1,
                                                   int *__restrict__ ipiv, int *__restrict__ info) {
  __shared__ double sv[16];
  __shared__ int si[16];
  __shared__ int piv_s, nz_s;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for (int j = j0; j < j1; ++j) {
    double *colj = a + long(j) * n;
    // first index of max |a(i,j)|, i >= j: per thread in increasing row order, then
    // (value, -index) maxima across lanes and waves
    double bv = -1.0;
    int bi = n;
    for (int i = j + t; i < n; i += 1024) {
      const double v = fabs(colj[i]);
      if (v > bv) { bv = v; bi = i; }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(bv, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if (lane == 0) { sv[wave] = bv; si[wave] = bi; }
    __syncthreads();
    if (wave == 0) {
      bv = lane < 16 ? sv[lane] : -1.0;
      bi = lane < 16 ? si[lane] : n;
      for (int o = 8; o > 0; o >>= 1) {
        const double ov = __shfl_xor(bv, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
      }
      if (lane == 0) { piv_s = bi; nz_s = bv != 0.0; }
    }
    __syncthreads();
    const int p = piv_s;
    const bool nz = nz_s != 0;
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
// Same dgetf2 steps with the panel held in registers: thread t owns panel rows
// j0 + t + 1024 q (q < PR), so the column loop touches global memory only to load and
// store the panel.  The pivot row travels through LDS (it is also the U row of the step).
template <int PR>
__global__ __launch_bounds__(PNT) void k_lu_panel_reg(double *__restrict__ a, int n, int j0, int j1,
                                                       int *__restrict__ ipiv, int *__restrict__ info) {
  __shared__ double bufP[NBR], bufJ[NBR];
  __shared__ double sv[16];
  __shared__ int si[16];
  __shared__ int piv_s, nz_s;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int w = j1 - j0;
  double v[PR][NBR];
#pragma unroll
  for (int q = 0; q < PR; ++q) {
    const int r = j0 + t + PNT * q;
#pragma unroll
    for (int cc = 0; cc < NBR; ++cc) v[q][cc] = (r < n && cc < w) ? a[long(j0 + cc) * n + r] : 0.0;
  }
#pragma unroll
  for (int jj = 0; jj < NBR; ++jj) {
    const int j = j0 + jj;
    if (jj < w) {
    double bv = -1.0;
    int bi = n;
#pragma unroll
    for (int q = 0; q < PR; ++q) {
      const int r = j0 + t + PNT * q;
      if (r >= j && r < n) {
        const double x = fabs(v[q][jj]);
        if (x > bv) { bv = x; bi = r; }
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(bv, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if (lane == 0) { sv[wave] = bv; si[wave] = bi; }
    __syncthreads();
    if (wave == 0) {
      bv = lane < PNT / 64 ? sv[lane] : -1.0;
      bi = lane < PNT / 64 ? si[lane] : n;
      for (int o = 4; o > 0; o >>= 1) {
        const double ov = __shfl_xor(bv, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
      }
      if (lane == 0) { piv_s = bi; nz_s = bv != 0.0; }
    }
    __syncthreads();
    const int p = piv_s;
    if (t == 0) ipiv[j] = p;
    if (nz_s) {
      const int tp = (p - j0) % PNT, qp = (p - j0) / PNT;
      // pivot row -> bufP, row j -> bufJ (row j is owned by thread jj, q = 0)
#pragma unroll
      for (int q = 0; q < PR; ++q)
        if (t == tp && q == qp)
#pragma unroll
          for (int cc = 0; cc < NBR; ++cc) bufP[cc] = v[q][cc];
      if (t == jj)
#pragma unroll
        for (int cc = 0; cc < NBR; ++cc) bufJ[cc] = v[0][cc];
      __syncthreads();
      if (p != j) {
        if (t == jj)
#pragma unroll
          for (int cc = 0; cc < NBR; ++cc) v[0][cc] = bufP[cc];
#pragma unroll
        for (int q = 0; q < PR; ++q)
          if (t == tp && q == qp)
#pragma unroll
            for (int cc = 0; cc < NBR; ++cc) v[q][cc] = bufJ[cc];
      }
      const double rinv = 1.0 / bufP[jj];
#pragma unroll
      for (int q = 0; q < PR; ++q) {
        const int r = j0 + t + PNT * q;
        if (r > j && r < n) {
          const double l = v[q][jj] * rinv;
          v[q][jj] = l;
#pragma unroll
          for (int cc = 0; cc < NBR; ++cc)
            if (cc > jj) v[q][cc] = fma(-l, bufP[cc], v[q][cc]);
        }
      }
    } else if (t == 0 && info[0] == 0) {
      info[0] = j + 1;
    }
    __syncthreads();
    }
  }
#pragma unroll
  for (int q = 0; q < PR; ++q) {
    const int r = j0 + t + PNT * q;
    if (r < n)
#pragma unroll
      for (int cc = 0; cc < NBR; ++cc)
        if (cc < w) a[long(j0 + cc) * n + r] = v[q][cc];
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
// reduced right-hand side, already permuted: b[i] = (f_I - K_ID x_D)[sigma[i]] with
// x_D = f_D / diag(A_cc) on the Dirichlet velocity dofs
__device__ inline double dirichlet_value(const OpView<double> &op, const double *f, int b, int c) {
  const int k = op.a_rowptr[c];   // Dirichlet rows hold only their diagonal
  return f[long(b) * op.n_udof + c] / op.a[(4 * b) * op.a_nnz + k];
}
__global__ void k_coarse_rhs(OpView<double> op, const double *__restrict__ f, const int *__restrict__ sigma,
                             double *__restrict__ bout, int n, int pin) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = sigma[i];
  const int N = op.n_udof, NA = op.n_active;
  double v;
  if (k < 3 * NA) {
    const int br = k / NA, r = k - br * NA;
    v = f[long(br) * N + r];
    for (int q = op.a_rowptr[r]; q < op.a_rowptr[r + 1]; ++q) {
      const int c = op.a_kcol[q];
      if (c < NA) continue;
      for (int bc = 0; bc < 3; ++bc) v -= op.a[(3 * br + bc) * op.a_nnz + q] * dirichlet_value(op, f, bc, c);
    }
  } else {
    const int pi = k - 3 * NA + pin;
    v = f[3L * N + pi];
    for (int q = op.b_rowptr[pi]; q < op.b_rowptr[pi + 1]; ++q) {
      const int c = op.b_kcol[q];
      if (c < NA) continue;
      for (int b = 0; b < 3; ++b) v -= op.b[b * op.b_nnz + q] * dirichlet_value(op, f, b, c);
    }
  }
  bout[i] = v;
}
// scatter the reduced solution back: kept dofs, Dirichlet dofs f_D/diag, pinned pressure 0
__global__ void k_coarse_out(OpView<double> op, const double *__restrict__ f, const double *__restrict__ xr,
                             double *__restrict__ u, int pin) {
  const long i = blockIdx.x * long(blockDim.x) + threadIdx.x;
  const int N = op.n_udof, NA = op.n_active, NP = op.n_pdof;
  if (i < 3L * N) {
    const int b = int(i / N), r = int(i - long(b) * N);
    u[i] = r < NA ? xr[cvel(b, r, NA)] : dirichlet_value(op, f, b, r);
  } else if (i < 3L * N + NP) {
    const int q = int(i - 3L * N);
    u[i] = (pin && q == 0) ? 0.0 : xr[cpre(q, NA, pin)];
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
    // the lane's row of the diagonal block, loaded before the dependent chain starts
    const int tr = K0 + min(t, kb - 1);
    double row[TB];
#pragma unroll
    for (int k = 0; k < TB; ++k) row[k] = a[long(K0 + min(k, kb - 1)) * n + tr];
    double x = t < kb ? b[K0 + t] : 0.0;
#pragma unroll
    for (int k = 0; k < TB; ++k) {
      if (k < kb) {
        const double xk = bcast(x, k);
        if (t > k && t < kb) x = fma(-xk, row[k], x);
      }
    }
    xs[t] = x;
    if (blockIdx.x == 0 && t < kb) y[K0 + t] = x;
  }
  __syncthreads();
  const int i = K0 + kb + blockIdx.x * 256 + t;
  if (i >= n) return;
  double v = b[i];
#pragma unroll 16
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
    const int tr = K0 + min(t, kb - 1);
    double row[TB];
#pragma unroll
    for (int k = 0; k < TB; ++k) row[k] = a[long(K0 + min(k, kb - 1)) * n + tr];
    double x = t < kb ? y[K0 + t] : 0.0;
#pragma unroll
    for (int k = TB - 1; k >= 0; --k) {
      if (k < kb) {
        // U(k,k) sits in lane k's row[k]
        const double xk = bcast(x, k) / bcast(row[k], k);
        if (t == k) x = xk;
        if (t < k) x = fma(-xk, row[k], x);
      }
    }
    xs[t] = x;
    if (blockIdx.x == 0 && t < kb) xo[K0 + t] = x;
  }
  __syncthreads();
  const int i = blockIdx.x * 256 + t;
  if (i >= K0) return;
  double v = y[i];
#pragma unroll 16
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

void launch_coarse_build(const DevLevel &L, Coarse &C, hipStream_t s) {
  const long n = C.n;
  k_fill_d<<<2048, 256, 0, s>>>(C.a, 0.0, n * n);
  PMG_HIP_CHECK(hipGetLastError());
  const int rows = L.n_udof + L.n_pdof;
  k_coarse_scatter<<<(rows + 127) / 128, 128, 0, s>>>(L.op64(), C.a, n, C.pin);
  PMG_HIP_CHECK(hipGetLastError());
}

void coarse_lu_factor(Coarse &C, hipStream_t s) {
  const int n = C.n;
  PMG_HIP_CHECK(hipMemsetAsync(C.info, 0, sizeof(int), s));
  for (int j0 = 0; j0 < n; j0 += NBR) {
    const int j1 = std::min(n, j0 + NBR);
    const int pr = (n - j0 + PNT - 1) / PNT;
    switch (pr) {
      case 1: k_lu_panel_reg<1><<<1, PNT, 0, s>>>(C.a, n, j0, j1, C.ipiv, C.info); break;
      case 2: k_lu_panel_reg<2><<<1, PNT, 0, s>>>(C.a, n, j0, j1, C.ipiv, C.info); break;
      case 3: k_lu_panel_reg<3><<<1, PNT, 0, s>>>(C.a, n, j0, j1, C.ipiv, C.info); break;
      case 4: k_lu_panel_reg<4><<<1, PNT, 0, s>>>(C.a, n, j0, j1, C.ipiv, C.info); break;
      case 5: k_lu_panel_reg<5><<<1, PNT, 0, s>>>(C.a, n, j0, j1, C.ipiv, C.info); break;
      case 6: k_lu_panel_reg<6><<<1, PNT, 0, s>>>(C.a, n, j0, j1, C.ipiv, C.info); break;
      default: k_lu_panel<<<1, 1024, 0, s>>>(C.a, n, j0, j1, C.ipiv, C.info); break;
    }
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
}

void coarse_lu_solve(Coarse &C, const OpView<double> &op, const double *f, double *u, hipStream_t s) {
  const int n = C.n;
  k_coarse_rhs<<<(n + 255) / 256, 256, 0, s>>>(op, f, C.sigma, C.b, n, C.pin);
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
    k_trsv_bwd<<<g, 256, 0, s>>>(C.a, C.y, C.x, n, K0);
    PMG_HIP_CHECK(hipGetLastError());
  }
  const long nf = 3L * op.n_udof + op.n_pdof;
  k_coarse_out<<<int((nf + 255) / 256), 256, 0, s>>>(op, f, C.x, u, C.pin);
  PMG_HIP_CHECK(hipGetLastError());
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
