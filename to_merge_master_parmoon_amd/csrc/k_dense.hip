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
#include "pmg_rows.hpp"
#include "coarse_rows.hpp"

namespace pmg {

namespace {

constexpr int NBR = 16;    // panel width of the register-resident panel factorization
constexpr int PNT = 512;   // threads of the register-resident panel factorization (the widest form)
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
// BT columns except the pinned one, B rows except the pinned one.  Row-major a[row*n+col]
// while factoring (row interchanges and panel rows are contiguous).
__global__ __launch_bounds__(256) void k_coarse_scatter(OpView<double> op, double *__restrict__ a, long n, int pin) {
  const int lane = threadIdx.x & 15;
  const int r = (blockIdx.x * 256 + threadIdx.x) >> 4;  // 16 lanes per operator row
  const int N = op.n_udof, NA = op.n_active, NP = op.n_pdof;
  if (r < NA) {
    for (int k = op.a_rowptr[r] + lane; k < op.a_rowptr[r + 1]; k += 16) {
      const int c = op.a_kcol[k];
      if (c >= NA) continue;
      for (int br = 0; br < 3; ++br)
        for (int bc = 0; bc < 3; ++bc) a[cvel(br, r, NA) * n + cvel(bc, c, NA)] = op.a[(3 * br + bc) * op.a_nnz + k];
    }
    for (int k = op.bt_rowptr[r] + lane; k < op.bt_rowptr[r + 1]; k += 16) {
      const int q = op.bt_kcol[k];
      if (pin && q == 0) continue;
      for (int b = 0; b < 3; ++b) a[cvel(b, r, NA) * n + cpre(q, NA, pin)] = op.bt[b * op.bt_nnz + k];
    }
  } else if (r >= N && r < N + NP) {
    const int i = r - N;
    if (pin && i == 0) return;
    for (int k = op.b_rowptr[i] + lane; k < op.b_rowptr[i + 1]; k += 16) {
      const int c = op.b_kcol[k];
      if (c >= NA) continue;
      for (int b = 0; b < 3; ++b) a[cpre(i, NA, pin) * n + cvel(b, c, NA)] = op.b[b * op.b_nnz + k];
    }
  }
}

// identity permutation over all n reduced unknowns (no interchanges yet)
__global__ void k_iota(int *__restrict__ sigma, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) sigma[i] = i;
}

__global__ void k_fill_d(double *x, double v, long n) {
  for (long i = blockIdx.x * long(blockDim.x) + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x) x[i] = v;
}

// Panel factorization of columns [j0, j1) (width <= NBR) with dgetf2 steps, the panel held
// in registers (thread t owns rows j0 + t + PNT q, q < PR); two barriers per column (pivot
// reduction, pivot row exchange), LDS double-buffered by column parity.  Then the panel's
// row interchanges are applied to every other column in one pass (the <= 2*NBR rows they
// involve are permuted at once, as dlaswp's sequential swaps would leave them), and the
// U block row of the previous step (rows pj0..pj1-1, kept in ubuf by k_lu_update) is
// copied into place.  Row-major matrix a[row*n + col].
template <int PR, int PT>
__global__ __launch_bounds__(PT) void k_lu_panel(double *__restrict__ a, int n, int j0, int j1,
                                                  int *__restrict__ ipiv, int *__restrict__ info,
                                                  int *__restrict__ swp, int *__restrict__ sigma) {
  __shared__ double sv[2][PT / 64];
  __shared__ int si[2][PT / 64];
  __shared__ double bufP[2][NBR], bufJ[2][NBR];
  __shared__ int pivs[NBR];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int w = j1 - j0;
  double v[PR][NBR];
#pragma unroll
  for (int q = 0; q < PR; ++q) {
    const int r = j0 + t + PT * q;
#pragma unroll
    for (int cc = 0; cc < NBR; ++cc) v[q][cc] = (r < n && cc < w) ? a[long(r) * n + j0 + cc] : 0.0;
  }
#pragma unroll
  for (int jj = 0; jj < NBR; ++jj) {
    if (jj < w) {
      const int j = j0 + jj, par = jj & 1;
      double bv = -1.0;
      int bi = n;
#pragma unroll
      for (int q = 0; q < PR; ++q) {
        const int r = j0 + t + PT * q;
        if (r >= j && r < n) {
          const double x = fabs(v[q][jj]);
          if (x > bv) {
            bv = x;
            bi = r;
          }
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(bv, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ov > bv || (ov == bv && oi < bi)) {
          bv = ov;
          bi = oi;
        }
      }
      if (lane == 0) {
        sv[par][wave] = bv;
        si[par][wave] = bi;
      }
      __syncthreads();
      bv = sv[par][0];
      bi = si[par][0];
#pragma unroll
      for (int w2 = 1; w2 < PT / 64; ++w2) {
        const double ov = sv[par][w2];
        const int oi = si[par][w2];
        if (ov > bv || (ov == bv && oi < bi)) {
          bv = ov;
          bi = oi;
        }
      }
      const int p = bi;
      const bool nz = bv != 0.0;
      if (t == 0) {
        ipiv[j] = p;
        pivs[jj] = p;
        if (!nz && info[0] == 0) info[0] = j + 1;
      }
      if (nz) {
        const int tp = (p - j0) % PT, qp = (p - j0) / PT;
#pragma unroll
        for (int q = 0; q < PR; ++q)
          if (t == tp && q == qp)
#pragma unroll
            for (int cc = 0; cc < NBR; ++cc) bufP[par][cc] = v[q][cc];
        if (t == jj)
#pragma unroll
          for (int cc = 0; cc < NBR; ++cc) bufJ[par][cc] = v[0][cc];
        __syncthreads();
        if (p != j) {
          if (t == jj)
#pragma unroll
            for (int cc = 0; cc < NBR; ++cc) v[0][cc] = bufP[par][cc];
#pragma unroll
          for (int q = 0; q < PR; ++q)
            if (t == tp && q == qp)
#pragma unroll
              for (int cc = 0; cc < NBR; ++cc) v[q][cc] = bufJ[par][cc];
        }
        const double rinv = 1.0 / bufP[par][jj];
#pragma unroll
        for (int q = 0; q < PR; ++q) {
          const int r = j0 + t + PT * q;
          if (r > j && r < n) {
            const double l = v[q][jj] * rinv;
            v[q][jj] = l;
#pragma unroll
            for (int cc = 0; cc < NBR; ++cc)
              if (cc > jj) v[q][cc] = fma(-l, bufP[par][cc], v[q][cc]);
          }
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < PR; ++q) {
    const int r = j0 + t + PT * q;
    if (r < n)
#pragma unroll
      for (int cc = 0; cc < NBR; ++cc)
        if (cc < w) a[long(r) * n + j0 + cc] = v[q][cc];
  }
  __syncthreads();
  if (wave == 0) {
    // rows involved, one slot per lane: j0..j1-1 (slots 0..w-1), then the pivot rows below
    // the panel in order of first use; src = row whose original value ends in the slot's
    // row after the sequential interchanges
    int row = lane < w ? j0 + lane : -1;
    int nk = w;
    for (int k = 0; k < w; ++k) {
      const int p = pivs[k];
      if (p >= j1 && __ballot(row == p) == 0) {
        if (lane == nk) row = p;
        ++nk;
      }
    }
    int src = row;
    for (int k = 0; k < w; ++k) {
      const int p = pivs[k];
      if (p == j0 + k) continue;
      const unsigned long long m = __ballot(row == p);
      const int f = __ffsll((long long)m) - 1;
      const int sk = __shfl(src, k, 64), sf = __shfl(src, f, 64);
      if (lane == k) src = sf;
      if (lane == f) src = sk;
    }
    // for k_lu_swap: the rows involved and their sources, then their count
    if (lane < 2 * NBR) {
      swp[lane] = row;
      swp[2 * NBR + lane] = src;
    }
    if (lane == 0) swp[4 * NBR] = nk;
    // the permutation vector (sigma[i] = original row now at row i) moves like a column
    const int sg = lane < nk ? sigma[src] : 0;
    if (lane < nk && src != row) sigma[row] = sg;
  }
}

// After a panel: its row interchanges applied to every column outside it (one thread per
// column, the <= 2*NBR rows involved permuted at once, as dlaswp's sequential swaps would
// leave them), and the U block row of the previous step (rows pj0..pj1-1, kept in ubuf by
// k_lu_update) copied into place for the columns >= pj1.  The two touch disjoint rows (the
// interchanges involve rows >= j0 = pj1 only).
__global__ __launch_bounds__(256) void k_lu_swap(double *__restrict__ a, int n, int j0, int j1,
                                                 const int *__restrict__ swp, const double *__restrict__ ubuf,
                                                 int pj0, int pj1) {
  __shared__ int rows_s[2 * NBR], src_s[2 * NBR];
  __shared__ int nk_s;
  const int t = threadIdx.x;
  if (t < 2 * NBR) {
    rows_s[t] = swp[t];
    src_s[t] = swp[2 * NBR + t];
  }
  if (t == 0) nk_s = swp[4 * NBR];
  __syncthreads();
  const int c = blockIdx.x * 256 + t;
  if (c >= n) return;
  if (pj1 > pj0 && c >= pj1)
    for (int kk = 0; kk < pj1 - pj0; ++kk) a[long(pj0 + kk) * n + c] = ubuf[long(kk) * n + c];
  if (c >= j0 && c < j1) return;
  const int nk = nk_s;
  double *cp = a + c;
  double x[2 * NBR];
#pragma unroll
  for (int k = 0; k < 2 * NBR; ++k) x[k] = k < nk ? cp[long(src_s[k]) * n] : 0.0;
#pragma unroll
  for (int k = 0; k < 2 * NBR; ++k)
    if (k < nk && src_s[k] != rows_s[k]) cp[long(rows_s[k]) * n] = x[k];
}

// Trailing update of one 64x64 tile of A22 = rows/columns >= j1: the block first solves its
// 16x64 tile of the U block row, U12 = L11^-1 A12 (unit lower, per-column fma sweep as
// dtrsm), then A22 -= L21 U12 as a per-entry fma chain over the panel in order.  A12 itself
// is not overwritten here (other blocks still read it); the blocks of the first row tile
// keep U12 in ubuf and the next panel kernel copies it into place.
__global__ __launch_bounds__(256) void k_lu_update(double *__restrict__ a, double *__restrict__ ubuf, int n,
                                                   int j0, int j1) {
  __shared__ double L11[NBR][NBR + 1];
  __shared__ double Us[NBR][64 + 1];
  __shared__ double Ls[NBR][64 + 1];
  const int w = j1 - j0;
  const int i0 = j1 + blockIdx.y * 64, c0 = j1 + blockIdx.x * 64;
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  {
    const int i = t / NBR, k = t % NBR;  // 256 = NBR * NBR; L11[k][i] = l(j0+i, j0+k)
    L11[k][i] = (k < w && i < w) ? a[long(j0 + i) * n + j0 + k] : 0.0;
  }
  for (int idx = t; idx < NBR * 64; idx += 256) {
    const int k = idx / 64, r = idx % 64;
    Us[k][r] = (k < w && c0 + r < n) ? a[long(j0 + k) * n + c0 + r] : 0.0;
    const int r2 = idx / NBR, k2 = idx % NBR;
    Ls[k2][r2] = (k2 < w && i0 + r2 < n) ? a[long(i0 + r2) * n + j0 + k2] : 0.0;
  }
  __syncthreads();
  if (t < 64) {
    double v[NBR];
#pragma unroll
    for (int i = 0; i < NBR; ++i) v[i] = Us[i][t];
#pragma unroll
    for (int k = 0; k < NBR; ++k) {
      if (k < w) {
        const double uk = v[k];
#pragma unroll
        for (int i = k + 1; i < NBR; ++i)
          if (i < w) v[i] = fma(-L11[k][i], uk, v[i]);
      }
    }
#pragma unroll
    for (int i = 0; i < NBR; ++i) Us[i][t] = v[i];
    if (blockIdx.y == 0 && c0 + t < n)
#pragma unroll
      for (int i = 0; i < NBR; ++i)
        if (i < w) ubuf[long(i) * n + c0 + t] = v[i];
  }
  __syncthreads();
  double acc[4][4];
#pragma unroll
  for (int ii = 0; ii < 4; ++ii)
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
      const int i = i0 + ty * 4 + ii, c = c0 + tx * 4 + cc;
      acc[ii][cc] = (i < n && c < n) ? a[long(i) * n + c] : 0.0;
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
      if (i < n && c < n) a[long(i) * n + c] = acc[ii][cc];
    }
}

// the row-major factors into the column-major copy read by the triangular sweeps
__global__ __launch_bounds__(256) void k_transpose(const double *__restrict__ a, double *__restrict__ f, int n) {
  __shared__ double tile[64][65];
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  const int t = threadIdx.x;
  for (int idx = t; idx < 64 * 64; idx += 256) {
    const int r = idx / 64, c = idx % 64;
    if (r0 + r < n && c0 + c < n) tile[r][c] = a[long(r0 + r) * n + c0 + c];
  }
  __syncthreads();
  for (int idx = t; idx < 64 * 64; idx += 256) {
    const int c = idx / 64, r = idx % 64;
    if (r0 + r < n && c0 + c < n) f[long(c0 + c) * n + r0 + r] = tile[r][c];
  }
}
// ------------------------------------------------------------------ per-visit solve
// reduced right-hand side, already permuted: b[i] = (f_I - K_ID x_D)[sigma[i]] with
// x_D = f_D / diag(A_cc) on the Dirichlet velocity dofs
__global__ __launch_bounds__(256) void k_coarse_rhs(OpView<double> op, const double *__restrict__ f,
                                                    const int *__restrict__ sigma, double *__restrict__ bout, int n,
                                                    int pin) {
  const int lane = threadIdx.x & 15;
  const int i = (blockIdx.x * 256 + threadIdx.x) >> 4;  // 16 lanes per reduced row
  if (i >= n) return;
  const double v = coarse_rhs_row(op, [&](long j) { return f[j]; }, sigma[i], pin, lane);
  if (lane == 0) bout[i] = v;
}
// scatter the reduced solution back: kept dofs, Dirichlet dofs f_D/diag, pinned pressure 0
__global__ void k_coarse_out(OpView<double> op, const double *__restrict__ f, const double *__restrict__ xr,
                             double *__restrict__ u, int pin) {
  const long i = blockIdx.x * long(blockDim.x) + threadIdx.x;
  const int N = op.n_udof, NA = op.n_active, NP = op.n_pdof;
  if (i < 3L * N) {
    const int b = int(i / N), r = int(i - long(b) * N);
    u[i] = r < NA ? xr[cvel(b, r, NA)] : coarse_dirichlet(op, [&](long j) { return f[j]; }, b, r);
  } else if (i < 3L * N + NP) {
    const int q = int(i - 3L * N);
    u[i] = (pin && q == 0) ? 0.0 : xr[cpre(q, NA, pin)];
  }
}

// Triangular sweeps of the coarse solve, one launch each.  Workgroup g handles the row
// block of 64 rows that is the g-th to be solved (forward: rows 64g..; backward: the g-th
// block from the bottom).  It accumulates the products with the solution segments of the
// blocks solved before it as each one is published (flag of that block = this launch's
// generation, acquire; each wave polls for itself, no workgroup barrier per segment),
// then solves its diagonal block in one wavefront and publishes its segment (release).  A workgroup only waits for workgroups with smaller ids, which are
// dispatched before it, so the chain always drains; every wait is bounded and sets *err
// when exceeded.  The generation lives in sync[0] and is advanced by the last workgroup to
// finish (sync[1] counts them), so no host reset is needed between launches.
constexpr long SPIN_LIMIT = 1L << 24;

__device__ inline bool wait_flag(const unsigned *flag, unsigned gen, int *err) {
  long spins = 0;
  while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != gen) {
    if (++spins > SPIN_LIMIT) {
      __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(2);
  }
  return true;
}

template <bool UPPER>
__global__ __launch_bounds__(256) void k_trsv_chain(const double *__restrict__ f, const double *__restrict__ rhs,
                                                    double *__restrict__ x, int n, unsigned *__restrict__ sync,
                                                    unsigned *__restrict__ flags, int *__restrict__ err) {
  __shared__ double part[4][TB];
  __shared__ unsigned gen_s;
  const int t = threadIdx.x, r = t & 63, qd = t >> 6;
  const int nblk = (n + TB - 1) / TB;
  const int g = blockIdx.x;
  const int rb = UPPER ? nblk - 1 - g : g;  // row block solved by this workgroup
  const int R0 = rb * TB, kb = min(TB, n - R0);
  if (t == 0) gen_s = __hip_atomic_load(&sync[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
  __syncthreads();
  const unsigned gen = gen_s;
  const int row = R0 + min(r, kb - 1);
  // the diagonal block's row of this lane (wave 0), loaded before any wait
  double row_v[TB];
  if (qd == 0) {
#pragma unroll
    for (int k = 0; k < TB; ++k) row_v[k] = f[long(R0 + min(k, kb - 1)) * n + row];
  }
  // earlier segments in solve order: wave qd takes columns 16qd..16qd+15 of each; the
  // factor values are loaded before the segment's flag is awaited
  double acc = 0.0;
  bool ok = true;
  for (int h = 0; h < g; ++h) {
    const int jb = UPPER ? nblk - 1 - h : h;
    const int J0 = jb * TB, jw = min(TB, n - J0);
    double fv[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) fv[q] = f[long(J0 + min(qd * 16 + q, jw - 1)) * n + row];
    if (r == 0 && ok) ok = wait_flag(&flags[jb], gen, err);
    ok = __shfl(ok ? 1 : 0, 0, 64) != 0;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int kk = qd * 16 + q;
      if (kk < jw) {
        const double xk = __hip_atomic_load(&x[J0 + kk], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        acc = fma(xk, fv[q], acc);
      }
    }
  }
  part[qd][r] = acc;
  __syncthreads();
  if (qd == 0) {
    double v = r < kb ? rhs[R0 + r] - ((part[0][r] + part[1][r]) + (part[2][r] + part[3][r])) : 0.0;
    if (!UPPER) {
#pragma unroll
      for (int k = 0; k < TB; ++k) {
        if (k < kb) {
          const double xk = bcast(v, k);
          if (r > k && r < kb) v = fma(-xk, row_v[k], v);
        }
      }
    } else {
      // U(k,k) sits in lane k's row_v[k].  Every lane forms the reciprocal of its own
      // diagonal entry before the chain; a step then takes the corrected quotient
      // q = x rc, e = fma(-q, u, x), q' = fma(e, rc, q), which with rc = RN(1/u) is the
      // correctly rounded x / u (Markstein): the value of the division, in three dependent
      // operations instead of the division's sequence.
      const double dg = r < kb ? f[long(R0 + r) * n + R0 + r] : 1.0;
      const double rc = 1.0 / dg;
#pragma unroll
      for (int k = TB - 1; k >= 0; --k) {
        if (k < kb) {
          const double xv = bcast(v, k), uk = bcast(dg, k), rk = bcast(rc, k);
          const double q = xv * rk;
          const double xk = fma(fma(-q, uk, xv), rk, q);
          if (r == k) v = xk;
          if (r < k) v = fma(-xk, row_v[k], v);
        }
      }
    }
    if (r < kb) x[R0 + r] = v;
    __threadfence();
  }
  __syncthreads();
  if (t == 0) {
    __hip_atomic_store(&flags[rb], gen, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned done = __hip_atomic_fetch_add(&sync[1], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (done == unsigned(nblk - 1)) {
      __hip_atomic_store(&sync[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&sync[0], gen, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// y[c*ys + r] = sum_k val[k] x[c*xs + col[k]] for the NC components, 16 lanes per row
template <typename T, int NC>
__global__ __launch_bounds__(256) void k_csr(int n_rows, const int *__restrict__ rows, const int *__restrict__ rp,
                                             const int *__restrict__ col, const double *__restrict__ val,
                                             const T *__restrict__ x, long xs, T *__restrict__ y, long ys,
                                             int zero_from) {
  const int lane = threadIdx.x & 15;
  const int grp = (blockIdx.x * 256 + threadIdx.x) >> 4, ngrp = gridDim.x * 16;
  for (int i = grp; i < n_rows; i += ngrp) {
    const int r = rows ? rows[i] : i;
    if (r >= zero_from) {
      if (lane == 0)
#pragma unroll
        for (int c = 0; c < NC; ++c) y[c * ys + r] = T(0);
      continue;
    }
    const int k0 = rp[r], k1 = rp[r + 1];
    T s[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) s[c] = 0;
    for (int k = k0 + lane; k < k1; k += 16) {
      const int j = col[k];
      const T v = T(val[k]);
#pragma unroll
      for (int c = 0; c < NC; ++c) s[c] += v * x[c * xs + j];
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) s[c] = red16(s[c]);
    if (lane == 0)
#pragma unroll
      for (int c = 0; c < NC; ++c) y[c * ys + r] = s[c];
  }
}

// The prolongation and the coarse-grid correction in one launch, for the fine rows r (all,
// or the listed ones): velocity (NC = 3) u_c[r] = f_c[r] on the Dirichlet rows [ds, ds + dc)
// and u_c[r] + a (P x)_c[r] elsewhere, and (ua non-null) ua[r] = (u_1, u_2, u_3, 0);
// pressure (NC = 1) u[r] += a (P x)[r].  The same values as the prolongation into a vector,
// the axpy and the sweep start that set the Dirichlet slice and build ua.
template <typename T, int NC>
__global__ __launch_bounds__(256) void k_csr_update(int n_rows, const int *__restrict__ rows,
                                                    const int *__restrict__ rp, const int *__restrict__ col,
                                                    const double *__restrict__ val, const T *__restrict__ x, long xs,
                                                    T *__restrict__ u, long us, T a, const T *__restrict__ f, int ds,
                                                    int dc, float4 *__restrict__ ua) {
  const int lane = threadIdx.x & 15;
  const int grp = (blockIdx.x * 256 + threadIdx.x) >> 4, ngrp = gridDim.x * 16;
  for (int i = grp; i < n_rows; i += ngrp) {
    const int r = rows ? rows[i] : i;
    const int k0 = rp[r], k1 = rp[r + 1];
    // the row's u (f on the Dirichlet slice) loaded beside its entries, not after the sum
    const bool dir = r >= ds && r < ds + dc;
    T y[NC];
    if (lane == 0)
#pragma unroll
      for (int c = 0; c < NC; ++c) y[c] = dir ? f[c * us + r] : u[c * us + r];
    T s[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) s[c] = 0;
    for (int k = k0 + lane; k < k1; k += 16) {
      const int j = col[k];
      const T v = T(val[k]);
#pragma unroll
      for (int c = 0; c < NC; ++c) s[c] += v * x[c * xs + j];
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) s[c] = red16(s[c]);
    if (lane == 0) {
      T w[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        T yy = y[c];
        yy += a * s[c];
        w[c] = dir ? y[c] : yy;
        u[c * us + r] = w[c];
      }
      if constexpr (NC == 3)
        if (ua) ua[r] = make_float4(float(w[0]), float(w[1]), float(w[2]), 0.0f);
    }
  }
}

}  // namespace

void launch_coarse_build(const DevLevel &L, Coarse &C, hipStream_t s) {
  const long n = C.n;
  k_fill_d<<<2048, 256, 0, s>>>(C.a, 0.0, n * n);
  PMG_HIP_CHECK(hipGetLastError());
  const int rows = L.n_udof + L.n_pdof;
  k_coarse_scatter<<<(rows * 16 + 255) / 256, 256, 0, s>>>(L.op64(), C.a, n, C.pin);
  PMG_HIP_CHECK(hipGetLastError());
  k_iota<<<int((n + 255) / 256), 256, 0, s>>>(C.sigma, int(n));
  PMG_HIP_CHECK(hipGetLastError());
}

void coarse_lu_factor(Coarse &C, hipStream_t s) {
  const int n = C.n;
  const int pr_max = 7;   // PR = 7 spills 16 VGPRs to scratch; 8 would spill ~1000
  if ((n + PNT - 1) / PNT > pr_max)
    throw DeviceError("coarse system of " + std::to_string(n) + " unknowns exceeds the dense LU limit (" +
                      std::to_string(pr_max * PNT) + ")");
  // panels of 256 threads while they fit (measured faster: 6.7 vs 6.9 ms for the 1284
  // system beside the patch set-up), 512 above
  const bool narrow = (n + 255) / 256 <= pr_max;
  PMG_HIP_CHECK(hipMemsetAsync(C.info, 0, sizeof(int), s));
  int pj0 = 0, pj1 = 0;
  for (int j0 = 0; j0 < n; j0 += NBR) {
    const int j1 = std::min(n, j0 + NBR);
    const int pt = narrow ? 256 : PNT;
    const int pr = (n - j0 + pt - 1) / pt;
#define PMG_PANEL(P, T) k_lu_panel<P, T><<<1, T, 0, s>>>(C.a, n, j0, j1, C.ipiv, C.info, C.swp, C.sigma)
    if (narrow) {
      switch (pr) {
        case 1: PMG_PANEL(1, 256); break;
        case 2: PMG_PANEL(2, 256); break;
        case 3: PMG_PANEL(3, 256); break;
        case 4: PMG_PANEL(4, 256); break;
        case 5: PMG_PANEL(5, 256); break;
        case 6: PMG_PANEL(6, 256); break;
        default: PMG_PANEL(7, 256); break;
      }
    } else {
      switch (pr) {
        case 1: PMG_PANEL(1, PNT); break;
        case 2: PMG_PANEL(2, PNT); break;
        case 3: PMG_PANEL(3, PNT); break;
        case 4: PMG_PANEL(4, PNT); break;
        case 5: PMG_PANEL(5, PNT); break;
        case 6: PMG_PANEL(6, PNT); break;
        default: PMG_PANEL(7, PNT); break;
      }
    }
#undef PMG_PANEL
    PMG_HIP_CHECK(hipGetLastError());
    k_lu_swap<<<(n + 255) / 256, 256, 0, s>>>(C.a, n, j0, j1, C.swp, C.ubuf, pj0, pj1);
    PMG_HIP_CHECK(hipGetLastError());
    if (j1 < n) {
      dim3 g((n - j1 + 63) / 64, (n - j1 + 63) / 64);
      k_lu_update<<<g, 256, 0, s>>>(C.a, C.ubuf, n, j0, j1);
      PMG_HIP_CHECK(hipGetLastError());
      pj0 = j0;
      pj1 = j1;
    } else {
      pj0 = pj1 = 0;
    }
  }
  dim3 gt((n + 63) / 64, (n + 63) / 64);
  k_transpose<<<gt, 256, 0, s>>>(C.a, C.f, n);
  PMG_HIP_CHECK(hipGetLastError());
}

void launch_coarse_rhs(const OpView<double> &op, const double *f, const int *sigma, double *b, int n, int pin,
                       hipStream_t s) {
  k_coarse_rhs<<<(n + 15) / 16, 256, 0, s>>>(op, f, sigma, b, n, pin);
  PMG_HIP_CHECK(hipGetLastError());
}

void coarse_lu_solve(Coarse &C, const OpView<double> &op, const double *f, double *u, hipStream_t s) {
  const int n = C.n;
  const int nblk = (n + TB - 1) / TB;
  k_coarse_rhs<<<(n + 15) / 16, 256, 0, s>>>(op, f, C.sigma, C.b, n, C.pin);
  PMG_HIP_CHECK(hipGetLastError());
  // L y = b (unit lower), then U x = y
  k_trsv_chain<false><<<nblk, 256, 0, s>>>(C.f, C.b, C.y, n, C.sync, C.flags, C.err);
  PMG_HIP_CHECK(hipGetLastError());
  k_trsv_chain<true><<<nblk, 256, 0, s>>>(C.f, C.y, C.x, n, C.sync + 2, C.flags + nblk, C.err);
  PMG_HIP_CHECK(hipGetLastError());
  const long nf = 3L * op.n_udof + op.n_pdof;
  k_coarse_out<<<int((nf + 255) / 256), 256, 0, s>>>(op, f, C.x, u, C.pin);
  PMG_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_csr_spmv(const Csr &m, int ncomp, const T *x, long xs, T *y, long ys, hipStream_t s, const int *rows,
                     int n_rows, int zero_from) {
  const int n = rows ? n_rows : m.n_rows;
  if (n <= 0) return;
  const int grid = std::min(8192, (n + 15) / 16);
  if (ncomp == 3)
    k_csr<T, 3><<<grid, 256, 0, s>>>(n, rows, m.rowptr, m.col, m.val, x, xs, y, ys, zero_from);
  else
    k_csr<T, 1><<<grid, 256, 0, s>>>(n, rows, m.rowptr, m.col, m.val, x, xs, y, ys, zero_from);
  PMG_HIP_CHECK(hipGetLastError());
}
template void launch_csr_spmv<double>(const Csr &, int, const double *, long, double *, long, hipStream_t,
                                      const int *, int, int);
template void launch_csr_spmv<float>(const Csr &, int, const float *, long, float *, long, hipStream_t, const int *,
                                     int, int);


template <typename T>
void launch_prolong_update(const Csr &pu, const Csr &pp, const T *xc, long nxc, T *u, long nu, T a, const T *f,
                           int ds, int dc, float *ua, hipStream_t s, const int *urows, int n_urows, const int *prows,
                           int n_prows) {
  const int nv = urows ? n_urows : pu.n_rows, np = prows ? n_prows : pp.n_rows;
  if (nv > 0)
    k_csr_update<T, 3><<<std::min(8192, (nv + 15) / 16), 256, 0, s>>>(nv, urows, pu.rowptr, pu.col, pu.val, xc, nxc,
                                                                       u, nu, a, f, ds, dc,
                                                                       reinterpret_cast<float4 *>(ua));
  if (np > 0)
    k_csr_update<T, 1><<<std::min(8192, (np + 15) / 16), 256, 0, s>>>(np, prows, pp.rowptr, pp.col, pp.val,
                                                                       xc + 3 * nxc, 0, u + 3 * nu, 0, a, nullptr,
                                                                       0, 0, nullptr);
  PMG_HIP_CHECK(hipGetLastError());
}
template void launch_prolong_update<double>(const Csr &, const Csr &, const double *, long, double *, long, double,
                                            const double *, int, int, float *, hipStream_t, const int *, int,
                                            const int *, int);
template void launch_prolong_update<float>(const Csr &, const Csr &, const float *, long, float *, long, float,
                                           const float *, int, int, float *, hipStream_t, const int *, int,
                                           const int *, int);

}  // namespace pmg
