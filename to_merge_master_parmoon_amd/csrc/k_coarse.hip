// k_coarse.hip - block-form FP64 coarse solve of the mixed-precision forms (gfx950).
//
// TNSE_MGLevel4::SolveExact (src/FE/NSE_MGLevel4.C:5728-5846) solves the coarsest level
// directly.  With the shared velocity block (A11 = A22 = A33, no coupling blocks on the
// active rows, the Galerkin Oseen form) the reduced coarse system (Dirichlet rows and the
// pinned pressure row eliminated, see k_dense.hip) is
//
//     [ A  0  0  BT1 ] [x1]   [r1]
//     [ 0  A  0  BT2 ] [x2] = [r2]        A: m x m (active velocity dofs of one component)
//     [ 0  0  A  BT3 ] [x3]   [r3]        BTb: m x q, Bb: q x m (q free pressure dofs)
//     [ B1 B2 B3  0  ] [p ]   [rp]
//
// and block elimination gives, with Ainv = A^{-1}, Wb = Ainv BTb, S = -sum_b Bb Wb:
//     p  = S^{-1} rp - sum_b Qb rb,       Qb = S^{-1} Bb Ainv
//     xb = Ainv rb - Wb p.
// Set-up (once per operator, FP64, on the second stream): LU with partial pivoting of A and
// of S, each by one workgroup (panels of 16 columns in registers, row interchanges, the U
// block row, and the trailing update on the matrix cores: v_mfma_f64_16x16x4f64); the
// solves with many right-hand sides [I | BT1 BT2 BT3] and I by blocked substitution (one
// workgroup per 16 columns, diagonal blocks in one wavefront, updates on the matrix
// cores); the products S, Bb Ainv and S^{-1}(...) by an FP64 MFMA GEMM.  Each visit is then
// two matrix-vector phases over explicit FP64 data (Ainv, [-Q | S^{-1}] and W: 5.6 MB at
// the 4^3 level) instead of two chains of dependent triangular sweeps; Ainv and W are the
// columns of X = A^{-1} [I | BT1 BT2 BT3], read in place.
//
// Lane layout of v_mfma_f64_16x16x4f64 (D = A B + C, A 16 x 4, B 4 x 16): lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16] and holds D[4 i + l / 16][l % 16] in its i-th
// result register.
#include "pmg_internal.hpp"
#include "pmg_rows.hpp"
#include "coarse_rows.hpp"

namespace pmg {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ inline d4 mfma16(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

// (bv, bi) := the larger value, the lower index among equals (the pivot rule of dgetf2)
__device__ inline void amax_take(double &bv, int &bi, double ov, int oi) {
  const bool take = ov > bv || (ov == bv && oi < bi);
  bv = take ? ov : bv;
  bi = take ? oi : bi;
}
// one step of the reduction with the partner lane that the DPP pattern CTRL selects
template <int CTRL>
__device__ inline void amax_dpp(double &bv, int &bi) {
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(bv), CTRL, 0xF, 0xF, false);
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(bv), CTRL, 0xF, 0xF, false);
  const int oi = __builtin_amdgcn_mov_dpp(bi, CTRL, 0xF, 0xF, false);
  amax_take(bv, bi, __hiloint2double(hi, lo), oi);
}
// the wavefront's largest value and its lowest index in every lane (all 64 lanes active):
// within each row of 16 lanes by DPP (quad xor 1, quad xor 2, half-row mirror, row
// mirror), then the four rows' results by readlane, taken in lane order
__device__ inline void wave_argmax(double &bv, int &bi) {
  amax_dpp<0xB1>(bv, bi);
  amax_dpp<0x4E>(bv, bi);
  amax_dpp<0x141>(bv, bi);
  amax_dpp<0x140>(bv, bi);
  const int h = __double2hiint(bv), l = __double2loint(bv);
  double v = __hiloint2double(__builtin_amdgcn_readlane(h, 0), __builtin_amdgcn_readlane(l, 0));
  int i = __builtin_amdgcn_readlane(bi, 0);
#pragma unroll
  for (int r = 1; r < 4; ++r)
    amax_take(v, i, __hiloint2double(__builtin_amdgcn_readlane(h, 16 * r), __builtin_amdgcn_readlane(l, 16 * r)),
              __builtin_amdgcn_readlane(bi, 16 * r));
  bv = v;
  bi = i;
}

constexpr int LU_NB = 16;          // panel width
constexpr int LU_T = 512;          // threads of the single-workgroup LU (one panel row each)
constexpr int TRSM_T = 256;        // threads of the substitution kernel
constexpr int TRSM_NR = 16;        // right-hand sides per workgroup

// ------------------------------------------------------------------ dense blocks
// A (m x m), R = [I | BT1 BT2 BT3] (m x (m + 3q)), B = [B1 B2 B3] (q x 3m), all row-major,
// zero-filled beforehand.  Reduced numbering as in k_dense.hip: active velocity dof r,
// free pressure dof i - pin.
__global__ __launch_bounds__(256) void k_cb_scatter(OpView<double> op, double *__restrict__ A, double *__restrict__ R,
                                                    double *__restrict__ B, int m, int q, int pin) {
  const int lane = threadIdx.x & 15;
  const int r = (blockIdx.x * 256 + threadIdx.x) >> 4;   // 16 lanes per operator row
  const int N = op.n_udof, NA = op.n_active, NP = op.n_pdof;
  const long ldr = long(m) + 3L * q;
  if (r < NA) {
    for (int k = op.a_rowptr[r] + lane; k < op.a_rowptr[r + 1]; k += 16) {
      const int c = op.a_kcol[k];
      if (c < NA) A[long(r) * m + c] = op.a[k];   // block A11 (= A22 = A33)
    }
    if (lane == 0) R[long(r) * ldr + r] = 1.0;
    for (int k = op.bt_rowptr[r] + lane; k < op.bt_rowptr[r + 1]; k += 16) {
      const int i = op.bt_kcol[k];
      if (pin && i == 0) continue;
      for (int b = 0; b < 3; ++b) R[long(r) * ldr + m + long(b) * q + (i - pin)] = op.bt[b * op.bt_nnz + k];
    }
  } else if (r >= N && r < N + NP) {
    const int i = r - N;
    if (pin && i == 0) return;
    for (int k = op.b_rowptr[i] + lane; k < op.b_rowptr[i + 1]; k += 16) {
      const int c = op.b_kcol[k];
      if (c >= NA) continue;
      for (int b = 0; b < 3; ++b) B[long(i - pin) * (3L * m) + long(b) * m + c] = op.b[b * op.b_nnz + k];
    }
  }
}

__global__ void k_cb_zero(double *x, long n) {
  for (long i = blockIdx.x * long(blockDim.x) + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x) x[i] = 0.0;
}

// ------------------------------------------------------------------ single-workgroup LU
// a (n x n, row-major, leading dimension ld) := L U with the row interchanges of dgetf2
// (pivot = first row of largest |a(i, j)|, i >= j); sigma[i] = original row now at row i;
// *info = 1 + the first column with a zero pivot (0: none).  n <= LU_T.  The panel (rows
// j0.., 16 columns) is factored in LDS, one row per thread.
__global__ __launch_bounds__(LU_T) void k_lu_wg(double *__restrict__ a, int n, int ld, int *__restrict__ sigma,
                                                 int *__restrict__ info) {
  __shared__ double Lp[LU_T][LU_NB + 1];     // the panel
  __shared__ double Us[LU_NB][LU_T];         // the U block row of the panel (columns j0 + 16 ..)
  __shared__ __align__(16) double wv[2][LU_T / 64];
  __shared__ __align__(16) int wi[2][LU_T / 64];
  __shared__ double WB[2][LU_T / 64][LU_NB];   // every wavefront's candidate pivot row
  __shared__ double JB[2][LU_NB];              // row j of the current step
  __shared__ int pivs[LU_NB];
  __shared__ int sig[LU_T];
  __shared__ int rows_s[2 * LU_NB], src_s[2 * LU_NB];
  __shared__ int nk_s, info_s;
  // the whole register file of its SIMDs (2 x 256 VGPRs): no wavefront of a concurrent
  // kernel shares this CU's issue slots with the pivot chain
  asm volatile("" ::: "v255");
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  if (t < n) sig[t] = t;
  if (t == 0) info_s = 0;
  for (int j0 = 0; j0 < n; j0 += LU_NB) {
    const int w = min(LU_NB, n - j0);
    const int r = j0 + t;
    const bool own = r < n;
    // the panel row of this thread in registers; per pivot step one barrier: the winning
    // row of every wavefront and row j publish themselves before it, so after it every
    // thread has the pivot row and the row it swaps with (no in-place exchange in LDS)
    double rv[LU_NB];
#pragma unroll
    for (int cc = 0; cc < LU_NB; ++cc) rv[cc] = (own && cc < w) ? a[long(r) * ld + j0 + cc] : 0.0;
#pragma unroll
    for (int jj = 0; jj < LU_NB; ++jj) {
      if (jj >= w) continue;   // (not break: a constant trip count lets the loop unroll fully)
      const int j = j0 + jj, par = jj & 1;
      double bv = (own && r >= j) ? fabs(rv[jj]) : -1.0;
      int bi = (own && r >= j) ? r : n;
      wave_argmax(bv, bi);
      if (lane == 0) {
        wv[par][wave] = bv;
        wi[par][wave] = bi;
      }
      if (own && r == bi)   // this wavefront's candidate pivot row
#pragma unroll
        for (int cc = 0; cc < LU_NB; ++cc) WB[par][wave][cc] = rv[cc];
      if (t == jj)          // row j, which the pivot row swaps with
#pragma unroll
        for (int cc = 0; cc < LU_NB; ++cc) JB[par][cc] = rv[cc];
      __syncthreads();
      {
        // every wavefront's candidate loaded first, then taken in wavefront order
        double cv[LU_T / 64];
        int ci[LU_T / 64];
#pragma unroll
        for (int w2 = 0; w2 < LU_T / 64; ++w2) {
          cv[w2] = wv[par][w2];
          ci[w2] = wi[par][w2];
        }
        bv = cv[0];
        bi = ci[0];
#pragma unroll
        for (int w2 = 1; w2 < LU_T / 64; ++w2) amax_take(bv, bi, cv[w2], ci[w2]);
      }
      const int p = bi;
      if (t == 0) {
        pivs[jj] = p;
        if (bv == 0.0 && info_s == 0) info_s = j + 1;
      }
      if (bv == 0.0) continue;   // uniform: dgetf2 records the zero pivot and moves on
      const int wp = (p - j0) >> 6;
      double prow[LU_NB];
#pragma unroll
      for (int cc = 0; cc < LU_NB; ++cc) prow[cc] = WB[par][wp][cc];
      if (t == jj) {
#pragma unroll
        for (int cc = 0; cc < LU_NB; ++cc) rv[cc] = prow[cc];
      } else if (r == p) {
#pragma unroll
        for (int cc = 0; cc < LU_NB; ++cc) rv[cc] = JB[par][cc];
      }
      if (own && r > j) {
        const double l = rv[jj] / prow[jj];
        rv[jj] = l;
#pragma unroll
        for (int cc = 0; cc < LU_NB; ++cc)
          if (cc > jj) rv[cc] = fma(-l, prow[cc], rv[cc]);
      }
    }
#pragma unroll
    for (int cc = 0; cc < LU_NB; ++cc) Lp[t][cc] = rv[cc];
    __syncthreads();
    // the factored panel to memory
    if (own)
#pragma unroll
      for (int cc = 0; cc < LU_NB; ++cc)
        if (cc < w) a[long(r) * ld + j0 + cc] = Lp[t][cc];
    if (wave == 0) {
      // rows touched by the panel's interchanges (slots 0..w-1: j0..j0+w-1, then the pivot
      // rows below the panel in order of first use) and the source row whose original
      // values end in each after the sequential interchanges
      int row = lane < w ? j0 + lane : -1;
      int nk = w;
      for (int k = 0; k < w; ++k) {
        const int p = pivs[k];
        if (p >= j0 + w && __ballot(row == p) == 0) {
          if (lane == nk) row = p;
          ++nk;
        }
      }
      int src = row;
      for (int k = 0; k < w; ++k) {
        const int p = pivs[k];
        if (p == j0 + k) continue;
        const unsigned long long msk = __ballot(row == p);
        const int f = __ffsll((long long)msk) - 1;
        const int sk = __shfl(src, k, 64), sf = __shfl(src, f, 64);
        if (lane == k) src = sf;
        if (lane == f) src = sk;
      }
      if (lane < 2 * LU_NB) {
        rows_s[lane] = row;
        src_s[lane] = src;
      }
      if (lane == 0) nk_s = nk;
      const int sg = (lane < nk) ? sig[src] : 0;
      __builtin_amdgcn_wave_barrier();
      if (lane < nk && src != row) sig[row] = sg;
    }
    __syncthreads();
    // interchanges on every column outside the panel: the <= 32 rows involved, staged in
    // registers (a thread's loads all in flight together), every row read before any is
    // written; only the rows whose content changes are read and written
    {
      const int nk = nk_s;
      // two halves of <= 256 columns (n <= LU_T); thread t moves column (t % 256) of the
      // half and every second row from row t / 256
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int c = 256 * h + (t & 255), k0 = t >> 8;
        const bool col = c < n && (c < j0 || c >= j0 + w);
        double xv[LU_NB];
#pragma unroll
        for (int q = 0; q < LU_NB; ++q) {
          const int k = k0 + 2 * q;
          const bool mv = col && k < nk && src_s[k] != rows_s[k];
          xv[q] = mv ? a[long(src_s[k]) * ld + c] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < LU_NB; ++q) {
          const int k = k0 + 2 * q;
          const bool mv = col && k < nk && src_s[k] != rows_s[k];
          if (mv) a[long(rows_s[k]) * ld + c] = xv[q];
        }
      }
      __syncthreads();
    }
    const int j1 = j0 + w, nt = n - j1;
    if (nt <= 0) break;
    // U block row: U12 = L11^{-1} A12 (unit lower), columns j1.. (one column per thread,
    // nt < LU_T), solved in place in LDS
    if (t < nt) {
      const int c = j1 + t;
#pragma unroll
      for (int k = 0; k < LU_NB; ++k) Us[k][t] = k < w ? a[long(j0 + k) * ld + c] : 0.0;
#pragma unroll 1
      for (int k = 0; k < LU_NB - 1; ++k) {
        const double uk = Us[k][t];
        for (int i = k + 1; i < LU_NB; ++i) Us[i][t] = fma(-Lp[i][k], uk, Us[i][t]);
      }
#pragma unroll
      for (int k = 0; k < LU_NB; ++k)
        if (k < w) a[long(j0 + k) * ld + c] = Us[k][t];
    } else {
#pragma unroll
      for (int k = 0; k < LU_NB; ++k) Us[k][t] = 0.0;
    }
    __syncthreads();
    // trailing update A22 -= L21 U12 on 16 x 16 tiles, TG tiles per wavefront at a time so
    // that their loads are in flight together (a tile alone is one memory round trip)
    const int nb = (nt + 15) / 16;
    const int col = lane & 15, rq = lane >> 4;
    constexpr int TG = 4;
    for (int tile0 = wave * TG; tile0 < nb * nb; tile0 += (LU_T / 64) * TG) {
      d4 acc[TG];
#pragma unroll
      for (int g = 0; g < TG; ++g) {
        const int tile = tile0 + g;
        const int ti = tile / nb, tc = tile - ti * nb;
        const int i0 = j1 + ti * 16, c0 = j1 + tc * 16;
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
          const int i = i0 + 4 * qq + rq, c = c0 + col;
          acc[g][qq] = (tile < nb * nb && i < n && c < n) ? a[long(i) * ld + c] : 0.0;
        }
      }
#pragma unroll
      for (int g = 0; g < TG; ++g) {
        const int tile = tile0 + g;
        if (tile >= nb * nb) continue;
        const int ti = tile / nb, tc = tile - ti * nb;
        const int i0 = j1 + ti * 16, c0 = j1 + tc * 16;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const double av = -Lp[i0 - j0 + col][4 * s + rq];
          const double bv = Us[4 * s + rq][c0 - j1 + col];
          acc[g] = mfma16(av, bv, acc[g]);
        }
      }
#pragma unroll
      for (int g = 0; g < TG; ++g) {
        const int tile = tile0 + g;
        const int ti = tile / nb, tc = tile - ti * nb;
        const int i0 = j1 + ti * 16, c0 = j1 + tc * 16;
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
          const int i = i0 + 4 * qq + rq, c = c0 + col;
          if (tile < nb * nb && i < n && c < n) a[long(i) * ld + c] = acc[g][qq];
        }
      }
    }
    __syncthreads();
  }
  __syncthreads();
  if (t < n) sigma[t] = sig[t];
  if (t == 0) *info = info_s;
}

// ------------------------------------------------------------------ many right-hand sides
// X[:, c0 .. c0+15] = (L U)^{-1} P R[:, c0 ..] for the factors of k_lu_wg (sigma: the row
// permutation P), one workgroup per 16 columns; R == nullptr: R = I (nrhs = n).  The slab
// lives in LDS; per 16-row block the diagonal solve runs in one wavefront and the rows
// below (forward) or above (backward) are updated on the matrix cores.  n <= 512.
__global__ __launch_bounds__(TRSM_T) void k_trsm_wg(const double *__restrict__ lu, int n, int ld,
                                                     const int *__restrict__ sigma, const double *__restrict__ R,
                                                     int ldr, int nrhs, double *__restrict__ X, int ldx) {
  __shared__ double xs[512][TRSM_NR + 1];
  __shared__ double Lb[512][TRSM_NR + 1];   // the factors' current 16-column block, all rows
  __shared__ double dg[16][17];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int c0 = blockIdx.x * TRSM_NR;
  for (int idx = t; idx < n * TRSM_NR; idx += TRSM_T) {
    const int i = idx / TRSM_NR, cc = idx - i * TRSM_NR, c = c0 + cc;
    double v = 0.0;
    if (c < nrhs) v = R ? R[long(sigma[i]) * ldr + c] : (sigma[i] == c ? 1.0 : 0.0);
    xs[i][cc] = v;
  }
  const int nb = (n + 15) / 16;
  const int col = lane & 15, rq = lane >> 4;
  // the block's rows [r0, r1) and columns k0 .. k0 + 15 of the factors into LDS at once (one
  // memory round trip per block instead of one per 16 x 16 tile)
  auto stage = [&](int r0, int r1, int k0, int kw) {
    for (int idx = t; idx < (r1 - r0) * 16; idx += TRSM_T) {
      const int i = r0 + (idx >> 4), k = idx & 15;
      Lb[i][k] = k < kw ? lu[long(i) * ld + k0 + k] : 0.0;
    }
  };
  // forward: unit lower L
  for (int kb = 0; kb < nb; ++kb) {
    const int k0 = kb * 16, kw = min(16, n - k0);
    __syncthreads();
    stage(k0, n, k0, kw);
    __syncthreads();
    {
      const int i = t >> 4, k = t & 15;
      dg[i][k] = (i < kw && k < kw) ? Lb[k0 + i][k] : 0.0;
    }
    __syncthreads();
    if (t < TRSM_NR) {
      double x[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) x[i] = i < kw ? xs[k0 + i][t] : 0.0;
#pragma unroll
      for (int k = 0; k < 16; ++k) {
#pragma unroll
        for (int i = k + 1; i < 16; ++i) x[i] = fma(-dg[i][k], x[k], x[i]);
        __asm__ __volatile__("" ::: "memory");
      }
#pragma unroll
      for (int i = 0; i < 16; ++i)
        if (i < kw) xs[k0 + i][t] = x[i];
    }
    __syncthreads();
    for (int tb = kb + 1 + wave; tb < nb; tb += TRSM_T / 64) {
      const int i0 = tb * 16;
      d4 acc;
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) {
        const int i = i0 + 4 * qq + rq;
        acc[qq] = i < n ? xs[i][col] : 0.0;
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int i = i0 + col, k = k0 + 4 * s + rq;
        const double av = (i < n && 4 * s + rq < kw) ? -Lb[i][4 * s + rq] : 0.0;
        const double bv = (4 * s + rq < kw) ? xs[k][col] : 0.0;
        acc = mfma16(av, bv, acc);
      }
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) {
        const int i = i0 + 4 * qq + rq;
        if (i < n) xs[i][col] = acc[qq];
      }
    }
  }
  // backward: upper U with its diagonal
  for (int kb = nb - 1; kb >= 0; --kb) {
    const int k0 = kb * 16, kw = min(16, n - k0);
    __syncthreads();
    stage(0, k0 + kw, k0, kw);
    __syncthreads();
    {
      const int i = t >> 4, k = t & 15;
      dg[i][k] = (i < kw && k < kw) ? Lb[k0 + i][k] : 1.0;
    }
    __syncthreads();
    if (t < TRSM_NR) {
      double x[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) x[i] = i < kw ? xs[k0 + i][t] : 0.0;
#pragma unroll
      for (int k = 15; k >= 0; --k) {
        x[k] = x[k] / dg[k][k];
#pragma unroll
        for (int i = 0; i < k; ++i) x[i] = fma(-dg[i][k], x[k], x[i]);
        __asm__ __volatile__("" ::: "memory");
      }
#pragma unroll
      for (int i = 0; i < 16; ++i)
        if (i < kw) xs[k0 + i][t] = x[i];
    }
    __syncthreads();
    for (int tb = wave; tb < kb; tb += TRSM_T / 64) {
      const int i0 = tb * 16;
      d4 acc;
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) acc[qq] = xs[i0 + 4 * qq + rq][col];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int i = i0 + col, k = k0 + 4 * s + rq;
        const double av = (4 * s + rq < kw) ? -Lb[i][4 * s + rq] : 0.0;
        const double bv = (4 * s + rq < kw) ? xs[k][col] : 0.0;
        acc = mfma16(av, bv, acc);
      }
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) xs[i0 + 4 * qq + rq][col] = acc[qq];
    }
  }
  __syncthreads();
  for (int idx = t; idx < n * TRSM_NR; idx += TRSM_T) {
    const int i = idx / TRSM_NR, cc = idx - i * TRSM_NR, c = c0 + cc;
    if (c < nrhs) X[long(i) * ldx + c] = xs[i][cc];
  }
}

// ------------------------------------------------------------------ FP64 GEMM
// C = alpha A B + beta C (row-major; A M x K, B K x N); one wavefront per 16 x 16 tile of C,
// four per workgroup, K in steps of 4 on the matrix cores.
__global__ __launch_bounds__(256) void k_gemm_f64(int M, int N, int K, double alpha, const double *__restrict__ A,
                                                  int lda, const double *__restrict__ B, int ldb, double beta,
                                                  double *__restrict__ C, int ldc) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tn = (N + 15) / 16;
  const int tile = blockIdx.x * 4 + wave;
  if (tile >= ((M + 15) / 16) * tn) return;
  const int i0 = (tile / tn) * 16, c0 = (tile % tn) * 16;
  const int col = lane & 15, rq = lane >> 4;
  const int ia = i0 + col, cb = c0 + col;
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < K; k += 16) {
    double av[4], bv[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int kk = k + 4 * s + rq;
      av[s] = (ia < M && kk < K) ? A[long(ia) * lda + kk] : 0.0;
      bv[s] = (cb < N && kk < K) ? B[long(kk) * ldb + cb] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) acc = mfma16(av[s], bv[s], acc);
  }
#pragma unroll
  for (int qq = 0; qq < 4; ++qq) {
    const int i = i0 + 4 * qq + rq;
    if (i < M && cb < N) {
      double *cp = C + long(i) * ldc + cb;
      *cp = beta == 0.0 ? alpha * acc[qq] : fma(alpha, acc[qq], beta * *cp);
    }
  }
}

// ------------------------------------------------------------------ per-visit phases
// phase 1: y_b = Ainv r_b (b = 0..2, one wavefront per row r of Ainv) and
// p = [-Q | S^{-1}] r (one wavefront per pressure row)
__global__ __launch_bounds__(256) void k_cb_phase1(const double *__restrict__ Ainv, int lda,
                                                   const double *__restrict__ P1, const double *__restrict__ rr,
                                                   double *__restrict__ y, double *__restrict__ p, int m, int q) {
  cb_phase1_row(Ainv, lda, P1, rr, y, p, m, q, blockIdx.x * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}

// phase 2: x_b = y_b - W_b p into the level vector u (active velocity dofs, one wavefront per
// row of W), the Dirichlet velocity dofs (f_D / diagonal) and the pressure (0 at the pinned
// dof) by the workgroups past the velocity rows
__global__ __launch_bounds__(256) void k_cb_phase2(OpView<double> op, const double *__restrict__ W, int ldw,
                                                   const double *__restrict__ y, const double *__restrict__ p,
                                                   const double *__restrict__ f, double *__restrict__ u, int m,
                                                   int q, int pin, int vel_blocks) {
  auto put = [&](long j, double v) { u[j] = v; };
  if (int(blockIdx.x) < vel_blocks) {
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r < m) cb_phase2_vel(op, W, ldw, y, p, put, m, q, r, threadIdx.x & 63);
    return;
  }
  cb_phase2_rest(op, p, [&](long j) { return f[j]; }, put, pin, long(blockIdx.x - vel_blocks) * 256 + threadIdx.x);
}

// The same three steps for the block-form coarse systems of several problems in one launch
// each (blockIdx.y = problem): the right-hand side read from, and the solution written to,
// the merged level's vectors through each problem's coarse dof map.
template <typename T>
__global__ __launch_bounds__(256) void k_cbb_rhs(const CbJob *__restrict__ jobs, const T *__restrict__ f) {
  const CbJob &J = jobs[blockIdx.y];
  const int lane = threadIdx.x & 15;
  const int i = (blockIdx.x * 256 + threadIdx.x) >> 4;
  if (i >= 3 * J.m + J.q) return;
  const int *__restrict__ idx = J.cidx;
  const double v = coarse_rhs_row(J.op, [&](long j) { return double(f[idx[j]]); }, i, J.pin, lane);
  if (lane == 0) J.rr[i] = v;
}
__global__ __launch_bounds__(256) void k_cbb_phase1(const CbJob *__restrict__ jobs) {
  const CbJob &J = jobs[blockIdx.y];
  cb_phase1_row(J.X, J.m + 3 * J.q, J.P1, J.rr, J.y, J.p, J.m, J.q, blockIdx.x * 4 + (threadIdx.x >> 6),
                threadIdx.x & 63);
}
template <typename T>
__global__ __launch_bounds__(256) void k_cbb_phase2(const CbJob *__restrict__ jobs, const T *__restrict__ f,
                                                    T *__restrict__ u) {
  const CbJob &J = jobs[blockIdx.y];
  const int *__restrict__ idx = J.cidx;
  auto put = [&](long j, double v) { u[idx[j]] = T(v); };
  const int vel_blocks = (J.m + 3) / 4;
  if (int(blockIdx.x) < vel_blocks) {
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r < J.m) cb_phase2_vel(J.op, J.X + J.m, J.m + 3 * J.q, J.y, J.p, put, J.m, J.q, r, threadIdx.x & 63);
    return;
  }
  cb_phase2_rest(J.op, J.p, [&](long j) { return double(f[idx[j]]); }, put, J.pin,
                 long(blockIdx.x - vel_blocks) * 256 + threadIdx.x);
}

}  // namespace

bool coarse_block_fits(int m, int q) { return m > 0 && q > 0 && m <= LU_T && q <= LU_T; }

void launch_cb_build(const DevLevel &L, CoarseBlock &B, hipStream_t s) {
  const long m = B.m, q = B.q;
  k_cb_zero<<<1024, 256, 0, s>>>(B.A, m * m);
  PMG_HIP_CHECK(hipGetLastError());
  k_cb_zero<<<1024, 256, 0, s>>>(B.X, m * (m + 3 * q));   // R, overwritten by the solve
  PMG_HIP_CHECK(hipGetLastError());
  k_cb_zero<<<1024, 256, 0, s>>>(B.B, q * 3 * m);
  PMG_HIP_CHECK(hipGetLastError());
  const int rows = L.n_udof + L.n_pdof;
  k_cb_scatter<<<(rows * 16 + 255) / 256, 256, 0, s>>>(L.op64(), B.A, B.X, B.B, B.m, B.q, B.pin);
  PMG_HIP_CHECK(hipGetLastError());
}

namespace {
void gemm(int M, int N, int K, double alpha, const double *A, int lda, const double *B, int ldb, double beta,
          double *C, int ldc, hipStream_t s) {
  const long tiles = long((M + 15) / 16) * ((N + 15) / 16);
  k_gemm_f64<<<int((tiles + 3) / 4), 256, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  PMG_HIP_CHECK(hipGetLastError());
}
}  // namespace

void coarse_block_factor(CoarseBlock &B, hipStream_t s) {
  const int m = B.m, q = B.q;
  const int ldx = m + 3 * q;          // X = [Ainv | W1 W2 W3]
  const int ldp = 3 * m + q;          // P1 = [-Q1 -Q2 -Q3 | S^{-1}]
  // A = P^T L U; X = A^{-1} [I | BT1 BT2 BT3] (in place: the solve reads its slab first)
  k_lu_wg<<<1, LU_T, 0, s>>>(B.A, m, m, B.sigA, B.info);
  PMG_HIP_CHECK(hipGetLastError());
  k_trsm_wg<<<(ldx + TRSM_NR - 1) / TRSM_NR, TRSM_T, 0, s>>>(B.A, m, m, B.sigA, B.X, ldx, ldx, B.X, ldx);
  PMG_HIP_CHECK(hipGetLastError());
  // S = -sum_b B_b W_b
  for (int b = 0; b < 3; ++b)
    gemm(q, q, m, -1.0, B.B + long(b) * m, 3 * m, B.X + m + long(b) * q, ldx, b == 0 ? 0.0 : 1.0, B.S, q, s);
  // V_b = B_b Ainv
  for (int b = 0; b < 3; ++b) gemm(q, m, m, 1.0, B.B + long(b) * m, 3 * m, B.X, ldx, 0.0, B.V + long(b) * m, 3 * m, s);
  // S = P^T L U; S^{-1} into the last q columns of P1; -Q = -S^{-1} V into the first 3m
  k_lu_wg<<<1, LU_T, 0, s>>>(B.S, q, q, B.sigS, B.info + 1);
  PMG_HIP_CHECK(hipGetLastError());
  k_trsm_wg<<<(q + TRSM_NR - 1) / TRSM_NR, TRSM_T, 0, s>>>(B.S, q, q, B.sigS, nullptr, 0, q, B.P1 + 3L * m, ldp);
  PMG_HIP_CHECK(hipGetLastError());
  gemm(q, 3 * m, q, -1.0, B.P1 + 3L * m, ldp, B.V, 3 * m, 0.0, B.P1, ldp, s);
}

void coarse_block_solve(CoarseBlock &B, const OpView<double> &op, const double *rr, const double *f, double *u,
                        hipStream_t s) {
  const int m = B.m, q = B.q;
  k_cb_phase1<<<(m + q + 3) / 4, 256, 0, s>>>(B.X, m + 3 * q, B.P1, rr, B.y, B.p, m, q);
  PMG_HIP_CHECK(hipGetLastError());
  const int vel_blocks = (m + 3) / 4;
  const long rest = 3L * (op.n_udof - op.n_active) + op.n_pdof;
  k_cb_phase2<<<vel_blocks + int((rest + 255) / 256), 256, 0, s>>>(op, B.X + m, m + 3 * q, B.y, B.p, f, u, m, q,
                                                                    B.pin, vel_blocks);
  PMG_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_coarse_block_batch(const CbJob *jobs, int n_jobs, const CbGrid &g, const T *f, T *u, hipStream_t s) {
  if (n_jobs <= 0) return;
  k_cbb_rhs<T><<<dim3(g.rhs, n_jobs), 256, 0, s>>>(jobs, f);
  PMG_HIP_CHECK(hipGetLastError());
  k_cbb_phase1<<<dim3(g.phase1, n_jobs), 256, 0, s>>>(jobs);
  PMG_HIP_CHECK(hipGetLastError());
  k_cbb_phase2<T><<<dim3(g.phase2, n_jobs), 256, 0, s>>>(jobs, f, u);
  PMG_HIP_CHECK(hipGetLastError());
}
template void launch_coarse_block_batch<float>(const CbJob *, int, const CbGrid &, const float *, float *,
                                               hipStream_t);
template void launch_coarse_block_batch<double>(const CbJob *, int, const CbGrid &, const double *, double *,
                                                hipStream_t);

CbGrid coarse_block_grid(const CbJob &J) {
  const long rest = 3L * (J.op.n_udof - J.op.n_active) + J.op.n_pdof;
  return CbGrid{(3 * J.m + J.q + 15) / 16, (J.m + J.q + 3) / 4, (J.m + 3) / 4 + int((rest + 255) / 256)};
}

}  // namespace pmg
