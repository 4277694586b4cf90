// k_vanka.hip - colored cell-Vanka smoother for Q2/P1disc hexahedra on gfx950.
//
// Restates TNSE_MGLevel4::CellVanka, colored variant (src/FE/NSE_MGLevel4.C:603-1649).
// The reference re-assembles every 85x85 patch matrix by linear search and calls
// dgetrf/dgetrs for it in every sweep; the matrices depend only on the operator values,
// so here they are assembled and LU-factored once per operator (k_patch_lu) and every
// sweep only applies the stored factors:
//   k_vanka_res   : per color, local residuals of all patches from the pre-color state
//                   (phase 1, :819-1064), written in pivot-permuted order;
//   k_vanka_solve : per color, one wavefront per patch: unit-lower and upper column sweeps
//                   streaming the factors from HBM, then u += damp * delta (phase 3,
//                   :1400-1609) - race free because same-color patches share no dof.
// The factorization and the sweeps perform the same fma/div sequence as the FP64 oracle
// (dgetf2 pivoting, reciprocal scaling, column-oriented dtrsm), so FP64 factors and
// local solves agree with it bit for bit.
#include "pmg_internal.hpp"
#include "pmg_rows.hpp"

namespace pmg {

typedef float v4f __attribute__((ext_vector_type(4)));   // MFMA f32 accumulator

namespace {


__device__ inline double bcast(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}
__device__ inline float bcast(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ inline double tfma(double a, double b, double c) { return fma(a, b, c); }
__device__ inline float tfma(float a, float b, float c) { return fmaf(a, b, c); }

template <typename T, typename F>
struct Ld;
template <typename T>
struct Ld<T, T> {
  static __device__ inline T get(const T *p, int i) { return p[i]; }
};
template <>
struct Ld<float, double> {
  static __device__ inline float get(const double *p, int i) { return float(p[i]); }
};
template <>
struct Ld<float, uint16_t> {
  static __device__ inline float get(const uint16_t *p, int i) {
    return __uint_as_float(uint32_t(p[i]) << 16);
  }
};

// One workgroup per patch position (cells in color order).  The local matrix is the
// System[col*85+row] of NSE_MGLevel4.C:1066-1290 (velocity rows j, j+27, j+54, pressure
// rows 81..84; Dirichlet rows keep only the diagonal blocks and no BT columns).  Its
// entries are gathered straight into the register tiles through the per-cell map of
// offsets into the row-blocked operator values (built once per level from the
// structure), then factored with dgetf2 semantics.
template <typename T>
__device__ __forceinline__ T patch_entry(const OpView<T> &op, const int *__restrict__ mp, int r, int c) {
  int mi = -1, li = 0, blk = 0;
  const T *base = op.av;
  bool dir_off = false;
  if (r < 3 * NU_LOC) {
    const int rr = r / NU_LOC, j = r - NU_LOC * rr;
    if (c < 3 * NU_LOC) {
      const int cc = c / NU_LOC, m = c - NU_LOC * cc;
      mi = PM_A + j * NU_LOC + m;
      li = PM_NA + j;
      blk = 3 * rr + cc;
      dir_off = rr != cc;
      if (op.a1) {
        // shared velocity block: A22 and A33 hold the values of A11, the coupling blocks
        // are zero (no gathers for them)
        if (rr != cc) mi = -1;
        blk = 0;
      }
    } else if (c < NLOC) {
      mi = PM_BT + j * NP_LOC + (c - 3 * NU_LOC);
      li = PM_MT + j;
      blk = rr;
      base = op.btv;
    }
  } else if (r < NLOC && c < 3 * NU_LOC) {
    const int pj = r - 3 * NU_LOC, cc = c / NU_LOC, m = c - NU_LOC * cc;
    mi = PM_B + pj * NU_LOC + m;
    li = PM_NB + pj;
    blk = cc;
    base = op.bv;
  }
  const int off = mi >= 0 ? mp[mi] : -1;
  int len = mp[li];
  bool valid = off >= 0;
  if (len < 0) {  // Dirichlet velocity row: A blocks r != c are dropped
    len = -len;
    valid = valid && !dir_off;
  }
  const T v = base[valid ? off + blk * len : 0];
  return valid ? v : T(0);
}

// factor entry in the storage precision (bf16: round to nearest even)
template <typename F>
struct StoreF {
  template <typename T>
  static __device__ inline F get(T v) { return F(v); }
};
template <>
struct StoreF<uint16_t> {
  static __device__ inline uint16_t get(float v) {
    const uint32_t u = __float_as_uint(v);
    return uint16_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
  }
};

// Step j = 16 JB + jl of dgetf2 on the register tiles, first half: pivot search in column
// j (first row of max |a(i, j)|, i >= j), then the swap of rows j and p over all columns.
// Returns whether the pivot is nonzero.
template <int JB, typename T>
__device__ __forceinline__ bool pivot_and_swap(T (&A)[6][6], int t, int tx, int ty, int jl, int *sigma, int &piv_s,
                                               int &nz_s, int &info_s) {
  constexpr int jb = JB;
  const int j = 16 * jb + jl;
  {
    // pivot: first row of max |a(i, j)|, i >= j
    if (tx == jl) {
      T bv = T(-1);
      int bi = NLOC;
#pragma unroll
      for (int a = jb; a < 6; ++a) {
        const int r = ty + 16 * a;
        const T v = fabs(A[a][jb]);
        if (r >= j && r < NLOC && v > bv) {
          bv = v;
          bi = r;
        }
      }
#pragma unroll
      for (int o = 8; o > 0; o >>= 1) {
        const T ov = __shfl_xor(bv, o, 16);
        const int oi = __shfl_xor(bi, o, 16);
        if (ov > bv || (ov == bv && oi < bi)) {
          bv = ov;
          bi = oi;
        }
      }
      if (ty == 0) {
        piv_s = bi;
        nz_s = bv != 0.0;
        if (bv == 0.0 && info_s == 0) info_s = j + 1;
      }
    }
    __syncthreads();
    const int p = piv_s;
    const bool nonzero = nz_s != 0;
    if (nonzero && p != j) {
      // swap rows j and p over all columns.  p >= j, so row p sits in slot pb >= jb of the
      // lanes ty == p % 16.  Each slot is handled in its own branch (pb is uniform) so
      // that every register access stays static.
      const int pl = p & 15, pb = p >> 4;
      if (t == 0) {
        const int s0 = sigma[j];
        sigma[j] = sigma[p];
        sigma[p] = s0;
      }
      const int src = ty == jl ? pl : (ty == pl ? jl : ty);
#pragma unroll
      for (int a = jb; a < 6; ++a) {
        if (a == pb) {
          if (pl != jl) {
#pragma unroll
            for (int b = 0; b < 6; ++b) {
              const T send = ty == jl ? A[jb][b] : A[a][b];
              const T recv = __shfl(send, src, 16);
              if (ty == jl) A[jb][b] = recv;
              if (ty == pl) A[a][b] = recv;
            }
          } else if (ty == jl) {
#pragma unroll
            for (int b = 0; b < 6; ++b) {
              const T rowj = A[jb][b];
              A[jb][b] = A[a][b];
              A[a][b] = rowj;
            }
          }
        }
      }
    }
    return nonzero;
  }
}

// one block of 16 columns (jb = JB) of the register-tiled dgetf2 in k_patch_lu
template <int JB, typename T>
__device__ __forceinline__ void lu_columns(T (&A)[6][6], int t, int tx, int ty, int *sigma, T *lbuf, T *ubuf,
                                           int &piv_s, int &nz_s, int &info_s) {
  constexpr int jb = JB;
  const int jl_end = jb == 5 ? NLOC - 80 : 16;
  for (int jl = 0; jl < jl_end; ++jl) {
    const int j = 16 * jb + jl;
    const bool nonzero = pivot_and_swap<JB, T>(A, t, tx, ty, jl, sigma, piv_s, nz_s, info_s);
    // scale column j below the diagonal by the reciprocal of the pivot, publish l and u
    if (tx == jl) {
      if (nonzero) {
        const T rc = T(1) / __shfl(A[jb][jb], jl, 16);
#pragma unroll
        for (int a = jb; a < 6; ++a)
          if (ty + 16 * a > j) A[a][jb] *= rc;
      }
#pragma unroll
      for (int a = jb; a < 6; ++a) lbuf[ty + 16 * a] = A[a][jb];
    }
    if (ty == jl) {
#pragma unroll
      for (int b = jb; b < 6; ++b) ubuf[tx + 16 * b] = A[jb][b];
    }
    __syncthreads();
    T lv[6], uv[6];
#pragma unroll
    for (int a = jb; a < 6; ++a) lv[a] = lbuf[ty + 16 * a];
#pragma unroll
    for (int b = jb; b < 6; ++b) uv[b] = ubuf[tx + 16 * b];
#pragma unroll
    for (int a = jb; a < 6; ++a)
#pragma unroll
      for (int b = jb; b < 6; ++b)
        if (ty + 16 * a > j && tx + 16 * b > j) A[a][b] = fma(-lv[a], uv[b], A[a][b]);
  }
}

// One block of 16 columns of Gauss-Jordan inversion on the same register tiles (FP32,
// mixed precision): the pivot search and row swaps of dgetf2, then row j is scaled by the
// reciprocal of the pivot and eliminated from every other row, with column j taking the
// corresponding column of the inverse (in-place Gauss-Jordan).  After the last column the
// tiles hold (P A)^{-1} for the row permutation P of dgetf2, so x = A^{-1} r = M (P r)
// with the pivot-permuted local residual the residual kernel writes.  Column j and row j
// are published unscaled through LDS; every thread forms the reciprocal itself.
template <int JB>
__device__ __forceinline__ void gj_columns(float (&A)[6][6], int t, int tx, int ty, int *sigma, float *lbuf,
                                           float *ubuf, int &piv_s, int &nz_s, int &info_s) {
  constexpr int jb = JB;
  const int jl_end = jb == 5 ? NLOC - 80 : 16;
  for (int jl = 0; jl < jl_end; ++jl) {
    const int j = 16 * jb + jl;
    const bool nonzero = pivot_and_swap<JB, float>(A, t, tx, ty, jl, sigma, piv_s, nz_s, info_s);
    if (tx == jl) {
#pragma unroll
      for (int a = 0; a < 6; ++a) lbuf[ty + 16 * a] = A[a][jb];
    }
    if (ty == jl) {
#pragma unroll
      for (int b = 0; b < 6; ++b) ubuf[tx + 16 * b] = A[jb][b];
    }
    __syncthreads();
    const float rc = nonzero ? 1.0f / lbuf[j] : 0.0f;
    float lv[6], uv[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) lv[a] = lbuf[ty + 16 * a];
#pragma unroll
    for (int b = 0; b < 6; ++b) {
      const int c = tx + 16 * b;
      uv[b] = c == j ? rc : ubuf[c] * rc;
    }
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int b = 0; b < 6; ++b) {
        const int r = ty + 16 * a, c = tx + 16 * b;
        A[a][b] = r == j ? uv[b] : fmaf(-lv[a], uv[b], c == j ? 0.0f : A[a][b]);
      }
  }
}

// ------------------------------------------------------------------ patch matrix + LU
// dgetf2 on a register-tiled copy: thread (tx, ty) = (t / 16, t % 16) holds rows ty + 16a
// and columns tx + 16b (a, b < 6; indices >= 85 are zero padding).  Every element receives
// the same operation sequence as in the column-sweep dgetf2 (swap, reciprocal scaling,
// fma(-l, u, a) in step order), so the factors are bit-identical.  Column j lives in the
// 16 lanes tx = j % 16 of one wave (pivot search and scaling by shuffles), row j in the
// lanes ty = j % 16 of every 16-lane group (row swaps by shuffles); l and u are published
// through LDS, two barriers per column.
// T: arithmetic of the factorization (FP64 for the plain path, FP32 for the mixed path,
// gathering from the FP32 operator copy), F: storage of the factors (T, or bf16).
template <typename T, typename F>
__global__ __launch_bounds__(256) void k_patch_lu(OpView<T> op, const int *__restrict__ cells,
                                                  const int *__restrict__ pmap, int ncells, F *__restrict__ lu,
                                                  uint8_t *__restrict__ invperm, int *__restrict__ info) {
  __shared__ int sigma[NLOC];
  __shared__ T lbuf[96], ubuf[96];
  __shared__ int piv_s, info_s, nz_s;
  const int pos = blockIdx.x;
  if (pos >= ncells) return;
  const int t = threadIdx.x;
  const int tx = t >> 4, ty = t & 15;
  const int *__restrict__ mp = pmap + long(cells[pos]) * PM_STRIDE;
  if (t < NLOC) sigma[t] = t;
  if (t == 0) info_s = 0;
  T A[6][6];
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int b = 0; b < 6; ++b) A[a][b] = patch_entry<T>(op, mp, ty + 16 * a, tx + 16 * b);
  __syncthreads();
  lu_columns<0>(A, t, tx, ty, sigma, lbuf, ubuf, piv_s, nz_s, info_s);
  lu_columns<1>(A, t, tx, ty, sigma, lbuf, ubuf, piv_s, nz_s, info_s);
  lu_columns<2>(A, t, tx, ty, sigma, lbuf, ubuf, piv_s, nz_s, info_s);
  lu_columns<3>(A, t, tx, ty, sigma, lbuf, ubuf, piv_s, nz_s, info_s);
  lu_columns<4>(A, t, tx, ty, sigma, lbuf, ubuf, piv_s, nz_s, info_s);
  lu_columns<5>(A, t, tx, ty, sigma, lbuf, ubuf, piv_s, nz_s, info_s);
  // packed store straight from the tiles: strictly lower part column by column at
  // [0, LU_L), upper part (with the diagonal) column by column at [LU_UOFF, LU_UOFF + LU_U)
  F *out = lu + long(pos) * LU_STRIDE;
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int b = 0; b < 6; ++b) {
      const int r = ty + 16 * a, c = tx + 16 * b;
      if (r < NLOC && c < NLOC) {
        if (r > c)
          out[lu_loff(c) + r - c - 1] = StoreF<F>::get(A[a][b]);
        else
          out[LU_UOFF + lu_uoff(c) + r] = StoreF<F>::get(A[a][b]);
      }
    }
  if (t < NLOC) invperm[long(pos) * PERM_STRIDE + sigma[t]] = uint8_t(t);
  if (t == 0) info[pos] = info_s;
}

// Storage of an explicit patch inverse M (85 x MINV_LD, 85 columns used) for k_vanka_apply,
// where lane i owns rows i and 64 + i and loads them as 16-byte vectors of PER entries:
// vector q of row r < 64 sits at (q * 64 + r) * PER, of row r >= 64 at
// 64 * MINV_LD + (q * 21 + r - 64) * PER, so every vector load of a wavefront reads one
// contiguous run.
template <typename F>
__host__ __device__ constexpr int minv_off(int r, int c) {
  constexpr int PER = 16 / int(sizeof(F));
  return r < 64 ? ((c / PER) * 64 + r) * PER + c % PER
                : 64 * MINV_LD + ((c / PER) * (NLOC - 64) + r - 64) * PER + c % PER;
}

// Patch matrix and its explicit inverse (mixed precision): the gathers of k_patch_lu, FP32
// Gauss-Jordan with the pivots of dgetf2, M = (P A)^{-1} stored in the layout of minv_off
// (the 3 padding columns zero) through an LDS copy of 16 rows at a time.
// (6 resident wavefronts per SIMD instead of the 4 that 99 VGPRs allow, at the price of a
// small spill: measured faster, 5.0 -> 4.0 ms for the three levels, and it leaves the
// coarse LU beside it less time to wait)
template <typename F>
__global__ __launch_bounds__(256, 6) void k_patch_inv(OpView<float> op, const int *__restrict__ cells,
                                                   const int *__restrict__ pmap, int ncells, F *__restrict__ minv,
                                                   uint8_t *__restrict__ invperm, int *__restrict__ info,
                                                   float *__restrict__ kref) {
  __shared__ int sigma[NLOC];
  __shared__ float lbuf[96], ubuf[96];
  __shared__ float stage[16][97];
  __shared__ int piv_s, info_s, nz_s;
  const int pos = blockIdx.x;
  if (pos >= ncells) return;
  const int t = threadIdx.x;
  const int tx = t >> 4, ty = t & 15;
  const int *__restrict__ mp = pmap + long(cells[pos]) * PM_STRIDE;
  if (t < NLOC) sigma[t] = t;
  if (t == 0) info_s = 0;
  float A[6][6];
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int b = 0; b < 6; ++b) A[a][b] = patch_entry<float>(op, mp, ty + 16 * a, tx + 16 * b);
  if (kref) {
    // the FP32 patch matrix for the refinement steps, row-major 85 x 85
    float *kr = kref + long(pos) * KREF_FULL;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int b = 0; b < 6; ++b) {
        const int r = ty + 16 * a, cc = tx + 16 * b;
        if (r < NLOC && cc < NLOC) kr[r * NLOC + cc] = A[a][b];
      }
  }
  __syncthreads();
  gj_columns<0>(A, t, tx, ty, sigma, lbuf, ubuf, piv_s, nz_s, info_s);
  gj_columns<1>(A, t, tx, ty, sigma, lbuf, ubuf, piv_s, nz_s, info_s);
  gj_columns<2>(A, t, tx, ty, sigma, lbuf, ubuf, piv_s, nz_s, info_s);
  gj_columns<3>(A, t, tx, ty, sigma, lbuf, ubuf, piv_s, nz_s, info_s);
  gj_columns<4>(A, t, tx, ty, sigma, lbuf, ubuf, piv_s, nz_s, info_s);
  gj_columns<5>(A, t, tx, ty, sigma, lbuf, ubuf, piv_s, nz_s, info_s);
  F *out = minv + long(pos) * MINV_STRIDE;
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    __syncthreads();
#pragma unroll
    for (int b = 0; b < 6; ++b) stage[ty][tx + 16 * b] = A[a][b];
    __syncthreads();
    // rows 16a .. 16a + 15 (those below 85), MINV_LD entries each; consecutive threads take
    // consecutive entries of a vector, then the same vector of the next row
    constexpr int PER = 16 / int(sizeof(F));
    const int nr = a == 5 ? NLOC - 80 : 16;
    for (int e = t; e < nr * MINV_LD; e += 256) {
      const int q = e / (nr * PER), w = e - q * nr * PER;
      const int rr = w / PER, c = q * PER + (w - rr * PER);
      out[minv_off<F>(16 * a + rr, c)] = StoreF<F>::get(stage[rr][c]);
    }
  }
  if (t < NLOC) invperm[long(pos) * PERM_STRIDE + sigma[t]] = uint8_t(t);
  if (t == 0) info[pos] = info_s;
}

// Block form of a patch with the shared velocity block (mixed precision).  The patch matrix
// is [A 0 0 BT1; 0 A 0 BT2; 0 0 A BT3; B1 B2 B3 0] with one 27 x 27 block A (patch_entry
// gives exactly these entries: zero coupling blocks, Dirichlet rows with the diagonal only
// and no BT).  A is inverted by FP32 Gauss-Jordan with partial pivoting (the column swaps
// undone at the end, so A^{-1} itself), then W_c = A^{-1} BT_c, the Schur complement
// S = -sum_c B_c W_c, S^{-1} (4 x 4, pivoted) and the pressure rows of the patch inverse
// P1 = [-S^{-1} B_c A^{-1} | S^{-1}]; A^{-1}, W and P1 are stored in the factor precision
// (SCH_* layout), so a local solve is p = P1 r and x_c = A^{-1} r_c - W_c p.  The local
// residual is used in its natural order (invperm = identity).
//
// FULL: the same steps, then the full 85 x 85 inverse expanded from them in the storage of
// k_patch_inv (minv_off, invperm = identity): with T_c = W_c S^{-1},
//   M(c, c') = delta(c, c') A^{-1} + T_c V_c',  M(c, p) = -T_c,  M(p, c') = -S^{-1} V_c',
//   M(p, p) = S^{-1},
// so the levels whose local solves use the full inverse get it without an 85 x 85
// Gauss-Jordan (the shared velocity block only, no refinement-step matrices).
template <typename F, bool FULL = false>
__global__ __launch_bounds__(128) void k_patch_schur(OpView<float> op, const int *__restrict__ cells,
                                                     const int *__restrict__ pmap, int ncells, F *__restrict__ sch,
                                                     uint8_t *__restrict__ invperm, int *__restrict__ info,
                                                     float *__restrict__ kref) {
  __shared__ float A[NU_LOC][NU_LOC + 1];
  __shared__ float BT[3][NU_LOC][NP_LOC];
  __shared__ float Bm[3][NP_LOC][NU_LOC];
  __shared__ float W[3][NU_LOC][NP_LOC];
  __shared__ float S[NP_LOC][NP_LOC];
  __shared__ float colk[NU_LOC], rowk[NU_LOC];
  __shared__ int piv[NU_LOC];
  __shared__ int info_s;
  const int pos = blockIdx.x;
  if (pos >= ncells) return;
  const int t = threadIdx.x;
  const int *__restrict__ mp = pmap + long(cells[pos]) * PM_STRIDE;
  if (t == 0) info_s = 0;
  {
    // the entries patch_entry gives for this form (the shared block A: value av[off] or 0;
    // BT_c and B_c: block c of the row, off + c |len|), with every load of a thread issued
    // before any entry is stored: two memory round trips in all (the map, then the values)
    // instead of two per entry
    constexpr int EA = NU_LOC * NU_LOC, IA = (EA + 127) / 128;
    constexpr int EB = 3 * NU_LOC * NP_LOC, IB = (EB + 127) / 128;
    int oa[IA], ob[IB], lb[IB], oc[IB], lc[IB];
#pragma unroll
    for (int i = 0; i < IA; ++i) {
      const int e = t + 128 * i;
      oa[i] = e < EA ? mp[PM_A + e] : -1;
    }
#pragma unroll
    for (int i = 0; i < IB; ++i) {
      const int e = t + 128 * i;
      const int c = e / (NU_LOC * NP_LOC), r = e - c * NU_LOC * NP_LOC, j = r / NP_LOC, k = r - NP_LOC * j;
      const bool in = e < EB;
      ob[i] = in ? mp[PM_BT + j * NP_LOC + k] : -1;
      lb[i] = in ? mp[PM_MT + j] : 0;
      oc[i] = in ? mp[PM_B + k * NU_LOC + j] : -1;
      lc[i] = in ? mp[PM_NB + k] : 0;
    }
    float va[IA], vb[IB], vc[IB];
#pragma unroll
    for (int i = 0; i < IA; ++i) {
      const float v = op.av[oa[i] >= 0 ? oa[i] : 0];
      va[i] = oa[i] >= 0 ? v : 0.0f;
    }
#pragma unroll
    for (int i = 0; i < IB; ++i) {
      const int c = (t + 128 * i) / (NU_LOC * NP_LOC);
      const float x = op.btv[ob[i] >= 0 ? ob[i] + c * abs(lb[i]) : 0];
      const float y = op.bv[oc[i] >= 0 ? oc[i] + c * abs(lc[i]) : 0];
      vb[i] = ob[i] >= 0 ? x : 0.0f;
      vc[i] = oc[i] >= 0 ? y : 0.0f;
    }
#pragma unroll
    for (int i = 0; i < IA; ++i) {
      const int e = t + 128 * i, j = e / NU_LOC, m = e - NU_LOC * j;
      if (e < EA) A[j][m] = va[i];
    }
#pragma unroll
    for (int i = 0; i < IB; ++i) {
      const int e = t + 128 * i;
      const int c = e / (NU_LOC * NP_LOC), r = e - c * NU_LOC * NP_LOC, j = r / NP_LOC, k = r - NP_LOC * j;
      if (e < EB) {
        BT[c][j][k] = vb[i];
        Bm[c][k][j] = vc[i];
      }
    }
  }
  __syncthreads();
  if (!FULL && kref) {
    // the FP32 block form for the refinement steps: A row-major, BT_c[j][k], B_c[k][j]
    float *kr = kref + long(pos) * KREF_BLK;
    for (int e = t; e < NU_LOC * NU_LOC; e += 128) kr[e] = A[e / NU_LOC][e % NU_LOC];
    for (int e = t; e < 3 * NU_LOC * NP_LOC; e += 128) {
      const int c = e / (NU_LOC * NP_LOC), q = e - c * NU_LOC * NP_LOC;
      kr[KREF_BT + e] = BT[c][q / NP_LOC][q % NP_LOC];
      kr[KREF_B + e] = Bm[c][q / NU_LOC][q % NU_LOC];
    }
  }
  for (int k = 0; k < NU_LOC; ++k) {
    if (t < 64) {
      float bv = -1.0f;
      int bi = NU_LOC;
      if (t >= k && t < NU_LOC) {
        bv = fabsf(A[t][k]);
        bi = t;
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const float ov = __shfl_xor(bv, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ov > bv || (ov == bv && oi < bi)) {
          bv = ov;
          bi = oi;
        }
      }
      if (t == 0) {
        piv[k] = bi;
        if (bv == 0.0f && info_s == 0) info_s = k + 1;
      }
    }
    __syncthreads();
    const int p = piv[k];
    if (p != k && t < NU_LOC) {
      const float x = A[k][t];
      A[k][t] = A[p][t];
      A[p][t] = x;
    }
    __syncthreads();
    if (t < NU_LOC) {
      colk[t] = A[t][k];
      rowk[t] = A[k][t];
    }
    __syncthreads();
    const float pv = rowk[k];
    const float rc = pv != 0.0f ? 1.0f / pv : 0.0f;
    for (int e = t; e < NU_LOC * NU_LOC; e += 128) {
      const int i = e / NU_LOC, c = e - NU_LOC * i;
      const float uk = c == k ? rc : rowk[c] * rc;
      A[i][c] = i == k ? uk : fmaf(-colk[i], uk, c == k ? 0.0f : A[i][c]);
    }
    __syncthreads();
  }
  // (P A)^{-1} -> A^{-1}: undo the row interchanges as column swaps, last first
  for (int k = NU_LOC - 1; k >= 0; --k) {
    const int p = piv[k];
    if (p != k && t < NU_LOC) {
      const float x = A[t][k];
      A[t][k] = A[t][p];
      A[t][p] = x;
    }
    __syncthreads();
  }
  // W = A^{-1} [BT_1 BT_2 BT_3] (27 x 12) on the matrix cores: wavefront w computes rows
  // 16 w .. 16 w + 15 as one 16 x 16 tile of v_mfma_f32_16x16x4f32 (K = 27 in 7 steps of 4,
  // padded with zeros); lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16] of each
  // 16 x 4 / 4 x 16 step and holds D[4 (l / 16) + i][l % 16] in result i
  {
    const int lane = t & 63, wv = t >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kk = 0; kk < 7; ++kk) {
      const int row = 16 * wv + lr, m = 4 * kk + lq;
      const float av = (row < NU_LOC && m < NU_LOC) ? A[row][m] : 0.0f;
      const float bv = (m < NU_LOC && lr < 3 * NP_LOC) ? BT[lr / NP_LOC][m][lr % NP_LOC] : 0.0f;
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = 16 * wv + 4 * lq + i;
      if (row < NU_LOC && lr < 3 * NP_LOC) W[lr / NP_LOC][row][lr % NP_LOC] = acc[i];
    }
  }
  __syncthreads();
  // S = -[B_1 B_2 B_3] W (4 x 4, K = 81 in 21 steps of 4) on the matrix cores, wavefront 0
  if (t < 64) {
    const int lr = t & 15, lq = t >> 4;
    v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 3
    for (int kk = 0; kk < 21; ++kk) {
      const int q = 4 * kk + lq, c = q / NU_LOC, j = q - NU_LOC * c;
      const float av = (lr < NP_LOC && q < 3 * NU_LOC) ? Bm[c][lr][j] : 0.0f;
      const float bv = (lr < NP_LOC && q < 3 * NU_LOC) ? W[c][j][lr] : 0.0f;
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = 4 * lq + i;
      if (row < NP_LOC && lr < NP_LOC) S[row][lr] = -acc[i];
    }
  }
  __syncthreads();
  if (t == 0) {
    // 4 x 4 Gauss-Jordan with partial pivoting, column swaps undone
    int sp[NP_LOC];
    for (int k = 0; k < NP_LOC; ++k) {
      int p = k;
      for (int i = k + 1; i < NP_LOC; ++i)
        if (fabsf(S[i][k]) > fabsf(S[p][k])) p = i;
      sp[k] = p;
      if (S[p][k] == 0.0f && info_s == 0) info_s = NU_LOC + k + 1;
      if (p != k)
        for (int c = 0; c < NP_LOC; ++c) {
          const float x = S[k][c];
          S[k][c] = S[p][c];
          S[p][c] = x;
        }
      const float rc = S[k][k] != 0.0f ? 1.0f / S[k][k] : 0.0f;
      float rowv[NP_LOC], colv[NP_LOC];
      for (int c = 0; c < NP_LOC; ++c) {
        rowv[c] = c == k ? rc : S[k][c] * rc;
        colv[c] = S[c][k];
      }
      for (int i = 0; i < NP_LOC; ++i)
        for (int c = 0; c < NP_LOC; ++c) S[i][c] = i == k ? rowv[c] : fmaf(-colv[i], rowv[c], c == k ? 0.0f : S[i][c]);
    }
    for (int k = NP_LOC - 1; k >= 0; --k)
      if (sp[k] != k)
        for (int i = 0; i < NP_LOC; ++i) {
          const float x = S[i][k];
          S[i][k] = S[i][sp[k]];
          S[i][sp[k]] = x;
        }
  }
  __syncthreads();
  // V_c = B_c A^{-1} (4 x 27 each), kept in BT's storage (no longer needed) as BT[c][j][k]
  for (int e = t; e < 3 * NP_LOC * NU_LOC; e += 128) {
    const int c = e / (NP_LOC * NU_LOC), r = e - c * NP_LOC * NU_LOC, k = r / NU_LOC, j = r - NU_LOC * k;
    float acc = 0.0f;
    for (int m = 0; m < NU_LOC; ++m) acc = fmaf(Bm[c][k][m], A[m][j], acc);
    BT[c][j][k] = acc;
  }
  __syncthreads();
  if constexpr (FULL) {
    // T_c = W_c S^{-1} into W's storage (each thread reads its row of W before the barrier)
    float tv[2][NP_LOC];
    for (int h = 0; h < 2; ++h) {
      const int e = t + 128 * h;
      if (e < 3 * NU_LOC) {
        const int c = e / NU_LOC, i = e - NU_LOC * c;
        for (int l = 0; l < NP_LOC; ++l) {
          float acc = 0.0f;
          for (int m = 0; m < NP_LOC; ++m) acc = fmaf(W[c][i][m], S[m][l], acc);
          tv[h][l] = acc;
        }
      }
    }
    __syncthreads();
    for (int h = 0; h < 2; ++h) {
      const int e = t + 128 * h;
      if (e < 3 * NU_LOC)
        for (int l = 0; l < NP_LOC; ++l) W[e / NU_LOC][e % NU_LOC][l] = tv[h][l];
    }
    __syncthreads();
    constexpr int PER = 16 / int(sizeof(F));
    F *out = sch + long(pos) * MINV_STRIDE;
    for (int o = t; o < NLOC * MINV_LD; o += 128) {
      // (row, column) of storage entry o (the inverse of minv_off)
      int r, cc;
      if (o < 64 * MINV_LD) {
        const int v = o / PER;
        r = v % 64;
        cc = (v / 64) * PER + o % PER;
      } else {
        const int o2 = o - 64 * MINV_LD, v = o2 / PER;
        r = 64 + v % (NLOC - 64);
        cc = (v / (NLOC - 64)) * PER + o2 % PER;
      }
      float v = 0.0f;
      if (cc < NLOC) {
        if (r < 3 * NU_LOC) {
          const int ci = r / NU_LOC, i = r - NU_LOC * ci;
          if (cc < 3 * NU_LOC) {
            const int cj = cc / NU_LOC, j = cc - NU_LOC * cj;
            float acc = ci == cj ? A[i][j] : 0.0f;
            for (int l = 0; l < NP_LOC; ++l) acc = fmaf(W[ci][i][l], BT[cj][j][l], acc);
            v = acc;
          } else {
            v = -W[ci][i][cc - 3 * NU_LOC];
          }
        } else {
          const int k = r - 3 * NU_LOC;
          if (cc < 3 * NU_LOC) {
            const int cj = cc / NU_LOC, j = cc - NU_LOC * cj;
            float acc = 0.0f;
            for (int l = 0; l < NP_LOC; ++l) acc = fmaf(S[k][l], BT[cj][j][l], acc);
            v = -acc;
          } else {
            v = S[k][cc - 3 * NU_LOC];
          }
        }
      }
      out[o] = StoreF<F>::get(v);
    }
    if (t < NLOC) invperm[long(pos) * PERM_STRIDE + t] = uint8_t(t);
    if (t == 0) info[pos] = info_s;
    return;
  }
  F *out = sch + long(pos) * SCH_STRIDE;
  for (int e = t; e < NU_LOC * NU_LOC; e += 128) out[SCH_AINV + e] = StoreF<F>::get(A[e / NU_LOC][e % NU_LOC]);
  for (int e = t; e < 3 * NU_LOC * NP_LOC; e += 128) {
    const int c = e / (NU_LOC * NP_LOC), r = e - c * NU_LOC * NP_LOC;
    out[SCH_W + e] = StoreF<F>::get(W[c][r / NP_LOC][r % NP_LOC]);
  }
  // P1 row k: -S^{-1} V (entries c * 27 + j), then S^{-1} (entries 81 + l)
  for (int e = t; e < NP_LOC * NLOC; e += 128) {
    const int k = e / NLOC, q = e - NLOC * k;
    float v;
    if (q < 3 * NU_LOC) {
      const int c = q / NU_LOC, j = q - NU_LOC * c;
      float acc = 0.0f;
      for (int l = 0; l < NP_LOC; ++l) acc = fmaf(S[k][l], BT[c][j][l], acc);
      v = -acc;
    } else {
      v = S[k][q - 3 * NU_LOC];
    }
    out[SCH_P1 + e] = StoreF<F>::get(v);
  }
  for (int e = SCH_END + t; e < SCH_STRIDE; e += 128) out[e] = StoreF<F>::get(0.0f);
  if (t < NLOC) invperm[long(pos) * PERM_STRIDE + t] = uint8_t(t);
  if (t == 0) info[pos] = info_s;
}

// ------------------------------------------------------------------ per-color residual
// (slot, local row) items, 16 lanes each: rows 0..26 produce the three velocity
// components of local dof j, rows 27..30 the pressure rows (NSE_MGLevel4.C:902-1061).
// items[] lists slot*32 + row sorted by the global row and each block takes one contiguous
// run of it, so neighbouring rows share cache lines.  Every item carries a descriptor
// (global row, first A or B entry, first BT entry, A/B length | BT length << 16; BT length
// 0 on Dirichlet rows) built once by k_res_desc, and a group loads the next item's
// descriptor while it works on the current one: an item costs two dependent memory round
// trips (row entries, then the gathers of u), not four.
__global__ void k_res_desc(const int *__restrict__ a_rowptr, const int *__restrict__ bt_rowptr,
                           const int *__restrict__ b_rowptr, int n_active, const int *__restrict__ items,
                           const int *__restrict__ rows, long n, int *__restrict__ desc) {
  const long i = blockIdx.x * long(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const int r = items[i] & 31, g = rows[i];
  int k0, len, q0 = 0, m = 0;
  if (r < NU_LOC) {
    k0 = a_rowptr[g];
    len = a_rowptr[g + 1] - k0;
    q0 = bt_rowptr[g];
    if (g < n_active) m = bt_rowptr[g + 1] - q0;
  } else {
    k0 = b_rowptr[g];
    len = b_rowptr[g + 1] - k0;
  }
  pmg_vi4 d;
  d.x = g;
  d.y = k0;
  d.z = q0;
  d.w = int(unsigned(len) | (unsigned(m) << 16));
  __builtin_memcpy(desc + 4 * i, &d, 16);
}

template <typename T, bool SH>
__global__ __launch_bounds__(256) void k_vanka_res(OpView<T> op, long n_items, int pos0,
                                                   const uint8_t *__restrict__ invperm,
                                                   const int *__restrict__ items, const int *__restrict__ desc,
                                                   const int *__restrict__ xg, const T *__restrict__ u,
                                                   const float *__restrict__ ua, const T *__restrict__ f,
                                                   T *__restrict__ rloc, T *__restrict__ uloc) {
  const int lane = threadIdx.x & 15, grp = threadIdx.x >> 4;
  const int N = op.n_udof, NA = op.n_active;
  const T *__restrict__ u1 = u;
  const T *__restrict__ u2 = u + N;
  const T *__restrict__ u3 = u + 2 * N;
  const T *__restrict__ up = u + 3 * N;
  int i0, i1;
  if (xg) {
    // workgroups are dealt to the 8 XCDs round-robin: block b runs on XCD b % 8 and takes
    // its share of group b % 8 (the grid is a multiple of 8)
    const int x = blockIdx.x & 7, k = blockIdx.x >> 3, K = gridDim.x >> 3;
    const long a = xg[x], n = xg[x + 1] - a;
    i0 = int(a + n * k / K);
    i1 = int(a + n * (k + 1) / K);
  } else {
    chunk_of_block(n_items, i0, i1);
  }
  int i = i0 + grp;
  int it_n = 0;
  pmg_vi4 d_n = {0, 0, 0, 0};
  if (i < i1) {
    it_n = items[i];
    __builtin_memcpy(&d_n, desc + 4L * i, 16);
  }
  for (; i < i1; i += 16) {
    const int it = it_n;
    const pmg_vi4 d = d_n;
    if (i + 16 < i1) {
      it_n = items[i + 16];
      __builtin_memcpy(&d_n, desc + 4L * (i + 16), 16);
    }
    const int slot = it >> 5, r = it & 31;
    const int grow = d.x, k0 = d.y, q0 = d.z;
    const int n = int(unsigned(d.w) & 0xffffu), m = int(unsigned(d.w) >> 16);
    const int pos = pos0 + slot;
    const uint8_t *ip = invperm + long(pos) * PERM_STRIDE;
    T *out = rloc + long(slot) * RLOC_STRIDE;
    if (r < NU_LOC) {
      const int UDOF = grow;
      const T f1 = f[UDOF], f2 = f[N + UDOF], f3 = f[2 * N + UDOF];
      // u at the row's own dof, loaded beside f, for the color's local solve (no same-color
      // patch writes it before that solve runs)
      T c1 = 0, c2 = 0, c3 = 0;
      if (uloc && lane == 0) {
        if (sizeof(T) == 4 && ua) {
          const float4 v = reinterpret_cast<const float4 *>(ua)[UDOF];
          c1 = T(v.x);
          c2 = T(v.y);
          c3 = T(v.z);
        } else {
          c1 = u1[UDOF];
          c2 = u2[UDOF];
          c3 = u3[UDOF];
        }
      }
      T s1 = 0, s2 = 0, s3 = 0;
      const int *__restrict__ kc = op.a_kcol + k0;
      if constexpr (SH && sizeof(T) == 4) {
        // eight A-row entries per lane per round of loads: a whole vertex row (125 entries)
        // in one round, so one dependent round trip less than four per round (measured:
        // 8^3 5.13 -> 4.86 us, 16^3 6.99 -> 6.61 us, 32^3 32.3 -> 31.3 us per color)
        if (ua)
          row_a1_aos<8>(kc, op.a1 + k0, n, lane, ua, s1, s2, s3);
        else
          row_a1<T>(kc, op.a1 + k0, n, lane, u1, u2, u3, s1, s2, s3);
      } else if constexpr (SH)
        row_a1<T>(kc, op.a1 + k0, n, lane, u1, u2, u3, s1, s2, s3);
      else if (UDOF < NA)
        row_a9<T>(kc, op.av + 9L * k0, n, lane, u1, u2, u3, s1, s2, s3);
      else
        row_a_diag<T, 1>(kc, op.av + 9L * k0, n, lane, u1, u2, u3, s1, s2, s3);
      if (m > 0) row_bt<T>(op.bt_kcol + q0, op.btv + 3L * q0, m, lane, up, s1, s2, s3);
      s1 = red16(s1);
      s2 = red16(s2);
      s3 = red16(s3);
      if (lane == 0) {
        out[ip[r]] = f1 - s1;
        out[ip[NU_LOC + r]] = f2 - s2;
        out[ip[2 * NU_LOC + r]] = f3 - s3;
        if (uloc) {
          T *uo = uloc + long(slot) * RLOC_STRIDE;
          uo[r] = c1;
          uo[NU_LOC + r] = c2;
          uo[2 * NU_LOC + r] = c3;
        }
      }
    } else {
      const int mm = r - NU_LOC;
      const int PDOF = grow;
      const T fp = f[3 * N + PDOF];
      const T cp = uloc && lane == 0 ? up[PDOF] : T(0);
      T s = 0;
      if constexpr (sizeof(T) == 4) {
        if (ua)
          row_b_aos<2>(op.b_kcol + k0, op.bv + 3L * k0, n, lane, ua, s);
        else
          row_b<T>(op.b_kcol + k0, op.bv + 3L * k0, n, lane, u1, u2, u3, s);
      } else {
        row_b<T>(op.b_kcol + k0, op.bv + 3L * k0, n, lane, u1, u2, u3, s);
      }
      s = red16(s);
      if (lane == 0) {
        out[ip[3 * NU_LOC + mm]] = fp - s;
        if (uloc) uloc[long(slot) * RLOC_STRIDE + 3 * NU_LOC + mm] = cp;
      }
    }
  }
}

// ------------------------------------------------------------------ per-color local solves
// One wavefront per patch; lane i owns local rows i and i+64 (i < 21).  The sweeps are a
// chain of 170 dependent steps, so nothing in the chain may wait on HBM: the wave first
// issues the loads of the whole factor (L part and U part, 16-byte vectors), stores the L
// part into LDS and runs the unit-lower sweep from LDS while the U part is still arriving
// in registers, then stores the U part over it and runs the upper sweep.
template <typename F>
struct PackedSizes {
  static constexpr int NVL = int((LU_L * sizeof(F) + 15) / 16);   // 16-byte vectors of L
  static constexpr int NVU = int((LU_U * sizeof(F) + 15) / 16);   // 16-byte vectors of U
  static constexpr int RL = (NVL + 63) / 64, RU = (NVU + 63) / 64;
  // every lane loads RL (RU) vectors unconditionally; the overshoot past the part stays
  // inside the patch stride or the LU_PAD tail of the allocation
  // the sweeps read at constant per-step offsets from each lane's base and discard the
  // entries of rows outside the step's range; PADL front entries keep those indices >= 0
  static constexpr int PADL = 64;
  static constexpr int LDS_BYTES = PADL * int(sizeof(F)) + 16 * 64 * (RU > RL ? RU : RL);
  // the highest entry a sweep reads, U(84,84) + 127 - 84 = lu_uoff(84) + 127, is inside
  static_assert(lu_uoff(NLOC - 1) + 127 < 16 * 64 * RU / int(sizeof(F)), "sweep reads past the LDS copy");
  static_assert(LU_UOFF * sizeof(F) + 16 * 64 * RU <= (LU_STRIDE + LU_PAD) * sizeof(F), "pad too small");
};

typedef unsigned int v4u __attribute__((ext_vector_type(4)));   // register-promotable 16 B

// FP32 sweeps with every factor entry a lane needs held in registers before the dependent
// chain starts: the unit-lower sweep keeps row lane (L(lane, k), k < 64) and row 64+lane
// (L(64+lane, k)) of L, the upper sweep the rows of U, each as one register per column k
// (uniform index in the unrolled loops).  Entries outside the step's range are zero, so a
// step is one broadcast and two fmas with no select and no LDS access; in the upper sweep
// a lane's value is final once its own pivot step is passed (later steps only touch rows
// above the pivot), so the quotients x / U(r,r) are formed for all lanes after the chain.
template <typename T, typename F>
__device__ __forceinline__ void sweeps_in_registers(v4u *s4, const F *sl0, const F *sl1, const v4u *rU, int lane,
                                                    bool two, T &x0, T &x1) {
  using PS = PackedSizes<F>;
  {
    T L0[64], L1[NLOC - 1];
#pragma unroll
    for (int k = 0; k < NLOC - 1; ++k) {
      const int o = lu_loff(k) - k - 1;
      if (k < 64) {
        const T a = Ld<T, F>::get(sl0, o);
        L0[k] = lane > k ? a : T(0);
      }
      const T b = Ld<T, F>::get(sl1, o);
      L1[k] = (two && lane + 64 > k) ? b : T(0);
    }
    // pin every entry in a register here (an empty asm that "modifies" it), so that the
    // LDS reads stay ahead of the chain instead of being sunk into it
#pragma unroll
    for (int k = 0; k < NLOC - 1; ++k) {
      if (k < 64) asm volatile("" : "+v"(L0[k]));
      asm volatile("" : "+v"(L1[k]));
    }
#pragma unroll
    for (int k = 0; k < NLOC - 1; ++k) {
      if (k < 64) {
        const T bk = bcast(x0, k);
        x0 = tfma(-bk, L0[k], x0);
        x1 = tfma(-bk, L1[k], x1);
      } else {
        const T bk = bcast(x1, k - 64);
        x1 = tfma(-bk, L1[k], x1);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PS::RU; ++q) s4[lane + 64 * q] = rU[q];
  __syncthreads();
  const T dg0 = Ld<T, F>::get(reinterpret_cast<const F *>(s4), lu_uoff(lane) + lane);
  const T dg1 = two ? Ld<T, F>::get(reinterpret_cast<const F *>(s4), lu_uoff(64 + lane) + 64 + lane) : T(1);
  const T rc0 = T(1) / dg0, rc1 = T(1) / dg1;
  T U0[NLOC], U1[NLOC - 64];
#pragma unroll
  for (int k = 0; k < NLOC; ++k) {
    const int o = lu_uoff(k);
    const T a = Ld<T, F>::get(sl0, o);
    U0[k] = lane < k ? a : T(0);
    if (k >= 64) {
      const T b = Ld<T, F>::get(sl1, o);
      U1[k - 64] = (two && lane + 64 < k) ? b : T(0);
    }
  }
#pragma unroll
  for (int k = 0; k < NLOC; ++k) {
    asm volatile("" : "+v"(U0[k]));
    if (k >= 64) asm volatile("" : "+v"(U1[k - 64]));
  }
#pragma unroll
  for (int k = NLOC - 1; k >= 0; --k) {
    if (k >= 64) {
      const T xk = bcast(x1 * rc1, k - 64);
      x0 = tfma(-xk, U0[k], x0);
      x1 = tfma(-xk, U1[k - 64], x1);
    } else {
      const T xk = bcast(x0 * rc0, k);
      x0 = tfma(-xk, U0[k], x0);
    }
  }
  x0 = x0 * rc0;
  x1 = x1 * rc1;
}

template <typename T, typename F, bool REG>
__global__ __launch_bounds__(64) void k_vanka_solve(const F *__restrict__ lu, const int *__restrict__ cells,
                                                    int ncells, int pos0, const int *__restrict__ slots,
                                                    const T *__restrict__ rloc,
                                                    const int *__restrict__ u_global,
                                                    const int *__restrict__ p_global, int N,
                                                    T *__restrict__ u, T damp) {
  using PS = PackedSizes<F>;
  extern __shared__ v4u s4raw[];
  v4u *s4 = s4raw + PS::PADL * sizeof(F) / 16;  // factor part, after the front pad
  const int lane = threadIdx.x;
  // each lane reads its rows at constant per-step offsets from these bases
  const F *sl0 = reinterpret_cast<const F *>(s4) + lane;
  const F *sl1 = sl0 + 64;
  if (int(blockIdx.x) >= ncells) return;
  const int slot = slots ? slots[blockIdx.x] : int(blockIdx.x);
  const int pos = pos0 + slot;
  const bool two = lane < NLOC - 64;  // lanes owning a second row 64 + lane
  // offsets in u of this lane's rows, looked up ahead of the sweeps (off the chain's end)
  const int cell = cells[pos];
  int o0, o1 = 0;
  {
    const int c = lane / NU_LOC, j = lane - c * NU_LOC;   // lane < 64 < 81: velocity
    o0 = c * N + u_global[long(cell) * NU_LOC + j];
  }
  if (two) {
    const int i = 64 + lane;
    if (i < 3 * NU_LOC) {
      const int c = i / NU_LOC, j = i - c * NU_LOC;
      o1 = c * N + u_global[long(cell) * NU_LOC + j];
    } else {
      o1 = 3 * N + p_global[long(cell) * NP_LOC + (i - 3 * NU_LOC)];
    }
  }
  const v4u *__restrict__ gL = reinterpret_cast<const v4u *>(lu + long(pos) * LU_STRIDE);
  const v4u *__restrict__ gU = reinterpret_cast<const v4u *>(lu + long(pos) * LU_STRIDE + LU_UOFF);
  {
    v4u rL[PS::RL];
#pragma unroll
    for (int q = 0; q < PS::RL; ++q) rL[q] = gL[lane + 64 * q];
#pragma unroll
    for (int q = 0; q < PS::RL; ++q) s4[lane + 64 * q] = rL[q];
  }
  v4u rU[PS::RU];
#pragma unroll
  for (int q = 0; q < PS::RU; ++q) rU[q] = gU[lane + 64 * q];
  const T *__restrict__ rs = rloc + long(slot) * RLOC_STRIDE;
  T x0 = rs[lane];
  T x1 = two ? rs[64 + lane] : T(0);
  // the values to update, loaded before the sweeps (no round trip after the chain)
  const T v0 = u[o0];
  const T v1 = two ? u[o1] : T(0);
  __syncthreads();
  if constexpr (REG) {
    sweeps_in_registers<T, F>(s4, sl0, sl1, rU, lane, two, x0, x1);
  } else {
    // unit lower sweep (dtrsm L, column oriented); L(r,k) at lu_loff(k) + r - k - 1.  Row
    // 64+lane is below every k < 64, row lane is above every k >= 64, so per step only one
    // of the two rows needs the "below the pivot" test.
#pragma unroll
    for (int k = 0; k < NLOC - 1; ++k) {
      const int o = lu_loff(k) - k - 1;
      if (k < 64) {
        const T bk = bcast(x0, k);
        const T l0 = Ld<T, F>::get(sl0, o), l1 = Ld<T, F>::get(sl1, o);
        const T y0 = tfma(-bk, l0, x0);
        x0 = lane > k ? y0 : x0;
        x1 = tfma(-bk, l1, x1);  // lanes without a second row carry an unused value
      } else {
        const T bk = bcast(x1, k - 64);
        const T l1 = Ld<T, F>::get(sl1, o);
        const T y1 = tfma(-bk, l1, x1);
        x1 = lane + 64 > k ? y1 : x1;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PS::RU; ++q) s4[lane + 64 * q] = rU[q];
    __syncthreads();
    // upper sweep (dtrsm U): x_k /= U(k,k), then eliminate column k above the diagonal;
    // U(r,k) at lu_uoff(k) + r.  Row lane is above every k >= 64, row 64+lane below every
    // k < 64.
    // Diagonal of U and its reciprocal per lane, computed in parallel before the chain.  The
    // quotient x_k / U(k,k) is then q = x r, e = fma(-q, U, x), q' = fma(e, r, q): with r the
    // correctly rounded reciprocal this is the correctly rounded quotient (Markstein), i.e.
    // the same value as the division of the column-oriented dtrsm, in three dependent steps.
    const T dg0 = Ld<T, F>::get(reinterpret_cast<const F *>(s4), lu_uoff(lane) + lane);
    const T dg1 = two ? Ld<T, F>::get(reinterpret_cast<const F *>(s4), lu_uoff(64 + lane) + 64 + lane) : T(1);
    const T rc0 = T(1) / dg0, rc1 = T(1) / dg1;
#pragma unroll
    for (int k = NLOC - 1; k >= 0; --k) {
      const int o = lu_uoff(k);
      if (k >= 64) {
        T q1 = x1 * rc1;
        if constexpr (sizeof(T) == 8) q1 = tfma(tfma(-q1, dg1, x1), rc1, q1);
        const T xk = bcast(q1, k - 64);
        const T u0 = Ld<T, F>::get(sl0, o), u1 = Ld<T, F>::get(sl1, o);
        x0 = tfma(-xk, u0, x0);  // every row lane < 64 <= k lies above the pivot
        const T y1 = tfma(-xk, u1, x1);
        x1 = (lane + 64 == k) ? xk : (lane + 64 < k ? y1 : x1);
      } else {
        T q0 = x0 * rc0;
        if constexpr (sizeof(T) == 8) q0 = tfma(tfma(-q0, dg0, x0), rc0, q0);
        const T xk = bcast(q0, k);
        const T u0 = Ld<T, F>::get(sl0, o);
        const T y0 = tfma(-xk, u0, x0);
        x0 = (lane == k) ? xk : (lane < k ? y0 : x0);
      }
    }
  }
  // u[UDOF] += damp * delta, p[PDOF] += damp * delta
  u[o0] = v0 + damp * x0;
  if (two) u[o1] = v1 + damp * x1;
}

// x = M (P r) for one patch per wavefront, M the explicit inverse of k_patch_inv: lane i
// owns rows i and 64 + i (i < 21), loads them as 16-byte vectors up front (no dependent
// chain: every product is independent; non-temporal, so the stream of inverses does not
// evict the vectors the next residual gathers from L2), the local residual is broadcast
// from the lanes holding it, then u[dof] += damp * x.
// Local solve with the explicit inverse M = (P A)^{-1}, one wavefront per patch: lane i
// holds rows i and 64 + i of M, the permuted local residual P r is broadcast by lane reads,
// x = M (P r) = A^{-1} r in the natural order, u += damp x (the dofs from pos_dofs, their
// current values from uloc, as the color's residual launch read them).
// REF: nref refinement steps: rr = r - K x with the FP32 patch matrix K (stored at set-up,
// staged in LDS once), permuted like the residual (row i to position invperm[i]) and
// multiplied by M again, x += M (P rr).
template <typename F, bool REF>
__global__ __launch_bounds__(64) void k_vanka_apply(const F *__restrict__ minv, const int *__restrict__ pos_dofs,
                                                    int ncells, int pos0, const int *__restrict__ slots,
                                                    const float *__restrict__ rloc, const float *__restrict__ uloc,
                                                    int N, float *__restrict__ u, float *__restrict__ ua, float damp,
                                                    const float *__restrict__ kref,
                                                    const uint8_t *__restrict__ invperm, int nref) {
  constexpr int NV = MINV_LD * int(sizeof(F)) / 16;   // 16-byte vectors per row
  constexpr int PER = 16 / int(sizeof(F));             // entries per vector
  constexpr int KLD = NLOC;   // odd row stride: lane i reading row i is conflict-free
  __shared__ __align__(16) float Kf[REF ? KREF_FULL : 4];
  __shared__ float xs[REF ? NLOC : 1], rp[REF ? RLOC_STRIDE : 1];
  const int lane = threadIdx.x;
  if (int(blockIdx.x) >= ncells) return;
  const int slot = slots ? slots[blockIdx.x] : int(blockIdx.x);
  const int pos = pos0 + slot;
  const bool two = lane < NLOC - 64;
  const F *__restrict__ mp = minv + long(pos) * MINV_STRIDE;
  v4u m0[NV], m1[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q)
    m0[q] = __builtin_nontemporal_load(reinterpret_cast<const v4u *>(mp + minv_off<F>(lane, q * PER)));
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    m1[q] = v4u{0u, 0u, 0u, 0u};
    if (two) m1[q] = __builtin_nontemporal_load(reinterpret_cast<const v4u *>(mp + minv_off<F>(64 + lane, q * PER)));
  }
  const int *__restrict__ pd = pos_dofs + long(pos) * POS_DOFS_STRIDE;
  // offsets in u (and in ua, -1: a pressure row) of this lane's rows
  int o0, o1 = 0, a0i, a1i = -1;
  {
    const int c = lane / NU_LOC, j = lane - c * NU_LOC;
    const int g = pd[j];
    o0 = c * N + g;
    a0i = 4 * g + c;
  }
  if (two) {
    const int i = 64 + lane;
    if (i < 3 * NU_LOC) {
      const int c = i / NU_LOC, j = i - c * NU_LOC;
      const int g = pd[j];
      o1 = c * N + g;
      a1i = 4 * g + c;
    } else {
      o1 = 3 * N + pd[i - 2 * NU_LOC];
    }
  }
  const float *__restrict__ rs = rloc + long(slot) * RLOC_STRIDE;
  const float r0 = rs[lane];
  const float r1 = two ? rs[64 + lane] : 0.0f;
  // the values to update, as the residual launch read them (natural local order)
  const float *__restrict__ us = uloc + long(slot) * RLOC_STRIDE;
  const float v0 = us[lane];
  const float v1 = two ? us[64 + lane] : 0.0f;
  if constexpr (REF) {
    if (nref > 0) {
      // the stored matrix has the LDS layout: a straight 16-byte copy, 8 loads in flight per
      // lane before their stores
      const v4u *__restrict__ kr = reinterpret_cast<const v4u *>(kref + long(pos) * KREF_FULL);
      v4u *kl = reinterpret_cast<v4u *>(Kf);
      constexpr int NQ = KREF_FULL / 4, PL = (NQ + 63) / 64;
#pragma unroll
      for (int b0 = 0; b0 < PL; b0 += 8) {
        v4u tq[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int q = lane + 64 * (b0 + k);
          if (b0 + k < PL && q < NQ) tq[k] = __builtin_nontemporal_load(kr + q);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int q = lane + 64 * (b0 + k);
          if (b0 + k < PL && q < NQ) kl[q] = tq[k];
        }
      }
    }
  }
  // (acc0, acc1) = rows lane and 64 + lane of M times the vector held as (q0, q1)
  auto apply = [&](float q0, float q1, float &acc0, float &acc1) {
    acc0 = 0.0f;
    acc1 = 0.0f;
#pragma unroll
    for (int j = 0; j < MINV_LD; ++j) {
      const float rj = j < 64 ? bcast(q0, j) : bcast(q1, j - 64);   // lanes >= 21 hold q1 = 0
      const int q = j / PER, k = j % PER;
      const unsigned w0 = m0[q][sizeof(F) == 2 ? k / 2 : k], w1 = m1[q][sizeof(F) == 2 ? k / 2 : k];
      float a0, a1;
      if constexpr (sizeof(F) == 2) {
        a0 = __uint_as_float(k % 2 ? (w0 & 0xffff0000u) : (w0 << 16));
        a1 = __uint_as_float(k % 2 ? (w1 & 0xffff0000u) : (w1 << 16));
      } else {
        a0 = __uint_as_float(w0);
        a1 = __uint_as_float(w1);
      }
      acc0 = fmaf(a0, rj, acc0);
      acc1 = fmaf(a1, rj, acc1);
    }
  };
  float acc0, acc1;
  apply(r0, r1, acc0, acc1);
  if constexpr (REF) {
    if (nref > 0) {
      const uint8_t *__restrict__ ip = invperm + long(pos) * PERM_STRIDE;
      const int p0 = ip[lane], p1 = two ? ip[64 + lane] : 0;
      // the natural-order residual of this lane's rows: r_nat[i] = (P r)[invperm[i]]
      rp[lane] = r0;
      if (two) rp[64 + lane] = r1;
      __syncthreads();
      const float rn0 = rp[p0], rn1 = two ? rp[p1] : 0.0f;
      for (int it = 0; it < nref; ++it) {
        __syncthreads();
        xs[lane] = acc0;
        if (two) xs[64 + lane] = acc1;
        __syncthreads();
        float e0 = rn0, e1 = rn1;
        for (int m = 0; m < NLOC; ++m) {
          const float xm = xs[m];
          e0 = fmaf(-Kf[lane * KLD + m], xm, e0);
          if (two) e1 = fmaf(-Kf[(64 + lane) * KLD + m], xm, e1);
        }
        rp[p0] = e0;
        if (two) rp[p1] = e1;
        __syncthreads();
        float d0, d1;
        apply(rp[lane], two ? rp[64 + lane] : 0.0f, d0, d1);
        acc0 += d0;
        acc1 += d1;
      }
    }
  }
  const float n0 = v0 + damp * acc0;
  u[o0] = n0;
  if (ua) ua[a0i] = n0;
  if (two) {
    const float n1 = v1 + damp * acc1;
    u[o1] = n1;
    if (ua && a1i >= 0) ua[a1i] = n1;
  }
}

// Local solve in the block form, one wavefront per patch: the patch's 1408 entries are
// staged in LDS (non-temporal 16-byte loads), then p = P1 r (P1 = [-S^{-1} B_c A^{-1} |
// S^{-1}], 16 lanes per pressure row) beside y_c = A^{-1} r_c (lanes 0..26, three
// components), then x_c = y_c - W_c p, and u += damp * (x, p) (with the interleaved velocity
// copy when ua is given).
// REF: nref refinement steps after the first solve: the local residual r - K x with the
// FP32 patch matrix K (A, BT_c and B_c as the set-up gathered them, stored beside the
// inverse data), solved with the stored data and added to x.
// AO (ua only): the velocity is written to the interleaved copy ua alone (one 16-byte store
// per dof instead of three 4-byte ones into u and three into ua); the sweep copies ua back
// into u after its last color.
// Both forms take the patch dofs from pos_dofs and their current values from uloc (written
// by the color's residual launch), so no load of the update waits on a chain of lookups.
template <typename F, bool REF, bool AO>
__global__ __launch_bounds__(64) void k_vanka_apply_schur(const F *__restrict__ sch,
                                                          const int *__restrict__ pos_dofs, int ncells, int pos0,
                                                          const int *__restrict__ slots,
                                                          const float *__restrict__ rloc,
                                                          const float *__restrict__ uloc, int N,
                                                          float *__restrict__ u, float *__restrict__ ua, float damp,
                                                          const float *__restrict__ kref, int nref) {
  constexpr int PER = 16 / int(sizeof(F));
  constexpr int NVEC = SCH_STRIDE / PER;
  __shared__ float M[SCH_STRIDE];
  __shared__ float r[RLOC_STRIDE], pp[NP_LOC];
  __shared__ __align__(16) float Kf[REF ? KREF_BLK : 4];   // A, BT_c, B_c (FP32), the stored layout
  __shared__ float xs[REF ? NLOC : 1], rr[REF ? NLOC : 1];
  const int lane = threadIdx.x;
  if (int(blockIdx.x) >= ncells) return;
  const int slot = slots ? slots[blockIdx.x] : int(blockIdx.x);
  const int pos = pos0 + slot;
  const F *__restrict__ src = sch + long(pos) * SCH_STRIDE;
  for (int q = lane; q < NVEC; q += 64) {
    const v4u w = __builtin_nontemporal_load(reinterpret_cast<const v4u *>(src + q * PER));
    if constexpr (sizeof(F) == 2) {
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        M[q * PER + 2 * h] = __uint_as_float(w[h] << 16);
        M[q * PER + 2 * h + 1] = __uint_as_float(w[h] & 0xffff0000u);
      }
    } else {
#pragma unroll
      for (int h = 0; h < 4; ++h) M[q * PER + h] = __uint_as_float(w[h]);
    }
  }
  const float *__restrict__ rs = rloc + long(slot) * RLOC_STRIDE;
  r[lane] = rs[lane];
  if (lane < NLOC - 64) r[64 + lane] = rs[64 + lane];
  const int *__restrict__ pd = pos_dofs + long(pos) * POS_DOFS_STRIDE;
  const float *__restrict__ us = uloc + long(slot) * RLOC_STRIDE;
  constexpr int KA = 0, KBT = KREF_BT, KB = KREF_B;
  if constexpr (REF) {
    if (nref > 0) {
      // A row-major with the odd stride 27 (conflict-free row reads), BT_c[j][k], B_c[k][j]:
      // the stored layout, copied in 16-byte pieces
      const v4u *__restrict__ kr = reinterpret_cast<const v4u *>(kref + long(pos) * KREF_BLK);
      v4u *kl = reinterpret_cast<v4u *>(Kf);
      constexpr int NQ = KREF_BLK / 4, PL = (NQ + 63) / 64;
      v4u tq[PL];
#pragma unroll
      for (int k = 0; k < PL; ++k)
        if (lane + 64 * k < NQ) tq[k] = __builtin_nontemporal_load(kr + lane + 64 * k);
#pragma unroll
      for (int k = 0; k < PL; ++k)
        if (lane + 64 * k < NQ) kl[lane + 64 * k] = tq[k];
    }
  }
  // the dofs to update and their values as the residual launch read them
  int g0 = 0, op_ = 0;
  float v0 = 0.0f, v1 = 0.0f, v2 = 0.0f, vp = 0.0f;
  if (lane < NU_LOC) {
    g0 = pd[lane];
    v0 = us[lane];
    v1 = us[NU_LOC + lane];
    v2 = us[2 * NU_LOC + lane];
  } else if (lane < NU_LOC + NP_LOC) {
    op_ = 3 * N + pd[lane];
    vp = us[2 * NU_LOC + lane];
  }
  __syncthreads();
  // x = M^{-1} rhs with the stored data; lane j < 27 returns x_c[j] (c = 0..2) in x[0..2],
  // lanes 27..30 return p[lane - 27] in x[0].  Two phases: p = P1 rhs (16 lanes per pressure
  // row) beside y_c = A^{-1} rhs_c (lane j, the three components with one read of row j of
  // A^{-1}), then x_c = y_c - W_c p.
  auto solve = [&](const float *rhs, float (&x)[3]) {
    {
      const int k = lane >> 4, part = lane & 15;
      float acc = 0.0f;
      for (int q = part; q < NLOC; q += 16) acc = fmaf(M[SCH_P1 + k * NLOC + q], rhs[q], acc);
      acc += __shfl_xor(acc, 8, 16);
      acc += __shfl_xor(acc, 4, 16);
      acc += __shfl_xor(acc, 2, 16);
      acc += __shfl_xor(acc, 1, 16);
      if (part == 0) pp[k] = acc;
    }
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    if (lane < NU_LOC) {
      for (int m = 0; m < NU_LOC; ++m) {
        const float w = M[SCH_AINV + lane * NU_LOC + m];
        a0 = fmaf(w, rhs[m], a0);
        a1 = fmaf(w, rhs[NU_LOC + m], a1);
        a2 = fmaf(w, rhs[2 * NU_LOC + m], a2);
      }
    }
    __syncthreads();
    x[0] = x[1] = x[2] = 0.0f;
    if (lane < NU_LOC) {
      const float y[3] = {a0, a1, a2};
      for (int c = 0; c < 3; ++c) {
        float acc = y[c];
        for (int k = 0; k < NP_LOC; ++k) acc = fmaf(-M[SCH_W + (c * NU_LOC + lane) * NP_LOC + k], pp[k], acc);
        x[c] = acc;
      }
    } else if (lane < NU_LOC + NP_LOC) {
      x[0] = pp[lane - NU_LOC];
    }
  };
  float x[3];
  solve(r, x);
  if constexpr (REF) {
    for (int it = 0; it < nref; ++it) {
      // x to LDS (natural local order), then rr = r - K x
      if (lane < NU_LOC) {
        xs[lane] = x[0];
        xs[NU_LOC + lane] = x[1];
        xs[2 * NU_LOC + lane] = x[2];
      } else if (lane < NU_LOC + NP_LOC) {
        xs[3 * NU_LOC + lane - NU_LOC] = x[0];
      }
      __syncthreads();
      if (lane < NU_LOC) {
        float a0 = r[lane], a1 = r[NU_LOC + lane], a2 = r[2 * NU_LOC + lane];
        for (int m = 0; m < NU_LOC; ++m) {
          const float w = Kf[KA + lane * NU_LOC + m];
          a0 = fmaf(-w, xs[m], a0);
          a1 = fmaf(-w, xs[NU_LOC + m], a1);
          a2 = fmaf(-w, xs[2 * NU_LOC + m], a2);
        }
        for (int k = 0; k < NP_LOC; ++k) {
          const float pk = xs[3 * NU_LOC + k];
          a0 = fmaf(-Kf[KBT + (0 * NU_LOC + lane) * NP_LOC + k], pk, a0);
          a1 = fmaf(-Kf[KBT + (1 * NU_LOC + lane) * NP_LOC + k], pk, a1);
          a2 = fmaf(-Kf[KBT + (2 * NU_LOC + lane) * NP_LOC + k], pk, a2);
        }
        rr[lane] = a0;
        rr[NU_LOC + lane] = a1;
        rr[2 * NU_LOC + lane] = a2;
      } else if (lane < NU_LOC + NP_LOC) {
        const int k = lane - NU_LOC;
        float a = r[3 * NU_LOC + k];
        for (int c = 0; c < 3; ++c)
          for (int j = 0; j < NU_LOC; ++j) a = fmaf(-Kf[KB + (c * NP_LOC + k) * NU_LOC + j], xs[c * NU_LOC + j], a);
        rr[3 * NU_LOC + k] = a;
      }
      __syncthreads();
      float dx[3];
      solve(rr, dx);
      x[0] += dx[0];
      x[1] += dx[1];
      x[2] += dx[2];
    }
  }
  if (lane < NU_LOC) {
    const float n0 = v0 + damp * x[0], n1 = v1 + damp * x[1], n2 = v2 + damp * x[2];
    if constexpr (AO) {
      reinterpret_cast<float4 *>(ua)[g0] = make_float4(n0, n1, n2, 0.0f);
    } else {
      u[g0] = n0;
      u[N + g0] = n1;
      u[2 * N + g0] = n2;
      if (ua) {
        ua[4 * g0] = n0;
        ua[4 * g0 + 1] = n1;
        ua[4 * g0 + 2] = n2;
      }
    }
  } else if (lane < NU_LOC + NP_LOC) {
    u[op_] = vp + damp * x[0];
  }
}

}  // namespace

void launch_patch_lu(const DevLevel &L, int *info, hipStream_t s) {
  if (L.n_cells <= 0) return;
  if (L.lu)
    k_patch_lu<double, double><<<L.n_cells, 256, 0, s>>>(L.op64(), L.cell_reorder, L.pmap, L.n_cells, L.lu, L.invperm,
                                                         info);
  else if (L.lu32)
    k_patch_lu<float, float><<<L.n_cells, 256, 0, s>>>(L.op32(), L.cell_reorder, L.pmap, L.n_cells, L.lu32, L.invperm,
                                                       info);
  else
    k_patch_lu<float, uint16_t><<<L.n_cells, 256, 0, s>>>(L.op32(), L.cell_reorder, L.pmap, L.n_cells, L.lu16,
                                                          L.invperm, info);
  PMG_HIP_CHECK(hipGetLastError());
}

void launch_res_desc(const DevLevel &L, const int *items, const int *rows, long n, int *desc, hipStream_t s) {
  if (n <= 0) return;
  k_res_desc<<<int((n + 255) / 256), 256, 0, s>>>(L.a_rowptr, L.bt_rowptr, L.b_rowptr, L.n_active, items, rows, n,
                                                  desc);
  PMG_HIP_CHECK(hipGetLastError());
}

void launch_patch_inv(const DevLevel &L, int *info, hipStream_t s, float *kref) {
  if (L.n_cells <= 0) return;
  if (L.a_shared && !kref) {
    // the full inverse expanded from the block form (no 85 x 85 elimination)
    if (L.minv32)
      k_patch_schur<float, true><<<L.n_cells, 128, 0, s>>>(L.op32(), L.cell_reorder, L.pmap, L.n_cells, L.minv32,
                                                           L.invperm, info, nullptr);
    else
      k_patch_schur<uint16_t, true><<<L.n_cells, 128, 0, s>>>(L.op32(), L.cell_reorder, L.pmap, L.n_cells,
                                                              L.minv16, L.invperm, info, nullptr);
    PMG_HIP_CHECK(hipGetLastError());
    return;
  }
  if (L.minv32)
    k_patch_inv<float><<<L.n_cells, 256, 0, s>>>(L.op32(), L.cell_reorder, L.pmap, L.n_cells, L.minv32, L.invperm,
                                                 info, kref);
  else
    k_patch_inv<uint16_t><<<L.n_cells, 256, 0, s>>>(L.op32(), L.cell_reorder, L.pmap, L.n_cells, L.minv16,
                                                    L.invperm, info, kref);
  PMG_HIP_CHECK(hipGetLastError());
}

void launch_patch_schur(const DevLevel &L, int *info, hipStream_t s, float *kref) {
  if (L.n_cells <= 0) return;
  if (L.minv32)
    k_patch_schur<float><<<L.n_cells, 128, 0, s>>>(L.op32(), L.cell_reorder, L.pmap, L.n_cells, L.minv32, L.invperm,
                                                   info, kref);
  else
    k_patch_schur<uint16_t><<<L.n_cells, 128, 0, s>>>(L.op32(), L.cell_reorder, L.pmap, L.n_cells, L.minv16,
                                                      L.invperm, info, kref);
  PMG_HIP_CHECK(hipGetLastError());
}

template <typename F>
void launch_vanka_apply_schur(const F *sch, const int *pos_dofs, int ncells, int pos0, const int *slots,
                              const float *rloc, const float *uloc, int n_udof, float *u, float *ua, float damp,
                              const float *kref, int nref, hipStream_t s, bool ua_only) {
  if (ncells <= 0) return;
  if (ua_only && !ua) throw DeviceError("launch_vanka_apply_schur: ua_only without the interleaved copy");
  if (!pos_dofs || !uloc) throw DeviceError("launch_vanka_apply_schur: patch dofs or their values missing");
  auto go = [&](auto ref_c, auto ao_c) {
    constexpr bool R = decltype(ref_c)::value, A = decltype(ao_c)::value;
    k_vanka_apply_schur<F, R, A><<<ncells, 64, 0, s>>>(sch, pos_dofs, ncells, pos0, slots, rloc, uloc, n_udof, u,
                                                       ua, damp, R ? kref : nullptr, R ? nref : 0);
  };
  using T1 = std::true_type;
  using F0 = std::false_type;
  if (nref > 0 && ua_only) go(T1{}, T1{});
  else if (nref > 0) go(T1{}, F0{});
  else if (ua_only) go(F0{}, T1{});
  else go(F0{}, F0{});
  PMG_HIP_CHECK(hipGetLastError());
}
template void launch_vanka_apply_schur<float>(const float *, const int *, int, int, const int *, const float *,
                                              const float *, int, float *, float *, float, const float *, int,
                                              hipStream_t, bool);
template void launch_vanka_apply_schur<uint16_t>(const uint16_t *, const int *, int, int, const int *, const float *,
                                                 const float *, int, float *, float *, float, const float *, int,
                                                 hipStream_t, bool);

template <typename F>
void launch_vanka_apply(const F *minv, const int *pos_dofs, int ncells, int pos0, const int *slots,
                        const float *rloc, const float *uloc, int n_udof, float *u, float *ua, float damp,
                        const float *kref, const uint8_t *invperm, int nref, hipStream_t s) {
  if (ncells <= 0) return;
  if (!pos_dofs || !uloc) throw DeviceError("launch_vanka_apply: patch dofs or their values missing");
  if (nref > 0)
    k_vanka_apply<F, true><<<ncells, 64, 0, s>>>(minv, pos_dofs, ncells, pos0, slots, rloc, uloc, n_udof, u, ua,
                                                 damp, kref, invperm, nref);
  else
    k_vanka_apply<F, false><<<ncells, 64, 0, s>>>(minv, pos_dofs, ncells, pos0, slots, rloc, uloc, n_udof, u, ua,
                                                  damp, nullptr, invperm, 0);
  PMG_HIP_CHECK(hipGetLastError());
}
template void launch_vanka_apply<float>(const float *, const int *, int, int, const int *, const float *,
                                        const float *, int, float *, float *, float, const float *, const uint8_t *,
                                        int, hipStream_t);
template void launch_vanka_apply<uint16_t>(const uint16_t *, const int *, int, int, const int *, const float *,
                                           const float *, int, float *, float *, float, const float *,
                                           const uint8_t *, int, hipStream_t);

template <typename T>
void launch_vanka_residual(const OpView<T> &op, long n_items, int pos0, const uint8_t *invperm,
                           const int *items, const int *desc, const int *xg, const T *u, const float *ua,
                           const T *f, T *rloc, T *uloc, hipStream_t s, int per_block) {
  if (n_items <= 0) return;
  // small colors (coarse levels) are latency bound: one work item per 16-lane group there,
  // two on the big levels
  if (per_block <= 0) per_block = n_items < 65536 ? 16 : 32;
  int grid = chunk_grid(n_items, per_block, 8192);
  if (xg) grid = (grid + 7) / 8 * 8;
  if (op.a1)
    k_vanka_res<T, true><<<grid, 256, 0, s>>>(op, n_items, pos0, invperm, items, desc, xg, u, ua, f, rloc, uloc);
  else
    k_vanka_res<T, false><<<grid, 256, 0, s>>>(op, n_items, pos0, invperm, items, desc, xg, u, ua, f, rloc, uloc);
  PMG_HIP_CHECK(hipGetLastError());
}
template void launch_vanka_residual<double>(const OpView<double> &, long, int, const uint8_t *, const int *,
                                            const int *, const int *, const double *, const float *, const double *,
                                            double *, double *, hipStream_t, int);
template void launch_vanka_residual<float>(const OpView<float> &, long, int, const uint8_t *, const int *,
                                           const int *, const int *, const float *, const float *, const float *,
                                           float *, float *, hipStream_t, int);

template <typename T, typename F>
void launch_vanka_solve(const F *lu, const int *cells, int ncells, int pos0, const int *slots, const T *rloc,
                        const int *u_global, const int *p_global, int n_udof, T *u, T damp, hipStream_t s) {
  if (ncells <= 0) return;
  // FP32: colors of up to 2048 patches (a few waves per SIMD at most) are bound by the
  // dependent sweep chain and take the register-resident sweeps; larger colors need the
  // occupancy of the leaner LDS form
  if (sizeof(T) == 4 && ncells <= 2048)
    k_vanka_solve<T, F, sizeof(T) == 4><<<ncells, 64, PackedSizes<F>::LDS_BYTES, s>>>(
        lu, cells, ncells, pos0, slots, rloc, u_global, p_global, n_udof, u, damp);
  else
    k_vanka_solve<T, F, false><<<ncells, 64, PackedSizes<F>::LDS_BYTES, s>>>(
        lu, cells, ncells, pos0, slots, rloc, u_global, p_global, n_udof, u, damp);
  PMG_HIP_CHECK(hipGetLastError());
}
template void launch_vanka_solve<double, double>(const double *, const int *, int, int, const int *, const double *,
                                                 const int *, const int *, int, double *, double, hipStream_t);
template void launch_vanka_solve<float, float>(const float *, const int *, int, int, const int *, const float *, const int *,
                                               const int *, int, float *, float, hipStream_t);
template void launch_vanka_solve<float, uint16_t>(const uint16_t *, const int *, int, int, const int *, const float *,
                                                  const int *, const int *, int, float *, float, hipStream_t);


}  // namespace pmg
