// k_batch.hip - kernels of the batched path (independent problems merged into one
// block-diagonal hierarchy, for gfx950): IntoL20 per problem, copies between a problem's
// own numbering and the merged numbering, and the FGMRES BLAS-1 with one scalar per
// problem.
//
// The per-problem reductions run over a fixed chunk table: every chunk is a contiguous
// piece (at most SEG_CHUNK entries) of one problem's range of the merged vector, one
// workgroup per chunk writes its partial sum, and a second launch adds each problem's
// partials in chunk order - deterministic, no atomics.
#include "pmg_internal.hpp"

namespace pmg {

namespace {

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// block of 256 threads; fixed order; valid in every thread
__device__ inline double block_sum(double v, double *sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  return (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// IntoL20Vector3D case -11 (MGComponents3D.C:1766-1786) on one problem's pressure block:
// seg = (offset, length) pairs, one workgroup per problem
template <typename T>
__global__ __launch_bounds__(256) void k_l20_seg(T *__restrict__ p, const int *__restrict__ seg) {
  __shared__ double sh[4];
  const int off = seg[2 * blockIdx.x], np = seg[2 * blockIdx.x + 1];
  double s = 0;
  for (int i = threadIdx.x; i < np / 4; i += 256) s += double(p[off + 4 * i]);
  s = block_sum(s, sh);
  const double m = s / double(np) * 4.0;
  for (int i = threadIdx.x; i < np / 4; i += 256) p[off + 4 * i] = T(p[off + 4 * i] - T(m));
}

template <typename T>
__global__ void k_gather_d(double *__restrict__ dst, const T *__restrict__ src, const int *__restrict__ idx, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = double(src[idx[i]]);
}
template <typename T>
__global__ void k_scatter_d(T *__restrict__ dst, const double *__restrict__ src, const int *__restrict__ idx, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[idx[i]] = T(src[i]);
}

// chunk c = (start, length, problem, unused)
__global__ __launch_bounds__(256) void k_seg_dot(const double *__restrict__ x, const double *__restrict__ y,
                                                 const int4 *__restrict__ ch, double *__restrict__ part) {
  __shared__ double sh[4];
  const int4 c = ch[blockIdx.x];
  const double *__restrict__ xc = x + c.x;
  const double *__restrict__ yc = y + c.x;
  double s = 0;
  // four of a thread's entries loaded before they are added in their order (a chunk is
  // latency bound when few chunks share the device)
  int i = threadIdx.x;
  for (; i + 3 * 256 < c.y; i += 4 * 256) {
    double a[4], b[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      a[j] = xc[i + j * 256];
      b[j] = yc[i + j * 256];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) s += a[j] * b[j];
  }
  for (; i < c.y; i += 256) s += xc[i] * yc[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// w -= h[problem * ldh] v ; partial of <w, vn> (vn may be w itself: the new norm)
__global__ __launch_bounds__(256) void k_seg_axpy_dot(double *w, const double *__restrict__ h, int ldh,
                                                      const double *__restrict__ v, const double *vn,
                                                      const int4 *__restrict__ ch, double *__restrict__ part) {
  __shared__ double sh[4];
  const int4 c = ch[blockIdx.x];
  const double a = -h[long(c.z) * ldh];
  const bool self = vn == w;
  double *__restrict__ wc = w + c.x;
  const double *__restrict__ vc = v + c.x;
  const double *__restrict__ nc = (vn && !self) ? vn + c.x : nullptr;
  double s = 0;
  // as k_seg_dot: four of a thread's entries in flight, added in their order
  int i = threadIdx.x;
  for (; i + 3 * 256 < c.y; i += 4 * 256) {
    double wv[4], vv[4], nv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      wv[j] = wc[i + j * 256];
      vv[j] = vc[i + j * 256];
      nv[j] = nc ? nc[i + j * 256] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double wi = wv[j] + a * vv[j];
      wc[i + j * 256] = wi;
      if (self)
        s += wi * wi;
      else if (nc)
        s += wi * nv[j];
    }
  }
  for (; i < c.y; i += 256) {
    const double wi = wc[i] + a * vc[i];
    wc[i] = wi;
    if (self)
      s += wi * wi;
    else if (nc)
      s += wi * nc[i];
  }
  if (!vn) return;
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// out[b * ldo] = sum of problem b's partials, in chunk order (eight loads in flight ahead
// of the additions, which keep that order)
__global__ void k_seg_reduce(const double *__restrict__ part, const int *__restrict__ cptr, int nb,
                             double *__restrict__ out, int ldo) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int k1 = cptr[b + 1];
  int k = cptr[b];
  double s = 0;
  for (; k + 8 <= k1; k += 8) {
    double t[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) t[j] = part[k + j];
#pragma unroll
    for (int j = 0; j < 8; ++j) s += t[j];
  }
  for (; k < k1; ++k) s += part[k];
  out[long(b) * ldo] = s;
}
// y = coef[problem] x; a zero coefficient writes zeros whatever x holds
__global__ __launch_bounds__(256) void k_seg_scale(double *__restrict__ y, const double *__restrict__ x,
                                                   const double *__restrict__ coef, const int4 *__restrict__ ch) {
  const int4 c = ch[blockIdx.x];
  const double a = coef[c.z];
  for (int i = threadIdx.x; i < c.y; i += 256) y[c.x + i] = a == 0.0 ? 0.0 : a * x[c.x + i];
}
// x += sum_j coef[problem * ldc + j] z_j for the problems with mask != 0
__global__ __launch_bounds__(256) void k_seg_lincomb(double *__restrict__ x, const double *__restrict__ z, long stride,
                                                     const double *__restrict__ coef, int ldc, int m,
                                                     const int *__restrict__ mask, const int4 *__restrict__ ch) {
  const int4 c = ch[blockIdx.x];
  if (!mask[c.z]) return;
  const double *cf = coef + long(c.z) * ldc;
  for (int i = threadIdx.x; i < c.y; i += 256) {
    const long e = long(c.x) + i;
    double t = 0.0;
    for (int j = 0; j < m; ++j) t += z[j * stride + e] * cf[j];
    x[e] += t;
  }
}

inline int blocks(long n, int bs = 256) { return int((n + bs - 1) / bs); }

}  // namespace

template <typename T>
void launch_l20_seg(T *p, const int *seg, int nseg, hipStream_t s) {
  if (nseg <= 0) return;
  k_l20_seg<T><<<nseg, 256, 0, s>>>(p, seg);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_gather_d(double *dst, const T *src, const int *idx, int n, hipStream_t s) {
  if (n <= 0) return;
  k_gather_d<T><<<blocks(n), 256, 0, s>>>(dst, src, idx, n);
  PMG_HIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_scatter_d(T *dst, const double *src, const int *idx, int n, hipStream_t s) {
  if (n <= 0) return;
  k_scatter_d<T><<<blocks(n), 256, 0, s>>>(dst, src, idx, n);
  PMG_HIP_CHECK(hipGetLastError());
}
template void launch_l20_seg<double>(double *, const int *, int, hipStream_t);
template void launch_l20_seg<float>(float *, const int *, int, hipStream_t);
template void launch_gather_d<double>(double *, const double *, const int *, int, hipStream_t);
template void launch_gather_d<float>(double *, const float *, const int *, int, hipStream_t);
template void launch_scatter_d<double>(double *, const double *, const int *, int, hipStream_t);
template void launch_scatter_d<float>(float *, const double *, const int *, int, hipStream_t);

void launch_seg_dot(const double *x, const double *y, const int4 *ch, int nch, double *part, hipStream_t s) {
  k_seg_dot<<<nch, 256, 0, s>>>(x, y, ch, part);
  PMG_HIP_CHECK(hipGetLastError());
}
void launch_seg_axpy_dot(double *w, const double *h, int ldh, const double *v, const double *vn, const int4 *ch,
                         int nch, double *part, hipStream_t s) {
  k_seg_axpy_dot<<<nch, 256, 0, s>>>(w, h, ldh, v, vn, ch, part);
  PMG_HIP_CHECK(hipGetLastError());
}
void launch_seg_reduce(const double *part, const int *cptr, int nb, double *out, int ldo, hipStream_t s) {
  k_seg_reduce<<<blocks(nb, 64), 64, 0, s>>>(part, cptr, nb, out, ldo);
  PMG_HIP_CHECK(hipGetLastError());
}
void launch_seg_scale(double *y, const double *x, const double *coef, const int4 *ch, int nch, hipStream_t s) {
  k_seg_scale<<<nch, 256, 0, s>>>(y, x, coef, ch);
  PMG_HIP_CHECK(hipGetLastError());
}
void launch_seg_lincomb(double *x, const double *z, long stride, const double *coef, int ldc, int m, const int *mask,
                        const int4 *ch, int nch, hipStream_t s) {
  if (m <= 0) return;
  k_seg_lincomb<<<nch, 256, 0, s>>>(x, z, stride, coef, ldc, m, mask, ch);
  PMG_HIP_CHECK(hipGetLastError());
}

}  // namespace pmg
