// blas.hip -- fused vector kernels for the Lanczos solvers.
//
// Replaces my_zaxpy / my_zdscal (my_l1_blas.f90:5-25) and the BLAS-1 calls
// zdotc / dznrm2 / zcopy / zscal of lanczos.f90, eigen_solver.f90 and
// get_expectation_values.f90.  Vectors are complex FP64 (double2) or complex
// FP32 (float2); every reduction accumulates in FP64, per block, and a
// single-block finalize kernel sums the block partials in a fixed order (so
// results are deterministic for a given grid).  Lanczos scalars (alpha, beta,
// 1/beta) live in a device array and kernels read them through indices, so a
// sequence of steps is enqueued without host round trips.
#include <algorithm>

#include "qs3x_internal.h"
#include "blas.h"

namespace qs3x {
namespace {

template <class R>
struct CT;
template <>
struct CT<double> {
  using T = double2;
};
template <>
struct CT<float> {
  using T = float2;
};

__device__ __forceinline__ void block_sum2(double re, double im, double *out_re, double *out_im) {
  __shared__ double red[2][kVecThreads / 64];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    re += __shfl_down(re, d, 64);
    im += __shfl_down(im, d, 64);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wid] = re;
    red[1][wid] = im;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0;
    for (int w = 0; w < kVecThreads / 64; ++w) {
      a += red[0][w];
      b += red[1][w];
    }
    *out_re = a;
    *out_im = b;
  }
  __syncthreads();
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <class R>
__global__ __launch_bounds__(kVecThreads) void k_randvec(typename CT<R>::T *v, int64_t n, uint64_t seed,
                                                         int64_t row0) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t k = mix64(seed ^ ((uint64_t)(row0 + i) * 0xD1B54A32D192ED03ull));
    uint64_t a = mix64(k), b = mix64(a);
    double re = (double)(a >> 11) * 0x1.0p-53 * 2.0 - 1.0;
    double im = (double)(b >> 11) * 0x1.0p-53 * 2.0 - 1.0;
    typename CT<R>::T o;
    o.x = (R)re;
    o.y = (R)im;
    v[i] = o;
  }
}

template <class RD, class RS>
__global__ __launch_bounds__(kVecThreads) void k_copy_scale(typename CT<RD>::T *dst, const typename CT<RS>::T *src,
                                                            int64_t n, const double *sc, int is, double s0) {
  const double s = (is >= 0) ? sc[is] : s0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    typename CT<RS>::T x = src[i];
    typename CT<RD>::T o;
    o.x = (RD)((double)x.x * s);
    o.y = (RD)((double)x.y * s);
    dst[i] = o;
  }
}

template <class R1, class R2>
__global__ __launch_bounds__(kVecThreads) void k_dot(const typename CT<R1>::T *x, const typename CT<R2>::T *y,
                                                     int64_t n, double *part) {
  double re = 0, im = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    auto a = x[i];
    auto b = y[i];
    re += (double)a.x * (double)b.x + (double)a.y * (double)b.y;
    im += (double)a.x * (double)b.y - (double)a.y * (double)b.x;
  }
  block_sum2(re, im, &part[2 * blockIdx.x], &part[2 * blockIdx.x + 1]);
}

// mode 0: sc[i0] = re, sc[i1] = im; mode 1: sc[i0] = re, sc[i1] = re + shift;
// mode 2: b = sqrt(re): sc[i0] = b, sc[i1] = 1/b (0 if b == 0);
// mode 3: sc[i0] = sqrt(re)
__global__ __launch_bounds__(kVecThreads) void k_finalize(const double *part, int nb, double *sc, int mode, int i0,
                                                          int i1, double shift) {
  double re = 0, im = 0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    re += part[2 * b];
    im += part[2 * b + 1];
  }
  __shared__ double ore, oim;
  block_sum2(re, im, &ore, &oim);
  if (threadIdx.x == 0) {
    if (mode == 0) {
      sc[i0] = ore;
      if (i1 >= 0) sc[i1] = oim;
    } else if (mode == 1) {
      sc[i0] = ore;
      if (i1 >= 0) sc[i1] = ore + shift;
    } else if (mode == 2) {
      double b = sqrt(ore > 0 ? ore : 0.0);
      sc[i0] = b;
      if (i1 >= 0) sc[i1] = b > 0 ? 1.0 / b : 0.0;
    } else {
      sc[i0] = sqrt(ore > 0 ? ore : 0.0);
    }
  }
}

// w = w - a*sa*u - b*sb*up (b term skipped when ib < 0); part = |w|^2.
template <class R>
__global__ __launch_bounds__(kVecThreads) void k_update3(typename CT<R>::T *w, const typename CT<R>::T *u,
                                                         const typename CT<R>::T *up, int64_t n, const double *sc,
                                                         int ia, int isa, int ib, int isb, double *part) {
  const R ca = (R)(sc[ia] * (isa >= 0 ? sc[isa] : 1.0));
  const R cb = (ib >= 0) ? (R)(sc[ib] * (isb >= 0 ? sc[isb] : 1.0)) : (R)0;
  double acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    auto x = w[i];
    auto a = u[i];
    x.x -= ca * a.x;
    x.y -= ca * a.y;
    if (ib >= 0) {
      auto b = up[i];
      x.x -= cb * b.x;
      x.y -= cb * b.y;
    }
    w[i] = x;
    acc += (double)x.x * (double)x.x + (double)x.y * (double)x.y;
  }
  if (part) block_sum2(acc, 0.0, &part[2 * blockIdx.x], &part[2 * blockIdx.x + 1]);
}

// psi (complex double) += z * sc[is] * u
template <class R>
__global__ __launch_bounds__(kVecThreads) void k_axpy_acc(double2 *psi, const typename CT<R>::T *u, int64_t n,
                                                          double z, const double *sc, int is) {
  const double c = z * (is >= 0 ? sc[is] : 1.0);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    auto a = u[i];
    double2 p = psi[i];
    p.x += c * (double)a.x;
    p.y += c * (double)a.y;
    psi[i] = p;
  }
}

// h_j partials: part[(j * nb + b) * 2] = sum conj(s_j u_j) w over this block's
// rows, for j in [j0, j0 + cnt), cnt <= kMaxGroup.
template <class R>
__global__ __launch_bounds__(kVecThreads) void k_mdot(const typename CT<R>::T *const *V, const double *scales,
                                                      const int *sidx, int j0, int cnt,
                                                      const typename CT<R>::T *w, int64_t n, double *part) {
  double re[kMaxGroup], im[kMaxGroup];
#pragma unroll
  for (int j = 0; j < kMaxGroup; ++j) re[j] = im[j] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    auto b = w[i];
#pragma unroll
    for (int j = 0; j < kMaxGroup; ++j) {
      if (j < cnt) {
        auto a = V[j0 + j][i];
        re[j] += (double)a.x * (double)b.x + (double)a.y * (double)b.y;
        im[j] += (double)a.x * (double)b.y - (double)a.y * (double)b.x;
      }
    }
  }
  const int nb = gridDim.x;
#pragma unroll
  for (int j = 0; j < kMaxGroup; ++j) {
    if (j < cnt) {
      const double s = scales ? scales[sidx ? sidx[j0 + j] : j0 + j] : 1.0;
      block_sum2(re[j] * s, im[j] * s, &part[2 * ((size_t)(j0 + j) * nb + blockIdx.x)],
                 &part[2 * ((size_t)(j0 + j) * nb + blockIdx.x) + 1]);
    }
  }
}

// h[j] = sum_b part[j][b] (complex), j < m; one block per j.
__global__ __launch_bounds__(kVecThreads) void k_finalize_multi(const double *part, int nb, double *h) {
  const int j = blockIdx.x;
  double re = 0, im = 0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    re += part[2 * ((size_t)j * nb + b)];
    im += part[2 * ((size_t)j * nb + b) + 1];
  }
  __shared__ double ore, oim;
  block_sum2(re, im, &ore, &oim);
  if (threadIdx.x == 0) {
    h[2 * j] = ore;
    h[2 * j + 1] = oim;
  }
}

// w -= sum_j h_j s_j u_j (FP64 arithmetic, rounded to R once); part = |w|^2
template <class R>
__global__ __launch_bounds__(kVecThreads) void k_maxpy(const typename CT<R>::T *const *V, const double *scales,
                                                       const int *sidx, int m, const double *h,
                                                       typename CT<R>::T *w, int64_t n, double *part) {
  __shared__ double sh[2 * kMaxBasis];
  for (int j = threadIdx.x; j < m; j += blockDim.x) {
    const double s = scales ? scales[sidx ? sidx[j] : j] : 1.0;
    sh[2 * j] = h[2 * j] * s;
    sh[2 * j + 1] = h[2 * j + 1] * s;
  }
  __syncthreads();
  double acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    auto x = w[i];
    double xr = x.x, xi = x.y;
    for (int j = 0; j < m; ++j) {
      auto a = V[j][i];
      const double hr = sh[2 * j], hi = sh[2 * j + 1];
      xr -= hr * (double)a.x - hi * (double)a.y;
      xi -= hr * (double)a.y + hi * (double)a.x;
    }
    x.x = (R)xr;
    x.y = (R)xi;
    w[i] = x;
    acc += (double)x.x * (double)x.x + (double)x.y * (double)x.y;
  }
  if (part) block_sum2(acc, 0.0, &part[2 * blockIdx.x], &part[2 * blockIdx.x + 1]);
}

}  // namespace

int vec_grid(int64_t n) {
  static int cus = 0;
  if (!cus) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device()) == hipSuccess) cus = p.multiProcessorCount;
    if (cus <= 0) cus = 256;
  }
  int64_t need = (n + kVecThreads - 1) / kVecThreads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)cus * 4));
}

#define GRID(n) dim3(vec_grid(n)), dim3(kVecThreads), 0, st

void vec_random(int prec, void *v, int64_t n, uint64_t seed, hipStream_t st) {
  if (prec == QS3X_FP64) hipLaunchKernelGGL(k_randvec<double>, GRID(n), (double2 *)v, n, seed, (int64_t)0);
  else hipLaunchKernelGGL(k_randvec<float>, GRID(n), (float2 *)v, n, seed, (int64_t)0);
  QS3X_HIP(hipGetLastError());
}

void vec_copy_scale(int pd, void *dst, int ps, const void *src, int64_t n, const double *sc, int is, double s0,
                    hipStream_t st) {
  if (pd == QS3X_FP64 && ps == QS3X_FP64)
    hipLaunchKernelGGL((k_copy_scale<double, double>), GRID(n), (double2 *)dst, (const double2 *)src, n, sc, is, s0);
  else if (pd == QS3X_FP64)
    hipLaunchKernelGGL((k_copy_scale<double, float>), GRID(n), (double2 *)dst, (const float2 *)src, n, sc, is, s0);
  else if (ps == QS3X_FP64)
    hipLaunchKernelGGL((k_copy_scale<float, double>), GRID(n), (float2 *)dst, (const double2 *)src, n, sc, is, s0);
  else
    hipLaunchKernelGGL((k_copy_scale<float, float>), GRID(n), (float2 *)dst, (const float2 *)src, n, sc, is, s0);
  QS3X_HIP(hipGetLastError());
}

int vec_dot(int px, const void *x, int py, const void *y, int64_t n, double *part, hipStream_t st) {
  const int nb = vec_grid(n);
  if (px == QS3X_FP64 && py == QS3X_FP64)
    hipLaunchKernelGGL((k_dot<double, double>), GRID(n), (const double2 *)x, (const double2 *)y, n, part);
  else if (px == QS3X_FP64)
    hipLaunchKernelGGL((k_dot<double, float>), GRID(n), (const double2 *)x, (const float2 *)y, n, part);
  else if (py == QS3X_FP64)
    hipLaunchKernelGGL((k_dot<float, double>), GRID(n), (const float2 *)x, (const double2 *)y, n, part);
  else
    hipLaunchKernelGGL((k_dot<float, float>), GRID(n), (const float2 *)x, (const float2 *)y, n, part);
  QS3X_HIP(hipGetLastError());
  return nb;
}

void finalize(const double *part, int nb, double *sc, int mode, int i0, int i1, double shift, hipStream_t st) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kVecThreads), 0, st, part, nb, sc, mode, i0, i1, shift);
  QS3X_HIP(hipGetLastError());
}

int vec_update3(int prec, void *w, const void *u, const void *up, int64_t n, const double *sc, int ia, int isa,
                int ib, int isb, double *part, hipStream_t st) {
  const int nb = vec_grid(n);
  if (prec == QS3X_FP64)
    hipLaunchKernelGGL(k_update3<double>, GRID(n), (double2 *)w, (const double2 *)u, (const double2 *)up, n, sc, ia,
                       isa, ib, isb, part);
  else
    hipLaunchKernelGGL(k_update3<float>, GRID(n), (float2 *)w, (const float2 *)u, (const float2 *)up, n, sc, ia, isa,
                       ib, isb, part);
  QS3X_HIP(hipGetLastError());
  return nb;
}

void vec_axpy_acc(int prec, double *psi, const void *u, int64_t n, double z, const double *sc, int is,
                  hipStream_t st) {
  if (prec == QS3X_FP64)
    hipLaunchKernelGGL(k_axpy_acc<double>, GRID(n), (double2 *)psi, (const double2 *)u, n, z, sc, is);
  else
    hipLaunchKernelGGL(k_axpy_acc<float>, GRID(n), (double2 *)psi, (const float2 *)u, n, z, sc, is);
  QS3X_HIP(hipGetLastError());
}

// h = V^H w for the m vectors of the device pointer array V (scaled by
// scales[j] if given) -> h (complex, device, 2m doubles).  part must hold
// 2 * m * vec_grid(n) doubles.
void vec_mdot(int prec, const void *const *V, const double *scales, const int *sidx, int m, const void *w, int64_t n,
              double *part, double *h, hipStream_t st) {
  const int nb = vec_grid(n);
  for (int j0 = 0; j0 < m; j0 += kMaxGroup) {
    const int cnt = std::min(kMaxGroup, m - j0);
    if (prec == QS3X_FP64)
      hipLaunchKernelGGL(k_mdot<double>, GRID(n), (const double2 *const *)V, scales, sidx, j0, cnt,
                         (const double2 *)w, n, part);
    else
      hipLaunchKernelGGL(k_mdot<float>, GRID(n), (const float2 *const *)V, scales, sidx, j0, cnt,
                         (const float2 *)w, n, part);
  }
  hipLaunchKernelGGL(k_finalize_multi, dim3(m), dim3(kVecThreads), 0, st, part, nb, h);
  QS3X_HIP(hipGetLastError());
}

int vec_maxpy(int prec, const void *const *V, const double *scales, const int *sidx, int m, const double *h,
              void *w, int64_t n, double *part, hipStream_t st) {
  QS3X_REQUIRE(m <= kMaxBasis, "vec_maxpy: too many basis vectors");
  const int nb = vec_grid(n);
  if (prec == QS3X_FP64)
    hipLaunchKernelGGL(k_maxpy<double>, GRID(n), (const double2 *const *)V, scales, sidx, m, h, (double2 *)w, n,
                       part);
  else
    hipLaunchKernelGGL(k_maxpy<float>, GRID(n), (const float2 *const *)V, scales, sidx, m, h, (float2 *)w, n, part);
  QS3X_HIP(hipGetLastError());
  return nb;
}

}  // namespace qs3x
