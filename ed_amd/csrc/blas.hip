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
#include "nt.h"

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

// Streaming kernels walk the vector in tiles of kVecThreads * U elements; in
// tile T a thread owns elements T * kVecThreads * U + u * kVecThreads + tid
// (u < U), so each load instruction is coalesced.  Full tiles are processed
// without bounds checks (all loads of a tile issue back to back, U times the
// vector count requests in flight); the one partial tile at the end, if any,
// is handled element by element by the block that owns it.
template <int U>
struct Tiling {
  int64_t n, nfull;
  __device__ explicit Tiling(int64_t n_) : n(n_), nfull(n_ / ((int64_t)kVecThreads * U)) {}
  __device__ int64_t base(int64_t T) const { return T * kVecThreads * U + threadIdx.x; }
  // this block owns the partial tail tile
  __device__ bool owns_tail() const {
    return nfull * kVecThreads * U < n && (int64_t)blockIdx.x == nfull % gridDim.x;
  }
};

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

constexpr int kU = 4;

template <class RD, class RS>
__global__ __launch_bounds__(kVecThreads) void k_copy_scale(typename CT<RD>::T *dst, const typename CT<RS>::T *src,
                                                            int64_t n, const double *sc, int is, double s0) {
  const double s = (is >= 0) ? sc[is] : s0;
  auto one = [&](typename CT<RS>::T x, int64_t i) {
    typename CT<RD>::T o;
    o.x = (RD)((double)x.x * s);
    o.y = (RD)((double)x.y * s);
    dst[i] = o;
  };
  Tiling<kU> tl(n);
  for (int64_t T = blockIdx.x; T < tl.nfull; T += gridDim.x) {
    const int64_t b = tl.base(T);
    typename CT<RS>::T x[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) x[u] = src[b + u * kVecThreads];
#pragma unroll
    for (int u = 0; u < kU; ++u) one(x[u], b + u * kVecThreads);
  }
  if (tl.owns_tail())
    for (int64_t i = tl.base(tl.nfull); i < n; i += kVecThreads) one(src[i], i);
}

template <class R1, class R2>
__global__ __launch_bounds__(kVecThreads) void k_dot(const typename CT<R1>::T *x, const typename CT<R2>::T *y,
                                                     int64_t n, double *part) {
  double re = 0, im = 0;
  auto acc = [&](typename CT<R1>::T a, typename CT<R2>::T b) {
    re += (double)a.x * (double)b.x + (double)a.y * (double)b.y;
    im += (double)a.x * (double)b.y - (double)a.y * (double)b.x;
  };
  Tiling<kU> tl(n);
  for (int64_t T = blockIdx.x; T < tl.nfull; T += gridDim.x) {
    const int64_t b = tl.base(T);
    typename CT<R1>::T a[kU];
    typename CT<R2>::T c[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      a[u] = x[b + u * kVecThreads];
      c[u] = y[b + u * kVecThreads];
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) acc(a[u], c[u]);
  }
  if (tl.owns_tail())
    for (int64_t i = tl.base(tl.nfull); i < n; i += kVecThreads) acc(x[i], y[i]);
  block_sum2(re, im, &part[2 * blockIdx.x], &part[2 * blockIdx.x + 1]);
}

// mode 0: sc[i0] = re, sc[i1] = im; mode 1: sc[i0] = re, sc[i1] = re + shift;
// mode 2: b = sqrt(re): sc[i0] = b, sc[i1] = 1/b (0 if b == 0);
// mode 3: sc[i0] = sqrt(re)
// A thread's partial sums, blocks b = thread, thread + blockDim, ... in that
// order, with the loads of kFinU blocks issued before they are added (the
// kernel is one block: its time is the latency of its loads).
constexpr int kFinU = 8;
template <int NS>
__device__ __forceinline__ void sum_partials(const double *part, int nb, double (&t)[NS]) {
#pragma unroll
  for (int k = 0; k < NS; ++k) t[k] = 0.0;
  for (int b0 = threadIdx.x; b0 < nb; b0 += kFinU * blockDim.x) {
    double v[kFinU][NS];
#pragma unroll
    for (int u = 0; u < kFinU; ++u) {
      const int b = b0 + u * blockDim.x;
#pragma unroll
      for (int k = 0; k < NS; ++k) v[u][k] = b < nb ? part[NS * b + k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kFinU; ++u) {
      if (b0 + u * (int)blockDim.x >= nb) break;
#pragma unroll
      for (int k = 0; k < NS; ++k) t[k] += v[u][k];
    }
  }
}

__global__ __launch_bounds__(kVecThreads) void k_finalize(const double *part, int nb, double *sc, int mode, int i0,
                                                          int i1, double shift) {
  double t2[2];
  sum_partials<2>(part, nb, t2);
  double re = t2[0], im = t2[1];
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

// Fused Lanczos step (see launch_hv_fused): part holds 4 sums per block,
// Re<v, Hv>, |P|^2, Re<v, P>, |v|^2.  alpha = the first; the next Lanczos
// vector is (P - alpha v) / beta with beta^2 = |P|^2 - 2 alpha Re<v, P> +
// alpha^2 |v|^2.  Writes sc[ia] = alpha, sc[io] = alpha + shift, sc[ib] =
// beta, sc[is] = 1 / beta (0 if beta = 0).
__global__ __launch_bounds__(kVecThreads) void k_finalize_fused(const double *part, int nb, double *sc, int ia, int io,
                                                                int ib, int is, double shift) {
  double t[4];
  sum_partials<4>(part, nb, t);
  __shared__ double o[4][2];
  block_sum2(t[0], t[1], &o[0][0], &o[1][0]);
  block_sum2(t[2], t[3], &o[2][0], &o[3][0]);
  if (threadIdx.x == 0) {
    const double a = o[0][0], pp = o[1][0], vp = o[2][0], vv = o[3][0];
    const double b2 = pp - 2.0 * a * vp + a * a * vv;
    const double b = sqrt(b2 > 0 ? b2 : 0.0);
    sc[ia] = a;
    if (io >= 0) sc[io] = a + shift;
    sc[ib] = b;
    sc[is] = b > 0 ? 1.0 / b : 0.0;
  }
}

// w = w - a*sa*u - b*sb*up (the b term only when HASB); part = |w|^2.
template <class R, bool HASB>
__global__ __launch_bounds__(kVecThreads) void k_update3(typename CT<R>::T *w, const typename CT<R>::T *u,
                                                         const typename CT<R>::T *up, int64_t n, const double *sc,
                                                         int ia, int isa, int ib, int isb, double *part) {
  using T = typename CT<R>::T;
  const R ca = (R)(sc[ia] * (isa >= 0 ? sc[isa] : 1.0));
  const R cb = HASB ? (R)(sc[ib] * (isb >= 0 ? sc[isb] : 1.0)) : (R)0;
  double acc = 0;
  auto one = [&](T x, T a, T b, int64_t i) {
    x.x -= ca * a.x;
    x.y -= ca * a.y;
    if (HASB) {
      x.x -= cb * b.x;
      x.y -= cb * b.y;
    }
    st_nt(&w[i], x);
    acc += (double)x.x * (double)x.x + (double)x.y * (double)x.y;
  };
  Tiling<kU> tl(n);
  for (int64_t T0 = blockIdx.x; T0 < tl.nfull; T0 += gridDim.x) {
    const int64_t b0 = tl.base(T0);
    T x[kU], a[kU], b[kU];
#pragma unroll
    for (int q = 0; q < kU; ++q) {
      x[q] = ld_nt(&w[b0 + q * kVecThreads]);
      a[q] = ld_nt(&u[b0 + q * kVecThreads]);
      if (HASB) b[q] = ld_nt(&up[b0 + q * kVecThreads]);
      else b[q] = a[q];
    }
#pragma unroll
    for (int q = 0; q < kU; ++q) one(x[q], a[q], b[q], b0 + q * kVecThreads);
  }
  if (tl.owns_tail())
    for (int64_t i = tl.base(tl.nfull); i < n; i += kVecThreads) one(w[i], u[i], HASB ? up[i] : u[i], i);
  if (part) block_sum2(acc, 0.0, &part[2 * blockIdx.x], &part[2 * blockIdx.x + 1]);
}

// k_update3 for complex float with two elements (16 bytes) per lane and
// load: the same arithmetic per element, half the load and store
// instructions.  Needs 16-byte aligned vectors; an odd last element is done
// by the block that owns the tail.
typedef float u3_f4 __attribute__((ext_vector_type(4)));
template <bool HASB>
__global__ __launch_bounds__(kVecThreads) void k_update3_f4(float2 *w, const float2 *u, const float2 *up, int64_t n,
                                                            const double *sc, int ia, int isa, int ib, int isb,
                                                            double *part) {
  const float ca = (float)(sc[ia] * (isa >= 0 ? sc[isa] : 1.0));
  const float cb = HASB ? (float)(sc[ib] * (isb >= 0 ? sc[isb] : 1.0)) : 0.f;
  double acc = 0;
  u3_f4 *w4 = (u3_f4 *)w;
  const u3_f4 *u4 = (const u3_f4 *)u, *up4 = (const u3_f4 *)up;
  auto one = [&](u3_f4 x, u3_f4 a, u3_f4 b, int64_t i) {
    x -= ca * a;
    if (HASB) x -= cb * b;
    __builtin_nontemporal_store(x, &w4[i]);
    acc += (double)x.x * (double)x.x + (double)x.y * (double)x.y;
    acc += (double)x.z * (double)x.z + (double)x.w * (double)x.w;
  };
  const int64_t n2 = n >> 1;
  Tiling<kU> tl(n2);
  for (int64_t T0 = blockIdx.x; T0 < tl.nfull; T0 += gridDim.x) {
    const int64_t b0 = tl.base(T0);
    u3_f4 x[kU], a[kU], b[kU];
#pragma unroll
    for (int q = 0; q < kU; ++q) {
      x[q] = __builtin_nontemporal_load(&w4[b0 + q * kVecThreads]);
      a[q] = __builtin_nontemporal_load(&u4[b0 + q * kVecThreads]);
      if (HASB) b[q] = __builtin_nontemporal_load(&up4[b0 + q * kVecThreads]);
      else b[q] = a[q];
    }
#pragma unroll
    for (int q = 0; q < kU; ++q) one(x[q], a[q], b[q], b0 + q * kVecThreads);
  }
  if (tl.owns_tail())
    for (int64_t i = tl.base(tl.nfull); i < n2; i += kVecThreads) one(w4[i], u4[i], HASB ? up4[i] : u4[i], i);
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    float2 x = w[n - 1];
    const float2 a = u[n - 1];
    x.x -= ca * a.x;
    x.y -= ca * a.y;
    if (HASB) {
      const float2 c = up[n - 1];
      x.x -= cb * c.x;
      x.y -= cb * c.y;
    }
    w[n - 1] = x;
    acc += (double)x.x * (double)x.x + (double)x.y * (double)x.y;
  }
  if (part) block_sum2(acc, 0.0, &part[2 * blockIdx.x], &part[2 * blockIdx.x + 1]);
}

// psi (complex double) += z * sc[is] * u
template <class R>
__global__ __launch_bounds__(kVecThreads) void k_axpy_acc(double2 *psi, const typename CT<R>::T *u, int64_t n,
                                                          double z, const double *sc, int is) {
  const double c = z * (is >= 0 ? sc[is] : 1.0);
  auto one = [&](typename CT<R>::T a, double2 p, int64_t i) {
    p.x += c * (double)a.x;
    p.y += c * (double)a.y;
    psi[i] = p;
  };
  Tiling<kU> tl(n);
  for (int64_t T = blockIdx.x; T < tl.nfull; T += gridDim.x) {
    const int64_t b = tl.base(T);
    typename CT<R>::T a[kU];
    double2 p[kU];
#pragma unroll
    for (int q = 0; q < kU; ++q) {
      a[q] = u[b + q * kVecThreads];
      p[q] = psi[b + q * kVecThreads];
    }
#pragma unroll
    for (int q = 0; q < kU; ++q) one(a[q], p[q], b + q * kVecThreads);
  }
  if (tl.owns_tail())
    for (int64_t i = tl.base(tl.nfull); i < n; i += kVecThreads) one(u[i], psi[i], i);
}

// Elements per thread per tile of the multi-vector kernels (FP64 vectors use
// twice the registers per element).
template <class R>
constexpr int multi_u() {
  return sizeof(R) == 8 ? 2 : 4;
}

// h_j partials: part[(j * nb + b) * 2] = sum conj(s_j u_j) w over this block's
// rows, for the CNT vectors j0 .. j0 + CNT - 1.
template <class R, int CNT>
__global__ __launch_bounds__(kVecThreads) void k_mdot(const typename CT<R>::T *const *V, const double *scales,
                                                      const int *sidx, int j0, const typename CT<R>::T *w,
                                                      int64_t n, double *part) {
  constexpr int UU = multi_u<R>();
  using T = typename CT<R>::T;
  const T *vp[CNT];
#pragma unroll
  for (int j = 0; j < CNT; ++j) vp[j] = V[j0 + j];
  double re[CNT], im[CNT];
#pragma unroll
  for (int j = 0; j < CNT; ++j) re[j] = im[j] = 0.0;
  auto acc = [&](int j, T a, T b) {
    re[j] += (double)a.x * (double)b.x + (double)a.y * (double)b.y;
    im[j] += (double)a.x * (double)b.y - (double)a.y * (double)b.x;
  };
  Tiling<UU> tl(n);
  for (int64_t T0 = blockIdx.x; T0 < tl.nfull; T0 += gridDim.x) {
    const int64_t b0 = tl.base(T0);
    T b[UU], a[CNT][UU];
#pragma unroll
    for (int q = 0; q < UU; ++q) {
      b[q] = w[b0 + q * kVecThreads];
#pragma unroll
      for (int j = 0; j < CNT; ++j) a[j][q] = vp[j][b0 + q * kVecThreads];
    }
#pragma unroll
    for (int j = 0; j < CNT; ++j)
#pragma unroll
      for (int q = 0; q < UU; ++q) acc(j, a[j][q], b[q]);
  }
  if (tl.owns_tail())
    for (int64_t i = tl.base(tl.nfull); i < n; i += kVecThreads) {
      const T b = w[i];
#pragma unroll
      for (int j = 0; j < CNT; ++j) acc(j, vp[j][i], b);
    }
  const int nb = gridDim.x;
#pragma unroll
  for (int j = 0; j < CNT; ++j) {
    const double s = scales ? scales[sidx ? sidx[j0 + j] : j0 + j] : 1.0;
    block_sum2(re[j] * s, im[j] * s, &part[2 * ((size_t)(j0 + j) * nb + blockIdx.x)],
               &part[2 * ((size_t)(j0 + j) * nb + blockIdx.x) + 1]);
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

// x -= sum_{j < G} c_{j0+j} V[j0+j][i] for the UU elements of a full tile
// (coefficients in LDS, FP64 arithmetic).
template <class R, int G, int UU>
__device__ __forceinline__ void maxpy_group(const typename CT<R>::T *const *V, const double *sh, int j0, int64_t b0,
                                            double *xr, double *xi) {
  using T = typename CT<R>::T;
  T a[G][UU];
#pragma unroll
  for (int j = 0; j < G; ++j) {
    const T *vj = V[j0 + j];
#pragma unroll
    for (int q = 0; q < UU; ++q) a[j][q] = vj[b0 + q * kVecThreads];
  }
#pragma unroll
  for (int j = 0; j < G; ++j) {
    const double hr = sh[2 * (j0 + j)], hi = sh[2 * (j0 + j) + 1];
#pragma unroll
    for (int q = 0; q < UU; ++q) {
      xr[q] -= hr * (double)a[j][q].x - hi * (double)a[j][q].y;
      xi[q] -= hr * (double)a[j][q].y + hi * (double)a[j][q].x;
    }
  }
}

// w -= sum_j h_j s_j u_j (FP64 arithmetic, rounded to R once); part = |w|^2.
// The m vectors are taken kMaxGroup at a time, then the remainder.
template <class R>
__global__ __launch_bounds__(kVecThreads) void k_maxpy(const typename CT<R>::T *const *V, const double *scales,
                                                       const int *sidx, int m, const double *h,
                                                       typename CT<R>::T *w, int64_t n, double *part) {
  constexpr int UU = multi_u<R>();
  using T = typename CT<R>::T;
  __shared__ double sh[2 * kMaxBasis];
  for (int j = threadIdx.x; j < m; j += blockDim.x) {
    const double s = scales ? scales[sidx ? sidx[j] : j] : 1.0;
    sh[2 * j] = h[2 * j] * s;
    sh[2 * j + 1] = h[2 * j + 1] * s;
  }
  __syncthreads();
  double acc = 0;
  Tiling<UU> tl(n);
  const int mfull = m - m % kMaxGroup;
  for (int64_t T0 = blockIdx.x; T0 < tl.nfull; T0 += gridDim.x) {
    const int64_t b0 = tl.base(T0);
    double xr[UU], xi[UU];
#pragma unroll
    for (int q = 0; q < UU; ++q) {
      const T x = w[b0 + q * kVecThreads];
      xr[q] = x.x;
      xi[q] = x.y;
    }
    for (int j0 = 0; j0 < mfull; j0 += kMaxGroup) maxpy_group<R, kMaxGroup, UU>(V, sh, j0, b0, xr, xi);
    switch (m - mfull) {
      case 1: maxpy_group<R, 1, UU>(V, sh, mfull, b0, xr, xi); break;
      case 2: maxpy_group<R, 2, UU>(V, sh, mfull, b0, xr, xi); break;
      case 3: maxpy_group<R, 3, UU>(V, sh, mfull, b0, xr, xi); break;
      case 4: maxpy_group<R, 4, UU>(V, sh, mfull, b0, xr, xi); break;
      case 5: maxpy_group<R, 5, UU>(V, sh, mfull, b0, xr, xi); break;
      case 6: maxpy_group<R, 6, UU>(V, sh, mfull, b0, xr, xi); break;
      case 7: maxpy_group<R, 7, UU>(V, sh, mfull, b0, xr, xi); break;
      default: break;
    }
#pragma unroll
    for (int q = 0; q < UU; ++q) {
      T x;
      x.x = (R)xr[q];
      x.y = (R)xi[q];
      w[b0 + q * kVecThreads] = x;
      acc += (double)x.x * (double)x.x + (double)x.y * (double)x.y;
    }
  }
  if (tl.owns_tail())
    for (int64_t i = tl.base(tl.nfull); i < n; i += kVecThreads) {
      T x = w[i];
      double xr = x.x, xi = x.y;
      for (int j = 0; j < m; ++j) {
        const T a = V[j][i];
        xr -= sh[2 * j] * (double)a.x - sh[2 * j + 1] * (double)a.y;
        xi -= sh[2 * j] * (double)a.y + sh[2 * j + 1] * (double)a.x;
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
  // enough tiles of kVecThreads * 4 elements to give every CU 16 blocks
  // (measured: streaming updates gain 4-8 % from 8 to 16 blocks per CU)
  int64_t need = (n + 4 * kVecThreads - 1) / (4 * kVecThreads);
  return (int)std::max<int64_t>(1, std::min<int64_t>(need, std::min<int64_t>((int64_t)cus * 16, kMaxBlocks)));
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

void finalize_fused(const double *part, int nb, double *sc, int ia, int io, int ib, int is, double shift,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_finalize_fused, dim3(1), dim3(kVecThreads), 0, st, part, nb, sc, ia, io, ib, is, shift);
  QS3X_HIP(hipGetLastError());
}

void finalize(const double *part, int nb, double *sc, int mode, int i0, int i1, double shift, hipStream_t st) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kVecThreads), 0, st, part, nb, sc, mode, i0, i1, shift);
  QS3X_HIP(hipGetLastError());
}

int vec_update3(int prec, void *w, const void *u, const void *up, int64_t n, const double *sc, int ia, int isa,
                int ib, int isb, double *part, hipStream_t st) {
  const int nb = vec_grid(n);
  if (prec == QS3X_FP64) {
    if (ib >= 0)
      hipLaunchKernelGGL((k_update3<double, true>), GRID(n), (double2 *)w, (const double2 *)u, (const double2 *)up, n,
                         sc, ia, isa, ib, isb, part);
    else
      hipLaunchKernelGGL((k_update3<double, false>), GRID(n), (double2 *)w, (const double2 *)u, (const double2 *)up,
                         n, sc, ia, isa, ib, isb, part);
  } else if (n >= 2 && (((uintptr_t)w | (uintptr_t)u | (uintptr_t)(ib >= 0 ? up : u)) & 15u) == 0) {
    const int nb2 = vec_grid(n / 2);
    if (ib >= 0)
      hipLaunchKernelGGL((k_update3_f4<true>), dim3(nb2), dim3(kVecThreads), 0, st, (float2 *)w, (const float2 *)u,
                         (const float2 *)up, n, sc, ia, isa, ib, isb, part);
    else
      hipLaunchKernelGGL((k_update3_f4<false>), dim3(nb2), dim3(kVecThreads), 0, st, (float2 *)w, (const float2 *)u,
                         (const float2 *)up, n, sc, ia, isa, ib, isb, part);
    QS3X_HIP(hipGetLastError());
    return nb2;
  } else {
    if (ib >= 0)
      hipLaunchKernelGGL((k_update3<float, true>), GRID(n), (float2 *)w, (const float2 *)u, (const float2 *)up, n, sc,
                         ia, isa, ib, isb, part);
    else
      hipLaunchKernelGGL((k_update3<float, false>), GRID(n), (float2 *)w, (const float2 *)u, (const float2 *)up, n,
                         sc, ia, isa, ib, isb, part);
  }
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
template <class R>
static void mdot_group(const void *const *V, const double *scales, const int *sidx, int j0, int cnt, const void *w,
                       int64_t n, double *part, hipStream_t st) {
  using T = typename CT<R>::T;
#define MD(C)                                                                                                \
  case C:                                                                                                    \
    hipLaunchKernelGGL((k_mdot<R, C>), GRID(n), (const T *const *)V, scales, sidx, j0, (const T *)w, n, part); \
    break
  switch (cnt) {
    MD(1); MD(2); MD(3); MD(4); MD(5); MD(6); MD(7); MD(8);
    default: throw Error{"vec_mdot: bad group size"};
  }
#undef MD
}

void vec_mdot(int prec, const void *const *V, const double *scales, const int *sidx, int m, const void *w, int64_t n,
              double *part, double *h, hipStream_t st) {
  static_assert(kMaxGroup == 8, "mdot_group dispatches groups of up to 8 vectors");
  const int nb = vec_grid(n);
  for (int j0 = 0; j0 < m; j0 += kMaxGroup) {
    const int cnt = std::min(kMaxGroup, m - j0);
    if (prec == QS3X_FP64) mdot_group<double>(V, scales, sidx, j0, cnt, w, n, part, st);
    else mdot_group<float>(V, scales, sidx, j0, cnt, w, n, part, st);
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
