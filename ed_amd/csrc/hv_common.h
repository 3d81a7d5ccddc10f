// hv_common.h -- definitions shared by the H.v kernels (hv.hip, hv_nn.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "qs3x_internal.h"

namespace qs3x {
namespace hvk {

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

struct HvArgs {
  const uint64_t *cfg;
  const double *list_r;
  const int32_t *index;
  const uint2 *bitmap;  // anchored index, 32 ranks per word: x = presence bits, y = row of the first
  int64_t nbitmap;      // words in bitmap
  const int64_t *list_s;
  const void *v1;
  void *v0;
  int64_t dim;
  int64_t ntiles;
  int L, bits, kbits, mode;
  int LX, LY, LZ, NOS;
  const int32_t *adj_ptr;
  const int32_t *adj;
  int nadj;
  const uint16_t *nn_bt;  // nearest-neighbour chain: bond type of (s, s+1), s = 1..L
  const double *bt;       // interleaved jxy2, jz4
  int nbt;
  const int64_t *binom;  // global [n][k]
  int kdim;
  const double *rclass;
  const double *explist;  // complex
  double ezz;
  const double *scale;  // device scalar or null
  double *dot_partials; // 2 per block or null
  unsigned long long *nnz_out;
  int nxcd;             // tile ranges per L2 domain (1: plain round-robin over blocks)
  int chunk_log2;       // nearest-neighbour kernels: interleaved chunks of 2^c tiles per domain (0: one range)
  // Fused Lanczos step (stencil sectors, launch_hv_fused).  The input vector
  // is v = c1 v1 - c2 vprev (v1: P of the previous step, vprev: its Lanczos
  // vector; vprev null: v = c1 v1) with c1 = fsc[fis], c2 = fsc[fia] fsc[fis];
  // the kernels write v to xout and P = H v - c3 vprev to v0 (c3 = fsc[fib]),
  // and 4 partial sums per block: Re<v, Hv>, |P|^2, Re<v, P>, |v|^2.
  const void *vprev;
  void *xout;
  const double *fsc;
  int fis, fia, fib;
};

// Persistent-grid tile schedule.  Workgroups are dispatched round-robin over
// the nxcd L2 domains (XCDs) of the device, so block b runs on domain b % nxcd;
// each domain is given one contiguous range of tiles, and its blocks stride
// through that range.  Rows that are neighbours in configuration space are
// then processed by the same XCD, whose L2 holds their vector elements.
struct TileRange {
  int64_t first, end, stride;
};
__device__ __forceinline__ TileRange tile_range(int64_t ntiles, int nxcd) {
  // every domain that owns a range must have at least one block
  const int nx = (int)min<int64_t>(nxcd > 0 ? nxcd : 1, (int64_t)gridDim.x);
  const int x = (int)(blockIdx.x % nx);
  const int64_t per = ((int64_t)gridDim.x - x + nx - 1) / nx;  // blocks on domain x
  const int64_t t0 = ntiles * x / nx, t1 = ntiles * (x + 1) / nx;
  return TileRange{t0 + (int64_t)(blockIdx.x / nx), t1, per};
}

// Interleaved variant for the nearest-neighbour kernels: chunks of 2^c
// tiles are dealt to the nxcd domains in turn, chunk g to domain g % nxcd.
// c is clog, lowered until every domain owns at least 8 chunks (a domain's
// blocks run on its XCD only, so a sector of few chunks would leave XCDs
// idle); c = 0 deals single tiles.  A block walks domain-local tile numbers
// v = first, first + stride, ... < end, and tile_at maps v to the global
// tile.  clog <= 0 or nxcd <= 1 is tile_range.
struct TileMap {
  int64_t first, end, stride;
  int nx, x, clog;
  bool chunked;
};
__device__ __forceinline__ TileMap tile_map(int64_t ntiles, int nxcd, int clog) {
  if (clog <= 0 || nxcd <= 1) {
    const TileRange r = tile_range(ntiles, nxcd);
    return TileMap{r.first, r.end, r.stride, 1, 0, 0, false};
  }
  const int nx = (int)min<int64_t>(nxcd, (int64_t)gridDim.x);
  const int x = (int)(blockIdx.x % nx);
  const int64_t per = ((int64_t)gridDim.x - x + nx - 1) / nx;
  int c = clog;
  while (c > 0 && (ntiles >> c) < (int64_t)nx * 8) --c;
  const int64_t csz = (int64_t)1 << c;
  const int64_t nchunks = (ntiles + csz - 1) >> c;
  const int64_t K = x < nchunks ? (nchunks - x + nx - 1) / nx : 0;  // chunks of domain x
  int64_t n = 0;
  if (K > 0) {
    const int64_t glast = (K - 1) * nx + x;
    n = (K - 1) * csz + (glast == nchunks - 1 ? ntiles - glast * csz : csz);
  }
  return TileMap{(int64_t)(blockIdx.x / nx), n, per, nx, x, c, true};
}
__device__ __forceinline__ int64_t tile_at(const TileMap &m, int64_t v) {
  if (!m.chunked) return v;
  return ((((v >> m.clog) * m.nx) + m.x) << m.clog) + (v & (((int64_t)1 << m.clog) - 1));
}

constexpr int kThreads = 256;

// Block reduction of a complex double into out[2*blockIdx.x].
__device__ inline void block_reduce_store(double re, double im, double *out) {
  __shared__ double red[2][kThreads / 64];
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
    for (int w = 0; w < kThreads / 64; ++w) {
      a += red[0][w];
      b += red[1][w];
    }
    out[2 * blockIdx.x] = a;
    out[2 * blockIdx.x + 1] = b;
  }
}

// Block reduction of four doubles into out[4*blockIdx.x .. +3] (fixed order).
__device__ inline void block_reduce_store4(const double (&v)[4], double *out) {
  __shared__ double red4[4][kThreads / 64];
  double t[4] = {v[0], v[1], v[2], v[3]};
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1)
#pragma unroll
    for (int k = 0; k < 4; ++k) t[k] += __shfl_down(t[k], d, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 4; ++k) red4[k][wid] = t[k];
  __syncthreads();
  if (threadIdx.x < 4) {
    double a = 0;
    for (int w = 0; w < kThreads / 64; ++w) a += red4[threadIdx.x][w];
    out[4 * blockIdx.x + threadIdx.x] = a;
  }
}

// Fused Lanczos step coefficients (see HvArgs): c1 = 1/beta, c2 = alpha/beta,
// c3 = beta of the previous step; first step (fib < 0): c1 = the start
// vector's scale, c2 = c3 = 0.
template <class R>
struct FusedCoef {
  R c1, c2, c3;
  bool hasv;
};
template <class R>
__device__ __forceinline__ FusedCoef<R> fused_coef(const HvArgs &a) {
  FusedCoef<R> f;
  f.c1 = (R)a.fsc[a.fis];
  f.hasv = a.fib >= 0 && a.vprev != nullptr;
  f.c2 = f.hasv ? (R)(a.fsc[a.fia] * a.fsc[a.fis]) : (R)0;
  f.c3 = f.hasv ? (R)a.fsc[a.fib] : (R)0;
  return f;
}
__device__ __forceinline__ float fz_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fz_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
// v = c1 p - c2 w, one rounding order for every kernel that forms it
template <class R>
__device__ __forceinline__ typename CT<R>::T fused_v(const FusedCoef<R> &f, typename CT<R>::T p,
                                                      typename CT<R>::T w) {
  typename CT<R>::T v;
  if (f.hasv) {
    v.x = fz_fma(f.c1, p.x, -(f.c2 * w.x));
    v.y = fz_fma(f.c1, p.y, -(f.c2 * w.y));
  } else {
    v.x = f.c1 * p.x;
    v.y = f.c1 * p.y;
  }
  return v;
}

using KFn = void (*)(HvArgs);

// Gap-lattice stencil H.v (hv_nn.hip: slab-window form; hv_col.hip: column
// form).  A task is a tile of kSegA values of n2 by B lines (n3) swept over
// the slabs n4 = c0 .. c1 - 1; a slab's window holds B + 2 lines of kSegW
// positions (a halo of one each side) in one slot of a kRing-slot LDS ring.
constexpr int kSegA = 62;
constexpr int kSegW = 64;
constexpr int kRing = 4;
struct StencilArgs {
  const int4 *lines;    // [(L + 1) * c + b]: first row, lowest a, highest a + 1 (lowest = highest + 1: empty)
  const int4 *tasks;    // a0, b0, c0, c1 in bits 0..15 of .w (c1 <= c0: nothing to do)
  int ntasks;
  int B;                // lines per tile
};
using SFn = void (*)(HvArgs, StencilArgs);

}  // namespace hvk

// Column-form stencil kernel (hv_col.hip) for B lines per tile and P window
// positions per lane (1, 2 or 4); null if not instantiated.
hvk::SFn stencil_col_kernel(int prec, bool realph, bool fused, int B, int P);
// Pair-form stencil kernel (hv_pair.hip; FP32, two adjacent window positions
// per lane, tiles 124 values of n2 wide) for B lines per tile and D slab
// windows in flight (2 or 4); null if not instantiated.
hvk::SFn stencil_pair_kernel(bool realph, bool fused, int B, int D);

// Nearest-neighbour chain kernels (hv_nn.hip): null when no specialisation
// exists for this N / precision.
// uniform: every bond of the chain has the same (Jxy, Jz).
// pairs: FP32 vectors v0 and v1 are 16-byte aligned, so the two-rows-per-lane
// kernel may be used (real phases, N_down <= 4).
hvk::KFn chain_nn_kernel(int N, int prec, bool realph, bool uniform, bool key32, bool pairs);
size_t chain_nn_lds_bytes(int N, int L, int nbt, int prec);
// Gap-lattice stencil H.v (hv_nn.hip): launches it and returns true where it
// applies (nearest-neighbour chain, uniform bonds, N_down = 4, option
// hv_stencil), with its grid and dynamic LDS; false leaves nothing launched.
bool launch_hv_stencil(qs3x_sector *s, int prec, const hvk::HvArgs &a, hipStream_t st, int *grid_out,
                       size_t *lds_out);
// The fused Lanczos step on the stencil (a.xout set, see HvArgs): both
// kernels write 4 partial sums per block into a.dot_partials (if set).
bool launch_hv_stencil_fused(qs3x_sector *s, int prec, const hvk::HvArgs &a, hipStream_t st, int *grid_out);

}  // namespace qs3x
