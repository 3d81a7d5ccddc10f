// batch.hip -- the DSF recurrence for many wave vectors in one block.
//
// The reference runs calcu_DSF_wavevector (get_expectation_values.f90:379-605)
// once per transferred wave vector q: S^-_q |psi0> into the k = q sector of
// N_down + 1, then itr_dsf Lanczos steps without reorthogonalisation.  Here
// up to 64 wave vectors run together:
//
// * One "all-k" sector holds every translation representative of the
//   target N_down (a chain), in the reference's row order.  The k = q sector
//   is the subset of rows whose stabiliser class has ro_q > 1e-15
//   (state_lists.f90:244-292), so a vector of the k = q sector is a vector of
//   the all-k sector that is zero on the other rows, and H_q maps it to one
//   that is again zero there (hops into excluded rows carry r_q = 0).
// * The 64 vectors are interleaved row-major, [row][q]: the hop structure of
//   a row (targets, ell, stabiliser orders) is computed once and applied to
//   all 64 vectors, and every gather reads one contiguous 64-element row.
// * Block H.v (nearest-neighbour chains): a wavefront takes 64 rows; with
//   lanes = rows it computes each row's hops (closed forms of hv_nn.hip, key
//   search for near-tie rows) into LDS; with lanes = q it then applies them:
//   y_q(row) = (diag) x_q(row) + sum_hops Jxy/2 r_q(ct)/r_q(cr) explist_q(ell) x_q(target).
// * S^-_q |psi0> (calcu_Sq_a, get_expectation_values.f90:281-377, pull form)
//   is built the same way: lanes = rows find the ground-state rows that feed
//   each target row, lanes = q weight them with exp(-i q R_i) and
//   conj(explist_q(ell)).
// * The recurrence keeps per-q scalars in one device array [slot][q].
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <vector>

#include "blas.h"
#include "nt.h"
#include "hv_common.h"
#include "qs3x_internal.h"
#include "rep.h"

namespace qs3x {

namespace {

using namespace hvk;

constexpr int kQ = 64;         // vectors per block, one per lane
constexpr int kBThreads = 256;  // 4 wavefronts
constexpr int kBWaves = kBThreads / 64;

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class F>
int guarded4(F &&f) {
  try {
    f();
    return 0;
  } catch (const Error &e) {
    set_error(e.msg);
    return 1;
  } catch (const std::exception &e) {
    set_error(std::string("qs3x: ") + e.what());
    return 2;
  }
}

// --------------------------------------------------------------- block H.v
struct BlockHvArgs {
  const uint64_t *cfg;
  const uint2 *bitmap;
  int64_t dim;
  int L, bits, kbits;
  const int64_t *binom;
  int kdim;
  const uint16_t *nn_bt;
  const double *bt;  // interleaved jxy2, jz4
  int nbt;
  double shift;       // diagonal constant in this precision (e_zz for FP64, 0 for FP32)
  const void *x;      // [dim][kQ]
  void *y;            // [dim][kQ]
  const double *sc;   // per-q scalars [slot][kQ]
  int iscale;         // slot of the scale of x (< 0: 1)
  const double *ratio;    // [N+1][N+1][kQ]: r_q(ct) / r_q(cr), 0 when r_q(cr) = 0
  const double *rowmask;  // [N+1][kQ]: 1 when r_q(c) > 0
  const void *eph;        // [L+1][kQ] complex in precision R: explist_q(ell, 1, 1); entry 0 = 1
  double *dot_partials;   // [grid][kQ] complex
  int nxcd;               // L2 domains for the chunk schedule (tile_range)
  const int32_t *rows;    // non-null: apply only these rows (ascending), nrows of them
  int64_t nrows;
  // Fused Lanczos step (the FUSED kernels): x holds P_{i-1} (the first step:
  // the start vectors), vprev v_{i-1} (null on the first step); a row writes
  // v_i = c1 P_{i-1} - c2 v_{i-1} to xout and P_i = H v_i - c3 v_{i-1} to y,
  // with per q c1 = sc[fis], c2 = sc[fia] sc[fis], c3 = sc[fib]; the partial
  // sums are [grid][kQ][4]: Re<v,Hv>, |P|^2, Re<v,P>, |v|^2.
  const void *vprev;
  void *xout;
  int fia, fis, fib;
};

// per-q coefficients of the fused step (the single-vector fused_coef per q)
template <class R>
__device__ __forceinline__ void fused_coef_q(const BlockHvArgs &a, int q, R &c1, R &c2, R &c3) {
  const bool hasv = a.vprev != nullptr && a.fib >= 0;
  c1 = (R)a.sc[(size_t)a.fis * kQ + q];
  c2 = hasv ? (R)(a.sc[(size_t)a.fia * kQ + q] * a.sc[(size_t)a.fis * kQ + q]) : (R)0;
  c3 = hasv ? (R)a.sc[(size_t)a.fib * kQ + q] : (R)0;
}
// v = c1 p - c2 w (w absent: c1 p), the rounding order of fused_v
template <class R>
__device__ __forceinline__ R fused_e(R c1, R c2, R p, R w, bool hasv) {
  return hasv ? fz_fma(c1, p, -(c2 * w)) : c1 * p;
}

// C(x, q) for 1 <= q <= N - 1 from the LDS table.
__device__ __forceinline__ int bnm(const int32_t *s_binom, int L, int q, int x) {
  return s_binom[(q - 1) * (L - 1) + x];
}

// Representative of a sorted chain configuration: anchored rank, stabiliser
// order, ell_x (0 when nn already is the representative).
template <int N, class KeyT>
__device__ __forceinline__ void chain_search_b(const int *nn, int L, int kb, const int32_t *s_binom, int &c,
                                               int &count, int &ell) {
  const KeyT fmask = (KeyT)((1ull << kb) - 1ull);
  int g[N];
#pragma unroll
  for (int q = 0; q < N - 1; ++q) g[q] = nn[q + 1] - nn[q];
  g[N - 1] = nn[0] + L - nn[N - 1];
  KeyT K[N];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    KeyT k = 0;
#pragma unroll
    for (int tt = 1; tt < N; ++tt) k = (KeyT)((k << kb) | (KeyT)g[(j - tt + N) % N]);
    K[j] = k;
  }
  KeyT kmax = K[0];
#pragma unroll
  for (int j = 1; j < N; ++j) kmax = K[j] > kmax ? K[j] : kmax;
  int mx = 0;
  count = 0;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const bool tie = (K[j] == kmax);
    count += tie ? 1 : 0;
    mx = (tie && nn[j] > mx) ? nn[j] : mx;
  }
  const bool isrep = (nn[0] == 1) && (K[0] == kmax);
  ell = isrep ? 0 : (L + 1 - mx);
  int f[N - 1];
  int sum = 0;
#pragma unroll
  for (int tt = 1; tt < N; ++tt) {
    f[tt - 1] = (int)((kmax >> (kb * (N - 1 - tt))) & fmask);
    sum += f[tt - 1];
  }
  int P = L - sum;
  c = s_binom[P - 1];
#pragma unroll
  for (int mm = 2; mm < N; ++mm) {
    P += f[N - mm];
    c += s_binom[(mm - 1) * (L - 1) + (P - 1)];
  }
}

// Hops of one row of a nearest-neighbour chain: for slot h (2m: spin m to
// the right, 2m + 1: to the left) the anchored rank of the image's
// representative (-1: no hop), its stabiliser order, ell_x and the bond site
// (bond (bsite, bsite + 1)).  Closed forms for clear rows (hv_nn.hip), the key
// search otherwise.
template <int N, class KeyT>
__device__ __forceinline__ void nn_row_hops(const int *n, int L, int kb, const int32_t *s_binom, int *hc, int *hcnt,
                                            int *hell, int *hbs) {
  constexpr int H = 2 * N;
  int g[N];
#pragma unroll
  for (int q = 0; q < N - 1; ++q) g[q] = n[q + 1] - n[q];
  const int W = L + 1 - n[N - 1];
  g[N - 1] = W;
  int M = 0;
#pragma unroll
  for (int q = 0; q < N - 1; ++q) M = max(M, g[q]);
  const bool clear = M <= W - 3;
  int crow = 0, c_ar = 0, c_al = 0;
#pragma unroll
  for (int q = 1; q < N; ++q) {
    const int base = (q - 1) * (L - 1) + n[q];
    crow += s_binom[base - 2];
    c_ar += s_binom[max(base - 3, (q - 1) * (L - 1))];
    c_al += s_binom[min(base - 1, (q - 1) * (L - 1) + L - 2)];
  }
#pragma unroll
  for (int h = 0; h < H; ++h) {
    const int m = h >> 1;
    const bool right = (h & 1) == 0;
    const bool ok = right ? (g[m] > 1) : ((m == 0 ? W : g[m - 1]) > 1);
    hbs[h] = right ? n[m] : (m == 0 ? L : n[m] - 1);
    hcnt[h] = 1;
    hell[h] = 0;
    if (!ok) {
      hc[h] = -1;
      continue;
    }
    if (clear) {
      if (m == 0) {
        hc[h] = right ? c_ar : c_al;
        hell[h] = right ? L - 1 : 1;
      } else if (m == 1) {
        hc[h] = crow + (right ? 1 : -1);
      } else {
        hc[h] = right ? crow + bnm(s_binom, L, m - 1, n[m] - 2) : crow - bnm(s_binom, L, m - 1, n[m] - 3);
      }
    } else {
      int nn[N];
      if (m == 0 && right) {
        nn[0] = 2;
#pragma unroll
        for (int q = 1; q < N; ++q) nn[q] = n[q];
      } else if (m == 0) {
#pragma unroll
        for (int q = 0; q < N - 1; ++q) nn[q] = n[q + 1];
        nn[N - 1] = L;
      } else {
#pragma unroll
        for (int q = 0; q < N; ++q) nn[q] = n[q] + ((q == m) ? (right ? 1 : -1) : 0);
      }
      int c, cnt, ell;
      chain_search_b<N, KeyT>(nn, L, kb, s_binom, c, cnt, ell);
      hc[h] = c;
      hcnt[h] = cnt;
      hell[h] = ell;
    }
  }
}

template <class R>
__device__ __forceinline__ typename CT<R>::T cload(const void *p, int64_t i) {
  return ((const typename CT<R>::T *)p)[i];
}

// y = H_q (s_q x) for the kQ interleaved vectors (see the file comment);
// FUSED: the fused Lanczos step of BlockHvArgs (every element gathered is
// formed as v_i from P_{i-1} and v_{i-1}).
template <int N, class R, class KeyT, bool FUSED>
__global__ __launch_bounds__(kBThreads) void k_hv_block(BlockHvArgs a) {
  using C = typename CT<R>::T;
  constexpr int H = 2 * N;
  extern __shared__ __align__(16) unsigned char smem[];
  const int L = a.L;
  const int nb = (N - 1) * (L - 1);
  // LDS: binom int32 [nb] | ratio R [(N+1)^2][kQ] | rowmask R [N+1][kQ] | jxy2 R [nbt] | jz4 R [nbt] |
  //      per wave: hop id int32 [64][H], hop aux int32 [64][H], row class int32 [64], anti R [64]
  int32_t *s_binom = (int32_t *)smem;
  R *s_ratio = (R *)(smem + (((size_t)nb * 4 + 15) & ~(size_t)15));
  R *s_mask = s_ratio + (N + 1) * (N + 1) * kQ;
  R *s_jxy2 = s_mask + (N + 1) * kQ;
  R *s_jz4 = s_jxy2 + a.nbt;
  int32_t *s_wave = (int32_t *)(((uintptr_t)(s_jz4 + a.nbt) + 15) & ~(uintptr_t)15);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const size_t wave_words = (size_t)64 * H * 2 + 64 + 64 + 64 * (sizeof(R) / 4);
  int32_t *w_id = s_wave + wid * wave_words;
  int32_t *w_aux = w_id + 64 * H;
  int32_t *w_cls = w_aux + 64 * H;
  int32_t *w_row = w_cls + 64;
  R *w_anti = (R *)(w_row + 64);
  for (int q = threadIdx.x; q < nb; q += kBThreads) {
    const int m = q / (L - 1) + 1, x = q % (L - 1);
    s_binom[q] = (int32_t)a.binom[(int64_t)x * a.kdim + m];
  }
  for (int q = threadIdx.x; q < (N + 1) * (N + 1) * kQ; q += kBThreads) s_ratio[q] = (R)a.ratio[q];
  for (int q = threadIdx.x; q < (N + 1) * kQ; q += kBThreads) s_mask[q] = (R)a.rowmask[q];
  for (int q = threadIdx.x; q < a.nbt; q += kBThreads) {
    s_jxy2[q] = (R)a.bt[2 * q];
    s_jz4[q] = (R)a.bt[2 * q + 1];
  }
  __syncthreads();

  const R scale = (!FUSED && a.iscale >= 0) ? (R)a.sc[(size_t)a.iscale * kQ + lane] : (R)1;
  R fc1 = (R)1, fc2 = (R)0, fc3 = (R)0;
  const bool hasv = FUSED && a.vprev != nullptr && a.fib >= 0;
  if constexpr (FUSED) fused_coef_q<R>(a, lane, fc1, fc2, fc3);
  // element i of this lane's vector: x, or (FUSED) v_i formed from P_{i-1} and v_{i-1}
  auto xload = [&](int64_t i) __attribute__((always_inline)) -> C {
    C p = cload<R>(a.x, i);
    if constexpr (FUSED) {
      const C w = hasv ? cload<R>(a.vprev, i) : C{(R)0, (R)0};
      p.x = fused_e<R>(fc1, fc2, p.x, w.x, hasv);
      p.y = fused_e<R>(fc1, fc2, p.y, w.y, hasv);
    }
    return p;
  };
  double acc4[4] = {0.0, 0.0, 0.0, 0.0};  // FUSED: Re<v,Hv>, |P|^2, Re<v,P>, |v|^2
  // phases of the closed-form anchor hops (ell = 1 and L - 1) of this lane's q
  const C e1 = cload<R>(a.eph, (size_t)1 * kQ + lane);
  const C eL1 = cload<R>(a.eph, (size_t)(L - 1) * kQ + lane);
  double dre = 0.0, dim_ = 0.0;
  // the rows applied: all of them, or the listed ones (entry i of the list
  // stands where row i would)
  const int64_t nit = a.rows ? a.nrows : a.dim;
  const int64_t nchunks = (nit + 63) / 64;
  // a block takes kBWaves consecutive chunks at a time, XCD-contiguous ranges
  const TileRange tr = tile_range((nchunks + kBWaves - 1) / kBWaves, a.nxcd);
  for (int64_t bt = tr.first; bt < tr.end; bt += tr.stride) {
    const int64_t ch = bt * kBWaves + wid;
    if (ch >= nchunks) continue;
    const int64_t base = ch * 64;
    // lanes = rows: hop structure of the chunk
    {
      const int64_t it = base + lane;
      const int64_t row = it < nit ? (a.rows ? (int64_t)a.rows[it] : it) : 0;
      int hc[H], hcnt[H], hell[H], hbs[H];
      int cls = 0;
      R anti = 0;
      if (it < nit) {
        const uint64_t w = a.cfg[row];
        cls = (int)(w >> 56);
        int n[N];
#pragma unroll
        for (int q = 0; q < N; ++q) n[q] = (int)((w >> (q * a.bits)) & ((1ull << a.bits) - 1ull));
        nn_row_hops<N, KeyT>(n, L, a.kbits, s_binom, hc, hcnt, hell, hbs);
      } else {
#pragma unroll
        for (int h = 0; h < H; ++h) {
          hc[h] = -1;
          hcnt[h] = 1;
          hell[h] = 0;
          hbs[h] = 1;
        }
      }
      uint2 wd[H];
#pragma unroll
      for (int h = 0; h < H; ++h) wd[h] = a.bitmap[(uint32_t)(hc[h] >= 0 ? hc[h] : 0) >> 5];
#pragma unroll
      for (int h = 0; h < H; ++h) {
        const uint32_t b = (uint32_t)hc[h] & 31u;
        const bool in = hc[h] >= 0 && ((wd[h].x >> b) & 1u);
        const int id = in ? (int)wd[h].y + __popc(wd[h].x & ((1u << b) - 1u)) : -1;
        const int bt = a.nn_bt[hbs[h]];
        if (hc[h] >= 0) anti += s_jz4[bt];
        w_id[lane * H + h] = id;
        w_aux[lane * H + h] = hcnt[h] | (hell[h] << 5) | (bt << 17);
      }
      w_cls[lane] = cls;
      w_row[lane] = (int32_t)row;
      w_anti[lane] = anti;
    }
    __builtin_amdgcn_wave_barrier();
    // lanes = vectors: apply the chunk's rows to the kQ vectors, two rows at a time
    const int nr = (int)min<int64_t>(64, nit - base);
    for (int r0 = 0; r0 < nr; r0 += 2) {
      constexpr int RR = 2;
      C xin[RR], y[RR][H];
      int id[RR][H], aux[RR][H];
#pragma unroll
      for (int k = 0; k < RR; ++k) {
        const int r = min(r0 + k, nr - 1);
        const int64_t rw = w_row[r];
        xin[k] = xload(rw * kQ + lane);
#pragma unroll
        for (int h = 0; h < H; ++h) {
          id[k][h] = w_id[r * H + h];
          aux[k][h] = w_aux[r * H + h];
        }
#pragma unroll
        for (int h = 0; h < H; ++h) y[k][h] = xload((int64_t)(id[k][h] >= 0 ? id[k][h] : rw) * kQ + lane);
      }
#pragma unroll
      for (int k = 0; k < RR; ++k) {
        const int r = r0 + k;
        if (r >= nr) break;
        const int cr = w_cls[r];
        R accx = 0, accy = 0;
#pragma unroll
        for (int h = 0; h < H; ++h) {
          if (id[k][h] < 0) continue;
          const int ct = aux[k][h] & 31, ell = (aux[k][h] >> 5) & 0xfff, bt = aux[k][h] >> 17;
          const R cf = s_jxy2[bt] * s_ratio[(ct * (N + 1) + cr) * kQ + lane];
          C e;
          if (ell == 0) {
            e.x = (R)1;
            e.y = (R)0;
          } else if (ell == 1) {
            e = e1;
          } else if (ell == L - 1) {
            e = eL1;
          } else {
            e = cload<R>(a.eph, (size_t)ell * kQ + lane);
          }
          const R pr = cf * e.x, pi = cf * e.y;
          accx += pr * y[k][h].x - pi * y[k][h].y;
          accy += pr * y[k][h].y + pi * y[k][h].x;
        }
        C x = xin[k];
        x.x *= scale;
        x.y *= scale;
        const R dg = (R)a.shift - (R)2 * w_anti[r];
        const R mk = s_mask[cr * kQ + lane];
        C out;
        out.x = mk * (dg * x.x + accx * scale);
        out.y = mk * (dg * x.y + accy * scale);
        const int64_t io = (int64_t)w_row[r] * kQ + lane;
        if constexpr (FUSED) {
          const C wo = hasv ? cload<R>(a.vprev, io) : C{(R)0, (R)0};
          C pn;
          pn.x = out.x - fc3 * wo.x;
          pn.y = out.y - fc3 * wo.y;
          st_nt(&((C *)a.y)[io], pn);
          st_nt(&((C *)a.xout)[io], x);
          acc4[0] += (double)x.x * (double)out.x + (double)x.y * (double)out.y;
          acc4[1] += (double)pn.x * (double)pn.x + (double)pn.y * (double)pn.y;
          acc4[2] += (double)x.x * (double)pn.x + (double)x.y * (double)pn.y;
          acc4[3] += (double)x.x * (double)x.x + (double)x.y * (double)x.y;
        } else {
          st_nt(&((C *)a.y)[io], out);  // written once per step
          dre += (double)x.x * (double)out.x + (double)x.y * (double)out.y;
          dim_ += (double)x.x * (double)out.y - (double)x.y * (double)out.x;
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  // per-q block partials: sum the 4 wavefronts in fixed order
  constexpr int NS = FUSED ? 4 : 2;
  __shared__ double red[NS][kBWaves][kQ];
  if constexpr (FUSED) {
#pragma unroll
    for (int k = 0; k < 4; ++k) red[k][wid][lane] = acc4[k];
  } else {
    red[0][wid][lane] = dre;
    red[1][wid][lane] = dim_;
  }
  __syncthreads();
  if (a.dot_partials && wid == 0) {
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      double t = 0;
      for (int w = 0; w < kBWaves; ++w) t += red[k][w][lane];
      a.dot_partials[NS * ((size_t)blockIdx.x * kQ + lane) + k] = t;
    }
  }
}

template <int N>
size_t hv_block_lds(int L, int nbt, int prec) {
  const size_t rs = prec == QS3X_FP64 ? 8 : 4;
  const size_t nb = (size_t)(N - 1) * (L - 1);
  size_t b = ((nb * 4 + 15) & ~(size_t)15) + rs * ((N + 1) * (N + 1) * kQ + (N + 1) * kQ + 2 * nbt);
  b = (b + 15) & ~(size_t)15;
  const size_t wave_words = (size_t)64 * 2 * N * 2 + 64 + 64 + 64 * (rs / 4);
  return b + 4 * wave_words * kBWaves;
}

using BFn = void (*)(BlockHvArgs);

template <int N, bool FUSED = false>
BFn pick_block(int prec, bool key32) {
  if (prec == QS3X_FP64)
    return key32 ? k_hv_block<N, double, uint32_t, FUSED> : k_hv_block<N, double, uint64_t, FUSED>;
  return key32 ? k_hv_block<N, float, uint32_t, FUSED> : k_hv_block<N, float, uint64_t, FUSED>;
}

BFn block_kernel(int N, int prec, bool key32, size_t *lds, int L, int nbt) {
  switch (N) {
#define BK(n)                                \
  case n:                                    \
    *lds = hv_block_lds<n>(L, nbt, prec);   \
    return pick_block<n>(prec, key32)
    BK(2); BK(3); BK(4); BK(5); BK(6); BK(7); BK(8);
#undef BK
    default: return nullptr;
  }
}

// The fused per-row block kernel: N_down = 4 only (the rows the block
// stencil leaves, the one place the fused step runs).
BFn block_kernel_fused(int N, int prec, bool key32, size_t *lds, int L, int nbt) {
  if (N != 4) return nullptr;
  *lds = hv_block_lds<4>(L, nbt, prec);
  return pick_block<4, true>(prec, key32);
}

// ---------------------------------------------- block gap-lattice stencil
// The clear rows of an N_down = 4 all-k nearest-neighbour chain sector (the
// rows whose hops all land on fixed lattice neighbours, see the stencil
// kernel of hv_nn.hip) for the kQ interleaved vectors.  A clear row (1, a, b,
// c) has stabiliser order 1 and so do its hop targets, so for every q its
// row mask and norm ratios are 1 and only the two anchor hops carry phases,
// explist_q(L - 1) for (a - 1, b - 1, c - 1) and explist_q(1) for (a + 1,
// b + 1, c + 1).  A lattice point of the sweep is one 128-byte chunk of a
// row: 16 wave vectors (FP32) or 8 (FP64), the task's vector group.  A task
// is a tile of kSA values of a by B lines swept over its slabs c0 .. c1 - 1;
// each slab's window (B + 2 lines of kSA + 2 points) goes into a four-slot
// LDS ring, slab c is computed from slots c - 1, c, c + 1 while the next two
// windows are in flight in registers, as in the single-vector stencil.  A
// row's chunk is read about (B + 2)(kSA + 2) / (B kSA) times per H.v instead
// of once per hop that reaches it, and no configuration word or index is
// read.  The groups of a tile are consecutive tasks of one XCD, so the four
// (eight) chunks of a row are requested at about the same time.
constexpr int kSA = 16;   // a-values per tile
constexpr int kSW = kSA + 2;
constexpr int kSRing = 4;

template <class R>
struct Chunk16;
template <>
struct Chunk16<float> {
  typedef float T __attribute__((ext_vector_type(4)));  // two complex
  static constexpr int nq = 2;
};
template <>
struct Chunk16<double> {
  typedef double T __attribute__((ext_vector_type(2)));  // one complex
  static constexpr int nq = 1;
};

struct BlockStencilArgs {
  const int4 *lines;  // line table of the sector ((L + 1) c + b: first row, lowest a, highest a + 1)
  const int4 *tasks;  // a0, b0, c0, c1 | group << 16 (c1 <= c0: nothing to do)
  int ntasks;
};

template <class R, int B>
constexpr size_t block_stencil_lds(int slabs) {
  return (size_t)kSRing * (B + 2) * kSW * 128 + (size_t)(slabs + 2) * (B + 2) * sizeof(int4);
}

// FUSED: the fused Lanczos step of BlockHvArgs.  The window holds v_i,
// formed from the loaded P_{i-1} and v_{i-1} as each window point goes into
// the ring; a thread's fetch slots are then its own output points first (in
// the order compute takes them) and the halo after them, so the v_{i-1} of
// its own rows stays in registers (two sets, slabs of even and odd c - c0)
// from the window's store until the slab is computed.
template <class R, int B, int NT, bool FUSED>
__global__ __launch_bounds__(NT) void k_hv_block_stencil(BlockHvArgs a, BlockStencilArgs sa) {
  using V = typename Chunk16<R>::T;
  constexpr int NQ = Chunk16<R>::nq;
  constexpr int GQ = 8 * NQ;        // wave vectors per group (one 128-byte chunk of a row)
  constexpr int RV = kQ / NQ;       // 16-byte chunks per row
  constexpr int NL = B + 2;         // window lines
  constexpr int NWIN = NL * kSW;    // window points
  constexpr int SLOT = NWIN * 8;    // 16-byte chunks per ring slot
  constexpr int PR = NT / 8;              // window points per round (8 chunks each)
  constexpr int PER = (NWIN + PR - 1) / PR;  // window points per thread
  constexpr int NOUT8 = B * kSA * 8;     // output chunks per slab
  constexpr int MI = NOUT8 / NT;         // output points per thread and slab
  constexpr int NS = FUSED ? 4 : 2;      // partial sums per q
  static_assert(NOUT8 % NT == 0, "whole rounds of output chunks");
  static_assert(kSRing * SLOT * 16 >= NT * NQ * NS * 8, "the partial-sum area must fit in the ring");
  extern __shared__ __align__(16) unsigned char smem[];
  V *s_ring = (V *)smem;
  int4 *s_tl = (int4 *)(smem + (size_t)kSRing * SLOT * sizeof(V));
  // per q: scale, row mask, explist_q(1) re, im, explist_q(L - 1) re, im; FUSED: c1, c2, c3
  __shared__ R s_qt[FUSED ? 9 : 6][kQ];
  __shared__ double s_acc[NS][kQ];  // per q: the block's partial sums
  const int L = a.L;
  const int tid = threadIdx.x, sub = tid & 7, ps = tid >> 3;
  const bool hasv = FUSED && a.vprev != nullptr && a.fib >= 0;
  for (int q = tid; q < kQ; q += NT) {
    s_qt[0][q] = (!FUSED && a.iscale >= 0) ? (R)a.sc[(size_t)a.iscale * kQ + q] : (R)1;
    s_qt[1][q] = (R)a.rowmask[1 * kQ + q];
    const R *e = (const R *)a.eph;
    s_qt[2][q] = e[2 * ((size_t)1 * kQ + q)];
    s_qt[3][q] = e[2 * ((size_t)1 * kQ + q) + 1];
    s_qt[4][q] = e[2 * ((size_t)(L - 1) * kQ + q)];
    s_qt[5][q] = e[2 * ((size_t)(L - 1) * kQ + q) + 1];
    if constexpr (FUSED) fused_coef_q<R>(a, q, s_qt[6][q], s_qt[7][q], s_qt[8][q]);
#pragma unroll
    for (int k = 0; k < NS; ++k) s_acc[k][q] = 0.0;
  }
  const R jxy2 = (R)a.bt[0], jz4 = (R)a.bt[1], shift = (R)a.shift;
  const V *__restrict__ x = (const V *)a.x;
  const V *__restrict__ vp = (const V *)a.vprev;
  V *__restrict__ y = (V *)a.y;
  V *__restrict__ xo = (V *)a.xout;
  const V zero = (V)(R)0;
  // fetch slot m of this thread: its window point (NWIN: none)
  auto wpoint = [&](int m) __attribute__((always_inline)) -> int {
    if constexpr (!FUSED) {
      return ps + PR * m;
    } else {
      if (m < MI) {
        const int pos = ps + PR * m;
        return (1 + pos / kSA) * kSW + pos % kSA + 1;
      }
      const int h = ps + PR * (m - MI);  // halo: lines 0 and NL - 1, then the end columns of lines 1 .. B
      if (h < kSW) return h;
      if (h < 2 * kSW) return (NL - 1) * kSW + (h - kSW);
      const int h2 = h - 2 * kSW;
      return h2 < 2 * B ? (1 + h2 / 2) * kSW + ((h2 & 1) ? kSW - 1 : 0) : NWIN;
    }
  };
  V preX[PER], preY[PER], wX[PER], wY[PER];
  V wown[2][FUSED ? MI : 1];  // FUSED: v_{i-1} of the thread's own points, slabs of even / odd c - c0
  // window points of slab te for vector group g
  auto fetch = [&](V (&pre)[PER], V (&w)[PER], int a0, int g, const int4 *te) __attribute__((always_inline)) {
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int p = wpoint(m);
      V v = zero, u = zero;
      if (p < NWIN) {
        const int j = p / kSW, i = p - j * kSW;
        const int4 e = te[j];
        const int av = a0 - 1 + i;
        if (av >= e.y && av < e.z) {
          const int64_t idx = (int64_t)(e.x + (av - e.y)) * RV + g * 8 + sub;
          v = x[idx];
          if (FUSED && hasv) u = vp[idx];
        }
      }
      pre[m] = v;
      if constexpr (FUSED) w[m] = u;
    }
  };
  // this thread's wave vectors of group g
  R qs[NQ], qm[NQ], e1r[NQ], e1i[NQ], eLr[NQ], eLi[NQ], fc1[NQ], fc2[NQ], fc3[NQ];
  double acc[NQ][NS];
  // par: the own-point set the slab's v_{i-1} goes to (-1: none)
  auto store = [&](const V (&pre)[PER], const V (&w)[PER], int sl, int par) __attribute__((always_inline)) {
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int p = wpoint(m);
      V v = pre[m];
      if constexpr (FUSED) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
          v[2 * k] = fused_e<R>(fc1[k], fc2[k], v[2 * k], w[m][2 * k], hasv);
          v[2 * k + 1] = fused_e<R>(fc1[k], fc2[k], v[2 * k + 1], w[m][2 * k + 1], hasv);
        }
        if (m < MI) {
          if (par == 0) wown[0][m] = w[m];
          if (par == 1) wown[1][m] = w[m];
        }
      }
      if (p < NWIN) s_ring[(size_t)sl * SLOT + p * 8 + sub] = v;
    }
  };
  auto compute = [&](int a0, int b0, int g, int c, const int4 *te0, const V (&wo)[FUSED ? MI : 1])
                     __attribute__((always_inline)) {
    const V *r0 = s_ring + (size_t)(c % kSRing) * SLOT;
    const V *rm = s_ring + (size_t)((c + kSRing - 1) % kSRing) * SLOT;
    const V *rp = s_ring + (size_t)((c + 1) % kSRing) * SLOT;
    const int W = L + 1 - c;
#pragma unroll
    for (int pb = 0; pb < NOUT8; pb += NT) {
      const int pos = (pb + tid) >> 3;
      const int j = 1 + pos / kSA, ai = pos % kSA;  // window line, a - a0
      const int4 e = te0[j];
      const int av = a0 + ai, b = b0 - 1 + j;
      const int g1 = av - 1, g2 = b - av, g3 = c - b;
      if (av >= e.y && av < e.z && max(g1, max(g2, g3)) <= W - 3) {
        const int r = e.x + (av - e.y);
        const int x0 = (j * kSW + ai + 1) * 8 + sub;
        V sm = rp[x0];                       // (a, b, c + 1)
        const V al = rp[x0 + (kSW + 1) * 8];  // anchor to site L: (a + 1, b + 1, c + 1), explist(1)
        V ar = zero;
        int nanti = 2;
        if (g2 > 1) {
          sm += r0[x0 + 8];        // (a + 1, b, c)
          sm += r0[x0 - kSW * 8];  // (a, b - 1, c)
          nanti += 2;
        }
        if (g1 > 1) {
          sm += r0[x0 - 8];                 // (a - 1, b, c)
          ar = rm[x0 - (kSW + 1) * 8];      // anchor to site 2: (a - 1, b - 1, c - 1), explist(L - 1)
          nanti += 2;
        }
        if (g3 > 1) {
          sm += r0[x0 + kSW * 8];  // (a, b + 1, c)
          sm += rm[x0];            // (a, b, c - 1)
          nanti += 2;
        }
        const V xv = r0[x0];
        const R dg = shift - (R)2 * jz4 * (R)nanti;
        V out, pn;
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
          const R sx = sm[2 * k] + e1r[k] * al[2 * k] - e1i[k] * al[2 * k + 1] + eLr[k] * ar[2 * k] -
                       eLi[k] * ar[2 * k + 1];
          const R sy = sm[2 * k + 1] + e1r[k] * al[2 * k + 1] + e1i[k] * al[2 * k] + eLr[k] * ar[2 * k + 1] +
                       eLi[k] * ar[2 * k];
          // FUSED: the window already holds v_i (scale 1)
          const R sq = FUSED ? (R)1 : qs[k];
          const R xr = xv[2 * k] * sq, xi = xv[2 * k + 1] * sq;
          const R orr = qm[k] * (dg * xr + jxy2 * sq * sx);
          const R oi = qm[k] * (dg * xi + jxy2 * sq * sy);
          out[2 * k] = orr;
          out[2 * k + 1] = oi;
          if constexpr (FUSED) {
            const V vo = wo[pb / NT];
            const R pr = orr - fc3[k] * vo[2 * k], pi = oi - fc3[k] * vo[2 * k + 1];
            pn[2 * k] = pr;
            pn[2 * k + 1] = pi;
            acc[k][0] += (double)xr * (double)orr + (double)xi * (double)oi;
            acc[k][1] += (double)pr * (double)pr + (double)pi * (double)pi;
            acc[k][2] += (double)xr * (double)pr + (double)xi * (double)pi;
            acc[k][3] += (double)xr * (double)xr + (double)xi * (double)xi;
          } else {
            acc[k][0] += (double)xr * (double)orr + (double)xi * (double)oi;
            acc[k][1] += (double)xr * (double)oi - (double)xi * (double)orr;
          }
        }
        const int64_t io = (int64_t)r * RV + g * 8 + sub;
        if constexpr (FUSED) {
          __builtin_nontemporal_store(pn, &y[io]);
          __builtin_nontemporal_store(xv, &xo[io]);
        } else {
          __builtin_nontemporal_store(out, &y[io]);
        }
      }
    }
  };
  __syncthreads();
  for (int t = blockIdx.x; t < sa.ntasks; t += gridDim.x) {
    const int4 task = sa.tasks[t];
    const int a0 = task.x, b0 = task.y, c0 = task.z, c1 = task.w & 0xffff, g = task.w >> 16;
    if (c1 <= c0) continue;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      const int q = g * GQ + sub * NQ + k;
      qs[k] = FUSED ? (R)1 : s_qt[0][q];
      qm[k] = s_qt[1][q];
      e1r[k] = s_qt[2][q];
      e1i[k] = s_qt[3][q];
      eLr[k] = s_qt[4][q];
      eLi[k] = s_qt[5][q];
      if constexpr (FUSED) {
        fc1[k] = s_qt[6][q];
        fc2[k] = s_qt[7][q];
        fc3[k] = s_qt[8][q];
      }
#pragma unroll
      for (int s2 = 0; s2 < NS; ++s2) acc[k][s2] = 0.0;
    }
    // line entries of slabs c0 - 1 .. c1 (entry (cs, j) at s_tl[(cs - c0 + 1) NL + j])
    const int nent = (c1 - c0 + 2) * NL;
    for (int q = tid; q < nent; q += NT) {
      const int cs = c0 - 1 + q / NL, b = b0 - 1 + q % NL;
      int4 e = make_int4(0, 0, 0, 0);
      if (cs >= 0 && cs <= L && b >= 0 && b <= L) e = sa.lines[(size_t)cs * (L + 1) + b];
      s_tl[q] = e;
    }
    __syncthreads();
    auto te = [&](int cs) { return s_tl + (cs - c0 + 1) * NL; };
    {
      V preZ[PER], wZ[PER];
      fetch(preX, wX, a0, g, te(c0 - 1));
      fetch(preY, wY, a0, g, te(c0));
      fetch(preZ, wZ, a0, g, te(c0 + 1));
      store(preX, wX, (c0 + kSRing - 1) % kSRing, -1);
      store(preY, wY, c0 % kSRing, 0);
      store(preZ, wZ, (c0 + 1) % kSRing, 1);
    }
    if (c0 + 2 <= c1) fetch(preX, wX, a0, g, te(c0 + 2));
    __syncthreads();
    for (int c = c0; c < c1; c += 2) {
      if (c + 3 <= c1) fetch(preY, wY, a0, g, te(c + 3));
      compute(a0, b0, g, c, te(c), wown[0]);
      if (c + 2 <= c1) store(preX, wX, (c + 2) % kSRing, 0);
      __syncthreads();
      if (c + 1 < c1) {
        if (c + 4 <= c1) fetch(preX, wX, a0, g, te(c + 4));
        compute(a0, b0, g, c + 1, te(c + 1), wown[1]);
        if (c + 3 <= c1) store(preY, wY, (c + 3) % kSRing, 1);
        __syncthreads();
      }
    }
    // the task's partial sums into the block's per-q totals, in a fixed order
    // (the ring is free until the next task's prologue)
    double *s_red = (double *)smem;  // [NT][NQ][NS]
#pragma unroll
    for (int k = 0; k < NQ; ++k)
#pragma unroll
      for (int s2 = 0; s2 < NS; ++s2) s_red[(tid * NQ + k) * NS + s2] = acc[k][s2];
    __syncthreads();
    if (tid < GQ) {
      const int sb = tid / NQ, k = tid % NQ;
#pragma unroll
      for (int s2 = 0; s2 < NS; ++s2) {
        double sr = 0.0;
        for (int p = 0; p < NT / 8; ++p) sr += s_red[((p * 8 + sb) * NQ + k) * NS + s2];
        s_acc[s2][g * GQ + tid] += sr;
      }
    }
    __syncthreads();
  }
  if (a.dot_partials) {
    for (int q = tid; q < kQ; q += NT)
#pragma unroll
      for (int s2 = 0; s2 < NS; ++s2) a.dot_partials[NS * ((size_t)blockIdx.x * kQ + q) + s2] = s_acc[s2][q];
  }
}

using BSFn = void (*)(BlockHvArgs, BlockStencilArgs);

// Block stencil kernel for B lines per tile and NT threads per block (256:
// B even, 2..14; 512: B = 4, 8, 12), or nullptr.
template <class R>
BSFn block_stencil_kernel_r(int B, int NT, size_t *lds, int slabs, bool fused) {
#define BSK(b, nt)                                                                            \
  if (B == b && NT == nt) {                                                                   \
    *lds = block_stencil_lds<R, b>(slabs);                                                    \
    return fused ? k_hv_block_stencil<R, b, nt, true> : k_hv_block_stencil<R, b, nt, false>; \
  }
  BSK(2, 256) BSK(4, 256) BSK(6, 256) BSK(8, 256) BSK(10, 256) BSK(12, 256) BSK(14, 256)
  BSK(4, 512) BSK(8, 512) BSK(12, 512)
#undef BSK
  return nullptr;
}

BSFn block_stencil_kernel(int prec, int B, int NT, size_t *lds, int slabs, bool fused = false) {
  return prec == QS3X_FP64 ? block_stencil_kernel_r<double>(B, NT, lds, slabs, fused)
                           : block_stencil_kernel_r<float>(B, NT, lds, slabs, fused);
}

// --------------------------------------------------------- block update
// w_q -= a_q s_q u_q + b_q s'_q up_q (b term when HASB); part[block][q] = |w_q|^2.
// The vectors are n 16-byte chunks ([row][kQ] complex: 2 wave vectors per
// chunk in FP32, 1 in FP64), swept in contiguous tiles of kBThreads x kU
// chunks per block as the single-vector update does.  A row is a whole
// number of chunks and a tile a whole number of rows' worth of lanes, so a
// thread always holds the same chunk column, i.e. the same wave vectors.
constexpr int kBU = 4;
template <class R, bool HASB>
__global__ __launch_bounds__(kBThreads) void k_update3_block(typename Chunk16<R>::T *w, const typename Chunk16<R>::T *u,
                                                             const typename Chunk16<R>::T *up, int64_t n,
                                                             const double *sc, int ia, int isa, int ib, int isb,
                                                             double *part) {
  using V = typename Chunk16<R>::T;
  constexpr int NQ = Chunk16<R>::nq;
  constexpr int RV = kQ / NQ;  // chunks per row
  static_assert(kBThreads % RV == 0, "a thread keeps its chunk column");
  const int col = threadIdx.x % RV;
  R ca[NQ], cb[NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    const int q = col * NQ + k;
    ca[k] = (R)(sc[(size_t)ia * kQ + q] * (isa >= 0 ? sc[(size_t)isa * kQ + q] : 1.0));
    cb[k] = HASB ? (R)(sc[(size_t)ib * kQ + q] * (isb >= 0 ? sc[(size_t)isb * kQ + q] : 1.0)) : (R)0;
  }
  double acc[NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) acc[k] = 0.0;
  auto one = [&](V x, V a, V b, int64_t i) __attribute__((always_inline)) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      x[2 * k] -= ca[k] * a[2 * k];
      x[2 * k + 1] -= ca[k] * a[2 * k + 1];
      if (HASB) {
        x[2 * k] -= cb[k] * b[2 * k];
        x[2 * k + 1] -= cb[k] * b[2 * k + 1];
      }
      acc[k] += (double)x[2 * k] * (double)x[2 * k] + (double)x[2 * k + 1] * (double)x[2 * k + 1];
    }
    __builtin_nontemporal_store(x, &w[i]);
  };
  constexpr int TILE = kBThreads * kBU;
  const int64_t nfull = n / TILE;
  for (int64_t t = blockIdx.x; t < nfull; t += gridDim.x) {
    const int64_t b0 = t * TILE + threadIdx.x;
    V x[kBU], a[kBU], b[kBU];
#pragma unroll
    for (int q = 0; q < kBU; ++q) {
      x[q] = __builtin_nontemporal_load(&w[b0 + q * kBThreads]);
      a[q] = __builtin_nontemporal_load(&u[b0 + q * kBThreads]);
      if (HASB) b[q] = __builtin_nontemporal_load(&up[b0 + q * kBThreads]);
      else b[q] = a[q];
    }
#pragma unroll
    for (int q = 0; q < kBU; ++q) one(x[q], a[q], b[q], b0 + q * kBThreads);
  }
  if (nfull * TILE < n && (int64_t)blockIdx.x == nfull % gridDim.x)
    for (int64_t i = nfull * TILE + threadIdx.x; i < n; i += kBThreads) one(w[i], u[i], HASB ? up[i] : u[i], i);
  __shared__ double red[kBThreads / RV][kQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) red[threadIdx.x / RV][col * NQ + k] = acc[k];
  __syncthreads();
  if (threadIdx.x < kQ) {
    double t = 0;
    for (int k = 0; k < kBThreads / RV; ++k) t += red[k][threadIdx.x];
    part[2 * ((size_t)blockIdx.x * kQ + threadIdx.x)] = t;
    part[2 * ((size_t)blockIdx.x * kQ + threadIdx.x) + 1] = 0.0;
  }
}

// Per-q finalize (one block per q), modes as k_finalize in blas.hip:
// 0: sc[i0] = re; 1: sc[i0] = re, sc[i1] = re + shift; 2: b = sqrt(re): sc[i0] = b, sc[i1] = 1/b (0 if b = 0)
__global__ __launch_bounds__(kBThreads) void k_finalize_q(const double *part, int nb, double *sc, int mode, int i0,
                                                          int i1, double shift) {
  const int q = blockIdx.x;
  double re = 0;
  for (int b = threadIdx.x; b < nb; b += kBThreads) re += part[2 * ((size_t)b * kQ + q)];
  __shared__ double red[kBThreads];
  red[threadIdx.x] = re;
  __syncthreads();
  for (int s = kBThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double v = red[0];
    if (mode == 0) {
      sc[(size_t)i0 * kQ + q] = v;
    } else if (mode == 1) {
      sc[(size_t)i0 * kQ + q] = v;
      if (i1 >= 0) sc[(size_t)i1 * kQ + q] = v + shift;
    } else {
      const double b = sqrt(v > 0 ? v : 0.0);
      sc[(size_t)i0 * kQ + q] = b;
      if (i1 >= 0) sc[(size_t)i1 * kQ + q] = b > 0 ? 1.0 / b : 0.0;
    }
  }
}

// Per-q finalize of the fused step (one block per q): from the four sums,
// alpha = Re<v,Hv>, beta^2 = |P|^2 - 2 alpha Re<v,P> + alpha^2 |v|^2 (the
// norm of P - alpha v); sc[ia] = alpha, sc[io] = alpha + shift, sc[ib] = beta,
// sc[is] = 1 / beta (0 if beta = 0).
__global__ __launch_bounds__(kBThreads) void k_finalize_q4(const double *part, int nb, double *sc, int ia, int io,
                                                           int ib, int is, double shift) {
  const int q = blockIdx.x;
  double t[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nb; b += kBThreads)
#pragma unroll
    for (int k = 0; k < 4; ++k) t[k] += part[4 * ((size_t)b * kQ + q) + k];
  __shared__ double red[4][kBThreads];
#pragma unroll
  for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = t[k];
  __syncthreads();
  for (int s = kBThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
#pragma unroll
      for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double al = red[0][0], pp = red[1][0], vp = red[2][0], vv = red[3][0];
    const double b2 = pp - 2.0 * al * vp + al * al * vv;
    const double b = sqrt(b2 > 0 ? b2 : 0.0);
    sc[(size_t)ia * kQ + q] = al;
    sc[(size_t)io * kQ + q] = al + shift;
    sc[(size_t)ib * kQ + q] = b;
    sc[(size_t)is * kQ + q] = b > 0 ? 1.0 / b : 0.0;
  }
}

// ------------------------------------------------------------ block S^-_q
// For target row lp (config m, NT down spins, an all-k chain row) and wave
// vector q: f_q(lp) = r_q(m) / sqrt(NOS) sum over the distinct rotations y of
// m that put a down spin on site 1, and over the down spins i of y with
// y \ {i} a ground-state representative l:
//     psi(l) / r_gs(l) * exp(-i q R_i) * conj(explist_q(ell_y)),
// as sq_accumulate / k_sq (observables.hip) evaluate calcu_Sq_a for S^-.
template <int NT, class R>
__global__ __launch_bounds__(kBThreads) void k_sminus_block(SecView gs, SecView tg, const double2 *__restrict__ psi,
                                                            const double *rq,    // [NT+1][kQ] r_q(c)
                                                            const double *eqr,   // [L][kQ] complex exp(-i q R_i)
                                                            const double *eph,   // [L+1][kQ] complex explist_q
                                                            typename CT<R>::T *F, double norm, double *part) {
  using C = typename CT<R>::T;
  constexpr int T = NT * NT;  // terms per row at most
  constexpr int CH = NT <= 4 ? 64 : (NT <= 6 ? 32 : 16);
  __shared__ int32_t s_l[kBWaves][CH][T];
  __shared__ int32_t s_aux[kBWaves][CH][T];  // site i | ell << 16
  __shared__ int32_t s_cnt[kBWaves][CH];
  __shared__ int32_t s_nt[kBWaves][CH];
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int L = tg.L;
  double acc2 = 0;
  const int64_t nchunks = (tg.dim + CH - 1) / CH;
  for (int64_t ch = (int64_t)blockIdx.x * kBWaves + wid; ch < nchunks; ch += (int64_t)gridDim.x * kBWaves) {
    const int64_t base = ch * CH;
    if (lane < CH) {
      const int64_t lp = base + lane;
      int nt = 0;
      int cnt = 0;
      if (lp < tg.dim) {
        int m[kMaxNod], y[kMaxNod], yp[kMaxNod], src[kMaxNod];
        sv_unpack(tg, lp, m);
        cnt = (int)(tg.cfg[lp] >> 56);
        for (int j = 0; j < NT; ++j) {
          for (int q = 0; q < NT; ++q) {
            const int s = m[(j + q) % NT] - m[j];
            y[q] = 1 + (s < 0 ? s + L : s);
          }
          bool seen = false;
          for (int j2 = 0; j2 < j && !seen; ++j2) {
            bool same = true;
            for (int q = 0; q < NT; ++q) {
              const int s = m[(j2 + q) % NT] - m[j2];
              yp[q] = 1 + (s < 0 ? s + L : s);
              same = same && (yp[q] == y[q]);
            }
            seen = same;
          }
          if (seen) continue;
          int ei = 0, isrep = 0;
          double rt = 0;
          (void)sv_lookup(tg, y, &ei, &isrep, &rt);
          const int ell = ei ? ei - 3 * (L + 1) : 0;
          for (int r = 0; r < NT; ++r) {
            int p = 0;
            for (int q = 0; q < NT; ++q)
              if (q != r) src[p++] = y[q];
            int gi = 0, grep = 0;
            double grt = 0;
            const int64_t l = sv_lookup(gs, src, &gi, &grep, &grt);
            if (l < 0 || !grep) continue;
            s_l[wid][lane][nt] = (int32_t)l;
            s_aux[wid][lane][nt] = y[r] | (ell << 16);
            ++nt;
          }
        }
      }
      s_nt[wid][lane] = nt;
      s_cnt[wid][lane] = cnt;
    }
    __builtin_amdgcn_wave_barrier();
    const int nr = (int)min<int64_t>(CH, tg.dim - base);
    for (int r = 0; r < nr; ++r) {
      const int nt = s_nt[wid][r];
      double ar = 0, ai = 0;
      for (int t = 0; t < nt; ++t) {
        const int64_t l = s_l[wid][r][t];
        const int i = s_aux[wid][r][t] & 0xffff, ell = s_aux[wid][r][t] >> 16;
        const double2 ps = psi[l];
        const double c = 1.0 / gs.list_r[l];
        const double wr = eqr[2 * ((size_t)(i - 1) * kQ + lane)], wi = eqr[2 * ((size_t)(i - 1) * kQ + lane) + 1];
        const double er = eph[2 * ((size_t)ell * kQ + lane)], eim = -eph[2 * ((size_t)ell * kQ + lane) + 1];
        const double br = c * (ps.x * wr - ps.y * wi), bi = c * (ps.x * wi + ps.y * wr);
        ar += br * er - bi * eim;
        ai += br * eim + bi * er;
      }
      const double rt = rq[(size_t)s_cnt[wid][r] * kQ + lane] * norm;
      const double fr = ar * rt, fi = ai * rt;
      C o;
      o.x = (R)fr;
      o.y = (R)fi;
      F[(base + r) * kQ + lane] = o;
      acc2 += fr * fr + fi * fi;
    }
    __builtin_amdgcn_wave_barrier();
  }
  __shared__ double red[kBWaves][kQ];
  red[wid][lane] = acc2;
  __syncthreads();
  if (wid == 0) {
    double s = 0;
    for (int k = 0; k < kBWaves; ++k) s += red[k][lane];
    part[2 * ((size_t)blockIdx.x * kQ + lane)] = s;
    part[2 * ((size_t)blockIdx.x * kQ + lane) + 1] = 0.0;
  }
}

using SFn = void (*)(SecView, SecView, const double2 *, const double *, const double *, const double *, void *,
                     double, double *);

template <int NT>
void launch_sminus(int prec, int grid, hipStream_t st, const SecView &gs, const SecView &tg, const double2 *psi,
                   const double *rq, const double *eqr, const double *eph, void *F, double norm, double *part) {
  if (prec == QS3X_FP64)
    hipLaunchKernelGGL((k_sminus_block<NT, double>), dim3(grid), dim3(kBThreads), 0, st, gs, tg, psi, rq, eqr, eph,
                       (double2 *)F, norm, part);
  else
    hipLaunchKernelGGL((k_sminus_block<NT, float>), dim3(grid), dim3(kBThreads), 0, st, gs, tg, psi, rq, eqr, eph,
                       (float2 *)F, norm, part);
}

// The last block-stencil tuning: rows (lines, threads, slabs, ms), the
// chosen shape last.
std::vector<std::array<double, 4>> &block_tune_table() {
  static std::vector<std::array<double, 4>> t;
  return t;
}

int cus() {
  static int c = 0;
  if (!c) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device()) == hipSuccess) c = p.multiProcessorCount;
    if (c <= 0) c = 256;
  }
  return c;
}

}  // namespace

}  // namespace qs3x

using namespace qs3x;

extern "C" {

int qs3x_block_tune_last(int32_t max_rows, int32_t *n_rows, double *rows) {
  return guarded4([&] {
    QS3X_REQUIRE(n_rows && (rows || max_rows == 0) && max_rows >= 0, "qs3x_block_tune_last: bad argument");
    const auto &t = block_tune_table();
    *n_rows = (int32_t)std::min<size_t>(t.size(), (size_t)max_rows);
    for (int32_t i = 0; i < *n_rows; ++i)
      for (int k = 0; k < 4; ++k) rows[4 * i + k] = t[i][k];
  });
}

int qs3x_sector_build_all_k(qs3x_model *m, int32_t nod, qs3x_sector **out) {
  return guarded4([&] {
    QS3X_REQUIRE(m && out, "qs3x_sector_build_all_k: NULL argument");
    QS3X_REQUIRE(m->chain, "qs3x_sector_build_all_k: chains only (LY = LZ = 1)");
    auto *s = new qs3x_sector();
    s->model = m;
    s->nod = nod;
    s->all_k = true;
    try {
      build_sector_device(s);
    } catch (...) {
      delete s;
      throw;
    }
    if (!s->chain) {
      delete s;
      throw Error{"qs3x_sector_build_all_k: the chain enumeration does not apply to this N_down / L"};
    }
    *out = s;
  });
}

int qs3x_dsf_sminus_batch(qs3x_sector *gs, int32_t nq, const int32_t *K, const double *pos, const double *psi,
                          int32_t itr_dsf, double rfield, int32_t prec, double *alpha, double *beta2,
                          double *norm1, double *stats) {
  return guarded4([&] {
    QS3X_REQUIRE(!is_u1(gs), "qs3x_dsf_sminus_batch: not available for U(1)-only sectors");
    QS3X_REQUIRE(gs && K && pos && psi && alpha && beta2 && norm1 && itr_dsf >= 1,
                 "qs3x_dsf_sminus_batch: bad argument");
    QS3X_REQUIRE(nq >= 1 && nq <= kQ, "qs3x_dsf_sminus_batch: 1 <= nq <= 64");
    QS3X_REQUIRE(prec == QS3X_FP64 || prec == QS3X_FP32, "qs3x_dsf_sminus_batch: bad precision");
    qs3x_model *m = gs->model;
    QS3X_REQUIRE(m->chain && m->nn_chain, "qs3x_dsf_sminus_batch: nearest-neighbour chains only");
    QS3X_REQUIRE(gs->chain && gs->mode == kIndexAnchored && gs->d_rclass.p,
                 "qs3x_dsf_sminus_batch: the ground-state sector must be a chain sector built on the device");
    const int L = m->NOS, NT = gs->nod + 1;
    QS3X_REQUIRE(NT >= 2 && NT <= 8, "qs3x_dsf_sminus_batch: target N_down must be 2..8");
    QS3X_REQUIRE(L < 4096, "qs3x_dsf_sminus_batch: L < 4096");
    hipStream_t st = stream();
    const double t0 = now_s();
    // all-k target sector
    qs3x_sector tgt;
    tgt.model = m;
    tgt.nod = NT;
    tgt.all_k = true;
    build_sector_device(&tgt);
    QS3X_REQUIRE(tgt.chain && tgt.d_bitmap.p, "qs3x_dsf_sminus_batch: the target is not an anchored chain sector");
    const int64_t dim = tgt.dim;
    // per-q tables (padding q >= nq: r_q = 0, so those vectors stay zero)
    std::vector<double> rq((size_t)(NT + 1) * kQ, 0.0), ratio((size_t)(NT + 1) * (NT + 1) * kQ, 0.0),
        mask((size_t)(NT + 1) * kQ, 0.0), eph((size_t)2 * (L + 1) * kQ, 0.0), eqr((size_t)2 * L * kQ, 0.0);
    for (int q = 0; q < kQ; ++q) {
      eph[2 * ((size_t)0 * kQ + q)] = 1.0;
      if (q >= nq) continue;
      const double rk[3] = {qs3x_wavevector(K[q], L), 0.0, 0.0};
      for (int c = 1; c <= NT; ++c) {
        const double ro = chain_ro(L, c, rk);
        rq[(size_t)c * kQ + q] = ro > 1.0e-15 ? std::sqrt(ro) : 0.0;
        mask[(size_t)c * kQ + q] = ro > 1.0e-15 ? 1.0 : 0.0;
      }
      for (int ct = 0; ct <= NT; ++ct)
        for (int cr = 0; cr <= NT; ++cr) {
          const double a = rq[(size_t)ct * kQ + q], b = rq[(size_t)cr * kQ + q];
          ratio[((size_t)ct * (NT + 1) + cr) * kQ + q] = b > 0 ? a / b : 0.0;
        }
      std::vector<double> ex;
      compute_explist(m, rk, ex);
      for (int l = 1; l <= L; ++l) {
        eph[2 * ((size_t)l * kQ + q)] = ex[2 * (l + 3 * (L + 1))];
        eph[2 * ((size_t)l * kQ + q) + 1] = ex[2 * (l + 3 * (L + 1)) + 1];
      }
      // exp(-i q . R_i), q the transfer from the ground-state sector
      const double qx = rk[0] - gs->rk[0], qy = -gs->rk[1], qz = -gs->rk[2];
      for (int i = 0; i < L; ++i) {
        const double x = qx * pos[3 * i] + qy * pos[3 * i + 1] + qz * pos[3 * i + 2];
        const std::complex<double> e = std::exp(std::complex<double>(0.0, -x));
        eqr[2 * ((size_t)i * kQ + q)] = e.real();
        eqr[2 * ((size_t)i * kQ + q) + 1] = e.imag();
      }
    }
    DevBuf<double> d_rq(rq.size()), d_ratio(ratio.size()), d_mask(mask.size()), d_eph(eph.size()),
        d_eqr(eqr.size()), d_psi((size_t)2 * gs->dim);
    // the phase table in the block H.v's precision
    DevBuf<unsigned char> d_ephR(prec == QS3X_FP64 ? sizeof(double) * eph.size() : sizeof(float) * eph.size());
    if (prec == QS3X_FP64) {
      QS3X_HIP(hipMemcpyAsync(d_ephR.p, eph.data(), sizeof(double) * eph.size(), hipMemcpyHostToDevice, st));
    } else {
      std::vector<float> ef(eph.begin(), eph.end());
      QS3X_HIP(hipMemcpyAsync(d_ephR.p, ef.data(), sizeof(float) * ef.size(), hipMemcpyHostToDevice, st));
      QS3X_HIP(hipStreamSynchronize(st));
    }
    d_rq.upload(rq.data(), rq.size());
    d_ratio.upload(ratio.data(), ratio.size());
    d_mask.upload(mask.data(), mask.size());
    d_eph.upload(eph.data(), eph.size());
    d_eqr.upload(eqr.data(), eqr.size());
    d_psi.upload(psi, (size_t)2 * gs->dim);
    const size_t es = prec == QS3X_FP64 ? 16 : 8;
    // Fused step (option block_fused, block-stencil sectors): buffers P, v
    // and their successors; otherwise u_{i-1}, u_i, w.
    bool want_fused = options().block_fused && options().block_stencil && NT == 4 && stencil_applies(&tgt);
    if (want_fused) {
      // the fourth buffer only where it fits beside the other three (1 GiB
      // spare); otherwise the separate block stencil and update
      size_t fr = 0, tot = 0;
      if (hipMemGetInfo(&fr, &tot) == hipSuccess && fr < 4 * es * (size_t)dim * kQ + ((size_t)1 << 30))
        want_fused = false;
    }
    DevBuf<unsigned char> U[4];
    for (int k = 0; k < (want_fused ? 4 : 3); ++k) U[k].alloc(es * (size_t)dim * kQ);
    const int steps = (int)std::min<int64_t>(itr_dsf, dim);
    const int NS = steps + 2;
    // per-q scalar slots [slot][kQ]: arec [NS] | aout [NS] | beta [NS] | scale [NS+1] | norm1, tmp
    const int A_REC = 0, A_OUT = NS, BETA = 2 * NS, SCALE = 3 * NS, NORM1 = 4 * NS + 2, TMP = 4 * NS + 3;
    DevBuf<double> sc((size_t)(4 * NS + 8) * kQ);
    const int gridS = cus() * 2, gridU = cus() * 8;
    // per-block partial sums: 2 per q (4 for the fused step), at most 2
    // blocks per CU for a stencil launch plus the per-row kernel's grid
    DevBuf<double> part((size_t)4 * kQ * std::max(gridU, gridS * 4));
    QS3X_HIP(hipStreamSynchronize(st));
    const double t1 = now_s();
    // f_q = S^-_q psi; norm1_q = |f_q|^2 (FP64); scale[0] = 1 / sqrt(norm1)
    {
      const SecView vg = make_view(gs), vt = make_view(&tgt);
      const double nrm = 1.0 / std::sqrt((double)L);
      switch (NT) {
        case 2: launch_sminus<2>(prec, gridS, st, vg, vt, (const double2 *)d_psi.p, d_rq.p, d_eqr.p, d_eph.p, U[0].p, nrm, part.p); break;
        case 3: launch_sminus<3>(prec, gridS, st, vg, vt, (const double2 *)d_psi.p, d_rq.p, d_eqr.p, d_eph.p, U[0].p, nrm, part.p); break;
        case 4: launch_sminus<4>(prec, gridS, st, vg, vt, (const double2 *)d_psi.p, d_rq.p, d_eqr.p, d_eph.p, U[0].p, nrm, part.p); break;
        case 5: launch_sminus<5>(prec, gridS, st, vg, vt, (const double2 *)d_psi.p, d_rq.p, d_eqr.p, d_eph.p, U[0].p, nrm, part.p); break;
        case 6: launch_sminus<6>(prec, gridS, st, vg, vt, (const double2 *)d_psi.p, d_rq.p, d_eqr.p, d_eph.p, U[0].p, nrm, part.p); break;
        case 7: launch_sminus<7>(prec, gridS, st, vg, vt, (const double2 *)d_psi.p, d_rq.p, d_eqr.p, d_eph.p, U[0].p, nrm, part.p); break;
        default: launch_sminus<8>(prec, gridS, st, vg, vt, (const double2 *)d_psi.p, d_rq.p, d_eqr.p, d_eph.p, U[0].p, nrm, part.p); break;
      }
      QS3X_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_finalize_q, dim3(kQ), dim3(kBThreads), 0, st, part.p, gridS, sc.p, 0, NORM1, -1, 0.0);
      hipLaunchKernelGGL(k_finalize_q, dim3(kQ), dim3(kBThreads), 0, st, part.p, gridS, sc.p, 2, TMP, SCALE, 0.0);
    }
    QS3X_HIP(hipStreamSynchronize(st));
    const double t2 = now_s();
    // recurrence
    int kb = 1;
    while ((1 << kb) <= L) ++kb;
    const bool key32 = (int64_t)(NT - 1) * kb <= 32;
    size_t lds = 0;
    BFn hv = block_kernel(NT, prec, key32, &lds, L, (int)m->bt_jxy2.size());
    QS3X_REQUIRE(hv, "qs3x_dsf_sminus_batch: no block kernel for this N_down");
    // 160 KB per CU less the kernel's static LDS (its reduction arrays) and a margin
    QS3X_REQUIRE(lds <= 156 * 1024, "qs3x_dsf_sminus_batch: L too large for the block kernel's LDS tables "
                                    "(over 156 KB)");
    if (lds > 64 * 1024)
      QS3X_HIP(hipFuncSetAttribute((const void *)hv, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int nbpc = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbpc, hv, kBThreads, lds) != hipSuccess || nbpc < 1) nbpc = 1;
    const int gridH = (int)std::min<int64_t>((int64_t)cus() * nbpc, ((dim + 63) / 64 + kBWaves - 1) / kBWaves);
    QS3X_REQUIRE(gridH * kQ * 2 <= (int64_t)part.n, "qs3x_dsf_sminus_batch: partial buffer");
    BlockHvArgs a{};
    a.cfg = tgt.d_cfg.p;
    a.bitmap = tgt.d_bitmap.p;
    a.dim = dim;
    a.L = L;
    a.bits = tgt.bits;
    a.kbits = kb;
    a.binom = tgt.d_binom.p;
    a.kdim = tgt.nbinom_k;
    a.nn_bt = m->d_nn_bt.p;
    a.bt = m->d_bt.p;
    a.nbt = (int)m->bt_jxy2.size();
    a.shift = prec == QS3X_FP64 ? m->ezz : 0.0;
    a.sc = sc.p;
    a.ratio = d_ratio.p;
    a.rowmask = d_mask.p;
    a.eph = d_ephR.p;
    a.dot_partials = part.p;
    a.nxcd = options().hv_xcd > 0 ? options().hv_xcd : 8;  // per-XCD ranges unless set
    // Clear rows through the block stencil where it applies, the others
    // (the sector's non-clear rows, about 1 % at L = 1000) through the
    // per-row block kernel; their partial sums follow the stencil's.
    BSFn hvs = nullptr, hvsF = nullptr;
    size_t lds_s = 0;
    int gridSt = 0, gridSl = 0, ntasks = 0, ntS = 256, gridStF = 0;
    DevBuf<int4> st_tasks;
    BlockStencilArgs sarg{};
    std::vector<int4> st_lines;
    const bool use_stencil = options().block_stencil && stencil_applies(&tgt);
    // one tile shape: kernel, task list, grid
    auto setup_stencil = [&](int B, int nt, int slabs) {
      ntS = nt;
      hvs = block_stencil_kernel(prec, B, ntS, &lds_s, slabs);
      QS3X_REQUIRE(hvs, "qs3x_dsf_sminus_batch: no block stencil kernel for these lines per tile and threads "
                        "(256 threads: 2..14 lines, even; 512: 4, 8 or 12)");
      QS3X_REQUIRE(lds_s <= 156 * 1024, "qs3x_dsf_sminus_batch: block stencil tile over 156 KB of LDS "
                                        "(fewer lines or slabs)");
      if (st_lines.empty()) stencil_tables(&tgt, &st_lines);
      const int groups = prec == QS3X_FP64 ? kQ / 8 : kQ / 16;
      std::vector<int4> tasks = stencil_tasks(st_lines, L, kSA, B, slabs, groups);
      ntasks = (int)tasks.size();
      st_tasks.alloc(std::max<size_t>(tasks.size(), 1));
      if (!tasks.empty()) st_tasks.upload(tasks.data(), tasks.size());
      if (lds_s > 64 * 1024)
        QS3X_HIP(hipFuncSetAttribute((const void *)hvs, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_s));
      // a multiple of the 8 XCDs: block g takes tasks 8 m + g % 8, its XCD's tiles
      auto grid_of = [&](BSFn k) {
        int bpc = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k, ntS, lds_s) != hipSuccess || bpc < 1) bpc = 1;
        const int gr = (int)std::min<int64_t>((int64_t)cus() * bpc, std::max(ntasks, 8));
        return std::max(8, gr / 8 * 8);
      };
      gridSt = grid_of(hvs);
      if (want_fused) {
        size_t lds_f = 0;
        hvsF = block_stencil_kernel(prec, B, ntS, &lds_f, slabs, true);
        QS3X_REQUIRE(hvsF && lds_f == lds_s, "qs3x_dsf_sminus_batch: no fused block stencil kernel");
        if (lds_s > 64 * 1024)
          QS3X_HIP(hipFuncSetAttribute((const void *)hvsF, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_s));
        gridStF = grid_of(hvsF);
      }
      gridSl = tgt.nslow > 0 ? (int)std::min<int64_t>((int64_t)cus() * nbpc,
                                                      ((tgt.nslow + 63) / 64 + kBWaves - 1) / kBWaves)
                             : 0;
      sarg = BlockStencilArgs{tgt.d_lines.p, st_tasks.p, ntasks};
      QS3X_REQUIRE((int64_t)(gridSt + gridSl) * kQ * 2 <= (int64_t)part.n,
                   "qs3x_dsf_sminus_batch: partial buffer (stencil)");
      QS3X_REQUIRE((int64_t)(gridStF + gridSl) * kQ * 4 <= (int64_t)part.n,
                   "qs3x_dsf_sminus_batch: partial buffer (fused stencil)");
    };
    if (use_stencil) {
      // defaults: the fastest measured at L = 1000, N_down = 4 in both
      // precisions (12 lines, 512 threads, 32 slabs: FP32 9.9 ms, FP64 20.5 ms
      // per launch, against 12.3 and 25.2 ms with 6 lines, 256 threads, 8
      // slabs); the fused step: 8 lines (12 spill registers at 512 threads;
      // FP32 20.4 ms, FP64 43.7 ms per step against 22.8 and 48.4 with 4)
      setup_stencil(options().block_stencil_lines > 0 ? options().block_stencil_lines : (want_fused ? 8 : 12),
                    options().block_stencil_threads > 0 ? options().block_stencil_threads : 512,
                    options().block_stencil_slabs > 0 ? options().block_stencil_slabs : 32);
    }
    const double sztot_new = (double)L / 2.0 - (double)NT;
    const double c = (prec == QS3X_FP32) ? m->ezz : 0.0;
    hipEvent_t e0, e1;
    QS3X_HIP(hipEventCreate(&e0));
    QS3X_HIP(hipEventCreate(&e1));
    // Option block_autotune: time one block stencil launch per tile shape
    // (lines, threads, slabs) on the S^- vectors and keep the fastest for the
    // recurrence (qs3x_block_tune_last reports the table).  With the fused
    // step the fused kernel is timed, reading the S^- vectors as both P and v
    // and writing two scratch buffers.
    block_tune_table().clear();
    if (use_stencil && options().block_autotune) {
      static const int kCand[][3] = {{12, 512, 32}, {12, 512, 16}, {12, 512, 64}, {8, 512, 32}, {4, 512, 32},
                                     {6, 256, 32}, {8, 256, 32}, {14, 256, 32}, {10, 256, 16}, {6, 256, 8}};
      double best_ms = 1e300;
      int bi = 0;
      for (int k = 0; k < (int)(sizeof(kCand) / sizeof(kCand[0])); ++k) {
        setup_stencil(kCand[k][0], kCand[k][1], kCand[k][2]);
        BlockHvArgs t = a;
        t.x = U[0].p;
        t.y = U[1].p;
        t.iscale = SCALE;
        BSFn kt = hvs;
        int gt = gridSt;
        if (hvsF) {
          t.iscale = -1;
          t.shift = 0.0;
          t.vprev = U[0].p;
          t.xout = U[2].p;
          t.fis = t.fia = t.fib = SCALE;
          kt = hvsF;
          gt = gridStF;
        }
        hipLaunchKernelGGL(kt, dim3(gt), dim3(ntS), lds_s, st, t, sarg);  // warm-up
        QS3X_HIP(hipGetLastError());
        QS3X_HIP(hipEventRecord(e0, st));
        for (int r = 0; r < 2; ++r) hipLaunchKernelGGL(kt, dim3(gt), dim3(ntS), lds_s, st, t, sarg);
        QS3X_HIP(hipEventRecord(e1, st));
        QS3X_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        QS3X_HIP(hipEventElapsedTime(&ms, e0, e1));
        const double per = ms / 2.0;
        block_tune_table().push_back({(double)kCand[k][0], (double)kCand[k][1], (double)kCand[k][2], per});
        if (per < 0.98 * best_ms) {
          best_ms = per;
          bi = k;
        }
      }
      setup_stencil(kCand[bi][0], kCand[bi][1], kCand[bi][2]);
      block_tune_table().push_back({(double)kCand[bi][0], (double)kCand[bi][1], (double)kCand[bi][2], best_ms});
    }
    float hv_ms_total = 0;
    const bool fused = hvsF != nullptr;
    size_t ldsF = 0;
    BFn hvF = fused ? block_kernel_fused(NT, prec, key32, &ldsF, L, (int)m->bt_jxy2.size()) : nullptr;
    QS3X_REQUIRE(!fused || (hvF && ldsF == lds), "qs3x_dsf_sminus_batch: no fused per-row block kernel");
    if (fused && ldsF > 64 * 1024)
      QS3X_HIP(hipFuncSetAttribute((const void *)hvF, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsF));
    // The fused step (lanczos.f90:50-92 for each q): step i reads P_{i-1}
    // and v_{i-1} (step 0: the S^- vectors f and scale[0] = 1/|f|), writes
    // v_i and P_i = (H - e_zz) v_i - beta_{i-1} v_{i-1}; alpha_i and beta_i
    // from its four per-q sums.  Both precisions apply H - e_zz (beta^2 =
    // |P|^2 - 2 alpha Re<v,P> + alpha^2 |v|^2 keeps its digits for alpha of
    // order one), and alpha + e_zz is the recurrence's alpha.  Buffers: even
    // steps read P from U[0] and v from U[3] and write v to U[1] and P to U[2];
    // odd steps the other way round.
    for (int i = 0; fused && i < steps; ++i) {
      const bool even = (i % 2) == 0;
      BlockHvArgs f = a;
      f.shift = 0.0;
      f.iscale = -1;
      f.x = even ? U[0].p : U[2].p;
      f.vprev = i == 0 ? nullptr : (even ? U[3].p : U[1].p);
      f.xout = even ? U[1].p : U[3].p;
      f.y = even ? U[2].p : U[0].p;
      f.fis = SCALE + i;
      f.fia = i > 0 ? A_REC + i - 1 : -1;
      f.fib = i > 0 ? BETA + i - 1 : -1;
      f.dot_partials = part.p;
      if (i == 1) QS3X_HIP(hipEventRecord(e0, st));
      hipLaunchKernelGGL(hvsF, dim3(gridStF), dim3(ntS), lds_s, st, f, sarg);
      QS3X_HIP(hipGetLastError());
      int nbh = gridStF;
      if (gridSl > 0) {
        BlockHvArgs b = f;
        b.rows = tgt.d_slow.p;
        b.nrows = tgt.nslow;
        b.dot_partials = part.p + (size_t)4 * kQ * gridStF;
        hipLaunchKernelGGL(hvF, dim3(gridSl), dim3(kBThreads), ldsF, st, b);
        QS3X_HIP(hipGetLastError());
        nbh += gridSl;
      }
      hipLaunchKernelGGL(k_finalize_q4, dim3(kQ), dim3(kBThreads), 0, st, part.p, nbh, sc.p, A_REC + i, A_OUT + i,
                         BETA + i, SCALE + i + 1, m->ezz - rfield * sztot_new);
      QS3X_HIP(hipGetLastError());
    }
    for (int i = 0; !fused && i < steps; ++i) {
      void *cur = U[i % 3].p, *prev = U[(i + 2) % 3].p, *w = U[(i + 1) % 3].p;
      a.x = cur;
      a.y = w;
      a.iscale = SCALE + i;
      if (i == 1) QS3X_HIP(hipEventRecord(e0, st));
      int nbh = gridH;
      if (hvs) {
        hipLaunchKernelGGL(hvs, dim3(gridSt), dim3(ntS), lds_s, st, a, sarg);
        QS3X_HIP(hipGetLastError());
        nbh = gridSt;
        if (gridSl > 0) {
          BlockHvArgs b = a;
          b.rows = tgt.d_slow.p;
          b.nrows = tgt.nslow;
          b.dot_partials = part.p + (size_t)2 * kQ * gridSt;
          hipLaunchKernelGGL(hv, dim3(gridSl), dim3(kBThreads), lds, st, b);
          QS3X_HIP(hipGetLastError());
          nbh += gridSl;
        }
      } else {
        hipLaunchKernelGGL(hv, dim3(gridH), dim3(kBThreads), lds, st, a);
        QS3X_HIP(hipGetLastError());
      }
      hipLaunchKernelGGL(k_finalize_q, dim3(kQ), dim3(kBThreads), 0, st, part.p, nbh, sc.p, 1, A_REC + i,
                         A_OUT + i, c - rfield * sztot_new);
      if (i + 1 == steps) break;
      {
        const int64_t nch = dim * (prec == QS3X_FP64 ? kQ : kQ / 2);
        const int ib = i > 0 ? BETA + i - 1 : -1, isb = i > 0 ? SCALE + i - 1 : -1;
        if (prec == QS3X_FP64) {
          using V = Chunk16<double>::T;
          if (i > 0)
            hipLaunchKernelGGL((k_update3_block<double, true>), dim3(gridU), dim3(kBThreads), 0, st, (V *)w,
                               (const V *)cur, (const V *)prev, nch, sc.p, A_REC + i, SCALE + i, ib, isb, part.p);
          else
            hipLaunchKernelGGL((k_update3_block<double, false>), dim3(gridU), dim3(kBThreads), 0, st, (V *)w,
                               (const V *)cur, (const V *)prev, nch, sc.p, A_REC + i, SCALE + i, ib, isb, part.p);
        } else {
          using V = Chunk16<float>::T;
          if (i > 0)
            hipLaunchKernelGGL((k_update3_block<float, true>), dim3(gridU), dim3(kBThreads), 0, st, (V *)w,
                               (const V *)cur, (const V *)prev, nch, sc.p, A_REC + i, SCALE + i, ib, isb, part.p);
          else
            hipLaunchKernelGGL((k_update3_block<float, false>), dim3(gridU), dim3(kBThreads), 0, st, (V *)w,
                               (const V *)cur, (const V *)prev, nch, sc.p, A_REC + i, SCALE + i, ib, isb, part.p);
        }
      }
      QS3X_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_finalize_q, dim3(kQ), dim3(kBThreads), 0, st, part.p, gridU, sc.p, 2, BETA + i,
                         SCALE + i + 1, 0.0);
    }
    QS3X_HIP(hipEventRecord(e1, st));
    QS3X_HIP(hipStreamSynchronize(st));
    if (steps > 1) QS3X_HIP(hipEventElapsedTime(&hv_ms_total, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const double t3 = now_s();
    std::vector<double> all(sc.n);
    sc.download(all.data(), all.size());
    for (int q = 0; q < nq; ++q) {
      norm1[q] = all[(size_t)NORM1 * kQ + q];
      for (int i = 0; i < itr_dsf; ++i) {
        alpha[(size_t)q * itr_dsf + i] = (i < steps) ? all[(size_t)(A_OUT + i) * kQ + q] : 1.0e300;
        const double b = (i >= 1 && i < steps) ? all[(size_t)(BETA + i - 1) * kQ + q] : 0.0;
        beta2[(size_t)q * itr_dsf + i] = b * b;
      }
    }
    if (stats) {
      stats[0] = t1 - t0;  // all-k sector and tables
      stats[1] = t2 - t1;  // S^-_q
      stats[2] = t3 - t2;  // recurrence
      stats[3] = steps > 1 ? hv_ms_total / (steps - 1) : 0.0;  // ms per block step (H.v + update), steps 2..
      stats[4] = (double)dim;
      stats[5] = (double)steps;
      stats[6] = hvs ? (fused ? 2.0 : 1.0) : 0.0;        // 1: clear rows through the block stencil, 2: fused step
      stats[7] = hvs ? (double)tgt.nslow : 0.0;          // rows through the per-row block kernel then
    }
  });
}

}  // extern "C"
