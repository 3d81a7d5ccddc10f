// hv_nn.hip -- H.v for chains whose every site has exactly the two bonds
// (s-1, s) and (s, s+1) (periodic nearest-neighbour XXZ chains, any mix of
// bond types).  Same operator and results as the general chain kernels of
// hv.cpp (ham_to_vec_wave_vector, ham2vec_v3.f90:37-72); the structure of a
// nearest-neighbour hop removes most of the representative search:
//
// A row is a representative anchored at site 1: down spins n_0 = 1 < n_1 <
// ... < n_{N-1} with gaps g_q = n_{q+1} - n_q and the wrap gap
// g_{N-1} = L + 1 - n_{N-1} >= every other gap.  A hop moves one down spin by
// one site, which changes two adjacent gaps by +-1 and keeps the order.
//  * spin m >= 1: the hopped configuration still contains site 1; if its wrap
//    gap is still strictly the largest it is its own representative
//    (stabiliser order 1, ell = 0) and its anchored rank changes by one
//    binomial, C(n_m - 2, m - 1) (Pascal's rule);
//  * the anchor's hop to site 2 always leaves the wrap gap (now g_{N-1} + 1)
//    strictly largest: the representative is the translate by L - 1, again
//    of stabiliser order 1;
//  * the anchor's hop to site L, and the rare hop that makes another gap reach
//    the wrap gap, take the full backward-gap key search of the general chain
//    kernel, in a loop that is uniform over the wavefront (each lane takes its
//    next pending hop per iteration).
// Rows are then looked up in the sector's bitmap index (16 bytes per 64
// anchored ranks: presence bits + row of the first present rank), which at
// L = 1000, N_down = 4 is 41 MB instead of the 663 MB int32 table, so the
// lookups stay in L2 / MALL.  All index reads of a row are issued together,
// then all gathers.
#include <algorithm>

#include "blas.h"
#include "hv_common.h"
#include "qs3x_internal.h"

namespace qs3x {
namespace {

using namespace hvk;

// C(x, q) for 1 <= q <= N - 1, 0 <= x <= L - 2, from the LDS table.
__device__ __forceinline__ int bnm(const int32_t *s_binom, int L, int q, int x) {
  return s_binom[(q - 1) * (L - 1) + x];
}

// Representative of a sorted chain configuration nn (down-spin sites, site 1
// not necessarily occupied): anchored rank c of the representative, its
// stabiliser order, and the explist index of ell (0 when nn already is the
// representative).  Backward-gap keys as in hv.hip's chain kernels.
template <int N, class KeyT>
__device__ __forceinline__ void chain_search(const int *nn, int L, int kb, const int32_t *s_binom, int &c,
                                             int &count, int &ei) {
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
  ei = isrep ? 0 : (L + 1 - mx) + 3 * (L + 1);
  int f[N - 1];
  int sum = 0;
#pragma unroll
  for (int tt = 1; tt < N; ++tt) {
    f[tt - 1] = (int)((kmax >> (kb * (N - 1 - tt))) & fmask);
    sum += f[tt - 1];
  }
  int P = L - sum;  // gap following the anchor
  c = s_binom[P - 1];
#pragma unroll
  for (int mm = 2; mm < N; ++mm) {
    P += f[N - mm];
    c += s_binom[(mm - 1) * (L - 1) + (P - 1)];
  }
}

template <int N, class R, bool REALPH, class KeyT>
__global__ __launch_bounds__(kThreads) void k_hv_chain_nn(HvArgs a) {
  static_assert(N >= 2, "N_down = 1 uses the general chain kernel");
  using C = typename CT<R>::T;
  constexpr int H = 2 * N;  // hop slots: 2m = spin m to the right, 2m + 1 = to the left
  extern __shared__ __align__(16) unsigned char smem[];
  const int L = a.L;
  // LDS: binom [N-1][L-1] int32 | rclass [N+1] R | jxy2 [nbt] R | jz4 [nbt] R | bond type [L+1] u16
  const int nb = (N - 1) * (L - 1);
  int32_t *s_binom = (int32_t *)smem;
  R *s_rcl = (R *)(smem + (((size_t)nb * 4 + 15) & ~(size_t)15));
  R *s_jxy2 = s_rcl + (N + 1);
  R *s_jz4 = s_jxy2 + a.nbt;
  uint16_t *s_bt = (uint16_t *)(s_jz4 + a.nbt);
  for (int q = threadIdx.x; q < nb; q += kThreads) {
    const int m = q / (L - 1) + 1, x = q % (L - 1);
    s_binom[q] = (int32_t)a.binom[(int64_t)x * a.kdim + m];
  }
  for (int q = threadIdx.x; q <= N; q += kThreads) s_rcl[q] = (R)a.rclass[q];
  for (int q = threadIdx.x; q < a.nbt; q += kThreads) {
    s_jxy2[q] = (R)a.bt[2 * q];
    s_jz4[q] = (R)a.bt[2 * q + 1];
  }
  for (int q = threadIdx.x; q <= L; q += kThreads) s_bt[q] = a.nn_bt[q];
  __syncthreads();

  const C *__restrict__ v1 = (const C *)a.v1;
  C *__restrict__ v0 = (C *)a.v0;
  const ulonglong2 *__restrict__ bm = a.bitmap;
  const R scale = a.scale ? (R)(*a.scale) : (R)1;
  const int kb = a.kbits;
  const R r1 = s_rcl[1];
  double dre = 0.0, dim_ = 0.0;

  for (int64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    const int64_t row = tile * kThreads + threadIdx.x;
    if (row >= a.dim) continue;
    const uint64_t w = a.cfg[row];
    const C xin = v1[row];
    int n[N];
#pragma unroll
    for (int q = 0; q < N; ++q) n[q] = (int)((w >> (q * a.bits)) & ((1ull << a.bits) - 1ull));
    int g[N];
#pragma unroll
    for (int q = 0; q < N - 1; ++q) g[q] = n[q + 1] - n[q];
    g[N - 1] = L + 1 - n[N - 1];
    const R rinv = (R)1 / s_rcl[(int)(w >> 56)];
    int crow = 0;  // anchored rank of the row
#pragma unroll
    for (int q = 1; q < N; ++q) crow += bnm(s_binom, L, q, n[q] - 2);

    R anti = 0;
    int hc[H];
    R hcf[H];
    int hei[REALPH ? 1 : H];
    uint32_t slow = 0;
#pragma unroll
    for (int m = 0; m < N; ++m) {
      const int s = n[m];
      {  // spin m to s + 1 over bond (s, s + 1)
        const int h = 2 * m;
        const bool ok = g[m] > 1;
        const int bt = s_bt[s];
        anti += ok ? s_jz4[bt] : (R)0;
        R cf = s_jxy2[bt] * rinv;
        int c, ei = 0;
        if (m == 0) {
          c = 0;
#pragma unroll
          for (int q = 1; q < N; ++q) c += bnm(s_binom, L, q, max(n[q] - 3, 0));
          ei = (L - 1) + 3 * (L + 1);
          cf *= r1;
        } else {
          int mo = 0;
#pragma unroll
          for (int q = 0; q < N - 1; ++q) mo = max(mo, g[q] + (q == m - 1 ? 1 : 0) - (q == m ? 1 : 0));
          const int wrap = g[N - 1] - (m == N - 1 ? 1 : 0);
          c = crow + (m == 1 ? 1 : bnm(s_binom, L, m - 1, n[m] - 2));
          if (wrap > mo) cf *= r1;
          else if (ok) slow |= 1u << h;
        }
        hc[h] = ok ? c : -1;
        hcf[h] = cf;
        if constexpr (!REALPH) hei[h] = ei;
      }
      {  // spin m to s - 1 over bond (s - 1, s); for the anchor s - 1 is site L
        const int h = 2 * m + 1;
        const bool ok = (m == 0 ? g[N - 1] : g[m - 1]) > 1;
        const int bt = s_bt[m == 0 ? L : s - 1];
        anti += ok ? s_jz4[bt] : (R)0;
        R cf = s_jxy2[bt] * rinv;
        int c = 0;
        if (m == 0) {
          if (ok) slow |= 1u << h;
        } else {
          int mo = 0;
#pragma unroll
          for (int q = 0; q < N - 1; ++q) mo = max(mo, g[q] - (q == m - 1 ? 1 : 0) + (q == m ? 1 : 0));
          const int wrap = g[N - 1] + (m == N - 1 ? 1 : 0);
          c = crow - (m == 1 ? 1 : bnm(s_binom, L, m - 1, n[m] - 3));
          if (wrap > mo) cf *= r1;
          else if (ok) slow |= 1u << h;
        }
        hc[h] = ok ? c : -1;
        hcf[h] = cf;
        if constexpr (!REALPH) hei[h] = 0;
      }
    }
    // hops that need the key search: one per lane per iteration
    while (__ballot(slow != 0u) != 0ull) {
      if (slow != 0u) {
        const int h = __builtin_ctz(slow);
        slow &= slow - 1u;
        const int m = h >> 1;
        const bool right = (h & 1) == 0;
        int nn[N];
        if (m == 0 && !right) {
#pragma unroll
          for (int q = 0; q < N - 1; ++q) nn[q] = n[q + 1];
          nn[N - 1] = L;
        } else {
#pragma unroll
          for (int q = 0; q < N; ++q) nn[q] = n[q] + ((q == m) ? (right ? 1 : -1) : 0);
        }
        int c, count, ei;
        chain_search<N, KeyT>(nn, L, kb, s_binom, c, count, ei);
        const R rc = s_rcl[count];
#pragma unroll
        for (int hh = 0; hh < H; ++hh) {
          if (hh == h) {
            hc[hh] = c;
            hcf[hh] *= rc;
            if constexpr (!REALPH) hei[hh] = ei;
          }
        }
      }
    }
    // every index read, then every gather
    ulonglong2 wd[H];
#pragma unroll
    for (int h = 0; h < H; ++h) wd[h] = bm[(hc[h] >= 0 ? hc[h] : 0) >> 6];
    int hid[H];
#pragma unroll
    for (int h = 0; h < H; ++h) {
      const uint64_t bit = 1ull << (hc[h] & 63);
      const bool in = hc[h] >= 0 && (wd[h].x & bit) != 0ull;
      hid[h] = in ? (int)(uint32_t)wd[h].y + __popcll(wd[h].x & (bit - 1ull)) : -1;
    }
    C y[H];
#pragma unroll
    for (int h = 0; h < H; ++h) y[h] = v1[hid[h] >= 0 ? (int64_t)hid[h] : row];
    R accx = 0, accy = 0;
#pragma unroll
    for (int h = 0; h < H; ++h) {
      const R cf = (hid[h] >= 0) ? hcf[h] : (R)0;
      if constexpr (REALPH) {
        accx += cf * y[h].x;
        accy += cf * y[h].y;
      } else {
        const R pr = (R)a.explist[2 * hei[h]] * cf, pi = (R)a.explist[2 * hei[h] + 1] * cf;
        accx += pr * y[h].x - pi * y[h].y;
        accy += pr * y[h].y + pi * y[h].x;
      }
    }
    C x = xin;
    x.x *= scale;
    x.y *= scale;
    const R dg = (sizeof(R) == 8 ? (R)a.ezz : (R)0) - (R)2 * anti;
    C out;
    out.x = dg * x.x + accx * scale;
    out.y = dg * x.y + accy * scale;
    v0[row] = out;
    if (a.dot_partials) {
      dre += (double)x.x * (double)out.x + (double)x.y * (double)out.y;
      dim_ += (double)x.x * (double)out.y - (double)x.y * (double)out.x;
    }
  }
  if (a.dot_partials) block_reduce_store(dre, dim_, a.dot_partials);
}

template <int N, class KeyT>
KFn pick_nn2(int prec, bool realph) {
  if (prec == QS3X_FP64) return realph ? k_hv_chain_nn<N, double, true, KeyT> : k_hv_chain_nn<N, double, false, KeyT>;
  return realph ? k_hv_chain_nn<N, float, true, KeyT> : k_hv_chain_nn<N, float, false, KeyT>;
}

template <int N>
KFn pick_nn(int prec, bool realph, bool key32) {
  return key32 ? pick_nn2<N, uint32_t>(prec, realph) : pick_nn2<N, uint64_t>(prec, realph);
}

}  // namespace

KFn chain_nn_kernel(int N, int prec, bool realph, bool key32) {
  switch (N) {
    case 2: return pick_nn<2>(prec, realph, key32);
    case 3: return pick_nn<3>(prec, realph, key32);
    case 4: return pick_nn<4>(prec, realph, key32);
    case 5: return pick_nn<5>(prec, realph, key32);
    case 6: return pick_nn<6>(prec, realph, key32);
    case 7: return pick_nn<7>(prec, realph, key32);
    case 8: return pick_nn<8>(prec, realph, key32);
    default: return nullptr;
  }
}

size_t chain_nn_lds_bytes(int N, int L, int nbt, int prec) {
  const size_t rs = (prec == QS3X_FP64) ? 8 : 4;
  const size_t nb = (size_t)(N - 1) * (L - 1);
  return ((nb * 4 + 15) & ~(size_t)15) + rs * (N + 1) + 2 * rs * nbt + 2 * (size_t)(L + 1);
}

}  // namespace qs3x
