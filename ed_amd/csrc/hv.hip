// hv.hip -- matrix-free H.v in the U(1) x translation basis (pull form).
//
// Replaces ham_to_vec_wave_vector (ham2vec_v3.f90:37-72):
//   v0(i) = sum_b [ antiparallel: -Jz_b/4 v1(i) + (Jxy_b/2) r(id)/r(i) explist(ell) v1(id)
//                   parallel:     +Jz_b/4 v1(i) ]
// where id is the row of the representative of the hopped configuration.
//
// MI355X layout: one row per lane, rows in ascending rank; a persistent grid
// stages the bond adjacency, the binomial columns and the norm classes in LDS
// once per workgroup.  Instead of scanning every bond of the bond file per row
// (NOxxz x 2 N compares), each down spin visits only its own bonds; the
// diagonal is e_zz - 2 sum_{antiparallel} Jz/4 with e_zz = sum_b Jz_b/4.
//
// Chain representative (LY = LZ = 1): with gaps g_q between cyclically
// consecutive down spins, rotating down spin j to site 1 gives an image whose
// largest elements are decided by the gaps preceding j, (g_{j-1}, g_{j-2}, ...).
// The colex-minimal image is the rotation whose backward gap sequence is
// lexicographically largest; packing each sequence into a 64-bit key makes the
// search N integer maxima.  Ties are exactly the stabiliser (its order gives
// r(id)); ell, the first translation in the reference's loop order reaching
// the minimum (state_lists.f90:384-416), is L + 1 - (largest tied site), or 0
// when the configuration already is the representative.  The representative's
// gaps are read back from the winning key, ranked among configurations that
// contain site 1, and looked up in a dense rank-addressed index (one 4-byte
// read) -- replacing findstate's binary search over list_s.
#include <algorithm>
#include <map>
#include <mutex>

#include "blas.h"
#include "hv_common.h"
#include "qs3x_internal.h"

namespace qs3x {
namespace {

using namespace hvk;

// ------------------------------------------------------------ chain kernel
// Template parameters: N = N_down; R = vector precision; REALPH = every phase
// the operator uses is 1 (k = 0); NORMLIST = norms read from list_r (sectors
// supplied as lists) instead of the stabiliser-class table; COUNT = count the
// in-sector hops instead of applying H; KeyT = uint32_t when the N-1 packed
// gaps fit in 32 bits, else uint64_t.
//
// Fast path: the representative of a row is anchored at site 1 with the gap
// ending at site 1 (the wrap gap) lexicographically first.  When a hop moves a
// down spin other than the one at site 1 and the new wrap gap is still
// strictly larger than every other gap, the hopped configuration is its own
// representative (stabiliser order 1, ell = 0) and its anchored rank is read
// directly from its sites.  Hops of the anchor and near-ties take the full
// key search.
template <int N, class R, bool REALPH, bool NORMLIST, bool COUNT, class KeyT>
__global__ __launch_bounds__(kThreads) void k_hv_chain(HvArgs a) {
  using C = typename CT<R>::T;
  extern __shared__ __align__(16) unsigned char smem[];
  const int L = a.L;
  // LDS: binom [N-1][L-1] int32 | rclass [N+1] R | jxy2 [nbt] R | jz4 [nbt] R |
  //      adj_ptr [L+2] int | adj [nadj] int
  const int nb = (N > 1) ? (N - 1) * (L - 1) : 0;
  int32_t *s_binom = (int32_t *)smem;
  R *s_rcl = (R *)(smem + (((size_t)nb * 4 + 15) & ~(size_t)15));
  R *s_jxy2 = s_rcl + (N + 1);
  R *s_jz4 = s_jxy2 + a.nbt;
  int *s_ptr = (int *)(s_jz4 + a.nbt);
  int *s_adj = s_ptr + (L + 2);
  for (int q = threadIdx.x; q < nb; q += kThreads) {
    int m = q / (L - 1) + 1, x = q % (L - 1);
    s_binom[q] = (int32_t)a.binom[(int64_t)x * a.kdim + m];
  }
  if (a.rclass)
    for (int q = threadIdx.x; q <= N; q += kThreads) s_rcl[q] = (R)a.rclass[q];
  for (int q = threadIdx.x; q < a.nbt; q += kThreads) {
    s_jxy2[q] = (R)a.bt[2 * q];
    s_jz4[q] = (R)a.bt[2 * q + 1];
  }
  for (int q = threadIdx.x; q < L + 2; q += kThreads) s_ptr[q] = a.adj_ptr[q];
  for (int q = threadIdx.x; q < a.nadj; q += kThreads) s_adj[q] = a.adj[q];
  __syncthreads();

  const C *__restrict__ v1 = (const C *)a.v1;
  C *__restrict__ v0 = (C *)a.v0;
  const R scale = a.scale ? (R)(*a.scale) : (R)1;
  const int kb = a.kbits;
  const KeyT fmask = (KeyT)((1ull << kb) - 1ull);
  double dre = 0.0, dim_ = 0.0;
  unsigned long long nnz = 0;

  for (int64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    const int64_t row = tile * kThreads + threadIdx.x;
    if (row >= a.dim) continue;
    const uint64_t w = a.cfg[row];
    int n[N];
#pragma unroll
    for (int q = 0; q < N; ++q) n[q] = (int)((w >> (q * a.bits)) & ((1ull << a.bits) - 1ull));
    const R ri = NORMLIST ? (R)a.list_r[row] : (COUNT ? (R)1 : s_rcl[(int)(w >> 56)]);
    const R rinv = (R)1 / ri;
    C x = v1[row];
    x.x *= scale;
    x.y *= scale;
    R anti = 0;
    R accx = 0, accy = 0;
#pragma unroll
    for (int m = 0; m < N; ++m) {
      const int s = n[m];
      const int e0 = s_ptr[s], e1 = s_ptr[s + 1];
      for (int e = e0; e < e1; ++e) {
        const int ent = s_adj[e];
        const int t = ent & 0xffff, btp = ent >> 16;
        bool in = false;
#pragma unroll
        for (int q = 0; q < N; ++q) in |= (n[q] == t);
        if (in) continue;
        anti += s_jz4[btp];
        // hopped configuration: remove n[m], insert t (sorted, branch-free)
        int r[N > 1 ? N - 1 : 1];
#pragma unroll
        for (int q = 0; q < N - 1; ++q) r[q] = (q < m) ? n[q] : n[q + 1];
        int nn[N];
#pragma unroll
        for (int q = 0; q < N; ++q) {
          const bool ltq = (q < N - 1) ? (r[q < N - 1 ? q : 0] < t) : false;
          const bool ltp = (q > 0) ? (r[q > 0 ? q - 1 : 0] < t) : true;
          nn[q] = ltq ? r[q < N - 1 ? q : 0] : (ltp ? t : r[q > 0 ? q - 1 : 0]);
        }
        int g[N];
#pragma unroll
        for (int q = 0; q < N - 1; ++q) g[q] = nn[q + 1] - nn[q];
        g[N - 1] = nn[0] + L - nn[N - 1];
        bool fast = false;
        if (m >= 1 && N > 1) {  // nn[0] = 1 still (site 1 stays occupied)
          int gm = 0;
#pragma unroll
          for (int q = 0; q < N - 1; ++q) gm = g[q] > gm ? g[q] : gm;
          fast = g[N - 1] > gm;
        }
        int32_t c = 0;
        int count = 1, ei = 0;
        if (fast) {
#pragma unroll
          for (int q = 1; q < N; ++q) c += s_binom[(q - 1) * (L - 1) + (nn[q] - 2)];
        } else {
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
          if (N > 1) {
            int f[N > 1 ? N - 1 : 1];
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
              P += f[N - mm];  // g[j*+mm-1] = field (N - (mm-1))
              c += s_binom[(mm - 1) * (L - 1) + (P - 1)];
            }
          }
        }
        const int id = a.index[c];
        if (id < 0) continue;
        if constexpr (COUNT) {
          ++nnz;
          continue;
        }
        const R rt = NORMLIST ? (R)a.list_r[id] : s_rcl[count];
        const R cf = s_jxy2[btp] * rt * rinv;
        C y = v1[id];
        if constexpr (REALPH) {
          accx += cf * y.x;
          accy += cf * y.y;
        } else {
          const R pr = (R)a.explist[2 * ei] * cf, pi = (R)a.explist[2 * ei + 1] * cf;
          accx += pr * y.x - pi * y.y;
          accy += pr * y.y + pi * y.x;
        }
      }
    }
    if constexpr (!COUNT) {
      const R dg = (sizeof(R) == 8 ? (R)a.ezz : (R)0) - (R)2 * anti;
      C out;
      out.x = (dg * x.x + accx * scale);
      out.y = (dg * x.y + accy * scale);
      v0[row] = out;
      if (a.dot_partials) {
        dre += (double)x.x * (double)out.x + (double)x.y * (double)out.y;
        dim_ += (double)x.x * (double)out.y - (double)x.y * (double)out.x;
      }
    }
  }
  if constexpr (COUNT) {
    for (int d = 32; d >= 1; d >>= 1) nnz += __shfl_down(nnz, d, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(a.nnz_out, nnz);
  } else {
    if (a.dot_partials) block_reduce_store(dre, dim_, a.dot_partials);
  }
}

// ------------------------------------------------ chain kernel, batched loads
// Same arithmetic as k_hv_chain, organised for memory-level parallelism: the
// up to N * DEG hops of a row are first reduced to (anchored rank, coefficient,
// phase index), then every index read is issued, then every gather, then the
// row is accumulated.  A row's dependent loads are two rounds instead of
// 2 * hops.  DEG is the largest bond-adjacency degree of any site.
template <int N, int DEG, class R, bool REALPH, class KeyT>
__global__ __launch_bounds__(kThreads) void k_hv_chain_mlp(HvArgs a) {
  using C = typename CT<R>::T;
  constexpr int H = N * DEG;
  extern __shared__ __align__(16) unsigned char smem[];
  const int L = a.L;
  const int nb = (N > 1) ? (N - 1) * (L - 1) : 0;
  int32_t *s_binom = (int32_t *)smem;
  R *s_rcl = (R *)(smem + (((size_t)nb * 4 + 15) & ~(size_t)15));
  R *s_jxy2 = s_rcl + (N + 1);
  R *s_jz4 = s_jxy2 + a.nbt;
  int *s_ptr = (int *)(s_jz4 + a.nbt);
  int *s_adj = s_ptr + (L + 2);
  for (int q = threadIdx.x; q < nb; q += kThreads) {
    int m = q / (L - 1) + 1, x = q % (L - 1);
    s_binom[q] = (int32_t)a.binom[(int64_t)x * a.kdim + m];
  }
  for (int q = threadIdx.x; q <= N; q += kThreads) s_rcl[q] = (R)a.rclass[q];
  for (int q = threadIdx.x; q < a.nbt; q += kThreads) {
    s_jxy2[q] = (R)a.bt[2 * q];
    s_jz4[q] = (R)a.bt[2 * q + 1];
  }
  for (int q = threadIdx.x; q < L + 2; q += kThreads) s_ptr[q] = a.adj_ptr[q];
  for (int q = threadIdx.x; q < a.nadj; q += kThreads) s_adj[q] = a.adj[q];
  __syncthreads();

  const C *__restrict__ v1 = (const C *)a.v1;
  C *__restrict__ v0 = (C *)a.v0;
  const R scale = a.scale ? (R)(*a.scale) : (R)1;
  const int kb = a.kbits;
  const KeyT fmask = (KeyT)((1ull << kb) - 1ull);
  double dre = 0.0, dim_ = 0.0;

  for (int64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    const int64_t row = tile * kThreads + threadIdx.x;
    if (row >= a.dim) continue;
    const uint64_t w = a.cfg[row];
    const C xin = v1[row];
    int n[N];
#pragma unroll
    for (int q = 0; q < N; ++q) n[q] = (int)((w >> (q * a.bits)) & ((1ull << a.bits) - 1ull));
    const R rinv = (R)1 / s_rcl[(int)(w >> 56)];
    R anti = 0;
    int32_t hc[H];
    R hcf[H];
    int hei[REALPH ? 1 : H];
#pragma unroll
    for (int h = 0; h < H; ++h) {
      hc[h] = -1;
      hcf[h] = 0;
      if constexpr (!REALPH) hei[h] = 0;
    }
    // phase 1: hop -> (anchored rank, coefficient, phase index)
#pragma unroll
    for (int m = 0; m < N; ++m) {
      const int s = n[m];
      const int e0 = s_ptr[s], deg = s_ptr[s + 1] - e0;
#pragma unroll
      for (int e = 0; e < DEG; ++e) {
        if (e >= deg) break;
        const int ent = s_adj[e0 + e];
        const int t = ent & 0xffff, btp = ent >> 16;
        bool in = false;
#pragma unroll
        for (int q = 0; q < N; ++q) in |= (n[q] == t);
        if (in) continue;
        anti += s_jz4[btp];
        int r[N > 1 ? N - 1 : 1];
#pragma unroll
        for (int q = 0; q < N - 1; ++q) r[q] = (q < m) ? n[q] : n[q + 1];
        int nn[N];
#pragma unroll
        for (int q = 0; q < N; ++q) {
          const bool ltq = (q < N - 1) ? (r[q < N - 1 ? q : 0] < t) : false;
          const bool ltp = (q > 0) ? (r[q > 0 ? q - 1 : 0] < t) : true;
          nn[q] = ltq ? r[q < N - 1 ? q : 0] : (ltp ? t : r[q > 0 ? q - 1 : 0]);
        }
        int g[N];
#pragma unroll
        for (int q = 0; q < N - 1; ++q) g[q] = nn[q + 1] - nn[q];
        g[N - 1] = nn[0] + L - nn[N - 1];
        bool fast = false;
        if (m >= 1 && N > 1) {
          int gm = 0;
#pragma unroll
          for (int q = 0; q < N - 1; ++q) gm = g[q] > gm ? g[q] : gm;
          fast = g[N - 1] > gm;
        }
        int32_t c = 0;
        int count = 1, ei = 0;
        if (fast) {
#pragma unroll
          for (int q = 1; q < N; ++q) c += s_binom[(q - 1) * (L - 1) + (nn[q] - 2)];
        } else {
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
          if (N > 1) {
            int f[N > 1 ? N - 1 : 1];
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
        }
        hc[m * DEG + e] = c;
        hcf[m * DEG + e] = s_jxy2[btp] * s_rcl[count] * rinv;
        if constexpr (!REALPH) hei[m * DEG + e] = ei;
      }
    }
    // phase 2: every index read
    int hid[H];
#pragma unroll
    for (int h = 0; h < H; ++h) hid[h] = (hc[h] >= 0) ? a.index[hc[h]] : -1;
    // phase 3: every gather
    C y[H];
#pragma unroll
    for (int h = 0; h < H; ++h) {
      if (hid[h] >= 0) y[h] = v1[hid[h]];
      else {
        y[h].x = 0;
        y[h].y = 0;
      }
    }
    // phase 4: accumulate
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

// ---------------------------------------------------------- general kernel
// Any LX x LY x LZ lattice: the representative is found by scanning the
// translation group in the reference's loop order (k, j, i) with a running
// colex minimum and strict <, exactly as representative_SQ does.
__device__ __forceinline__ int shift_site(int s, int i, int j, int k, int LX, int LY, int LZ) {
  int s0 = s - 1;
  int x = s0 % LX, y = (s0 / LX) % LY, z = s0 / (LX * LY);
  x = (x + i) % LX;
  y = (y + j) % LY;
  z = (z + k) % LZ;
  return 1 + x + LX * (y + LY * z);
}

// U(1)-only H.v (source_only_u1/ham2vec_v3.f90:35-65).  Rows are all
// C(NOS, N) configurations in colex order: row r holds the configuration of
// 0-based rank r = sum_q C(p_q, q) over its 0-based down sites p_1 < ... < p_N
// (list_fly / inv_list of source_only_u1/state_lists.f90, shifted to 0-based
// sites and ranks).  A hop's target row is the rank of the flipped
// configuration: no representative, norm or phase.  Up to 64 sites the
// occupation is one 64-bit word, so membership tests are bit tests and the
// target rank is a sum over the set bits of the flipped word; with WIDE (any
// NOS) membership is a scan of the N sorted sites and the target rank a
// merge of the remaining sites with the new one.
template <class R, bool COUNT, bool WIDE>
__global__ __launch_bounds__(kThreads) void k_hv_u1(HvArgs a, int N) {
  using C = typename CT<R>::T;
  const C *__restrict__ v1 = (const C *)a.v1;
  C *__restrict__ v0 = (C *)a.v0;
  const R scale = a.scale ? (R)(*a.scale) : (R)1;
  const int64_t *__restrict__ B = a.binom;  // B[n * kdim + k] = C(n, k)
  const int kd = a.kdim;
  double dre = 0.0, dim_ = 0.0;
  unsigned long long nnz = 0;
  for (int64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    const int64_t row = tile * kThreads + threadIdx.x;
    if (row >= a.dim) continue;
    // unrank: for q = N..1 the largest p with C(p, q) <= r (C(p, q) rises with p)
    uint64_t occ = 0;
    int p[kMaxNod];
    int64_t r = row;
    int hi = a.NOS - 1;
    for (int q = N; q >= 1; --q) {
      int lo = q - 1, h = hi;
      while (lo < h) {
        const int mid = (lo + h + 1) >> 1;
        if (B[(int64_t)mid * kd + q] <= r) lo = mid;
        else h = mid - 1;
      }
      p[q - 1] = lo;
      r -= B[(int64_t)lo * kd + q];
      if (!WIDE) occ |= 1ull << lo;
      hi = lo - 1;
    }
    C x = v1[row];
    x.x *= scale;
    x.y *= scale;
    R anti = 0, accx = 0, accy = 0;
    for (int m = 0; m < N; ++m) {
      const int s1 = p[m] + 1;  // 1-based site
      for (int e = a.adj_ptr[s1]; e < a.adj_ptr[s1 + 1]; ++e) {
        const int ent = a.adj[e];
        const int t = (ent & 0xffff) - 1, btp = ent >> 16;
        bool occupied;
        if constexpr (WIDE) {
          occupied = false;
          for (int k = 0; k < N; ++k) occupied |= (p[k] == t);
        } else {
          occupied = (occ >> t) & 1ull;
        }
        if (occupied) continue;  // parallel bond
        anti += (R)a.bt[2 * btp + 1];
        if constexpr (COUNT) {
          ++nnz;
          continue;
        }
        int64_t id = 0;
        if constexpr (WIDE) {
          // the sites p minus p[m], with t inserted in order
          int q = 1;
          bool ins = false;
          for (int k = 0; k < N; ++k) {
            if (k == m) continue;
            if (!ins && t < p[k]) {
              id += B[(int64_t)t * kd + q++];
              ins = true;
            }
            id += B[(int64_t)p[k] * kd + q++];
          }
          if (!ins) id += B[(int64_t)t * kd + q];
        } else {
          uint64_t w = occ ^ (1ull << p[m]) ^ (1ull << t);
          for (int q = 1; w; ++q, w &= w - 1) id += B[(int64_t)__builtin_ctzll(w) * kd + q];
        }
        const R cf = (R)a.bt[2 * btp];
        const C y = v1[id];
        accx += cf * y.x;
        accy += cf * y.y;
      }
    }
    if constexpr (!COUNT) {
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
  }
  if constexpr (COUNT) {
    for (int d = 32; d >= 1; d >>= 1) nnz += __shfl_down(nnz, d, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(a.nnz_out, nnz);
  } else {
    if (a.dot_partials) block_reduce_store(dre, dim_, a.dot_partials);
  }
}

// Representative of a sorted configuration nn (1-based sites) on an
// LX x LY x LZ lattice: the colex-minimal translation image, as
// representative_SQ (state_lists.f90:384-416) finds it by trying every
// translation (i, j, k), i = 1..LX innermost, keeping the first strictly
// smaller image.  A minimal image has a down spin in the plane z = 0, one in
// the plane y = 0 and one in x = 0 (else the translation by -1 along that
// axis lowers every site index), so every translation that reaches the
// minimum shifts some spin's x, some spin's y and some spin's z to 0: the
// N^d distinct such shifts are the only candidates (d the lattice
// dimension), and among those reaching the minimum the reference keeps the
// first in its loop order, or none (ell = 0) when nn itself is minimal.
// Falls back to all LX LY LZ translations when the candidates would be as
// many.  best: the representative; ei, ei_best as the scan sets them.
__device__ __forceinline__ void representative_candidates(const int *nn, int N, int LX, int LY, int LZ, int *nd,
                                                          int *best, int &ei, int &ei_best) {
  for (int q = 0; q < N; ++q) best[q] = nn[q];
  ei = 0;
  ei_best = 0;
  int cx[kMaxNod], cy[kMaxNod], cz[kMaxNod];
  int nx = 0, ny = 0, nz = 0;
  for (int q = 0; q < N; ++q) {
    const int s0 = nn[q] - 1;
    const int x = s0 % LX, y = (s0 / LX) % LY, z = s0 / (LX * LY);
    const int tx = (LX - x) % LX, ty = (LY - y) % LY, tz = (LZ - z) % LZ;
    bool dx = false, dy = false, dz = false;
    for (int r = 0; r < nx; ++r) dx |= cx[r] == tx;
    for (int r = 0; r < ny; ++r) dy |= cy[r] == ty;
    for (int r = 0; r < nz; ++r) dz |= cz[r] == tz;
    if (!dx) cx[nx++] = tx;
    if (!dy) cy[ny++] = ty;
    if (!dz) cz[nz++] = tz;
  }
  const bool full = (int64_t)nx * ny * nz >= (int64_t)LX * LY * LZ;
  const int kz = full ? LZ : nz, ky = full ? LY : ny, kx = full ? LX : nx;
  int64_t best_order = -1;  // loop order of the kept translation (-1: nn itself)
  for (int c3 = 0; c3 < kz; ++c3)
    for (int c2 = 0; c2 < ky; ++c2)
      for (int c1 = 0; c1 < kx; ++c1) {
        // translation by (i, j, k) single shifts, i = 1..LX (i = LX: none)
        const int i = full ? c1 + 1 : (cx[c1] == 0 ? LX : cx[c1]);
        const int j = full ? c2 + 1 : (cy[c2] == 0 ? LY : cy[c2]);
        const int k = full ? c3 + 1 : (cz[c3] == 0 ? LZ : cz[c3]);
        if (i == LX && j == LY && k == LZ) continue;  // nn itself
        for (int q = 0; q < N; ++q) nd[q] = shift_site(nn[q], i, j, k, LX, LY, LZ);
        for (int q = 1; q < N; ++q) {
          int v = nd[q], z = q - 1;
          while (z >= 0 && nd[z] > v) {
            nd[z + 1] = nd[z];
            --z;
          }
          nd[z + 1] = v;
        }
        int cmp = 0;  // image against best, colex (largest site first)
        for (int l = N - 1; l >= 0 && cmp == 0; --l) cmp = nd[l] < best[l] ? -1 : (nd[l] > best[l] ? 1 : 0);
        const int64_t order = ((int64_t)(k - 1) * LY + (j - 1)) * LX + (i - 1);
        if (cmp < 0 || (cmp == 0 && best_order >= 0 && order < best_order)) {
          if (cmp < 0)
            for (int q = 0; q < N; ++q) best[q] = nd[q];
          best_order = order;
          ei = i + (LX + 1) * (j + (LY + 1) * k);
          ei_best = 1;
        }
      }
}

template <class R, bool REALPH, bool COUNT>
__global__ __launch_bounds__(kThreads) void k_hv_general(HvArgs a, int N) {
  using C = typename CT<R>::T;
  const C *__restrict__ v1 = (const C *)a.v1;
  C *__restrict__ v0 = (C *)a.v0;
  const R scale = a.scale ? (R)(*a.scale) : (R)1;
  double dre = 0.0, dim_ = 0.0;
  unsigned long long nnz = 0;
  for (int64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    const int64_t row = tile * kThreads + threadIdx.x;
    if (row >= a.dim) continue;
    const uint64_t w = a.cfg[row];
    int n[kMaxNod], nn[kMaxNod], nd[kMaxNod], best[kMaxNod];
    for (int q = 0; q < N; ++q) n[q] = (int)((w >> (q * a.bits)) & ((1ull << a.bits) - 1ull));
    const R ri = (R)a.list_r[row];
    C x = v1[row];
    x.x *= scale;
    x.y *= scale;
    R anti = 0, accx = 0, accy = 0;
    for (int m = 0; m < N; ++m) {
      const int s = n[m];
      for (int e = a.adj_ptr[s]; e < a.adj_ptr[s + 1]; ++e) {
        const int ent = a.adj[e];
        const int t = ent & 0xffff, btp = ent >> 16;
        bool in = false;
        for (int q = 0; q < N; ++q) in |= (n[q] == t);
        if (in) continue;
        anti += (R)a.bt[2 * btp + 1];
        // hopped configuration, sorted
        int p = 0;
        for (int q = 0; q < N; ++q)
          if (q != m) nn[p++] = n[q];
        int pos = N - 1;
        while (pos > 0 && nn[pos - 1] > t) {
          nn[pos] = nn[pos - 1];
          --pos;
        }
        nn[pos] = t;
        int ei = 0, ei_best = 0;
        representative_candidates(nn, N, a.LX, a.LY, a.LZ, nd, best, ei, ei_best);
        int64_t rank = best[0];
        for (int i2 = 2; i2 <= N; ++i2) rank += a.binom[(int64_t)(best[i2 - 1] - 1) * a.kdim + i2];
        int64_t id = -1;
        if (a.mode == kIndexFullRank) {
          id = a.index[rank - 1];
        } else {
          int64_t lo = 0, hi = a.dim - 1;
          while (lo <= hi) {
            int64_t mid = lo + (hi - lo) / 2;
            int64_t v = a.list_s[mid];
            if (v < rank) lo = mid + 1;
            else if (v > rank) hi = mid - 1;
            else { id = mid; break; }
          }
        }
        if (id < 0) continue;
        if constexpr (COUNT) {
          ++nnz;
          continue;
        }
        const R cf = (R)a.bt[2 * btp] * (R)a.list_r[id] / ri;
        C y = v1[id];
        if (REALPH || !ei_best) {
          accx += cf * y.x;
          accy += cf * y.y;
        } else {
          const R pr = (R)a.explist[2 * ei] * cf, pi = (R)a.explist[2 * ei + 1] * cf;
          accx += pr * y.x - pi * y.y;
          accy += pr * y.y + pi * y.x;
        }
      }
    }
    if constexpr (!COUNT) {
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
  }
  if constexpr (COUNT) {
    for (int d = 32; d >= 1; d >>= 1) nnz += __shfl_down(nnz, d, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(a.nnz_out, nnz);
  } else {
    if (a.dot_partials) block_reduce_store(dre, dim_, a.dot_partials);
  }
}

// ----------------------------------------------------------------- dispatch

template <int N, class KeyT>
KFn pick_chain_k(int prec, bool realph, bool normlist, bool count) {
  if (count) return k_hv_chain<N, double, true, false, true, uint64_t>;
  if (prec == QS3X_FP64) {
    if (normlist)
      return realph ? k_hv_chain<N, double, true, true, false, uint64_t>
                    : k_hv_chain<N, double, false, true, false, uint64_t>;
    return realph ? k_hv_chain<N, double, true, false, false, KeyT> : k_hv_chain<N, double, false, false, false, KeyT>;
  }
  return realph ? k_hv_chain<N, float, true, false, false, KeyT> : k_hv_chain<N, float, false, false, false, KeyT>;
}

template <int N>
KFn pick_chain(int prec, bool realph, bool normlist, bool count, bool key32) {
  return key32 ? pick_chain_k<N, uint32_t>(prec, realph, normlist, count)
               : pick_chain_k<N, uint64_t>(prec, realph, normlist, count);
}

KFn chain_kernel(int N, int prec, bool realph, bool normlist, bool count, bool key32) {
  switch (N) {
    case 1: return pick_chain<1>(prec, realph, normlist, count, key32);
    case 2: return pick_chain<2>(prec, realph, normlist, count, key32);
    case 3: return pick_chain<3>(prec, realph, normlist, count, key32);
    case 4: return pick_chain<4>(prec, realph, normlist, count, key32);
    case 5: return pick_chain<5>(prec, realph, normlist, count, key32);
    case 6: return pick_chain<6>(prec, realph, normlist, count, key32);
    case 7: return pick_chain<7>(prec, realph, normlist, count, key32);
    case 8: return pick_chain<8>(prec, realph, normlist, count, key32);
    case 9: return pick_chain<9>(prec, realph, normlist, count, key32);
    case 10: return pick_chain<10>(prec, realph, normlist, count, key32);
    case 11: return pick_chain<11>(prec, realph, normlist, count, key32);
    case 12: return pick_chain<12>(prec, realph, normlist, count, key32);
    default: return nullptr;
  }
}

template <int N, int DEG, class KeyT>
KFn pick_mlp3(int prec, bool realph) {
  if (prec == QS3X_FP64)
    return realph ? k_hv_chain_mlp<N, DEG, double, true, KeyT> : k_hv_chain_mlp<N, DEG, double, false, KeyT>;
  return realph ? k_hv_chain_mlp<N, DEG, float, true, KeyT> : k_hv_chain_mlp<N, DEG, float, false, KeyT>;
}

template <int N>
KFn pick_mlp2(int prec, bool realph, int deg, bool key32) {
  if (deg <= 2)
    return key32 ? pick_mlp3<N, 2, uint32_t>(prec, realph) : pick_mlp3<N, 2, uint64_t>(prec, realph);
  return key32 ? pick_mlp3<N, 4, uint32_t>(prec, realph) : pick_mlp3<N, 4, uint64_t>(prec, realph);
}

// Batched-load chain kernels exist for N <= 8 and adjacency degree <= 4.
KFn chain_mlp_kernel(int N, int prec, bool realph, int deg, bool key32) {
  if (deg > 4) return nullptr;
  switch (N) {
    case 1: return pick_mlp2<1>(prec, realph, deg, key32);
    case 2: return pick_mlp2<2>(prec, realph, deg, key32);
    case 3: return pick_mlp2<3>(prec, realph, deg, key32);
    case 4: return pick_mlp2<4>(prec, realph, deg, key32);
    case 5: return pick_mlp2<5>(prec, realph, deg, key32);
    case 6: return pick_mlp2<6>(prec, realph, deg, key32);
    case 7: return pick_mlp2<7>(prec, realph, deg, key32);
    case 8: return pick_mlp2<8>(prec, realph, deg, key32);
    default: return nullptr;
  }
}

// Resident blocks per CU for a kernel at this LDS size (cached per kernel and size).
int blocks_per_cu(KFn f, size_t lds) {
  static std::vector<std::pair<std::pair<void *, size_t>, int>> cache;
  for (auto &e : cache)
    if (e.first.first == (void *)f && e.first.second == lds) return e.second;
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, kThreads, lds) != hipSuccess || nb <= 0) nb = 8;
  cache.push_back({{(void *)f, lds}, nb});
  return nb;
}

int persistent_grid(int64_t ntiles) {
  static int cus = 0;
  if (!cus) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device()) == hipSuccess) cus = p.multiProcessorCount;
    if (cus <= 0) cus = 256;
  }
  int64_t g = std::min<int64_t>((int64_t)cus * 8, kMaxBlocks);
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, ntiles));
}

// Persistent grid that is resident all at once: CUs x resident blocks per CU.
int resident_grid(KFn f, size_t lds, int64_t ntiles) {
  static int cus = 0;
  if (!cus) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device()) == hipSuccess) cus = p.multiProcessorCount;
    if (cus <= 0) cus = 256;
  }
  int64_t g = std::min<int64_t>((int64_t)cus * blocks_per_cu(f, lds), kMaxBlocks);
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, ntiles));
}

}  // namespace

static HvArgs make_args(qs3x_sector *s, void *v0, const void *v1, const double *scale,
                        double *partials, int prec_of_args) {
  qs3x_model *m = s->model;
  HvArgs a{};
  a.cfg = s->d_cfg.p;
  a.list_r = s->d_list_r.p;
  a.index = s->d_index.p;
  a.bitmap = s->d_bitmap.p;
  a.nbitmap = (int64_t)s->d_bitmap.n;
  a.chunk_log2 = hv_chunk_log2();
  a.nn_bt = m->d_nn_bt.p;
  a.list_s = s->d_list_s.p;
  a.v1 = v1;
  a.v0 = v0;
  a.dim = s->dim;
  a.ntiles = (s->dim + kThreads - 1) / kThreads;
  a.L = m->NOS;
  a.bits = s->bits;
  int kb = 1;
  while ((1 << kb) <= m->NOS) ++kb;
  a.kbits = kb;
  a.mode = s->mode;
  a.LX = m->LX; a.LY = m->LY; a.LZ = m->LZ; a.NOS = m->NOS;
  a.adj_ptr = m->d_adj_ptr.p;
  a.adj = m->d_adj.p;
  a.nadj = (int)m->adj.size();
  a.bt = m->d_bt.p;
  a.nbt = (int)m->bt_jxy2.size();
  a.binom = s->d_binom.p;
  a.kdim = s->nbinom_k;
  a.rclass = s->d_rclass.p;
  a.explist = s->d_explist.p;
  a.ezz = m->ezz;
  a.scale = scale;
  a.dot_partials = partials;
  a.nxcd = hv_xcd_for(prec_of_args);
  return a;
}

static size_t chain_lds_bytes(qs3x_sector *s, int prec) {
  const int N = s->nod, L = s->model->NOS;
  const size_t rs = (prec == QS3X_FP64) ? 8 : 4;
  size_t nb = (N > 1) ? (size_t)(N - 1) * (L - 1) : 0;
  return ((nb * 4 + 15) & ~(size_t)15) + rs * (N + 1) + 2 * rs * s->model->bt_jxy2.size() + 4 * (size_t)(L + 2) +
         4 * s->model->adj.size();
}

// the N-1 packed gap fields of a key fit in 32 bits
static bool key32(qs3x_sector *s) {
  int kb = 1;
  while ((1 << kb) <= s->model->NOS) ++kb;
  return (int64_t)(s->nod - 1) * kb <= 32;
}

// dynamic LDS the nearest-neighbour kernels may use: 160 KB per CU on gfx950
// less their static queues (2 KB) and a margin
constexpr size_t kNnMaxLds = 156 * 1024;

static bool use_chain(qs3x_sector *s, int prec) {
  return s->chain && s->mode == kIndexAnchored && chain_lds_bytes(s, prec) <= kNnMaxLds;
}

// Kernels whose tables exceed the default 64 KB of dynamic LDS (long chains)
// must say so before their launch.  The attribute is set once per kernel (to
// the largest size asked so far), not on every launch of a Lanczos loop.
static void allow_lds(KFn f, size_t lds) {
  if (lds <= 64 * 1024) return;
  static std::mutex mu;
  static std::map<KFn, size_t> done;
  std::lock_guard<std::mutex> lk(mu);
  auto it = done.find(f);
  if (it != done.end() && it->second >= lds) return;
  QS3X_HIP(hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  done[f] = lds;
}

bool hv_fused_available(const qs3x_sector *s, int prec) {
  const int f = options().lanczos_fused;
  if (f == 0 || (f == 1 && prec != QS3X_FP64)) return false;
  if (!options().hv_stencil || s->dim == 0 || !stencil_applies(s)) return false;
  const qs3x_sector::HvTune &tn = s->tune[prec == QS3X_FP64 ? 0 : 1];
  const int pick = tn.set ? tn.kernel : options().hv_kernel;
  return pick <= 0 && s->d_bitmap.p != nullptr;
}

bool launch_hv_fused(qs3x_sector *s, int prec, void *pout, void *xout, const void *pprev, const void *vprev,
                     const double *fsc, int fis, int fia, int fib, double *part4, int *nblocks_out, hipStream_t st) {
  if (!hv_fused_available(s, prec)) return false;
  HvArgs a = make_args(s, pout, pprev, nullptr, part4, prec);
  a.ezz = 0.0;  // H - e_zz in both precisions (the caller adds e_zz to alpha)
  a.vprev = vprev;
  a.xout = xout;
  a.fsc = fsc;
  a.fis = fis;
  a.fia = fia;
  a.fib = fib;
  int g = 0;
  if (!launch_hv_stencil_fused(s, prec, a, st, &g)) return false;
  if (nblocks_out) *nblocks_out = g;
  s->last_family = 6;
  return true;
}

void launch_hv(qs3x_sector *s, int prec, void *v0, const void *v1, const double *scale_dev,
               double *dot_partials, int *nblocks_out, hipStream_t st) {
  if (s->dim == 0) {
    if (nblocks_out) *nblocks_out = 0;
    return;
  }
  HvArgs a = make_args(s, v0, v1, scale_dev, dot_partials, prec);
  const qs3x_sector::HvTune &tn = s->tune[prec == QS3X_FP64 ? 0 : 1];
  if (tn.set) a.nxcd = tn.xcd;
  if (tn.set && tn.chunk >= 0) a.chunk_log2 = tn.chunk;
  // a tuned configuration divides the resident grid (never enlarges it, so the
  // callers' partial-sum buffers keep their size)
  auto tuned = [&](int g) { return tn.set ? std::max(1, g * tn.grid_q / 4) : g; };
  int grid = tuned(persistent_grid(a.ntiles));
  if (nblocks_out) *nblocks_out = grid;
  if (s->mode == kIndexU1) {
    const bool wide = s->model->NOS > 64;
    if (prec == QS3X_FP64) {
      if (wide) hipLaunchKernelGGL((k_hv_u1<double, false, true>), dim3(grid), dim3(kThreads), 0, st, a, s->nod);
      else hipLaunchKernelGGL((k_hv_u1<double, false, false>), dim3(grid), dim3(kThreads), 0, st, a, s->nod);
    } else {
      if (wide) hipLaunchKernelGGL((k_hv_u1<float, false, true>), dim3(grid), dim3(kThreads), 0, st, a, s->nod);
      else hipLaunchKernelGGL((k_hv_u1<float, false, false>), dim3(grid), dim3(kThreads), 0, st, a, s->nod);
    }
    QS3X_HIP(hipGetLastError());
    s->last_family = 4;
    s->last_lds = 0;
    return;
  }
  const bool normlist = s->d_rclass.p == nullptr;  // sectors from caller lists carry list_r only
  const int pick = tn.set ? tn.kernel : options().hv_kernel;
  // The nearest-neighbour kernels need only their own LDS tables (binomials,
  // bond types; not the adjacency of the general chain kernels), up to what
  // a CU holds beside their 2 KB of static LDS.
  if (pick <= 0 && s->chain && s->mode == kIndexAnchored && !normlist && s->model->nn_chain && s->d_bitmap.p) {
    int sg = 0;
    size_t slds = 0;
    if (!(tn.set && tn.stencil == 0) && launch_hv_stencil(s, prec, a, st, &sg, &slds)) {
      if (nblocks_out) *nblocks_out = sg;
      s->last_family = 5;
      s->last_lds = (int64_t)slds;
      return;
    }
    const size_t lds = chain_nn_lds_bytes(s->nod, s->model->NOS, (int)s->model->bt_jxy2.size(), prec);
    const bool uniform = s->model->bt_jxy2.size() == 1;
    const bool pairs = (((uintptr_t)a.v0 | (uintptr_t)a.v1) & 15u) == 0 && !(tn.set && tn.pairs == 0);
    KFn f = (lds <= kNnMaxLds) ? chain_nn_kernel(s->nod, prec, s->real_phase, uniform, key32(s), pairs) : nullptr;
    if (f) {
      allow_lds(f, lds);
      grid = tuned(resident_grid(f, lds, a.ntiles));
      if (nblocks_out) *nblocks_out = grid;
      hipLaunchKernelGGL(f, dim3(grid), dim3(kThreads), lds, st, a);
      QS3X_HIP(hipGetLastError());
      s->last_family = 0;
      s->last_lds = (int64_t)lds;
      return;
    }
  }
  if (pick <= 1 && use_chain(s, prec) && !normlist) {
    KFn f = chain_mlp_kernel(s->nod, prec, s->real_phase, s->model->max_deg, key32(s));
    if (f) {
      allow_lds(f, chain_lds_bytes(s, prec));
      grid = tuned(resident_grid(f, chain_lds_bytes(s, prec), a.ntiles));
      if (nblocks_out) *nblocks_out = grid;
      hipLaunchKernelGGL(f, dim3(grid), dim3(kThreads), chain_lds_bytes(s, prec), st, a);
      QS3X_HIP(hipGetLastError());
      s->last_family = 1;
      s->last_lds = (int64_t)chain_lds_bytes(s, prec);
      return;
    }
  }
  if (pick <= 2 && use_chain(s, prec) && !(normlist && prec != QS3X_FP64)) {
    KFn f = chain_kernel(s->nod, prec, s->real_phase, normlist, false, key32(s));
    if (f) {
      allow_lds(f, chain_lds_bytes(s, prec));
      grid = tuned(resident_grid(f, chain_lds_bytes(s, prec), a.ntiles));
      if (nblocks_out) *nblocks_out = grid;
      hipLaunchKernelGGL(f, dim3(grid), dim3(kThreads), chain_lds_bytes(s, prec), st, a);
      QS3X_HIP(hipGetLastError());
      s->last_family = 2;
      s->last_lds = (int64_t)chain_lds_bytes(s, prec);
      return;
    }
  }
  if (prec == QS3X_FP64) {
    if (s->real_phase)
      hipLaunchKernelGGL((k_hv_general<double, true, false>), dim3(grid), dim3(kThreads), 0, st, a, s->nod);
    else
      hipLaunchKernelGGL((k_hv_general<double, false, false>), dim3(grid), dim3(kThreads), 0, st, a, s->nod);
  } else {
    if (s->real_phase)
      hipLaunchKernelGGL((k_hv_general<float, true, false>), dim3(grid), dim3(kThreads), 0, st, a, s->nod);
    else
      hipLaunchKernelGGL((k_hv_general<float, false, false>), dim3(grid), dim3(kThreads), 0, st, a, s->nod);
  }
  QS3X_HIP(hipGetLastError());
  s->last_family = 3;
  s->last_lds = 0;
}

int64_t count_nnz(qs3x_sector *s) {
  if (s->dim == 0) return 0;
  DevBuf<unsigned long long> d_n(1);
  unsigned long long zero = 0, nz = 0;
  d_n.upload(&zero, 1);
  HvArgs a = make_args(s, nullptr, nullptr, nullptr, nullptr, QS3X_FP64);
  a.nnz_out = d_n.p;
  const int grid = persistent_grid(a.ntiles);
  hipStream_t st = stream();
  const bool normlist = s->d_rclass.p == nullptr;
  KFn f = (s->mode != kIndexU1 && use_chain(s, QS3X_FP64))
              ? chain_kernel(s->nod, QS3X_FP64, true, normlist, true, false) : nullptr;
  if (s->mode == kIndexU1)
    if (s->model->NOS > 64)
      hipLaunchKernelGGL((k_hv_u1<double, true, true>), dim3(grid), dim3(kThreads), 0, st, a, s->nod);
    else
      hipLaunchKernelGGL((k_hv_u1<double, true, false>), dim3(grid), dim3(kThreads), 0, st, a, s->nod);
  else if (f) {
    allow_lds(f, chain_lds_bytes(s, QS3X_FP64));
    hipLaunchKernelGGL(f, dim3(grid), dim3(kThreads), chain_lds_bytes(s, QS3X_FP64), st, a);
  }
  else
    hipLaunchKernelGGL((k_hv_general<double, true, true>), dim3(grid), dim3(kThreads), 0, st, a, s->nod);
  QS3X_HIP(hipGetLastError());
  d_n.download(&nz, 1);
  return (int64_t)nz;
}

}  // namespace qs3x
