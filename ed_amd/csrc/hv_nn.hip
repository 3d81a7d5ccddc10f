// hv_nn.hip -- H.v for chains whose every site has exactly the two bonds
// (s-1, s) and (s, s+1) (periodic nearest-neighbour XXZ chains, any mix of
// bond types).  Same operator and results as the general chain kernels of
// hv.hip (ham_to_vec_wave_vector, ham2vec_v3.f90:37-72); the structure of a
// nearest-neighbour hop removes the representative search from almost every
// row:
//
// A row is a representative anchored at site 1: down spins n_0 = 1 < n_1 <
// ... < n_{N-1} with gaps g_q = n_{q+1} - n_q and the wrap gap
// W = g_{N-1} = L + 1 - n_{N-1} >= every other gap.  A hop moves one down spin
// by one site, changing two cyclically adjacent gaps by +-1.  When the largest
// other gap M satisfies M <= W - 3 (a "clear" row), every hop's image has a
// unique largest gap, so its representative has stabiliser order 1 and is
// known in closed form:
//  * spin m >= 1 to n_m +- 1: the image contains site 1 and is its own
//    representative (ell = 0); its anchored rank is the row's plus or minus
//    one binomial, C(n_m - 2, m - 1) or C(n_m - 3, m - 1) (Pascal's rule);
//  * the anchor to site 2: representative = image translated by -1
//    (ell = L - 1), anchored rank sum_q C(n_q - 3, q);
//  * the anchor to site L: representative = image translated by +1
//    (ell = 1), anchored rank sum_q C(n_q - 1, q).
// Rows with M > W - 3 (under 1 % of the rows at L = 1000) are queued per
// wavefront in LDS and processed 64 at a time with the backward-gap key search
// of the general chain kernel on every hop, so the search runs on full
// wavefronts and the clear rows never execute it.
//
// Rows are looked up in the sector's bitmap index (8 bytes per 32 anchored
// ranks: presence bits + row of the first present rank; 41 MB at L = 1000,
// N_down = 4 instead of the 663 MB int32 table, so the lookups stay in L2 /
// MALL).  All index reads of a row are issued together, then all gathers.
//
// The kernels are bound by their count of vector-memory instructions (a
// wavefront load costs the texture path the same whether a lane moves 8 or
// 16 bytes), so they save loads where they can: spin-1 neighbours come from
// the adjacent lanes, and in FP32 one cooperative load reads the index words
// of a whole wavefront, and k_hv_chain_nn2 gives each lane two adjacent rows
// so that its streams and most gathers move 16 bytes per lane.
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "blas.h"
#include "hv_common.h"
#include "nt.h"
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

// Per-wavefront LDS queue of rows that need the key search.
constexpr int kQueue = 128;

// Shared kernel state (LDS tables and per-launch constants).
template <int N, class R>
struct NnCtx {
  int L, kb;
  const int32_t *s_binom;
  const R *s_rcl, *s_cfc;  // r(class); r(1) / r(class)
  const R *s_jxy2, *s_jz4;
  const uint16_t *s_bt;
};

// Index reads, gathers, accumulation, store and dot partial of RPT rows whose
// H hop slots hold (anchored rank or -1, coefficient, phase index).  Rows with
// store[k] false are skipped (their slots must be -1).  All index reads of the
// RPT rows are issued before any gather.  DIRECT1 (clear rows): the hops of
// spin 1 (slots 2, 3) land on anchored rank crow +- 1, which is itself a
// representative, so their rows are row +- 1 and need no index read; the
// other hops need no presence test.
template <int N, class R, bool REALPH, int RPT, bool DIRECT1 = false>
__device__ __forceinline__ void finish_rows(const HvArgs &a, const int64_t *row, const typename CT<R>::T *xin,
                                            const bool *store, const int (*hc)[2 * N], const R (*hcf)[2 * N],
                                            const int (*hei)[2 * N], const R *anti, R scale, double &dre,
                                            double &dim_) {
  using C = typename CT<R>::T;
  constexpr int H = 2 * N;
  const C *__restrict__ v1 = (const C *)a.v1;
  uint2 wd[RPT][H];
  // The cooperative read pays in FP32 with two rows per lane (-12 % at
  // L = 1000, N = 4); in FP64 its cross-lane moves cost more than the loads
  // it saves (+10 %).
  constexpr bool COOP = DIRECT1 && sizeof(R) == 4 && N <= 5;
  if constexpr (!COOP) {
#pragma unroll
    for (int k = 0; k < RPT; ++k)
#pragma unroll
      for (int h = 0; h < H; ++h) wd[k][h] = a.bitmap[(uint32_t)(hc[k][h] >= 0 ? hc[k][h] : 0) >> 5];
  } else {
    // Cooperative index read.  The 64 rows of a wavefront are consecutive, so
    // for each hop slot their anchored ranks are nearly consecutive and fall
    // in a few words of the bitmap (2 or fewer for 90 % of the wavefronts at
    // L = 1000, N = 4).  One load fetches kSpan words for every (row, slot)
    // pair, starting at the word of the first lane that needs the slot; each
    // lane takes its word from the lane that read it, and only lanes whose
    // word lies outside that window read it themselves.
    constexpr int NB = H - 2;  // slots that read the bitmap (all but 2 and 3)
    constexpr int kSpan = 4;
    static_assert(RPT * NB * kSpan <= 64, "one word per lane");
    const int lane = threadIdx.x & 63;
    uint32_t base[RPT][H];  // wavefront-uniform
#pragma unroll
    for (int k = 0; k < RPT; ++k)
#pragma unroll
      for (int h = 0; h < H; ++h) {
        if (h == 2 || h == 3) continue;
        const uint64_t m = __ballot(hc[k][h] >= 0);
        const int first = m ? __builtin_ctzll(m) : 0;
        base[k][h] = (uint32_t)__builtin_amdgcn_readlane((int)((uint32_t)max(hc[k][h], 0) >> 5), first);
      }
    uint32_t cb = 0;
#pragma unroll
    for (int k = 0; k < RPT; ++k)
#pragma unroll
      for (int h = 0; h < H; ++h) {
        if (h == 2 || h == 3) continue;
        const int slot = k * NB + (h < 2 ? h : h - 2);
        cb = ((lane >> 2) == slot) ? base[k][h] : cb;
      }
    const uint32_t cwi = min(cb + (uint32_t)(lane & (kSpan - 1)), (uint32_t)(a.nbitmap - 1));
    const uint2 cw = a.bitmap[cwi];
#pragma unroll
    for (int k = 0; k < RPT; ++k)
#pragma unroll
      for (int h = 0; h < H; ++h) {
        if (h == 2 || h == 3) continue;
        const int slot = k * NB + (h < 2 ? h : h - 2);
        const uint32_t q = (uint32_t)max(hc[k][h], 0) >> 5;
        const uint32_t j = q - base[k][h];
        const bool inr = j < (uint32_t)kSpan;
        const int src = slot * kSpan + (inr ? (int)j : 0);
        uint2 w;
        w.x = (uint32_t)__shfl((int)cw.x, src, 64);
        w.y = (uint32_t)__shfl((int)cw.y, src, 64);
        if (hc[k][h] >= 0 && !inr) w = a.bitmap[q];
        wd[k][h] = w;
      }
  }
  // DIRECT1: a wavefront holds 64 consecutive rows, so rows row + 1 and row - 1
  // are the neighbouring lanes' rows and their elements are those lanes'
  // inputs.  Only lane 63 (row + 1) and lane 0 (row - 1) read memory, in one
  // load.  The vector-memory data path is the kernel's bound, so two loads
  // per row fewer is what counts.
  C ynb[RPT][2];
  if constexpr (DIRECT1) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      C up, dn;
      up.x = __shfl_down(xin[k].x, 1, 64);
      up.y = __shfl_down(xin[k].y, 1, 64);
      dn.x = __shfl_up(xin[k].x, 1, 64);
      dn.y = __shfl_up(xin[k].y, 1, 64);
      const bool e2 = lane == 63 && hc[k][2] >= 0, e3 = lane == 0 && hc[k][3] >= 0;
      C ye = up;
      if (e2 || e3) ye = v1[(uint64_t)(uint32_t)(row[k] + (e2 ? 1 : -1))];
      ynb[k][0] = e2 ? ye : up;
      ynb[k][1] = e3 ? ye : dn;
    }
  }
  // A hop that lands nowhere gathers the row's own element with coefficient 0.
  // Rows fit in 32 bits (the index table is int32), so the gather index is
  // zero-extended rather than selected in 64 bits.
  uint32_t hid[RPT][H];
  bool hin[RPT][H];
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    // rows that are not stored may lie past the end of the sector
    const uint32_t self = store[k] ? (uint32_t)row[k] : 0u;
#pragma unroll
    for (int h = 0; h < H; ++h) {
      const bool ok = hc[k][h] >= 0;
      uint32_t id;
      bool in = ok;
      if (DIRECT1 && (h == 2 || h == 3)) {
        id = (uint32_t)row[k] + (h == 2 ? 1u : ~0u);
      } else {
        const uint32_t b = (uint32_t)hc[k][h] & 31u;
        id = wd[k][h].y + (uint32_t)__popc(wd[k][h].x & ((1u << b) - 1u));
        // A clear row's images all have stabiliser order 1, so each one that
        // is reached is a representative of the sector: no presence test.
        if (!DIRECT1) in = ok & (((wd[k][h].x >> b) & 1u) != 0u);
      }
      hin[k][h] = in;
      hid[k][h] = in ? id : self;
    }
  }
  C y[RPT][H];
#pragma unroll
  for (int k = 0; k < RPT; ++k)
#pragma unroll
    for (int h = 0; h < H; ++h)
      y[k][h] = (DIRECT1 && (h == 2 || h == 3)) ? ynb[k][h - 2] : v1[(uint64_t)hid[k][h]];
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    if (!store[k]) continue;
    R accx = 0, accy = 0;
#pragma unroll
    for (int h = 0; h < H; ++h) {
      const R cf = hin[k][h] ? hcf[k][h] : (R)0;
      if constexpr (REALPH) {
        accx += cf * y[k][h].x;
        accy += cf * y[k][h].y;
      } else {
        const R pr = (R)a.explist[2 * hei[k][h]] * cf, pi = (R)a.explist[2 * hei[k][h] + 1] * cf;
        accx += pr * y[k][h].x - pi * y[k][h].y;
        accy += pr * y[k][h].y + pi * y[k][h].x;
      }
    }
    C x = xin[k];
    x.x *= scale;
    x.y *= scale;
    const R dg = (sizeof(R) == 8 ? (R)a.ezz : (R)0) - (R)2 * anti[k];
    C out;
    out.x = dg * x.x + accx * scale;
    out.y = dg * x.y + accy * scale;
    st_nt(&((C *)a.v0)[row[k]], out);  // written once, read by the next pass
    if (a.dot_partials) {
      dre += (double)x.x * (double)out.x + (double)x.y * (double)out.y;
      dim_ += (double)x.x * (double)out.y - (double)x.y * (double)out.x;
    }
  }
}

// Hop slots of a row through the key search: target anchored rank (-1: no
// hop), coefficient, explist index; anti = the antiparallel bonds' Jz/4 sum.
template <int N, class R, bool UNI, class KeyT>
__device__ __forceinline__ void slow_row_hops(const HvArgs &a, const NnCtx<N, R> &k, uint64_t w, int *hc, R *hcf,
                                              int *hei, R &anti) {
  constexpr int H = 2 * N;
  const int L = k.L;
  int n[N];
#pragma unroll
  for (int q = 0; q < N; ++q) n[q] = (int)((w >> (q * a.bits)) & ((1ull << a.bits) - 1ull));
  int g[N];
#pragma unroll
  for (int q = 0; q < N - 1; ++q) g[q] = n[q + 1] - n[q];
  g[N - 1] = L + 1 - n[N - 1];
  const R rinv = k.s_cfc[(int)(w >> 56)] / k.s_rcl[1];
  anti = 0;
#pragma unroll
  for (int h = 0; h < H; ++h) {
    const int m = h >> 1;
    const bool right = (h & 1) == 0;
    const bool ok = right ? (g[m] > 1) : ((m == 0 ? g[N - 1] : g[m - 1]) > 1);
    const int bsite = right ? n[m] : (m == 0 ? L : n[m] - 1);  // bond (bsite, bsite + 1)
    const int bt = UNI ? 0 : k.s_bt[bsite];
    anti += ok ? k.s_jz4[bt] : (R)0;
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
    int c = -1, count = 1, ei = 0;
    if (ok) chain_search<N, KeyT>(nn, L, k.kb, k.s_binom, c, count, ei);
    hc[h] = ok ? c : -1;
    hcf[h] = k.s_jxy2[bt] * k.s_rcl[count] * rinv;
    hei[h] = ei;
  }
}

// A queued row: every hop through the key search.
template <int N, class R, bool REALPH, bool UNI, class KeyT>
__device__ __forceinline__ void slow_row(const HvArgs &a, const NnCtx<N, R> &k, int64_t row, R scale, double &dre,
                                         double &dim_) {
  using C = typename CT<R>::T;
  constexpr int H = 2 * N;
  const uint64_t w = a.cfg[row];
  const C xin = ((const C *)a.v1)[row];
  R anti = 0;
  int hc[H], hei[H];
  R hcf[H];
  slow_row_hops<N, R, UNI, KeyT>(a, k, w, hc, hcf, hei, anti);
  const bool store = true;
  finish_rows<N, R, REALPH, 1>(a, &row, &xin, &store, &hc, &hcf, &hei, &anti, scale, dre, dim_);
}

// The fused Lanczos step of HvArgs for a queued row: every element it reads
// is formed as v = c1 P - c2 v' from the two input vectors.
template <int N, class R, bool REALPH, class KeyT>
__device__ __forceinline__ void slow_row_fused(const HvArgs &a, const NnCtx<N, R> &k, const FusedCoef<R> &fc,
                                               int64_t row, double (&acc4)[4]) {
  using C = typename CT<R>::T;
  constexpr int H = 2 * N;
  const C *__restrict__ v1 = (const C *)a.v1;
  const C *__restrict__ vp = (const C *)a.vprev;
  const uint64_t w = a.cfg[row];
  R anti = 0;
  int hc[H], hei[H];
  R hcf[H];
  slow_row_hops<N, R, true, KeyT>(a, k, w, hc, hcf, hei, anti);
  uint2 wd[H];
#pragma unroll
  for (int h = 0; h < H; ++h) wd[h] = a.bitmap[(uint32_t)(hc[h] >= 0 ? hc[h] : 0) >> 5];
  C vo, p[H + 1], q[H + 1];
  vo.x = vo.y = (R)0;
  bool hin[H];
#pragma unroll
  for (int h = 0; h < H; ++h) {
    const uint32_t b = (uint32_t)hc[h] & 31u;
    hin[h] = hc[h] >= 0 && ((wd[h].x >> b) & 1u);
    const uint32_t id = hin[h] ? wd[h].y + (uint32_t)__popc(wd[h].x & ((1u << b) - 1u)) : (uint32_t)row;
    p[h] = v1[id];
    q[h].x = q[h].y = (R)0;
    if (fc.hasv) q[h] = vp[id];
  }
  p[H] = v1[row];
  q[H].x = q[H].y = (R)0;
  if (fc.hasv) q[H] = vp[row];
  vo = q[H];
  const C x = fused_v<R>(fc, p[H], q[H]);
  R accx = 0, accy = 0;
#pragma unroll
  for (int h = 0; h < H; ++h) {
    const C y = fused_v<R>(fc, p[h], q[h]);
    const R cf = hin[h] ? hcf[h] : (R)0;
    if constexpr (REALPH) {
      accx += cf * y.x;
      accy += cf * y.y;
    } else {
      const R pr = (R)a.explist[2 * hei[h]] * cf, pi = (R)a.explist[2 * hei[h] + 1] * cf;
      accx += pr * y.x - pi * y.y;
      accy += pr * y.y + pi * y.x;
    }
  }
  const R dg = (sizeof(R) == 8 ? (R)a.ezz : (R)0) - (R)2 * anti;
  C out, pn;
  out.x = dg * x.x + accx;
  out.y = dg * x.y + accy;
  pn.x = out.x - fc.c3 * vo.x;
  pn.y = out.y - fc.c3 * vo.y;
  st_nt(&((C *)a.xout)[row], x);
  st_nt(&((C *)a.v0)[row], pn);
  acc4[0] += (double)x.x * (double)out.x + (double)x.y * (double)out.y;
  acc4[1] += (double)pn.x * (double)pn.x + (double)pn.y * (double)pn.y;
  acc4[2] += (double)x.x * (double)pn.x + (double)x.y * (double)pn.y;
  acc4[3] += (double)x.x * (double)x.x + (double)x.y * (double)x.y;
}

// Hop slots of a clear row (see the file comment); returns false, leaving
// the slots untouched, when the row needs the key search.
template <int N, class R, bool REALPH, bool UNI>
__device__ __forceinline__ bool clear_slots(const NnCtx<N, R> &k, int bits, uint64_t w, R jxy2u, R jz4u, int ei_ar,
                                            int ei_al, int *hc, R *hcf, int *hei, R &anti) {
  constexpr int H = 2 * N;
  const int L = k.L;
  const int32_t *s_binom = k.s_binom;
  int n[N];
#pragma unroll
  for (int q = 0; q < N; ++q) n[q] = (int)((w >> (q * bits)) & ((1ull << bits) - 1ull));
  int g[N];
#pragma unroll
  for (int q = 0; q < N - 1; ++q) g[q] = n[q + 1] - n[q];
  const int W = L + 1 - n[N - 1];
  g[N - 1] = W;
  int M = 0;
#pragma unroll
  for (int q = 0; q < N - 1; ++q) M = max(M, g[q]);
  if (!(M <= W - 3)) return false;
  const R cfc = k.s_cfc[(int)(w >> 56)];  // r(1) / r(row)
  int crow = 0, c_ar = 0, c_al = 0;
#pragma unroll
  for (int q = 1; q < N; ++q) {
    const int base = (q - 1) * (L - 1) + n[q];
    crow += s_binom[base - 2];
    c_ar += s_binom[max(base - 3, (q - 1) * (L - 1))];  // C(n_q - 3, q); clamped when n_1 = 2
    c_al += s_binom[base - 1];
  }
  R an = 0;
  int nanti = 0;
#pragma unroll
  for (int h = 0; h < H; ++h) {
    const int m = h >> 1;
    const bool right = (h & 1) == 0;
    const bool ok = right ? (g[m] > 1) : ((m == 0 ? W : g[m - 1]) > 1);
    int c;
    if (m == 0) c = right ? c_ar : c_al;
    else if (m == 1) c = crow + (right ? 1 : -1);
    else c = right ? crow + s_binom[(m - 2) * (L - 1) + n[m] - 2] : crow - s_binom[(m - 2) * (L - 1) + n[m] - 3];
    hc[h] = ok ? c : -1;
    if constexpr (UNI) {
      nanti += ok ? 1 : 0;
      hcf[h] = jxy2u * cfc;
    } else {
      const int bt = k.s_bt[right ? n[m] : (m == 0 ? L : n[m] - 1)];
      an += ok ? k.s_jz4[bt] : (R)0;
      hcf[h] = k.s_jxy2[bt] * cfc;
    }
    if constexpr (!REALPH) hei[h] = (m == 0) ? (right ? ei_ar : ei_al) : 0;
  }
  anti = UNI ? jz4u * (R)nanti : an;
  return true;
}

// LDS tables of the nearest-neighbour kernels (one block-wide pass, then a
// barrier).
template <int N, class R, bool UNI>
__device__ __forceinline__ NnCtx<N, R> nn_setup(const HvArgs &a, unsigned char *smem) {
  const int L = a.L;
  // LDS: binom [N-1][L-1] int32 | rclass [N+1] R | r1/rclass [N+1] R | jxy2 [nbt] R | jz4 [nbt] R |
  //      bond type [L+1] u16
  const int nb = (N - 1) * (L - 1);
  int32_t *s_binom = (int32_t *)smem;
  R *s_rcl = (R *)(smem + (((size_t)nb * 4 + 15) & ~(size_t)15));
  R *s_cfc = s_rcl + (N + 1);
  R *s_jxy2 = s_cfc + (N + 1);
  R *s_jz4 = s_jxy2 + a.nbt;
  uint16_t *s_bt = (uint16_t *)(s_jz4 + a.nbt);
  for (int q = threadIdx.x; q < nb; q += kThreads) {
    const int m = q / (L - 1) + 1, x = q % (L - 1);
    s_binom[q] = (int32_t)a.binom[(int64_t)x * a.kdim + m];
  }
  for (int q = threadIdx.x; q <= N; q += kThreads) {
    s_rcl[q] = (R)a.rclass[q];
    s_cfc[q] = (R)(a.rclass[1] / a.rclass[q]);  // 0 < r(q): classes present in the sector
  }
  for (int q = threadIdx.x; q < a.nbt; q += kThreads) {
    s_jxy2[q] = (R)a.bt[2 * q];
    s_jz4[q] = (R)a.bt[2 * q + 1];
  }
  if (!UNI)
    for (int q = threadIdx.x; q <= L; q += kThreads) s_bt[q] = a.nn_bt[q];
  __syncthreads();

  return NnCtx<N, R>{L, a.kbits, s_binom, s_rcl, s_cfc, s_jxy2, s_jz4, s_bt};
}

// RPT rows per lane per iteration: two in FP32 up to N_down = 5 (more
// independent loads in flight per wavefront; 2.5 % faster at L = 1000, N = 4),
// one otherwise (the second row's registers cost occupancy: FP64 2.4 % slower).
// FP32 with two rows per lane is held to 5 wavefronts per SIMD (82 VGPRs
// instead of 97, which would allow 4): the kernel is bound by its
// vector-memory instruction stream, and the fifth wavefront is worth more
// than the registers.
template <int N, class R, bool REALPH, bool UNI, class KeyT>
__global__ __launch_bounds__(kThreads) void
__attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 && N <= 5 ? 5 : 1)))
k_hv_chain_nn(HvArgs a) {
  static_assert(N >= 2, "N_down = 1 uses the general chain kernel");
  using C = typename CT<R>::T;
  constexpr int H = 2 * N;  // hop slots: 2m = spin m to the right, 2m + 1 = to the left
  constexpr int RPT = (sizeof(R) == 4 && N <= 5) ? 2 : 1;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int32_t s_queue[kThreads / 64][kQueue];
  const int L = a.L;
  const NnCtx<N, R> k = nn_setup<N, R, UNI>(a, smem);
  const C *__restrict__ v1 = (const C *)a.v1;
  const R scale = a.scale ? (R)(*a.scale) : (R)1;
  const R jxy2u = k.s_jxy2[0], jz4u = k.s_jz4[0];
  const int ei_ar = (L - 1) + 3 * (L + 1), ei_al = 1 + 3 * (L + 1);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int32_t *queue = s_queue[wid];
  int qhead = 0, qcount = 0;  // wave-uniform
  double dre = 0.0, dim_ = 0.0;
  const int64_t nsuper = (a.dim + (int64_t)kThreads * RPT - 1) / ((int64_t)kThreads * RPT);

  const TileMap tm = tile_map(nsuper, a.nxcd, a.chunk_log2);
  for (int64_t v = tm.first; v < tm.end; v += tm.stride) {
    const int64_t t = tile_at(tm, v);
    int64_t row[RPT];
    bool valid[RPT], clear[RPT];
    uint64_t w[RPT];
    C xin[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      row[r] = (t * RPT + r) * kThreads + threadIdx.x;
      valid[r] = row[r] < a.dim;
      w[r] = valid[r] ? a.cfg[row[r]] : 0ull;
      if (valid[r]) xin[r] = v1[row[r]];
      else xin[r].x = xin[r].y = (R)0;
    }
    int hc[RPT][H], hei[RPT][REALPH ? 1 : H];
    R hcf[RPT][H], anti[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      clear[r] = valid[r] && clear_slots<N, R, REALPH, UNI>(k, a.bits, w[r], jxy2u, jz4u, ei_ar, ei_al, hc[r],
                                                             hcf[r], REALPH ? nullptr : hei[r], anti[r]);
      if (!clear[r]) {
#pragma unroll
        for (int h = 0; h < H; ++h) {
          hc[r][h] = -1;
          hcf[r][h] = (R)0;
          if constexpr (!REALPH) hei[r][h] = 0;
        }
        anti[r] = (R)0;
      }
    }
    if constexpr (REALPH) {
      int hz[RPT][H];
      finish_rows<N, R, REALPH, RPT, true>(a, row, xin, clear, hc, hcf, hz, anti, scale, dre, dim_);
    } else {
      finish_rows<N, R, REALPH, RPT, true>(a, row, xin, clear, hc, hcf, hei, anti, scale, dre, dim_);
    }
    // queue the rows that need the key search (order: slice, then lane)
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const bool needs = valid[r] && !clear[r];
      const uint64_t mask = __ballot(needs);
      if (mask != 0ull) {
        const int before = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
        if (needs) queue[(qhead + qcount + before) & (kQueue - 1)] = (int32_t)row[r];
        qcount += __popcll(mask);
        __builtin_amdgcn_wave_barrier();
        if (qcount >= 64) {
          const int64_t rr = queue[(qhead + lane) & (kQueue - 1)];
          __builtin_amdgcn_wave_barrier();
          qhead = (qhead + 64) & (kQueue - 1);
          qcount -= 64;
          slow_row<N, R, REALPH, UNI, KeyT>(a, k, rr, scale, dre, dim_);
        }
      }
    }
  }
  if (qcount > 0) {
    __builtin_amdgcn_wave_barrier();
    if (lane < qcount) {
      const int64_t rr = queue[(qhead + lane) & (kQueue - 1)];
      slow_row<N, R, REALPH, UNI, KeyT>(a, k, rr, scale, dre, dim_);
    }
  }
  if (a.dot_partials) block_reduce_store(dre, dim_, a.dot_partials);
}

typedef float nn_f4 __attribute__((ext_vector_type(4)));
typedef float nn_f4a8 __attribute__((ext_vector_type(4))) __attribute__((aligned(8)));
typedef unsigned long long nn_u64x2 __attribute__((ext_vector_type(2)));

// FP32, real phases, N_down <= 4: each lane holds two adjacent rows, row0 = 2 * lane and
// row0 + 1 of its wavefront's 128 consecutive rows.  A vector-memory
// instruction costs the same whether a lane moves 8 or 16 bytes (measured on
// L2-resident data: 16-19 cycles per wavefront load per CU for both), and
// this kernel is bound by that instruction stream, so the rows' configuration
// words, inputs and outputs move as 16-byte pairs, and a hop whose targets for
// the two rows are adjacent rows (the common case: the targets of consecutive
// rows are consecutive) gathers both with one 16-byte load.  The index read
// is the cooperative one of finish_rows with one window of kSpan words per
// hop slot covering both rows.  Requires 16-byte aligned vectors.
template <int N, bool REALPH, bool UNI, class KeyT>
__global__ __launch_bounds__(kThreads) void __attribute__((amdgpu_waves_per_eu(5))) k_hv_chain_nn2(HvArgs a) {
  static_assert(N >= 2 && N <= 4, "two-row kernel: 2 <= N_down <= 4");
  using R = float;
  using C = float2;
  constexpr int H = 2 * N;
  constexpr int NB = H - 2;  // hop slots that read the bitmap (all but 2 and 3)
  constexpr int kSpan = 8;
  static_assert(NB * kSpan <= 64, "one bitmap word per lane");
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int32_t s_queue[kThreads / 64][kQueue];
  const int L = a.L;
  const NnCtx<N, R> k = nn_setup<N, R, UNI>(a, smem);
  const C *__restrict__ v1 = (const C *)a.v1;
  C *__restrict__ v0 = (C *)a.v0;
  const R scale = a.scale ? (R)(*a.scale) : (R)1;
  const R jxy2u = k.s_jxy2[0], jz4u = k.s_jz4[0];
  const int ei_ar = (L - 1) + 3 * (L + 1), ei_al = 1 + 3 * (L + 1);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int32_t *queue = s_queue[wid];
  int qhead = 0, qcount = 0;  // wave-uniform
  double dre = 0.0, dim_ = 0.0;
  const int64_t dim = a.dim;
  const int64_t ntile = (dim + 2 * kThreads - 1) / (2 * kThreads);

  const TileMap tm = tile_map(ntile, a.nxcd, a.chunk_log2);
  for (int64_t v = tm.first; v < tm.end; v += tm.stride) {
    const int64_t t = tile_at(tm, v);
    const int64_t r0 = t * (2 * kThreads) + wid * 128 + 2 * lane;
    const int64_t row[2] = {r0, r0 + 1};
    const bool valid[2] = {r0 < dim, r0 + 1 < dim};
    uint64_t w[2];
    C xin[2];
    if (valid[1]) {
      const nn_u64x2 cw = *(const nn_u64x2 *)(a.cfg + r0);
      const nn_f4 xv = *(const nn_f4 *)(v1 + r0);
      w[0] = cw.x;
      w[1] = cw.y;
      xin[0] = make_float2(xv.x, xv.y);
      xin[1] = make_float2(xv.z, xv.w);
    } else {
      const int64_t rr = min(r0, dim - 1);
      w[0] = a.cfg[rr];
      xin[0] = v1[rr];
      w[1] = 0ull;
      xin[1] = make_float2(0.f, 0.f);
    }
    int hc[2][H], hei[2][REALPH ? 1 : H];
    R hcf[2][H], anti[2];
    bool clear[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      clear[r] = valid[r] && clear_slots<N, R, REALPH, UNI>(k, a.bits, w[r], jxy2u, jz4u, ei_ar, ei_al, hc[r],
                                                             hcf[r], REALPH ? nullptr : hei[r], anti[r]);
      if (!clear[r]) {
#pragma unroll
        for (int h = 0; h < H; ++h) {
          hc[r][h] = -1;
          hcf[r][h] = (R)0;
          if constexpr (!REALPH) hei[r][h] = 0;
        }
        anti[r] = (R)0;
      }
    }
    // cooperative index read: window per hop slot starts at the word of the
    // first lane whose row 0 (else row 1) has the hop
    uint32_t base[H];
#pragma unroll
    for (int h = 0; h < H; ++h) {
      if (h == 2 || h == 3) continue;
      const uint64_t m0 = __ballot(hc[0][h] >= 0), m1 = __ballot(hc[1][h] >= 0);
      const uint64_t m = m0 ? m0 : m1;
      const int first = m ? __builtin_ctzll(m) : 0;
      const int qf = (int)((uint32_t)max(m0 ? hc[0][h] : hc[1][h], 0) >> 5);
      base[h] = (uint32_t)__builtin_amdgcn_readlane(qf, first);
    }
    uint32_t cb = 0;
#pragma unroll
    for (int h = 0; h < H; ++h) {
      if (h == 2 || h == 3) continue;
      cb = ((lane / kSpan) == (h < 2 ? h : h - 2)) ? base[h] : cb;
    }
    const uint2 cw = a.bitmap[min(cb + (uint32_t)(lane % kSpan), (uint32_t)(a.nbitmap - 1))];
    uint32_t hid[2][H];
    bool hin[2][H];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int h = 0; h < H; ++h) {
        hin[r][h] = hc[r][h] >= 0;
        hid[r][h] = 0u;
        if (h == 2 || h == 3) continue;
        const uint32_t q = (uint32_t)max(hc[r][h], 0) >> 5;
        const uint32_t j = q - base[h];
        const bool inr = j < (uint32_t)kSpan;
        const int src = (h < 2 ? h : h - 2) * kSpan + (inr ? (int)j : 0);
        uint2 wd;
        wd.x = (uint32_t)__shfl((int)cw.x, src, 64);
        wd.y = (uint32_t)__shfl((int)cw.y, src, 64);
        if (hc[r][h] >= 0 && !inr) wd = a.bitmap[q];
        // a clear row's reachable images are all representatives of the sector
        const uint32_t b = (uint32_t)hc[r][h] & 31u;
        hid[r][h] = wd.y + (uint32_t)__popc(wd.x & ((1u << b) - 1u));
      }
    // spin-1 hops (slots 2, 3) of a clear row land on rows row +- 1: the other
    // row of the pair, or the neighbouring lane's; lanes 0 and 63 read the
    // wavefront's edge rows in one load
    C y[2][H];
    {
      C up0, dn1;
      up0.x = __shfl_down(xin[0].x, 1, 64);
      up0.y = __shfl_down(xin[0].y, 1, 64);
      dn1.x = __shfl_up(xin[1].x, 1, 64);
      dn1.y = __shfl_up(xin[1].y, 1, 64);
      const bool e_hi = lane == 63 && hin[1][2], e_lo = lane == 0 && hin[0][3];
      C ye = up0;
      if (e_hi || e_lo) ye = v1[e_hi ? r0 + 2 : r0 - 1];
      y[0][2] = xin[1];
      y[0][3] = e_lo ? ye : dn1;
      y[1][2] = e_hi ? ye : up0;
      y[1][3] = xin[0];
    }
#pragma unroll
    for (int h = 0; h < H; ++h) {
      if (h == 2 || h == 3) continue;
      const bool pair = hin[0][h] && hin[1][h] && hid[1][h] == hid[0][h] + 1u;
      if (pair) {
        const nn_f4a8 g = *(const nn_f4a8 *)(v1 + hid[0][h]);
        y[0][h] = make_float2(g.x, g.y);
        y[1][h] = make_float2(g.z, g.w);
      } else {
        y[0][h] = hin[0][h] ? v1[hid[0][h]] : make_float2(0.f, 0.f);
        y[1][h] = hin[1][h] ? v1[hid[1][h]] : make_float2(0.f, 0.f);
      }
    }
    C out[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      R accx = 0, accy = 0;
#pragma unroll
      for (int h = 0; h < H; ++h) {
        const R cf = hin[r][h] ? hcf[r][h] : (R)0;
        if constexpr (REALPH) {
          accx += cf * y[r][h].x;
          accy += cf * y[r][h].y;
        } else {
          const R pr = (R)a.explist[2 * hei[r][h]] * cf, pi = (R)a.explist[2 * hei[r][h] + 1] * cf;
          accx += pr * y[r][h].x - pi * y[r][h].y;
          accy += pr * y[r][h].y + pi * y[r][h].x;
        }
      }
      C x = xin[r];
      x.x *= scale;
      x.y *= scale;
      const R dg = -(R)2 * anti[r];
      out[r].x = dg * x.x + accx * scale;
      out[r].y = dg * x.y + accy * scale;
      if (a.dot_partials && clear[r]) {
        dre += (double)x.x * (double)out[r].x + (double)x.y * (double)out[r].y;
        dim_ += (double)x.x * (double)out[r].y - (double)x.y * (double)out[r].x;
      }
    }
    if (clear[0] && clear[1]) {
      nn_f4 o;
      o.x = out[0].x;
      o.y = out[0].y;
      o.z = out[1].x;
      o.w = out[1].y;
      __builtin_nontemporal_store(o, (nn_f4 *)(v0 + r0));
    } else {
      if (clear[0]) st_nt(&v0[r0], out[0]);
      if (clear[1]) st_nt(&v0[r0 + 1], out[1]);
    }
    // queue the rows that need the key search
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const bool needs = valid[r] && !clear[r];
      const uint64_t mask = __ballot(needs);
      if (mask != 0ull) {
        const int before = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
        if (needs) queue[(qhead + qcount + before) & (kQueue - 1)] = (int32_t)row[r];
        qcount += __popcll(mask);
        __builtin_amdgcn_wave_barrier();
        if (qcount >= 64) {
          const int64_t rr = queue[(qhead + lane) & (kQueue - 1)];
          __builtin_amdgcn_wave_barrier();
          qhead = (qhead + 64) & (kQueue - 1);
          qcount -= 64;
          slow_row<N, R, REALPH, UNI, KeyT>(a, k, rr, scale, dre, dim_);
        }
      }
    }
  }
  if (qcount > 0) {
    __builtin_amdgcn_wave_barrier();
    if (lane < qcount) {
      const int64_t rr = queue[(qhead + lane) & (kQueue - 1)];
      slow_row<N, R, REALPH, UNI, KeyT>(a, k, rr, scale, dre, dim_);
    }
  }
  if (a.dot_partials) block_reduce_store(dre, dim_, a.dot_partials);
}

// ---------------------------------------------------------------------------
// Gap-lattice stencil H.v: nearest-neighbour chains with uniform bonds and
// N_down = 4.  A row is the anchored representative (1, a, b, c) (a = n_2,
// b = n_3, c = n_4); rows ascend in colex order, i.e. by (c, b, a), so the rows
// of one "line" (b, c) are consecutive in a.  For a clear row (see the file
// comment) every hop lands on a fixed neighbour of (a, b, c), a
// representative of stabiliser order 1:
//   spin a: (a +- 1, b, c)   spin b: (a, b +- 1, c)   spin c: (a, b, c +- 1)
//   anchor to site 2: (a - 1, b - 1, c - 1), phase explist(L - 1)
//   anchor to site L: (a + 1, b + 1, c + 1), phase explist(1)
// so H.v is an 8-point stencil on the (a, b, c) lattice.  A task is a tile of
// kSegA values of a (a0 ..) by B lines (b0 ..) swept over the slabs c0 .. c1 - 1.
// A slab's window -- the tile plus a halo of one in a and b: B + 2 line
// segments of kSegA + 2 = 64 positions, each one wavefront load of consecutive
// rows -- goes into one slot of a four-slot LDS ring: slab c is computed from
// slots c - 1, c, c + 1 while slab c + 2 is in flight in registers, with one
// barrier per slab.  Every element of v1 is read about (B + 2) 64 / (B kSegA)
// times instead of once per hop that reaches it, and no configuration word or
// index is read.  Rows that are not clear go through the per-wavefront queue
// and the key search of slow_row, as in the kernels above.
// kSegA a-values per tile, kSegW window positions per line (a halo of one
// each side), kRing LDS slots: hv_common.h
constexpr int kMaxLines = 8;  // lines per tile of this (slab-window) form

// FUSED: the fused Lanczos step of HvArgs -- the window holds v = c1 P - c2 v'
// formed from two loaded vectors when it is stored to LDS, and a row writes v
// and P_new = H v - c3 v' (v' of its own row read at the start of the slab).
// BT: lines per tile (sa.B == BT), which sizes the window and round registers.
// D: slab windows in flight per block (2 or 3) while a slab is computed.
template <class R, bool REALPH, bool FUSED, int BT, int D>
__global__ __launch_bounds__(kThreads) void k_hv_stencil4(HvArgs a, StencilArgs sa) {
  using C = typename CT<R>::T;
  extern __shared__ __align__(16) unsigned char smem[];
  const int L = a.L;
  constexpr int B = BT, NL = B + 2;
  // LDS: ring [kRing][NL][kSegW] elements | line entries of the task's slabs
  // c0 - 1 .. c1: [slabs + 2][NL] int4
  C *s_ring = (C *)smem;
  int4 *s_tl = (int4 *)(smem + (size_t)kRing * NL * kSegW * sizeof(C));
  const C *__restrict__ v1 = (const C *)a.v1;
  const C *__restrict__ vp = (const C *)a.vprev;
  C *__restrict__ v0 = (C *)a.v0;
  C *__restrict__ xo = (C *)a.xout;
  FusedCoef<R> fc{(R)1, (R)0, (R)0, false};
  if constexpr (FUSED) fc = fused_coef<R>(a);
  const R scale = (!FUSED && a.scale) ? (R)(*a.scale) : (R)1;
  const R jxy2 = (R)a.bt[0], jz4 = (R)a.bt[1];
  const int ei_ar = (L - 1) + 3 * (L + 1), ei_al = 1 + 3 * (L + 1);
  R ph_ar_r = (R)1, ph_ar_i = (R)0, ph_al_r = (R)1, ph_al_i = (R)0;
  if constexpr (!REALPH) {
    ph_ar_r = (R)a.explist[2 * ei_ar];
    ph_ar_i = (R)a.explist[2 * ei_ar + 1];
    ph_al_r = (R)a.explist[2 * ei_al];
    ph_al_i = (R)a.explist[2 * ei_al + 1];
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double dre = 0.0, dim_ = 0.0;
  double acc4[4] = {0.0, 0.0, 0.0, 0.0};  // FUSED: Re<v,Hv>, |P|^2, Re<v,P>, |v|^2
  const int W1 = L + 1;
  constexpr int kPer = (BT + 2 + 3) / 4;  // window lines per wavefront
  constexpr int kRounds = (BT * kSegA + kThreads - 1) / kThreads;  // output rounds per slab

  // Registers of slab windows: wavefront w holds lines w, w + 4, ...; lane =
  // position a0 - 1 + lane.  Line entries come from LDS, so the loads issue at
  // once.  Two windows are in flight (sets X and Y, alternating) while a slab
  // is computed, so a block keeps about two slabs of loads outstanding.
  C pre[D][kPer], pw[D][kPer];
  auto fetch = [&](C (&pre)[kPer], C (&pw)[kPer], int a0, const int4 *te) __attribute__((always_inline)) {
#pragma unroll
    for (int m = 0; m < kPer; ++m) {
      const int j = wid + 4 * m;
      C v, w;
      v.x = v.y = w.x = w.y = (R)0;
      if (j < NL) {
        const int4 e = te[j];
        const int av = a0 - 1 + lane;
        if (av >= e.y && av < e.z) {
          v = v1[e.x + (av - e.y)];
          if (FUSED && fc.hasv) w = vp[e.x + (av - e.y)];
        }
      }
      pre[m] = v;
      if constexpr (FUSED) pw[m] = w;
    }
  };
  auto store = [&](const C (&pre)[kPer], const C (&pw)[kPer], int sl) __attribute__((always_inline)) {
#pragma unroll
    for (int m = 0; m < kPer; ++m) {
      const int j = wid + 4 * m;
      if (j < NL) s_ring[((size_t)sl * NL + j) * kSegW + lane] = FUSED ? fused_v<R>(fc, pre[m], pw[m]) : pre[m];
    }
  };
  // The rows of slab c that this thread computes (clear rows only: the
  // others are in the sector's slow-row list), in output round k.
  auto own_row = [&](int a0, int b0, int c, const int4 *te0, int k, int &r, int &j, int &ai)
                     __attribute__((always_inline)) -> bool {
    const int q = k * kThreads + (int)threadIdx.x;
    j = 1 + q / kSegA;
    ai = q % kSegA;  // window line, a - a0
    r = 0;
    if (q >= B * kSegA) return false;
    const int4 e = te0[j];
    const int av = a0 + ai;
    const int b = b0 - 1 + j;
    const int g1 = av - 1, g2 = b - av, g3 = c - b;
    r = e.x + (av - e.y);
    return av >= e.y && av < e.z && max(g1, max(g2, g3)) <= L + 1 - c - 3;
  };
  // FUSED: v' of slab c's own rows.  Issued right after the slab's window is
  // stored, in the set of that window, so at compute time they are older
  // than the window loads in flight (vector-memory loads retire in order)
  // and the wait for them does not wait for the prefetch.
  C vos[D][kRounds];
  auto load_vo = [&](C (&vo)[kRounds], int a0, int b0, int c, const int4 *te0) __attribute__((always_inline)) {
#pragma unroll
    for (int k = 0; k < kRounds; ++k) {
      int r, j, ai;
      const bool ok = own_row(a0, b0, c, te0, k, r, j, ai);
      vo[k].x = vo[k].y = (R)0;
      if (fc.hasv && ok) vo[k] = vp[r];
    }
  };
  // slab c from slots c - 1, c, c + 1 of the ring
  auto compute = [&](int a0, int b0, int c, const int4 *te0, const C (&vo)[kRounds]) __attribute__((always_inline)) {
    const int s0 = c % kRing, sm = (c + kRing - 1) % kRing, sp = (c + 1) % kRing;
    const C *r0 = s_ring + (size_t)s0 * NL * kSegW, *rm = s_ring + (size_t)sm * NL * kSegW,
            *rp = s_ring + (size_t)sp * NL * kSegW;
    int rr[kRounds], rj[kRounds], rai[kRounds];
    bool rok[kRounds];
#pragma unroll
    for (int k = 0; k < kRounds; ++k) rok[k] = own_row(a0, b0, c, te0, k, rr[k], rj[k], rai[k]);
#pragma unroll
    for (int k = 0; k < kRounds; ++k) {
      if (!rok[k]) continue;
      const int j = rj[k], ai = rai[k], r = rr[k];
      const int av = a0 + ai;
      const int b = b0 - 1 + j;
      const int g1 = av - 1, g2 = b - av, g3 = c - b;
      const int x0 = j * kSegW + ai + 1;  // window index of (a, b) in a slot
      // The 8 neighbours are summed without tests: a hop that does not exist
      // (g1, g2 or g3 = 1: two down spins on one site) lands on a window
      // point that holds no row -- a = 1, a = b or b = c lie outside every
      // line's range of a -- and the window holds 0 there.  Only the
      // diagonal counts the antiparallel bonds.
      C x = r0[x0];
      const C yc1 = rp[x0];              // (a, b, c + 1): spin c to the right (W > 1)
      const C yal = rp[x0 + kSegW + 1];  // anchor to site L: (a + 1, b + 1, c + 1)
      const C ya1 = r0[x0 + 1];          // (a + 1, b, c)         g2 > 1
      const C yb0 = r0[x0 - kSegW];      // (a, b - 1, c)         g2 > 1
      const C ya0 = r0[x0 - 1];          // (a - 1, b, c)         g1 > 1
      const C yar = rm[x0 - kSegW - 1];  // anchor to site 2: (a - 1, b - 1, c - 1), g1 > 1
      const C yb1 = r0[x0 + kSegW];      // (a, b + 1, c)         g3 > 1
      const C yc0 = rm[x0];              // (a, b, c - 1)         g3 > 1
      R sx = (yc1.x + ya1.x) + (yb0.x + ya0.x) + (yb1.x + yc0.x);
      R sy = (yc1.y + ya1.y) + (yb0.y + ya0.y) + (yb1.y + yc0.y);
      const R alx = yal.x, aly = yal.y, arx = yar.x, ary = yar.y;
      const int nanti = 2 + 2 * ((g1 > 1) + (g2 > 1) + (g3 > 1));
      if constexpr (REALPH) {
        sx += arx + alx;
        sy += ary + aly;
      } else {
        sx += ph_ar_r * arx - ph_ar_i * ary + ph_al_r * alx - ph_al_i * aly;
        sy += ph_ar_r * ary + ph_ar_i * arx + ph_al_r * aly + ph_al_i * alx;
      }
      x.x *= scale;
      x.y *= scale;
      const R dg = (sizeof(R) == 8 ? (R)a.ezz : (R)0) - (R)2 * jz4 * (R)nanti;
      C out;
      out.x = dg * x.x + jxy2 * scale * sx;
      out.y = dg * x.y + jxy2 * scale * sy;
      if constexpr (FUSED) {
        C pn;
        pn.x = out.x - fc.c3 * vo[k].x;
        pn.y = out.y - fc.c3 * vo[k].y;
        st_nt(&xo[r], x);
        st_nt(&v0[r], pn);
        acc4[0] += (double)x.x * (double)out.x + (double)x.y * (double)out.y;
        acc4[1] += (double)pn.x * (double)pn.x + (double)pn.y * (double)pn.y;
        acc4[2] += (double)x.x * (double)pn.x + (double)x.y * (double)pn.y;
        acc4[3] += (double)x.x * (double)x.x + (double)x.y * (double)x.y;
      } else {
        st_nt(&v0[r], out);
        if (a.dot_partials) {
          dre += (double)x.x * (double)out.x + (double)x.y * (double)out.y;
          dim_ += (double)x.x * (double)out.y - (double)x.y * (double)out.x;
        }
      }
    }
  };

  __syncthreads();
  for (int t = blockIdx.x; t < sa.ntasks; t += gridDim.x) {
    const int4 task = sa.tasks[t];
    const int a0 = task.x, b0 = task.y, c0 = task.z, c1 = task.w & 0xffff;
    if (c1 <= c0) continue;
    // line entries of slabs c0 - 1 .. c1 (entry (cs, j) at s_tl[(cs - c0 + 1) NL + j])
    const int nent = (c1 - c0 + 2) * NL;
    for (int q = threadIdx.x; q < nent; q += kThreads) {
      const int cs = c0 - 1 + q / NL, b = b0 - 1 + q % NL;
      int4 e = make_int4(0, 0, 0, 0);
      if (cs >= 0 && cs <= L && b >= 0 && b <= L) e = sa.lines[(size_t)cs * W1 + b];
      s_tl[q] = e;
    }
    __syncthreads();
    auto te = [&](int cs) { return s_tl + (cs - c0 + 1) * NL; };
    // prologue: slabs c0 - 1, c0, c0 + 1 loaded together into their slots,
    // then c0 + 2 in flight (X)
    {
      C preZ[kPer], pwZ[kPer];
      fetch(pre[0], pw[0], a0, te(c0 - 1));
      fetch(pre[1 % D], pw[1 % D], a0, te(c0));
      fetch(preZ, pwZ, a0, te(c0 + 1));
      store(pre[0], pw[0], (c0 + kRing - 1) % kRing);
      store(pre[1 % D], pw[1 % D], c0 % kRing);
      store(preZ, pwZ, (c0 + 1) % kRing);
    }
    if constexpr (FUSED) {
      load_vo(vos[0], a0, b0, c0, te(c0));
      if (c0 + 1 < c1) load_vo(vos[1 % D], a0, b0, c0 + 1, te(c0 + 1));
    }
    // slab s travels in set (s - c0) % D: slabs c0 + 2 .. c0 + D in flight
#pragma unroll
    for (int k = 2; k <= D; ++k)
      if (c0 + k <= c1) fetch(pre[k % D], pw[k % D], a0, te(c0 + k));
    __syncthreads();
    // slab c: slab c + D + 1 goes in flight in the set that slab c + 1 left,
    // slab c is computed, and slab c + 2 (D slabs ago in flight) is stored
    for (int c = c0; c < c1; c += D) {
#pragma unroll
      for (int u = 0; u < D; ++u) {
        const int cc = c + u;
        if (cc < c1) {
          if (cc + D + 1 <= c1) fetch(pre[(u + 1) % D], pw[(u + 1) % D], a0, te(cc + D + 1));
          compute(a0, b0, cc, te(cc), vos[u]);
          if (cc + 2 <= c1) {
            store(pre[(u + 2) % D], pw[(u + 2) % D], (cc + 2) % kRing);
            if (FUSED && cc + 2 < c1) load_vo(vos[(u + 2) % D], a0, b0, cc + 2, te(cc + 2));
          }
          __syncthreads();
        }
      }
    }
  }
  if constexpr (FUSED) {
    if (a.dot_partials) block_reduce_store4(acc4, a.dot_partials);
  } else {
    if (a.dot_partials) block_reduce_store(dre, dim_, a.dot_partials);
  }
}

// Rows of an N_down = 4 stencil sector that are not clear (the stencil kernel
// skips them): a count per chunk of kSlowChunk rows, then the rows in order.
constexpr int kSlowChunk = 4096;
constexpr int kSlowGridMax = 2048;  // blocks of the slow-row kernel (with the stencil's, within kMaxBlocks)
__device__ __forceinline__ bool stencil_slow(uint64_t w, int bits, int L) {
  const uint64_t m = (1ull << bits) - 1ull;
  const int av = (int)((w >> bits) & m), b = (int)((w >> (2 * bits)) & m), c = (int)((w >> (3 * bits)) & m);
  return max(av - 1, max(b - av, c - b)) > L + 1 - c - 3;
}
__global__ void k_slow_count(const uint64_t *__restrict__ cfg, int64_t dim, int bits, int L, int *counts) {
  __shared__ int s_n;
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  int n = 0;
  const int64_t r0 = (int64_t)blockIdx.x * kSlowChunk;
  for (int q = threadIdx.x; q < kSlowChunk; q += blockDim.x) {
    const int64_t r = r0 + q;
    if (r < dim && stencil_slow(cfg[r], bits, L)) ++n;
  }
  atomicAdd(&s_n, n);
  __syncthreads();
  if (threadIdx.x == 0) counts[blockIdx.x] = s_n;
}
// blockDim.x = 256; offsets: exclusive prefix of the counts
__global__ void k_slow_write(const uint64_t *__restrict__ cfg, int64_t dim, int bits, int L, const int *offsets,
                             int32_t *out) {
  __shared__ int s_wave[4];
  int base = offsets[blockIdx.x];
  const int64_t r0 = (int64_t)blockIdx.x * kSlowChunk;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int q0 = 0; q0 < kSlowChunk; q0 += 256) {
    const int64_t r = r0 + q0 + threadIdx.x;
    const bool sl = r < dim && stencil_slow(cfg[r], bits, L);
    const uint64_t mask = __ballot(sl);
    const int before = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
    if (lane == 0) s_wave[wid] = __popcll(mask);
    __syncthreads();
    int wbase = base;
    for (int w = 0; w < wid; ++w) wbase += s_wave[w];
    if (sl) out[wbase + before] = (int32_t)r;
    base += s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    __syncthreads();
  }
}
// H.v of the listed rows through the key search of slow_row (one row per thread).
template <class R, bool REALPH, class KeyT, bool FUSED>
__global__ __launch_bounds__(kThreads) void k_hv_slow_rows(HvArgs a, const int32_t *__restrict__ rows, int nrows,
                                                           const int32_t *__restrict__ binom32) {
  constexpr int N = 4;
  __shared__ R s_tab[2 * (N + 1) + 2];
  for (int q = threadIdx.x; q <= N; q += kThreads) {
    s_tab[q] = (R)a.rclass[q];
    s_tab[N + 1 + q] = (R)(a.rclass[1] / a.rclass[q]);
  }
  if (threadIdx.x == 0) {
    s_tab[2 * (N + 1)] = (R)a.bt[0];
    s_tab[2 * (N + 1) + 1] = (R)a.bt[1];
  }
  __syncthreads();
  const NnCtx<N, R> k{a.L, a.kbits, binom32, s_tab, s_tab + (N + 1), s_tab + 2 * (N + 1), s_tab + 2 * (N + 1) + 1,
                      nullptr};
  if constexpr (FUSED) {
    const FusedCoef<R> fc = fused_coef<R>(a);
    double acc4[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i - (int)threadIdx.x < nrows; i += gridDim.x * kThreads) {
      if (i < nrows) slow_row_fused<N, R, REALPH, KeyT>(a, k, fc, rows[i], acc4);
    }
    if (a.dot_partials) block_reduce_store4(acc4, a.dot_partials);
  } else {
    const R scale = a.scale ? (R)(*a.scale) : (R)1;
    double dre = 0.0, dim_ = 0.0;
    for (int i = blockIdx.x * kThreads + threadIdx.x; i - (int)threadIdx.x < nrows; i += gridDim.x * kThreads) {
      if (i < nrows) slow_row<N, R, REALPH, true, KeyT>(a, k, rows[i], scale, dre, dim_);
    }
    if (a.dot_partials) block_reduce_store(dre, dim_, a.dot_partials);
  }
}

// Line table of an anchored N_down = 4 chain sector: for every row, the first
// row of its line (b, c) writes (row, a) and the last row writes a + 1.
__global__ void k_stencil_lines(const uint64_t *__restrict__ cfg, int64_t dim, int bits, int L, int4 *lines) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim) return;
  const uint64_t m = (1ull << bits) - 1ull;
  auto key = [&](uint64_t w) { return (int64_t)((w >> (3 * bits)) & m) * (L + 1) + (int64_t)((w >> (2 * bits)) & m); };
  const uint64_t w = cfg[i];
  const int64_t ln = key(w);
  const int av = (int)((w >> bits) & m);
  if (i == 0 || key(cfg[i - 1]) != ln) {
    lines[ln].x = (int)i;
    lines[ln].y = av;
  }
  if (i == dim - 1 || key(cfg[i + 1]) != ln) lines[ln].z = av + 1;
}

template <int N, bool UNI, class KeyT>
KFn pick_nn3(int prec, bool realph, bool pairs) {
  if (prec == QS3X_FP64)
    return realph ? k_hv_chain_nn<N, double, true, UNI, KeyT> : k_hv_chain_nn<N, double, false, UNI, KeyT>;
  // the two-row kernel pays with real phases up to N_down = 4; with complex
  // phases or N_down = 5 its registers cost a wavefront per SIMD (FP32 at
  // L = 1000, N = 4, k = 1: 0.551 against 0.513 ms; L = 300, N = 5: 0.917
  // against 0.874 ms)
  if constexpr (N <= 4) {
    if (pairs && realph) return k_hv_chain_nn2<N, true, UNI, KeyT>;
  }
  return realph ? k_hv_chain_nn<N, float, true, UNI, KeyT> : k_hv_chain_nn<N, float, false, UNI, KeyT>;
}

template <int N>
KFn pick_nn(int prec, bool realph, bool uniform, bool key32, bool pairs) {
  if (uniform)
    return key32 ? pick_nn3<N, true, uint32_t>(prec, realph, pairs)
                 : pick_nn3<N, true, uint64_t>(prec, realph, pairs);
  return key32 ? pick_nn3<N, false, uint32_t>(prec, realph, pairs) : pick_nn3<N, false, uint64_t>(prec, realph, pairs);
}

}  // namespace

KFn chain_nn_kernel(int N, int prec, bool realph, bool uniform, bool key32, bool pairs) {
  switch (N) {
    case 2: return pick_nn<2>(prec, realph, uniform, key32, pairs);
    case 3: return pick_nn<3>(prec, realph, uniform, key32, pairs);
    case 4: return pick_nn<4>(prec, realph, uniform, key32, pairs);
    case 5: return pick_nn<5>(prec, realph, uniform, key32, pairs);
    case 6: return pick_nn<6>(prec, realph, uniform, key32, pairs);
    case 7: return pick_nn<7>(prec, realph, uniform, key32, pairs);
    case 8: return pick_nn<8>(prec, realph, uniform, key32, pairs);
    default: return nullptr;
  }
}

size_t chain_nn_lds_bytes(int N, int L, int nbt, int prec) {
  const size_t rs = (prec == QS3X_FP64) ? 8 : 4;
  const size_t nb = (size_t)(N - 1) * (L - 1);
  return ((nb * 4 + 15) & ~(size_t)15) + 2 * rs * (N + 1) + 2 * rs * nbt + 2 * (size_t)(L + 1);
}

// ---------------------------------------------------------------- stencil
namespace {

using SlowFn = void (*)(HvArgs, const int32_t *, int, const int32_t *);

template <int BT, int D>
SFn stencil_kernel_b(int prec, bool realph, bool fused) {
  if (fused) {
    if (prec == QS3X_FP64)
      return realph ? k_hv_stencil4<double, true, true, BT, D> : k_hv_stencil4<double, false, true, BT, D>;
    return realph ? k_hv_stencil4<float, true, true, BT, D> : k_hv_stencil4<float, false, true, BT, D>;
  }
  if (prec == QS3X_FP64)
    return realph ? k_hv_stencil4<double, true, false, BT, D> : k_hv_stencil4<double, false, false, BT, D>;
  return realph ? k_hv_stencil4<float, true, false, BT, D> : k_hv_stencil4<float, false, false, BT, D>;
}
template <int D>
SFn stencil_kernel_d(int prec, bool realph, bool fused, int B) {
  switch (B) {
    case 2: return stencil_kernel_b<2, D>(prec, realph, fused);
    case 4: return stencil_kernel_b<4, D>(prec, realph, fused);
    case 6: return stencil_kernel_b<6, D>(prec, realph, fused);
    default: return stencil_kernel_b<8, D>(prec, realph, fused);
  }
}
// Slab-window form: lines per tile 2, 4, 6 or 8, windows in flight 2 or 3.
// Column form (hv_col.hip): P window positions per lane (1, 2 or 4), lines
// per tile 2, 4, 6, 8 (and 12, 14 with P = 1), two windows in flight.
SFn stencil_kernel(int prec, bool realph, bool fused, int B, int depth, bool col, int P) {
  if (col) return stencil_col_kernel(prec, realph, fused, B, P);
  return depth >= 3 ? stencil_kernel_d<3>(prec, realph, fused, B) : stencil_kernel_d<2>(prec, realph, fused, B);
}
template <bool FU>
SlowFn slow_kernel_t(int prec, bool realph, bool key32) {
  if (prec == QS3X_FP64) {
    if (realph)
      return key32 ? k_hv_slow_rows<double, true, uint32_t, FU> : k_hv_slow_rows<double, true, uint64_t, FU>;
    return key32 ? k_hv_slow_rows<double, false, uint32_t, FU> : k_hv_slow_rows<double, false, uint64_t, FU>;
  }
  if (realph) return key32 ? k_hv_slow_rows<float, true, uint32_t, FU> : k_hv_slow_rows<float, true, uint64_t, FU>;
  return key32 ? k_hv_slow_rows<float, false, uint32_t, FU> : k_hv_slow_rows<float, false, uint64_t, FU>;
}
SlowFn slow_kernel(int prec, bool realph, bool key32, bool fused = false) {
  return fused ? slow_kernel_t<true>(prec, realph, key32) : slow_kernel_t<false>(prec, realph, key32);
}

// Band width and task length: options hv_stencil_lines / hv_stencil_slabs,
// or by default the fastest measured at L = 1000, N_down = 4 (n4-major task
// order): 10 lines in both precisions (the slab-window form takes 8), 16
// slabs per task in FP32 (pair form: H.v 0.197 ms, fused Lanczos step 0.404
// ms) and 8 in FP64 (column form: H.v 0.348 ms, fused step 0.711 ms).
int stencil_lines(int prec) {
  if (options().hv_stencil_lines > 0) return options().hv_stencil_lines;
  return 10;
}
int stencil_depth(int prec, bool fused) {
  if (options().hv_stencil_depth > 0) return options().hv_stencil_depth;
  return 2;
}
// Window positions per lane of the column form: option hv_stencil_span, or 1.
int stencil_span(int prec) {
  if (options().hv_stencil_span > 0) return options().hv_stencil_span;
  return 1;
}
int stencil_slabs(int prec) {
  if (options().hv_stencil_slabs > 0) return options().hv_stencil_slabs;
  return prec == QS3X_FP64 ? 8 : 16;
}

}  // namespace

// Tasks of a gap-lattice stencil sweep: tiles of segA values of a by B lines,
// each swept over the slabs where it holds rows, in tasks of up to `slabs`
// slabs, each task repeated for `groups` consecutive vector groups (group in
// bits 16.. of .w, the end slab in bits 0..15).  Tasks are dealt to the 8
// XCDs by tile: task 8 m + x belongs to XCD x, whose tiles are one contiguous
// range in (b, a) order holding about an eighth of the tasks, so the halo
// lines and columns that neighbouring tiles share stay in one L2, and the
// groups of one tile run at about the same time there.
std::vector<int4> stencil_tasks(const std::vector<int4> &lines, int L, int segA, int B, int slabs, int groups) {
  const int W1 = L + 1;
  const int na = (L + segA - 1) / segA + 1, nb = (L + B - 1) / B + 1;
  std::vector<int> cmin((size_t)na * nb, 0x7fffffff), cmax((size_t)na * nb, -1);
  for (int c = 0; c <= L; ++c)
    for (int b = 3; b < c; ++b) {
      const int4 e = lines[(size_t)c * W1 + b];
      if (e.z <= e.y) continue;
      const int tb = (b - 3) / B;
      for (int ta = (e.y - 2) / segA; ta <= (e.z - 1 - 2) / segA; ++ta) {
        const size_t t = (size_t)tb * na + ta;
        cmin[t] = std::min(cmin[t], c);
        cmax[t] = std::max(cmax[t], c);
      }
    }
  std::vector<std::vector<int4>> per_tile;
  size_t total = 0;
  for (int tb = 0; tb < nb; ++tb)
    for (int ta = 0; ta < na; ++ta) {
      const size_t t = (size_t)tb * na + ta;
      if (cmax[t] < cmin[t]) continue;
      std::vector<int4> tl;
      for (int c0 = cmin[t]; c0 <= cmax[t]; c0 += slabs)
        for (int g = 0; g < groups; ++g)
          tl.push_back(make_int4(2 + ta * segA, 3 + tb * B, c0, std::min(c0 + slabs, cmax[t] + 1) | (g << 16)));
      total += tl.size();
      per_tile.push_back(tl);
    }
  const int nx = 8;
  std::vector<std::vector<int4>> grp(nx);
  {
    // contiguous runs of tiles holding about total / nx tasks each
    size_t acc = 0;
    int x = 0;
    std::vector<std::vector<const std::vector<int4> *>> gt(nx);
    for (auto &tl : per_tile) {
      if (x < nx - 1 && acc >= total * (size_t)(x + 1) / nx) ++x;
      gt[x].push_back(&tl);
      acc += tl.size();
    }
    // depth d of every tile in turn, keeping the groups of one task together
    for (int g = 0; g < nx; ++g) {
      size_t depth = 0;
      for (auto *tl : gt[g]) depth = std::max(depth, tl->size());
      for (size_t d = 0; d < depth; d += groups)
        for (auto *tl : gt[g])
          for (int k = 0; k < groups && d + k < tl->size(); ++k) grp[g].push_back((*tl)[d + k]);
    }
  }
  size_t mx = 0;
  for (auto &g : grp) mx = std::max(mx, g.size());
  std::vector<int4> tasks(mx * nx, make_int4(0, 0, 0, 0));
  for (int x = 0; x < nx; ++x)
    for (size_t m = 0; m < grp[x].size(); ++m) tasks[m * nx + x] = grp[x][m];
  return tasks;
}

// Tasks of a single-vector stencil sweep in n4-major order: the slabs are cut
// into chunks [k S, (k + 1) S) at the same places for every tile, and the
// chunks are taken in turn, all tiles of chunk k before any of chunk k + 1.
// The blocks of the persistent grid, which take tasks in list order, then
// work on a few consecutive chunks at any time -- a front a few tens of MB
// wide moving through the vectors, where tile-major order has them spread
// over the whole of both vectors.  Within a chunk the tiles are dealt to the
// 8 XCDs as contiguous runs in (n3, n2) order (task 8 m + x on XCD x), so
// tiles that share halo lines share an L2.
std::vector<int4> stencil_tasks_cmajor(const std::vector<int4> &lines, int L, int segA, int B, int slabs) {
  const int W1 = L + 1;
  const int na = (L + segA - 1) / segA + 1, nb = (L + B - 1) / B + 1;
  std::vector<int> cmin((size_t)na * nb, 0x7fffffff), cmax((size_t)na * nb, -1);
  for (int c = 0; c <= L; ++c)
    for (int b = 3; b < c; ++b) {
      const int4 e = lines[(size_t)c * W1 + b];
      if (e.z <= e.y) continue;
      const int tb = (b - 3) / B;
      for (int ta = (e.y - 2) / segA; ta <= (e.z - 1 - 2) / segA; ++ta) {
        const size_t t = (size_t)tb * na + ta;
        cmin[t] = std::min(cmin[t], c);
        cmax[t] = std::max(cmax[t], c);
      }
    }
  const int nx = 8;
  std::vector<int4> tasks;
  for (int k = 0; k * slabs <= L; ++k) {
    const int lo = k * slabs, hi = lo + slabs;
    std::vector<int4> ch;  // this chunk's tasks, tiles in (b, a) order
    for (int tb = 0; tb < nb; ++tb)
      for (int ta = 0; ta < na; ++ta) {
        const size_t t = (size_t)tb * na + ta;
        if (cmax[t] < cmin[t] || cmax[t] < lo || cmin[t] >= hi) continue;
        const int c0 = std::max(lo, cmin[t]), c1 = std::min(hi, cmax[t] + 1);
        ch.push_back(make_int4(2 + ta * segA, 3 + tb * B, c0, c1));
      }
    if (ch.empty()) continue;
    const size_t per = (ch.size() + nx - 1) / nx;  // tasks per XCD in this chunk
    for (size_t m = 0; m < per; ++m)
      for (int x = 0; x < nx; ++x) {
        const size_t i = (size_t)x * per + m;
        tasks.push_back(i < ch.size() ? ch[i] : make_int4(0, 0, 0, 0));
      }
  }
  return tasks;
}

namespace {

// Plan for one precision of the single-vector stencil (groups = 1).
// form: 0 slab window, 1 column, 2 pair; P: window positions per lane of the
// column form (tiles 64 P - 2 wide; the slab-window form has P = 1); the pair
// form's tiles are 124 wide in windows of 128.
void stencil_plan(int prec, qs3x_sector::StencilPlan &pl, const std::vector<int4> &lines, int L, int B, int slabs,
                  int P, int form) {
  const int segA = form == 2 ? 124 : kSegW * P - 2;
  const int segW = form == 2 ? 128 : kSegW * P;
  std::vector<int4> tasks = options().hv_stencil_order == 1 ? stencil_tasks_cmajor(lines, L, segA, B, slabs)
                                                             : stencil_tasks(lines, L, segA, B, slabs, 1);
  const size_t es = prec == QS3X_FP64 ? 16 : 8;
  pl.nlines = 0;
  for (const int4 &e : lines) pl.nlines += e.z > e.y ? 1 : 0;
  pl.B = B;
  pl.slabs = slabs;
  pl.P = P;
  pl.form = form;
  pl.order = options().hv_stencil_order;
  pl.lds = (size_t)kRing * (B + 2) * segW * es + (size_t)(slabs + 2) * (B + 2) * sizeof(int4);
  pl.ntasks = (int)tasks.size();
  pl.tasks.alloc(std::max<size_t>(tasks.size(), 1));
  if (!tasks.empty()) pl.tasks.upload(tasks.data(), tasks.size());
  pl.grid = pl.grid_fused = 0;  // at the first launch
}

// The rows the stencil kernel skips, in row order (deterministic).
void stencil_slow_list(qs3x_sector *s) {
  const int L = s->model->NOS;
  const int nch = (int)((s->dim + kSlowChunk - 1) / kSlowChunk);
  DevBuf<int> counts(std::max(nch, 1)), offs(std::max(nch, 1));
  hipStream_t st = stream();
  hipLaunchKernelGGL(k_slow_count, dim3(nch), dim3(256), 0, st, s->d_cfg.p, s->dim, s->bits, L, counts.p);
  QS3X_HIP(hipGetLastError());
  std::vector<int> hc(nch), ho(nch);
  counts.download(hc.data(), nch);
  int64_t tot = 0;
  for (int i = 0; i < nch; ++i) {
    ho[i] = (int)tot;
    tot += hc[i];
  }
  s->nslow = tot;
  s->d_slow.alloc(std::max<int64_t>(tot, 1));
  offs.upload(ho.data(), nch);
  hipLaunchKernelGGL(k_slow_write, dim3(nch), dim3(256), 0, st, s->d_cfg.p, s->dim, s->bits, L, offs.p, s->d_slow.p);
  QS3X_HIP(hipGetLastError());
  QS3X_HIP(hipStreamSynchronize(st));
}

}  // namespace

bool stencil_applies(const qs3x_sector *s) {
  return s->nod == 4 && s->chain && s->mode == kIndexAnchored && s->model->nn_chain &&
         s->model->bt_jxy2.size() == 1 && s->d_rclass.p != nullptr && s->dim < ((int64_t)1 << 31) &&
         binom_host(s->model->NOS - 2, 3) < ((int64_t)1 << 31) && s->model->NOS < 32768;
}

void stencil_tables(qs3x_sector *s, std::vector<int4> *lines_out) {
  const int L = s->model->NOS;
  if (!s->d_lines.p) {
    const size_t n = (size_t)(L + 1) * (L + 1);
    s->d_lines.alloc(n);
    QS3X_HIP(hipMemsetAsync(s->d_lines.p, 0, n * sizeof(int4), stream()));
    hipLaunchKernelGGL(k_stencil_lines, dim3((unsigned)((s->dim + 255) / 256)), dim3(256), 0, stream(),
                       s->d_cfg.p, s->dim, s->bits, L, s->d_lines.p);
    QS3X_HIP(hipGetLastError());
    std::vector<int32_t> b32((size_t)3 * (L - 1));
    for (int q = 1; q <= 3; ++q)
      for (int x = 0; x < L - 1; ++x) b32[(size_t)(q - 1) * (L - 1) + x] = (int32_t)binom_host(x, q);
    s->d_binom32.alloc(b32.size());
    s->d_binom32.upload(b32.data(), b32.size());
    stencil_slow_list(s);
  }
  if (lines_out) {
    lines_out->assign((size_t)(L + 1) * (L + 1), make_int4(0, 0, 0, 0));
    s->d_lines.download(lines_out->data(), lines_out->size());
  }
}

static bool launch_stencil_impl(qs3x_sector *s, int prec, const HvArgs &a, hipStream_t st, int *grid_out,
                                size_t *lds_out, bool fused) {
  if (!options().hv_stencil || !stencil_applies(s)) return false;
  const int L = s->model->NOS;
  qs3x_sector::StencilPlan &pl = s->stencil[prec == QS3X_FP64 ? 0 : 1];
  // tile height, task slabs and windows in flight: the sector's autotuned
  // values, else the options, else the per-precision defaults
  const qs3x_sector::HvTune &tn = s->tune[prec == QS3X_FP64 ? 0 : 1];
  // form: 0 slab window, 1 column, 2 pair (FP32 sectors under 2^28 rows; else
  // the column form)
  int form = (tn.set && tn.st_col >= 0) ? tn.st_col : options().hv_stencil_kernel;
  if (form == 2 && (prec == QS3X_FP64 || s->dim >= ((int64_t)1 << 28))) form = 1;
  const bool col = form >= 1;
  int want_B = (tn.set && tn.st_lines > 0) ? tn.st_lines : stencil_lines(prec);
  if (form == 0) want_B = std::min(want_B, kMaxLines);
  if (form == 2) want_B = std::min(want_B, 10);
  int want_P = form == 1 ? ((tn.set && tn.st_span > 0) ? tn.st_span : stencil_span(prec)) : 1;
  if (want_P > 1 && (want_B > 8 || prec == QS3X_FP64)) want_P = 1;  // wide tiles: FP32, up to 8 lines
  const int want_slabs = (tn.set && tn.st_slabs > 0) ? tn.st_slabs : stencil_slabs(prec);
  const int depth = (tn.set && tn.st_depth > 0) ? tn.st_depth : stencil_depth(prec, fused);
  if (pl.tried && (pl.slabs != want_slabs || pl.B != want_B || pl.P != want_P || pl.form != form ||
                   pl.order != options().hv_stencil_order))
    pl.tried = false;  // another tile, kernel form or task order
  if (!pl.tried) {
    pl.tried = true;
    pl.ok = false;
    std::vector<int4> lines;
    stencil_tables(s, &lines);
    stencil_plan(prec, pl, lines, L, want_B, want_slabs, want_P, form);
    // the plan's tables went up on the library stream; the launch below may
    // be on the caller's stream, which nothing orders after it
    QS3X_HIP(hipStreamSynchronize(stream()));
    pl.ok = true;
  }
  if (!pl.ok) return false;
  QS3X_REQUIRE(pl.lds <= 156 * 1024, "stencil H.v: the tile's LDS ring (" + std::to_string(pl.lds) +
                                         " bytes) exceeds 156 KB; use fewer lines per tile or a smaller span");
  int kb = 1;
  while ((1 << kb) <= L) ++kb;
  SFn f = pl.form == 2 ? stencil_pair_kernel(s->real_phase, fused, pl.B, depth >= 4 ? 4 : 2)
                       : stencil_kernel(prec, s->real_phase, fused, pl.B, depth, col, pl.P);
  if (!f && pl.form == 2) f = stencil_pair_kernel(s->real_phase, fused, pl.B, 2);
  QS3X_REQUIRE(f != nullptr, "stencil H.v: no kernel for this tile height and span");
  SlowFn fs = slow_kernel(prec, s->real_phase, 3 * kb <= 32, fused);
  // partial sums per block: 2 (complex <v, Hv>) or 4 in the fused step
  const size_t np = fused ? 4 : 2;
  if (pl.lds > 64 * 1024) {
    static std::mutex mu;
    static std::map<SFn, size_t> done;
    std::lock_guard<std::mutex> lk(mu);
    auto it = done.find(f);
    if (it == done.end() || it->second < pl.lds) {
      QS3X_HIP(hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl.lds));
      done[f] = pl.lds;
    }
  }
  // one row per thread (a latency-bound key search: more rows per thread
  // serialise their dependent loads), up to kSlowGridMax blocks
  const int slow_grid =
      s->nslow > 0 ? (int)std::min<int64_t>((s->nslow + kThreads - 1) / kThreads, kSlowGridMax) : 0;
  int &pgrid = fused ? pl.grid_fused : pl.grid;
  const void *&pfor = fused ? pl.grid_fused_for : pl.grid_for;
  if (pgrid == 0 || pfor != (const void *)f) {
    pfor = (const void *)f;
    // resident blocks per CU, from the kernel's registers and LDS (160 KB per
    // CU, 512 VGPRs per lane slot on each of 4 SIMDs, at most 8 wavefronts
    // per SIMD), a multiple of the 8 XCDs so that block g (on XCD g % 8) takes
    // only tasks 8 m + g % 8, its XCD's tiles; with the slow-row grid within
    // the callers' partial-sum buffers
    int dev = 0, cus = 0;
    QS3X_HIP(hipGetDevice(&dev));
    QS3X_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    hipFuncAttributes at{};
    QS3X_HIP(hipFuncGetAttributes(&at, (const void *)f));
    const int vg = std::max(at.numRegs, 1);
    const int waves_simd = std::min(8, 512 / ((vg + 7) & ~7));
    const int by_regs = std::max(1, waves_simd * 4 / (kThreads / 64));
    const int by_lds = std::max(1, (int)((160 * 1024) / (pl.lds + at.sharedSizeBytes)));
    const int per_cu = std::min(by_regs, by_lds);
    int grid = std::min<int64_t>(pl.ntasks, (int64_t)cus * per_cu);
    grid = std::min(grid, kMaxBlocks - kSlowGridMax);
    pgrid = std::max(8, (grid / 8) * 8);
  }
  StencilArgs sa{s->d_lines.p, pl.tasks.p, pl.ntasks, pl.B};
  // The slow rows (a latency-bound tail of about 1 % of the rows) run after
  // the stencil on st, or with option hv_stencil_side on a side stream forked
  // from and joined back into st; the two write disjoint rows and partial sums.
  // (the pair form writes every row of its tiles and leaves the rows that are
  // not clear to the slow-row kernel to overwrite: never on a side stream)
  if (slow_grid > 0 && (!options().hv_stencil_side || pl.form == 2)) {
    HvArgs b = a;
    if (b.dot_partials) b.dot_partials += np * (size_t)pgrid;
    hipLaunchKernelGGL(f, dim3(pgrid), dim3(kThreads), pl.lds, st, a, sa);
    QS3X_HIP(hipGetLastError());
    hipLaunchKernelGGL(fs, dim3(slow_grid), dim3(kThreads), 0, st, b, s->d_slow.p, (int)s->nslow, s->d_binom32.p);
    QS3X_HIP(hipGetLastError());
  } else if (slow_grid > 0) {
    static hipStream_t side[64] = {};
    static hipEvent_t fork[64] = {}, join[64] = {};
    const int dev = device();
    QS3X_REQUIRE(dev >= 0 && dev < 64, "launch_hv_stencil: device ordinal out of range");
    if (!side[dev]) {
      QS3X_HIP(hipStreamCreateWithFlags(&side[dev], hipStreamNonBlocking));
      QS3X_HIP(hipEventCreateWithFlags(&fork[dev], hipEventDisableTiming));
      QS3X_HIP(hipEventCreateWithFlags(&join[dev], hipEventDisableTiming));
    }
    QS3X_HIP(hipEventRecord(fork[dev], st));
    QS3X_HIP(hipStreamWaitEvent(side[dev], fork[dev], 0));
    HvArgs b = a;
    if (b.dot_partials) b.dot_partials += np * (size_t)pgrid;
    hipLaunchKernelGGL(fs, dim3(slow_grid), dim3(kThreads), 0, side[dev], b, s->d_slow.p, (int)s->nslow,
                       s->d_binom32.p);
    QS3X_HIP(hipGetLastError());
    QS3X_HIP(hipEventRecord(join[dev], side[dev]));
    hipLaunchKernelGGL(f, dim3(pgrid), dim3(kThreads), pl.lds, st, a, sa);
    QS3X_HIP(hipGetLastError());
    QS3X_HIP(hipStreamWaitEvent(st, join[dev], 0));
  } else {
    hipLaunchKernelGGL(f, dim3(pgrid), dim3(kThreads), pl.lds, st, a, sa);
    QS3X_HIP(hipGetLastError());
  }
  if (grid_out) *grid_out = pgrid + slow_grid;
  if (lds_out) *lds_out = pl.lds;
  return true;
}

bool launch_hv_stencil(qs3x_sector *s, int prec, const HvArgs &a, hipStream_t st, int *grid_out, size_t *lds_out) {
  return launch_stencil_impl(s, prec, a, st, grid_out, lds_out, false);
}

bool launch_hv_stencil_fused(qs3x_sector *s, int prec, const HvArgs &a, hipStream_t st, int *grid_out) {
  size_t lds = 0;
  return launch_stencil_impl(s, prec, a, st, grid_out, &lds, true);
}

}  // namespace qs3x
