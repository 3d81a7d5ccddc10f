// rep.h -- device-side representative search and row lookup with a runtime
// N_down, shared by the observable kernels (observables.hip).  The H.v hot
// path has compile-time specialisations of the same logic in hv.hip.
//
// representative_SQ (state_lists.f90:384-416) + inv_list (:131-139) +
// findstate (:364-382): a sorted configuration -> (row of its representative
// or -1, flat explist index of ell, norm r of the target row).
#pragma once
#include "qs3x_internal.h"

namespace qs3x {

struct SecView {
  const uint64_t *cfg;
  const double *list_r;
  const int64_t *list_s;
  const int32_t *index;
  const int64_t *binom;  // [n][k], row length kdim
  const double *rclass;  // chain: norm by stabiliser order (null: read list_r)
  const double *explist; // complex
  int64_t dim;
  int N, bits, kbits, kdim, mode, chain;
  int L, LX, LY, LZ;
};

inline SecView make_view(const qs3x_sector *s) {
  SecView v{};
  v.cfg = s->d_cfg.p;
  v.list_r = s->d_list_r.p;
  v.list_s = s->d_list_s.p;
  v.index = s->d_index.p;
  v.binom = s->d_binom.p;
  v.rclass = s->d_rclass.p;
  v.explist = s->d_explist.p;
  v.dim = s->dim;
  v.N = s->nod;
  v.bits = s->bits;
  int kb = 1;
  while ((1 << kb) <= s->model->NOS) ++kb;
  v.kbits = kb;
  v.kdim = s->nbinom_k;
  v.mode = s->mode;
  v.chain = (s->chain && s->mode == kIndexAnchored) ? 1 : 0;
  v.L = s->model->NOS;
  v.LX = s->model->LX;
  v.LY = s->model->LY;
  v.LZ = s->model->LZ;
  return v;
}

__device__ __forceinline__ int64_t sv_binom(const SecView &v, int n, int k) {
  return (n < 0 || k < 0) ? 0 : v.binom[(int64_t)n * v.kdim + k];
}

__device__ __forceinline__ void sv_unpack(const SecView &v, int64_t row, int *n) {
  const uint64_t w = v.cfg[row];
  for (int q = 0; q < v.N; ++q) n[q] = (int)((w >> (q * v.bits)) & ((1ull << v.bits) - 1ull));
}

__device__ __forceinline__ double sv_rowr(const SecView &v, int64_t row) { return v.list_r[row]; }

__device__ __forceinline__ int sv_shift_site(int s, int i, int j, int k, const SecView &v) {
  int s0 = s - 1;
  int x = s0 % v.LX, y = (s0 / v.LX) % v.LY, z = s0 / (v.LX * v.LY);
  x = (x + i) % v.LX;
  y = (y + j) % v.LY;
  z = (z + k) % v.LZ;
  return 1 + x + v.LX * (y + v.LY * z);
}

// nn: sorted configuration (N = v.N sites).  Returns the row of its
// representative (or -1); *ei = flat explist index of ell; *isrep = 1 when
// nn itself is the representative; *rt = norm of the target row.
__device__ int64_t sv_lookup(const SecView &v, const int *nn, int *ei, int *isrep, double *rt) {
  const int N = v.N;
  int64_t id = -1;
  if (v.chain) {
    const int L = v.L;
    int g[kMaxNod];
    for (int q = 0; q + 1 < N; ++q) g[q] = nn[q + 1] - nn[q];
    g[N - 1] = nn[0] + L - nn[N - 1];
    uint64_t kmax = 0, k0 = 0;
    uint64_t K[kMaxNod];
    for (int j = 0; j < N; ++j) {
      uint64_t k = 0;
      for (int t = 1; t < N; ++t) k = (k << v.kbits) | (uint64_t)g[(j - t + N) % N];
      K[j] = k;
      if (j == 0 || k > kmax) kmax = k;
    }
    k0 = K[0];
    int count = 0, mx = 0;
    for (int j = 0; j < N; ++j)
      if (K[j] == kmax) {
        ++count;
        mx = nn[j] > mx ? nn[j] : mx;
      }
    const bool rep = (nn[0] == 1) && (k0 == kmax);
    *isrep = rep ? 1 : 0;
    const int ell = rep ? 0 : (L + 1 - mx);
    *ei = rep ? 0 : (ell + 3 * (L + 1));
    int64_t c = 0;
    if (N > 1) {
      const uint64_t fm = (1ull << v.kbits) - 1ull;
      int f[kMaxNod];
      int sum = 0;
      for (int t = 1; t < N; ++t) {
        f[t - 1] = (int)((kmax >> (v.kbits * (N - 1 - t))) & fm);
        sum += f[t - 1];
      }
      int P = L - sum;
      c = sv_binom(v, P - 1, 1);
      for (int mm = 2; mm < N; ++mm) {
        P += f[N - mm];
        c += sv_binom(v, P - 1, mm);
      }
    }
    id = v.index[c];
    if (id >= 0) *rt = v.rclass ? v.rclass[count] : v.list_r[id];
    return id;
  }
  // general lattice: scan the translation group in loop order
  int best[kMaxNod], nd[kMaxNod];
  for (int q = 0; q < N; ++q) best[q] = nn[q];
  int e = 0;
  bool improved = false;
  for (int k = 1; k <= v.LZ; ++k)
    for (int j = 1; j <= v.LY; ++j)
      for (int i = 1; i <= v.LX; ++i) {
        for (int q = 0; q < N; ++q) nd[q] = sv_shift_site(nn[q], i, j, k, v);
        for (int q = 1; q < N; ++q) {
          int t = nd[q], z = q - 1;
          while (z >= 0 && nd[z] > t) {
            nd[z + 1] = nd[z];
            --z;
          }
          nd[z + 1] = t;
        }
        for (int l = N - 1; l >= 0; --l) {
          if (nd[l] > best[l]) break;
          if (nd[l] < best[l]) {
            e = i + (v.LX + 1) * (j + (v.LY + 1) * k);
            improved = true;
            for (int q = 0; q <= l; ++q) best[q] = nd[q];
            break;
          }
        }
      }
  *isrep = improved ? 0 : 1;
  *ei = e;
  int64_t rank = best[0];
  for (int i2 = 2; i2 <= N; ++i2) rank += sv_binom(v, best[i2 - 1] - 1, i2);
  if (v.mode == kIndexFullRank) {
    id = v.index[rank - 1];
  } else {
    int64_t lo = 0, hi = v.dim - 1;
    while (lo <= hi) {
      int64_t mid = lo + (hi - lo) / 2;
      int64_t x = v.list_s[mid];
      if (x < rank) lo = mid + 1;
      else if (x > rank) hi = mid - 1;
      else {
        id = mid;
        break;
      }
    }
  }
  if (id >= 0) *rt = v.list_r[id];
  return id;
}

// j_flip_ni semantics for an antiparallel pair: move the down spin from
// `from` to `to` in the sorted configuration n (N sites) -> nn sorted.
__device__ __forceinline__ void sv_move(const int *n, int N, int from, int to, int *nn) {
  int p = 0;
  for (int q = 0; q < N; ++q)
    if (n[q] != from) nn[p++] = n[q];
  int pos = N - 1;
  while (pos > 0 && nn[pos - 1] > to) {
    nn[pos] = nn[pos - 1];
    --pos;
  }
  nn[pos] = to;
}

}  // namespace qs3x
