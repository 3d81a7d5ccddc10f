// sector.hip -- on-device construction of (N_down, k) sectors.
//
// Replaces allocate_lists_omp_SQ / checkstate_SQ (state_lists.f90:66-112,
// 225-292).  The reference tests all C(NOS, N) ranks with an O(NOS) translation
// scan each and sorts the survivors.  Here:
//
//  * chain lattices (LY = LZ = 1): every colex-minimal configuration contains
//    site 1 (a rotation bringing the smallest down spin to site 1 lowers every
//    element), so only the C(L-1, N-1) configurations with site 1 occupied are
//    candidates.  They are enumerated by their colex rank over sites 2..L,
//    which is ascending full rank, so the output needs no sort.  A candidate is
//    its orbit's colex minimum iff its backward gap sequence (the gap ending at
//    site 1 first) is lexicographically maximal among the N rotations that
//    bring a down spin to site 1; the number of rotations attaining the maximum
//    is the stabiliser order, which fixes the norm through a host table that
//    evaluates the reference's own phase sum (state_lists.f90:84-110).
//  * other lattices: every rank is a candidate and the translation group is
//    scanned in the reference's loop order (k, j, i), as checkstate_SQ does.
//
// Compaction is a three-pass prefix sum (count per tile, scan of tile counts,
// recount and scatter), so list_s comes out in ascending rank order.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>

#include "qs3x_internal.h"

namespace qs3x {

// ------------------------------------------------------------------ host side
int64_t binom_host(int64_t n, int64_t k) {
  // f_combination (state_lists.f90:17-33): multiply then divide; C(n<k) = 0.
  if (n < k || k < 0) return 0;
  int64_t c = 1;
  for (int64_t i = 1; i <= k; ++i) {
    c = c * (n - k + i);
    c = c / i;
  }
  return c;
}

static std::vector<int64_t> binom_table(int nmax, int kmax) {
  // row-major [n][k], n = 0..nmax, k = 0..kmax; Pascal's rule (exact in int64
  // for every table the engine builds: C(1000,5) < 2^63).
  std::vector<int64_t> t((size_t)(nmax + 1) * (kmax + 1), 0);
  for (int n = 0; n <= nmax; ++n) {
    t[(size_t)n * (kmax + 1) + 0] = 1;
    for (int k = 1; k <= kmax && k <= n; ++k)
      t[(size_t)n * (kmax + 1) + k] =
          t[(size_t)(n - 1) * (kmax + 1) + k - 1] + ((k <= n - 1) ? t[(size_t)(n - 1) * (kmax + 1) + k] : 0);
  }
  return t;
}

void compute_explist(const qs3x_model *m, const double rk[3], std::vector<double> &ex) {
  // explist(i,j,k) = exp(i (rkx i + rky j + rkz k)) (input_param.f90:88-89),
  // evaluated as (cos x, sin x) of the same left-to-right sum; column-major.
  const int LX = m->LX, LY = m->LY, LZ = m->LZ;
  ex.assign((size_t)2 * (LX + 1) * (LY + 1) * (LZ + 1), 0.0);
  for (int k = 0; k <= LZ; ++k)
    for (int j = 0; j <= LY; ++j)
      for (int i = 0; i <= LX; ++i) {
        double x = rk[0] * (double)i + rk[1] * (double)j + rk[2] * (double)k;
        size_t q = (size_t)i + (size_t)(LX + 1) * ((size_t)j + (size_t)(LY + 1) * k);
        ex[2 * q] = std::cos(x);
        ex[2 * q + 1] = std::sin(x);
      }
}

// The reference's norm for a configuration whose stabiliser is the set
// {T^p, T^2p, ..., T^L} of a chain (p = L / count): ro = |sum e^{-i(pkx i +
// pky + pkz)}|^2 NOS / count, summed in loop order (state_lists.f90:84-110).
double chain_ro(int L, int count, const double rk[3]) {
  if (count <= 0 || L % count) return 0.0;
  const int p = L / count;
  const double p1 = rk[2] * 1.0;
  const double p2 = rk[1] * 1.0 + p1;
  double re = 0.0, im = 0.0;
  for (int i = p; i <= L; i += p) {
    double y = -(rk[0] * (double)i + p2);
    re += std::cos(y);
    im += std::sin(y);
  }
  double a = std::hypot(re, im);
  return (a * a) * (double)L / (double)count;
}

static int bits_for(int v) {
  int b = 1;
  while ((1 << b) <= v) ++b;
  return b;
}

// ---------------------------------------------------------------- device side
namespace {

constexpr int kTileThreads = 256;
constexpr int kPerThread = 8;
constexpr int kTile = kTileThreads * kPerThread;

__device__ __forceinline__ int64_t B(const int64_t *__restrict__ bt, int kdim, int n, int k) {
  return (n < 0 || k < 0) ? 0 : bt[(int64_t)n * kdim + k];
}

// colex unrank of c into a k-subset b[0..k-1] of {1..M} (b ascending):
// c = sum_m C(b_m - 1, m), m = 1..k.
__device__ void colex_unrank(int64_t c, int k, int M, const int64_t *__restrict__ bt, int kdim,
                             int *b) {
  int hi = M;
  for (int m = k; m >= 1; --m) {
    // largest x in [m, hi] with C(x-1, m) <= c
    int lo = m, h = hi;
    while (lo < h) {
      int mid = (lo + h + 1) >> 1;
      if (B(bt, kdim, mid - 1, m) <= c) lo = mid;
      else h = mid - 1;
    }
    b[m - 1] = lo;
    c -= B(bt, kdim, lo - 1, m);
    hi = lo - 1;
  }
}

// Chain candidate test.  a[0..N-1] ascending with a[0] = 1.  Returns the
// stabiliser order if a is the colex-minimal member of its orbit, else 0.
__device__ int chain_rep_count(const int *a, int N, int L) {
  int g[kMaxNod];
  for (int q = 0; q + 1 < N; ++q) g[q] = a[q + 1] - a[q];
  g[N - 1] = a[0] + L - a[N - 1];
  int count = 1;
  for (int j = 1; j < N; ++j) {
    int cmp = 0;  // sign of S_j - S_0 lexicographically
    for (int t = 1; t < N && cmp == 0; ++t) {
      int x = g[(j - t + N) % N], y = g[(N - t) % N];
      cmp = (x > y) - (x < y);
    }
    if (cmp > 0) return 0;
    if (cmp == 0) ++count;
  }
  return count;
}

__device__ __forceinline__ uint64_t pack_cfg(const int *a, int N, int bits, int cls) {
  uint64_t w = 0;
  for (int q = 0; q < N; ++q) w |= (uint64_t)a[q] << (q * bits);
  return w | ((uint64_t)cls << 56);
}

// Exclusive scan over a 256-thread block; returns this thread's offset and the
// block total via *total.
__device__ int block_exclusive_scan(int v, int *total) {
  __shared__ int wsum[kTileThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  int base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kTileThreads / 64; ++w) {
    if (w < wid) base += wsum[w];
    tot += wsum[w];
  }
  __syncthreads();
  *total = tot;
  return base + x - v;
}

template <bool FILL>
__global__ __launch_bounds__(kTileThreads) void k_chain_candidates(
    int64_t ncand, int N, int L, const int64_t *__restrict__ bt, int kdim,
    const double *__restrict__ ro_by_count, int64_t *__restrict__ tile_count,
    const int64_t *__restrict__ tile_offset, int64_t *__restrict__ list_s,
    double *__restrict__ list_r, uint64_t *__restrict__ cfg, int32_t *__restrict__ index, int bits) {
  const int64_t tile = blockIdx.x;
  const int64_t c0 = tile * kTile + (int64_t)threadIdx.x * kPerThread;
  int keep_mask = 0, cnt[kPerThread];
  int a[kMaxNod], b[kMaxNod];
  int nk = 0;
  for (int e = 0; e < kPerThread; ++e) {
    int64_t c = c0 + e;
    cnt[e] = 0;
    if (c >= ncand) continue;
    a[0] = 1;
    if (N > 1) {
      colex_unrank(c, N - 1, L - 1, bt, kdim, b);
      for (int q = 0; q < N - 1; ++q) a[q + 1] = b[q] + 1;
    }
    int count = chain_rep_count(a, N, L);
    if (count > 0 && ro_by_count[count] > 1.0e-15) {
      keep_mask |= 1 << e;
      cnt[e] = count;
      ++nk;
    }
  }
  int total;
  int off = block_exclusive_scan(nk, &total);
  if (!FILL) {
    if (threadIdx.x == 0) tile_count[tile] = total;
    return;
  }
  int64_t pos = tile_offset[tile] + off;
  for (int e = 0; e < kPerThread; ++e) {
    int64_t c = c0 + e;
    if (c >= ncand) continue;
    if (!(keep_mask >> e & 1)) {
      index[c] = -1;
      continue;
    }
    a[0] = 1;
    if (N > 1) {
      colex_unrank(c, N - 1, L - 1, bt, kdim, b);
      for (int q = 0; q < N - 1; ++q) a[q + 1] = b[q] + 1;
    }
    int64_t rank = 1;  // inv_list (state_lists.f90:131-139)
    for (int i = 2; i <= N; ++i) rank += B(bt, kdim, a[i - 1] - 1, i);
    list_s[pos] = rank;
    list_r[pos] = sqrt(ro_by_count[cnt[e]]);
    cfg[pos] = pack_cfg(a, N, bits, cnt[e]);
    index[c] = (int32_t)pos;
    ++pos;
  }
}

// General lattices: every rank is a candidate; translations scanned in the
// reference's loop order.  Writes coef (phase sum) and count for kept rows so
// the host can evaluate the norm with the reference's own abs().
struct Lat {
  int LX, LY, LZ, NOS;
};

__device__ __forceinline__ int shift_site(int s, int i, int j, int k, const Lat &l) {
  int s0 = s - 1;
  int x = s0 % l.LX, y = (s0 / l.LX) % l.LY, z = s0 / (l.LX * l.LY);
  x = (x + i) % l.LX;
  y = (y + j) % l.LY;
  z = (z + k) % l.LZ;
  return 1 + x + l.LX * (y + l.LY * z);
}

__device__ void sort_small(int *a, int n) {
  for (int q = 1; q < n; ++q) {
    int t = a[q], j = q - 1;
    while (j >= 0 && a[j] > t) {
      a[j + 1] = a[j];
      --j;
    }
    a[j + 1] = t;
  }
}

// Unrank a full rank t (1-based) into n[0..N-1] (list_fly, state_lists.f90:294-308).
__device__ void unrank_full(int64_t t, int N, int NOS, const int64_t *__restrict__ bt, int kdim,
                            int *n) {
  int64_t s = t;
  int hi = NOS;
  for (int i = N; i >= 2; --i) {
    // smallest m in [i, hi] with C(m, i) >= s
    int lo = i, h = hi;
    while (lo < h) {
      int mid = (lo + h) >> 1;
      if (B(bt, kdim, mid, i) >= s) h = mid;
      else lo = mid + 1;
    }
    n[i - 1] = lo;
    s -= B(bt, kdim, lo - 1, i);
    hi = lo - 1;
  }
  n[0] = (int)s;
}

template <bool FILL>
__global__ __launch_bounds__(kTileThreads) void k_general_candidates(
    int64_t ncand, int N, Lat lat, const int64_t *__restrict__ bt, int kdim,
    const double *__restrict__ phase /* complex, LX*LY*LZ in loop order */,
    int64_t *__restrict__ tile_count, const int64_t *__restrict__ tile_offset,
    int64_t *__restrict__ list_s, double *__restrict__ coef_out /* 3 per row */,
    uint64_t *__restrict__ cfg, int32_t *__restrict__ index, int bits, int use_index) {
  const int64_t tile = blockIdx.x;
  const int64_t c0 = tile * kTile + (int64_t)threadIdx.x * kPerThread;
  int n[kMaxNod], np[kMaxNod];
  int keep_mask = 0, nk = 0;
  double cre[kPerThread], cim[kPerThread];
  int ccount[kPerThread];
  for (int e = 0; e < kPerThread; ++e) {
    int64_t c = c0 + e;
    if (c >= ncand) continue;
    unrank_full(c + 1, N, lat.NOS, bt, kdim, n);
    double re = 0.0, im = 0.0;
    int count = 0;
    bool rep = true;
    int g = 0;
    for (int k = 1; k <= lat.LZ && rep; ++k)
      for (int j = 1; j <= lat.LY && rep; ++j)
        for (int i = 1; i <= lat.LX; ++i, ++g) {
          for (int q = 0; q < N; ++q) np[q] = shift_site(n[q], i, j, k, lat);
          sort_small(np, N);
          int l = N - 1, cmp = 0;
          for (; l >= 0; --l) {
            if (np[l] < n[l]) { cmp = -1; break; }
            if (np[l] > n[l]) { cmp = 1; break; }
          }
          if (cmp < 0) { rep = false; break; }
          if (cmp == 0) {
            // phase index in loop order: g = (i-1) + LX*((j-1) + LY*(k-1))
            re = re + phase[2 * g];
            im = im + phase[2 * g + 1];
            ++count;
          }
        }
    if (!rep) continue;
    double ro = (re * re + im * im) * (double)lat.NOS / (double)count;
    if (ro > 1.0e-15) {
      keep_mask |= 1 << e;
      cre[e] = re; cim[e] = im; ccount[e] = count;
      ++nk;
    }
  }
  int total;
  int off = block_exclusive_scan(nk, &total);
  if (!FILL) {
    if (threadIdx.x == 0) tile_count[tile] = total;
    return;
  }
  int64_t pos = tile_offset[tile] + off;
  for (int e = 0; e < kPerThread; ++e) {
    int64_t c = c0 + e;
    if (c >= ncand) continue;
    if (!(keep_mask >> e & 1)) {
      if (use_index) index[c] = -1;
      continue;
    }
    unrank_full(c + 1, N, lat.NOS, bt, kdim, n);
    list_s[pos] = c + 1;
    coef_out[3 * pos + 0] = cre[e];
    coef_out[3 * pos + 1] = cim[e];
    coef_out[3 * pos + 2] = (double)ccount[e];
    cfg[pos] = pack_cfg(n, N, bits, 0);
    if (use_index) index[c] = (int32_t)pos;
    ++pos;
  }
}

// Exclusive scan of n int64 values in place (single block, serial carry).
__global__ __launch_bounds__(1024) void k_scan_tiles(int64_t *__restrict__ v, int64_t n,
                                                     int64_t *__restrict__ total) {
  __shared__ int64_t wsum[16];
  __shared__ int64_t carry_s;
  if (threadIdx.x == 0) carry_s = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int64_t base = 0; base < n; base += 1024) {
    int64_t i = base + threadIdx.x;
    int64_t x0 = (i < n) ? v[i] : 0;
    int64_t x = x0;
    for (int d = 1; d < 64; d <<= 1) {
      int64_t y = __shfl_up(x, d, 64);
      if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    int64_t pre = 0, tot = 0;
    for (int w = 0; w < 16; ++w) {
      if (w < wid) pre += wsum[w];
      tot += wsum[w];
    }
    int64_t carry = carry_s;
    if (i < n) v[i] = carry + pre + x - x0;
    __syncthreads();
    if (threadIdx.x == 0) carry_s = carry + tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry_s;
}

// Unrank list_s into packed configurations (sector_from_lists) and fill the
// index table from them.
__global__ void k_cfg_from_ranks(int64_t dim, int N, int NOS, const int64_t *__restrict__ list_s,
                                 const int64_t *__restrict__ bt, int kdim, int bits,
                                 uint64_t *__restrict__ cfg) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim) return;
  int n[kMaxNod];
  unrank_full(list_s[i], N, NOS, bt, kdim, n);
  cfg[i] = pack_cfg(n, N, bits, 0);
}

// Chain index from configurations: the anchored colex rank of n (which must
// contain site 1) -> row, and the stabiliser order into the class bits.
__global__ void k_chain_index_from_cfg(int64_t dim, int N, int L, int bits,
                                       const int64_t *__restrict__ bt, int kdim,
                                       uint64_t *__restrict__ cfg, int32_t *__restrict__ index,
                                       int *__restrict__ bad) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim) return;
  int a[kMaxNod];
  uint64_t w = cfg[i];
  for (int q = 0; q < N; ++q) a[q] = (int)((w >> (q * bits)) & ((1ull << bits) - 1));
  if (a[0] != 1) {
    atomicOr(bad, 1);
    return;
  }
  int count = chain_rep_count(a, N, L);
  if (count == 0) {
    atomicOr(bad, 2);
    return;
  }
  int64_t c = 0;
  for (int m = 1; m < N; ++m) c += B(bt, kdim, a[m] - 2, m);
  index[c] = (int32_t)i;
  cfg[i] = (w & ((1ull << 56) - 1)) | ((uint64_t)count << 56);
}

__global__ void k_full_index_from_ranks(int64_t dim, const int64_t *__restrict__ list_s,
                                        int32_t *__restrict__ index) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim) return;
  index[list_s[i] - 1] = (int32_t)i;
}

// Bitmap words from the dense anchored index: word w covers ranks 32w..32w+31.
__global__ void k_bitmap_from_index(const int32_t *__restrict__ index, int64_t ncand, uint2 *__restrict__ bm,
                                    int64_t nwords) {
  const int64_t wi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (wi >= nwords) return;
  uint32_t bits = 0;
  int32_t first = 0;
  const int64_t c0 = wi * 32;
  for (int b = 0; b < 32 && c0 + b < ncand; ++b) {
    const int32_t id = index[c0 + b];
    if (id >= 0) {
      if (bits == 0) first = id;
      bits |= 1u << b;
    }
  }
  uint2 o;
  o.x = bits;
  o.y = (uint32_t)first;
  bm[wi] = o;
}

__global__ void k_fill_i32(int32_t *__restrict__ p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

// packed: the sector keeps one packed configuration word per row (every
// sector but the U(1)-only block, which unranks rows instead).
static void upload_common(qs3x_sector *s, int kmax, bool packed = true) {
  qs3x_model *m = s->model;
  s->nbinom_n = m->NOS + 1;
  s->nbinom_k = kmax + 1;
  auto bt = binom_table(m->NOS, kmax);
  s->d_binom.alloc(bt.size());
  s->d_binom.upload(bt.data(), bt.size());
  compute_explist(m, s->rk, s->h_explist);
  s->d_explist.alloc(s->h_explist.size());
  s->d_explist.upload(s->h_explist.data(), s->h_explist.size());
  // Only explist(0,0,0) = 1 exactly and, on a chain, the slice (i,1,1) are
  // used; the operator is real when all of those are exactly (1, 0).
  bool real = true;
  const int LX = m->LX, LY = m->LY, LZ = m->LZ;
  for (int k = 0; k <= LZ && real; ++k)
    for (int j = 0; j <= LY && real; ++j)
      for (int i = 0; i <= LX; ++i) {
        size_t q = (size_t)i + (size_t)(LX + 1) * ((size_t)j + (size_t)(LY + 1) * k);
        if (s->h_explist[2 * q] != 1.0 || s->h_explist[2 * q + 1] != 0.0) {
          real = false;
          break;
        }
      }
  s->real_phase = real;
  s->bits = bits_for(m->NOS);
  QS3X_REQUIRE(!packed || s->nod * s->bits <= 56,
               "sector: N_down * ceil(log2(NOS+1)) must be <= 56 bits for the packed rows");
}

static bool chain_fast_ok(const qs3x_model *m, int nod) {
  if (!m->chain || nod < 1 || nod > kMaxChainNod) return false;
  int kb = bits_for(m->NOS);
  // gap keys fit in 64 bits; anchored ranks (and the dense index) fit in int32
  return (int64_t)(nod - 1) * kb <= 64 && binom_host(m->NOS - 1, nod - 1) < (int64_t)INT32_MAX;
}

static void build_bitmap(qs3x_sector *s);

void build_sector_u1(qs3x_sector *s) {
  qs3x_model *m = s->model;
  const int N = s->nod;
  QS3X_REQUIRE(N >= 1 && N <= kMaxNod, "U(1) sector: N_down must be in 1..16");
  QS3X_REQUIRE(N <= m->NOS, "U(1) sector: N_down > NOS");
  // rows are 32-bit gather targets in the H.v kernels' int64 arithmetic, so
  // only the row count bounds the block
  QS3X_REQUIRE(m->NOS <= 65535, "U(1) sector: NOS must be below 65536 (16-bit site fields of the adjacency)");
  s->rk[0] = s->rk[1] = s->rk[2] = 0.0;
  upload_common(s, N + 1, false);
  s->mode = kIndexU1;
  s->chain = false;
  s->real_phase = true;
  // THS = combination(NOS, NOD) (source_only_u1/main.f90:27)
  s->dim = binom_host(m->NOS, N);
  QS3X_REQUIRE(s->dim > 0, "U(1) sector: C(NOS, N_down) overflows int64");
}

void build_sector_device(qs3x_sector *s) {
  qs3x_model *m = s->model;
  const int N = s->nod;
  QS3X_REQUIRE(N >= 1 && N <= kMaxNod, "sector: N_down must be in 1..16 (N_down = 0 is the "
                                       "ferromagnetic state, see qs3x_fm_energy)");
  QS3X_REQUIRE(N <= m->NOS, "sector: N_down > NOS");
  upload_common(s, N + 1);
  hipStream_t st = stream();
  DevBuf<int64_t> d_total(1);
  int64_t total = 0;
  if (chain_fast_ok(m, N)) {
    s->chain = true;
    s->mode = kIndexAnchored;
    const int L = m->NOS;
    int64_t ncand = binom_host(L - 1, N - 1);
    QS3X_REQUIRE(ncand < (int64_t)INT32_MAX * 8, "sector: too many candidates");
    // norm table by stabiliser order
    s->h_rclass.assign(N + 1, 0.0);
    std::vector<double> ro(N + 1, 0.0);
    for (int c = 1; c <= N; ++c) {
      // an all-k sector keeps every representative (norms are per wave vector)
      ro[c] = s->all_k ? 1.0 : chain_ro(L, c, s->rk);
      s->h_rclass[c] = std::sqrt(ro[c]);
    }
    DevBuf<double> d_ro(N + 1);
    d_ro.upload(ro.data(), N + 1);
    s->d_rclass.alloc(N + 1);
    s->d_rclass.upload(s->h_rclass.data(), N + 1);
    int64_t ntiles = (ncand + kTile - 1) / kTile;
    DevBuf<int64_t> d_tiles(ntiles);
    hipLaunchKernelGGL(k_chain_candidates<false>, dim3((unsigned)ntiles), dim3(kTileThreads), 0, st,
                       ncand, N, L, s->d_binom.p, s->nbinom_k, d_ro.p, d_tiles.p, nullptr, nullptr,
                       nullptr, nullptr, nullptr, s->bits);
    QS3X_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, st, d_tiles.p, ntiles, d_total.p);
    d_total.download(&total, 1);
    QS3X_REQUIRE(total < (int64_t)INT32_MAX, "sector: dimension exceeds 2^31-1");
    s->dim = total;
    s->d_list_s.alloc(total);
    s->d_list_r.alloc(total);
    s->d_cfg.alloc(total);
    s->index_len = ncand;
    s->d_index.alloc(ncand);
    hipLaunchKernelGGL(k_chain_candidates<true>, dim3((unsigned)ntiles), dim3(kTileThreads), 0, st,
                       ncand, N, L, s->d_binom.p, s->nbinom_k, d_ro.p, nullptr, d_tiles.p,
                       s->d_list_s.p, s->d_list_r.p, s->d_cfg.p, s->d_index.p, s->bits);
    QS3X_HIP(hipGetLastError());
    build_bitmap(s);
    QS3X_HIP(hipStreamSynchronize(st));
    return;
  }
  // general lattice
  QS3X_REQUIRE(!s->all_k, "sector: an all-k sector needs the chain enumeration (LY = LZ = 1, C(L-1, N_down-1) "
                          "< 2^31, packed keys within 64 bits)");
  s->chain = false;
  const int64_t ncand = binom_host(m->NOS, N);
  QS3X_REQUIRE(ncand > 0 && ncand < ((int64_t)1 << 40), "sector: C(NOS, N) too large for the "
                                                       "general-lattice build");
  Lat lat{m->LX, m->LY, m->LZ, m->NOS};
  // phase table exp(-i(pkx i + (pky j + pkz k))) in loop order, as in
  // checkstate_SQ (p1 = pkz k, p2 = pky j + p1).
  std::vector<double> ph((size_t)2 * m->LX * m->LY * m->LZ);
  {
    size_t g = 0;
    for (int k = 1; k <= m->LZ; ++k) {
      double p1 = s->rk[2] * (double)k;
      for (int j = 1; j <= m->LY; ++j) {
        double p2 = s->rk[1] * (double)j + p1;
        for (int i = 1; i <= m->LX; ++i, ++g) {
          double y = -(s->rk[0] * (double)i + p2);
          ph[2 * g] = std::cos(y);
          ph[2 * g + 1] = std::sin(y);
        }
      }
    }
  }
  DevBuf<double> d_ph(ph.size());
  d_ph.upload(ph.data(), ph.size());
  const bool use_index = ncand <= ((int64_t)1 << 31) - 2;
  s->mode = use_index ? kIndexFullRank : kIndexSearch;
  int64_t ntiles = (ncand + kTile - 1) / kTile;
  DevBuf<int64_t> d_tiles(ntiles);
  hipLaunchKernelGGL(k_general_candidates<false>, dim3((unsigned)ntiles), dim3(kTileThreads), 0, st,
                     ncand, N, lat, s->d_binom.p, s->nbinom_k, d_ph.p, d_tiles.p, nullptr, nullptr,
                     nullptr, nullptr, nullptr, s->bits, 0);
  QS3X_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, st, d_tiles.p, ntiles, d_total.p);
  d_total.download(&total, 1);
  QS3X_REQUIRE(total < (int64_t)INT32_MAX, "sector: dimension exceeds 2^31-1");
  s->dim = total;
  s->d_list_s.alloc(total);
  s->d_list_r.alloc(total);
  s->d_cfg.alloc(total);
  if (use_index) {
    s->index_len = ncand;
    s->d_index.alloc(ncand);
  }
  DevBuf<double> d_coef((size_t)3 * std::max<int64_t>(total, 1));
  hipLaunchKernelGGL(k_general_candidates<true>, dim3((unsigned)ntiles), dim3(kTileThreads), 0, st,
                     ncand, N, lat, s->d_binom.p, s->nbinom_k, d_ph.p, nullptr, d_tiles.p,
                     s->d_list_s.p, d_coef.p, s->d_cfg.p, s->d_index.p, s->bits, use_index ? 1 : 0);
  QS3X_HIP(hipGetLastError());
  // Norms on the host with the reference's abs (hypot) of the bit-identical
  // phase sum: ro = |coef|^2 NOS / count, list_r = sqrt(ro).
  std::vector<double> coef((size_t)3 * total), lr(total);
  if (total) d_coef.download(coef.data(), coef.size());
  for (int64_t r = 0; r < total; ++r) {
    double a = std::hypot(coef[3 * r], coef[3 * r + 1]);
    lr[r] = std::sqrt((a * a) * (double)m->NOS / coef[3 * r + 2]);
  }
  if (total) s->d_list_r.upload(lr.data(), total);
  QS3X_HIP(hipStreamSynchronize(st));
}

static void build_bitmap(qs3x_sector *s) {
  const int64_t ncand = s->index_len;
  const int64_t nwords = (ncand + 31) / 32;
  s->d_bitmap.alloc(std::max<int64_t>(nwords, 1));
  if (nwords)
    hipLaunchKernelGGL(k_bitmap_from_index, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, stream(),
                       s->d_index.p, ncand, s->d_bitmap.p, nwords);
  QS3X_HIP(hipGetLastError());
}

void sector_from_lists(qs3x_sector *s, const int64_t *list_s, const double *list_r,
                       const double *explist) {
  qs3x_model *m = s->model;
  const int N = s->nod;
  QS3X_REQUIRE(N >= 1 && N <= kMaxNod, "sector_from_lists: N_down must be in 1..16");
  upload_common(s, N + 1);
  const size_t nex = (size_t)2 * (m->LX + 1) * (m->LY + 1) * (m->LZ + 1);
  if (explist) {
    s->h_explist.assign(explist, explist + nex);
    s->d_explist.upload(explist, nex);
    bool real = true;
    for (size_t q = 0; q < nex / 2; ++q)
      if (explist[2 * q] != 1.0 || explist[2 * q + 1] != 0.0) real = false;
    s->real_phase = real;
  }
  hipStream_t st = stream();
  const int64_t dim = s->dim;
  for (int64_t i = 1; i < dim; ++i)
    QS3X_REQUIRE(list_s[i] > list_s[i - 1], "sector_from_lists: list_s must be strictly ascending");
  if (dim)
    QS3X_REQUIRE(list_s[0] >= 1 && list_s[dim - 1] <= binom_host(m->NOS, N),
                 "sector_from_lists: ranks must lie in [1, C(NOS, N_down)]");
  s->d_list_s.alloc(dim);
  s->d_list_r.alloc(dim);
  s->d_cfg.alloc(dim);
  s->d_list_s.upload(list_s, dim);
  s->d_list_r.upload(list_r, dim);
  const int thr = 256;
  const unsigned blocks = (unsigned)((dim + thr - 1) / thr);
  if (dim)
    hipLaunchKernelGGL(k_cfg_from_ranks, dim3(blocks), dim3(thr), 0, st, dim, N, m->NOS,
                       s->d_list_s.p, s->d_binom.p, s->nbinom_k, s->bits, s->d_cfg.p);
  // A chain sector whose rows all contain site 1 uses the anchored index.
  s->chain = false;
  if (chain_fast_ok(m, N)) {
    const int64_t ncand = binom_host(m->NOS - 1, N - 1);
    s->index_len = ncand;
    s->d_index.alloc(ncand);
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((ncand + thr - 1) / thr)), dim3(thr), 0, st,
                       s->d_index.p, ncand, -1);
    DevBuf<int> d_bad(1);
    int bad = 0;
    d_bad.upload(&bad, 1);
    if (dim)
      hipLaunchKernelGGL(k_chain_index_from_cfg, dim3(blocks), dim3(thr), 0, st, dim, N, m->NOS,
                         s->bits, s->d_binom.p, s->nbinom_k, s->d_cfg.p, s->d_index.p, d_bad.p);
    d_bad.download(&bad, 1);
    if (bad == 0) {
      s->chain = true;
      s->mode = kIndexAnchored;
      build_bitmap(s);
      QS3X_HIP(hipStreamSynchronize(st));
      return;
    }
    // not representative lists: fall through to a general index
    if (dim)
      hipLaunchKernelGGL(k_cfg_from_ranks, dim3(blocks), dim3(thr), 0, st, dim, N, m->NOS,
                         s->d_list_s.p, s->d_binom.p, s->nbinom_k, s->bits, s->d_cfg.p);
  }
  const int64_t ncand = binom_host(m->NOS, N);
  if (ncand <= ((int64_t)1 << 31) - 2) {
    s->mode = kIndexFullRank;
    s->index_len = ncand;
    s->d_index.alloc(ncand);
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((ncand + thr - 1) / thr)), dim3(thr), 0, st,
                       s->d_index.p, ncand, -1);
    if (dim)
      hipLaunchKernelGGL(k_full_index_from_ranks, dim3(blocks), dim3(thr), 0, st, dim,
                         s->d_list_s.p, s->d_index.p);
  } else {
    s->mode = kIndexSearch;
    s->index_len = 0;
  }
  QS3X_HIP(hipGetLastError());
  QS3X_HIP(hipStreamSynchronize(st));
}

}  // namespace qs3x
