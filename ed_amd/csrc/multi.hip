// multi.hip -- plain Lanczos for many sectors at once (one launch per N_down
// per step, one hipGraph per step).
//
// lanczos_routines (lanczos.f90:14-152) runs one sector at a time.  A scan
// over every (N_down, k) sector of a chain is then a long sequence of small
// solves whose steps are a handful of short kernels each, so launch latency,
// not bandwidth, sets the time.  Here the sectors of a group advance together:
// * their vectors are concatenated; every sector is padded to whole tiles of
//   256 rows, so a tile belongs to one sector, found from the tile number by
//   a binary search over the sectors' first tiles;
// * one H.v launch per N_down covers every tile of every sector of that N_down
//   (per-hop backward-gap key search over the bond adjacency, as the per-hop
//   chain kernel of hv.hip; per-sector norms and phases from a descriptor);
// * one update launch covers all sectors, with per-sector alpha, beta and
//   scale in a [slot][sector] array; dots go to per-tile partials that a
//   segmented finalize (one block per sector) sums in a fixed order;
// * the slot of step i is read from a device step counter, so one captured
//   hipGraph is the whole step and is replayed between convergence checks;
// * a sector stops (its tiles are skipped) when the host's check finds it
//   converged or broken down, at its step bound (dim in FP64, as the
//   reference; 2 dim + 8 in FP32, see qs3x_lanczos) or at maxitr.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "blas.h"
#include "hv_common.h"
#include "qs3x_internal.h"

namespace qs3x {
namespace {

using namespace hvk;

constexpr int kMT = 256;  // rows per tile = threads per block
constexpr int kMaxCls = kMaxChainNod + 1;

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class F>
int guarded5(F &&f) {
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

struct SecDesc {
  const uint64_t *cfg;
  const uint2 *bitmap;
  const double *explist;  // complex
  int64_t dim;
  int64_t voff;   // first row in the group vector
  int64_t tile0;  // first tile in the group numbering
  double rcl[kMaxCls];  // norm by stabiliser order
};

// slot layout of the per-sector scalars: sc[(base + step) * S + s]
struct Slots {
  int a_rec, a_out, beta, scale, tmp;
};

__device__ __forceinline__ int find_sector(const int64_t *tile0, int s0, int s1, int64_t t) {
  // last s in [s0, s1) with tile0[s] <= t
  int lo = s0, hi = s1 - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tile0[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Block reduction of (re, im) into out[2 * idx] (all kMT threads call it).
__device__ __forceinline__ void tile_reduce(double re, double im, double *out, int64_t idx) {
  __shared__ double red[2][kMT / 64];
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
    for (int w = 0; w < kMT / 64; ++w) {
      a += red[0][w];
      b += red[1][w];
    }
    out[2 * idx] = a;
    out[2 * idx + 1] = b;
  }
  __syncthreads();
}

struct MultiArgs {
  const SecDesc *sec;
  const int64_t *tile0;  // [S + 1]
  const int *active;     // [S]
  int s0, s1;            // sectors of this launch (one N_down)
  int64_t t0, t1;        // their tiles
  int L, bits, kbits, kdim;
  const int64_t *binom;
  const int32_t *adj_ptr, *adj;
  int nadj;
  const double *bt;
  int nbt;
  double shift;  // e_zz in FP64, 0 in FP32
  const void *v1;
  void *v0;
  const double *sc;
  int S;
  int scale_base;  // slot base of the scale of v1 (slot = scale_base + step)
  const int *step;
  double *part;  // [tiles] complex
  // Hop cache (null: none): per hop slot h and row of the launch (row
  // (t - t0) kMT + thread), the target's index (-1: absent) and a code
  // (0: no hop; else bit 31 | bond type << 16 | stabiliser order << 12 |
  // ell), written by step 0 and read by every later step instead of the
  // representative search.
  int32_t *rec_id;
  uint32_t *rec_code;
  int64_t rec_rows;
};

template <int N, int DEG, class R, class KeyT>
__global__ __launch_bounds__(kMT) void k_hv_multi(MultiArgs a) {
  using C = typename CT<R>::T;
  constexpr int H = N * DEG;
  extern __shared__ __align__(16) unsigned char smem[];
  const int L = a.L;
  const int nb = (N > 1) ? (N - 1) * (L - 1) : 0;
  int32_t *s_binom = (int32_t *)smem;
  R *s_jxy2 = (R *)(smem + (((size_t)nb * 4 + 15) & ~(size_t)15));
  R *s_jz4 = s_jxy2 + a.nbt;
  int *s_ptr = (int *)(s_jz4 + a.nbt);
  int *s_adj = s_ptr + (L + 2);
  for (int q = threadIdx.x; q < nb; q += kMT) {
    const int m = q / (L - 1) + 1, x = q % (L - 1);
    s_binom[q] = (int32_t)a.binom[(int64_t)x * a.kdim + m];
  }
  for (int q = threadIdx.x; q < a.nbt; q += kMT) {
    s_jxy2[q] = (R)a.bt[2 * q];
    s_jz4[q] = (R)a.bt[2 * q + 1];
  }
  for (int q = threadIdx.x; q < L + 2; q += kMT) s_ptr[q] = a.adj_ptr[q];
  for (int q = threadIdx.x; q < a.nadj; q += kMT) s_adj[q] = a.adj[q];
  __syncthreads();
  const int kb = a.kbits;
  const KeyT fmask = (KeyT)((1ull << kb) - 1ull);
  const int step = *a.step;

  for (int64_t t = a.t0 + blockIdx.x; t < a.t1; t += gridDim.x) {
    const int s = find_sector(a.tile0, a.s0, a.s1, t);
    if (!a.active[s]) continue;  // block-uniform
    const SecDesc &d = a.sec[s];
    const int64_t lr = (t - d.tile0) * kMT + threadIdx.x;
    const bool valid = lr < d.dim;
    const C *v1 = (const C *)a.v1 + d.voff;
    C *v0 = (C *)a.v0 + d.voff;
    const R scale = (R)a.sc[(size_t)(a.scale_base + step) * a.S + s];
    double dre = 0, dim_ = 0;
    if (valid) {
      const uint64_t w = d.cfg[lr];
      int n[N];
#pragma unroll
      for (int q = 0; q < N; ++q) n[q] = (int)((w >> (q * a.bits)) & ((1ull << a.bits) - 1ull));
      const R rinv = (R)(1.0 / d.rcl[(int)(w >> 56)]);
      // The row's own gap keys: a hop that moves a spin by one site without
      // wrapping (a chain bond; the target is empty) keeps the order of the
      // down spins and changes two gaps by one, g[m] -= d and g[m - 1] += d,
      // so each rotation's key moves by +-(a field's unit) twice:
      // K'[j] = KB[j] + d (unit(j, m - 1) - unit(j, m)), the units of gaps that
      // key j leaves out being 0.  (Checked against the direct build on
      // 1,009,321 keys of 8 chain shapes.)
      KeyT KB[N];
      {
        int g0[N];
#pragma unroll
        for (int q = 0; q < N - 1; ++q) g0[q] = n[q + 1] - n[q];
        g0[N - 1] = n[0] + L - n[N - 1];
#pragma unroll
        for (int j = 0; j < N; ++j) {
          KeyT k = 0;
#pragma unroll
          for (int tt = 1; tt < N; ++tt) k = (KeyT)((k << kb) | (KeyT)g0[(j - tt + N) % N]);
          KB[j] = k;
        }
      }
      R anti = 0;
      int hc[H], hcnt[H], hei[H], hbt[H];
      R hj[H];
      int hid[H];
#pragma unroll
      for (int h = 0; h < H; ++h) {
        hc[h] = -1;
        hcnt[h] = 1;
        hei[h] = 0;
        hj[h] = 0;
        hbt[h] = -1;
      }
      const int64_t rrow = (t - a.t0) * kMT + threadIdx.x;
      if (a.rec_id && step > 0) {
        // the hops recorded at step 0, in the same slot order (so anti sums
        // in the same order as the search below)
#pragma unroll
        for (int h = 0; h < H; ++h) {
          const uint32_t cd = a.rec_code[(size_t)h * a.rec_rows + rrow];
          hid[h] = a.rec_id[(size_t)h * a.rec_rows + rrow];
          if (cd) {
            const int btp = (int)((cd >> 16) & 0x7fffu), ell = (int)(cd & 0xfffu);
            anti += s_jz4[btp];
            hj[h] = s_jxy2[btp];
            hcnt[h] = (int)((cd >> 12) & 0xfu);
            hei[h] = ell ? ell + 3 * (L + 1) : 0;
          }
        }
      } else {
#pragma unroll
        for (int m = 0; m < N; ++m) {
          const int st = n[m];
          const int e0 = s_ptr[st], deg = s_ptr[st + 1] - e0;
#pragma unroll
          for (int e = 0; e < DEG; ++e) {
            if (e >= deg) break;
            const int ent = s_adj[e0 + e];
            const int tg = ent & 0xffff, btp = ent >> 16;
            bool in = false;
#pragma unroll
            for (int q = 0; q < N; ++q) in |= (n[q] == tg);
            if (in) continue;
            anti += s_jz4[btp];
            int nn[N];
            KeyT K[N];
            const int dd = tg - st;
            if (dd == 1 || dd == -1) {
#pragma unroll
              for (int q = 0; q < N; ++q) nn[q] = n[q] + (q == m ? dd : 0);
#pragma unroll
              for (int j = 0; j < N; ++j) {
                const int ta = (j - (m - 1 + N) % N + N) % N, tb = (j - m + N) % N;  // field of g[m - 1], g[m]
                const KeyT ua = ta ? (KeyT)1 << (kb * (N - 1 - ta)) : (KeyT)0;
                const KeyT ub = tb ? (KeyT)1 << (kb * (N - 1 - tb)) : (KeyT)0;
                K[j] = dd > 0 ? (KeyT)(KB[j] + ua - ub) : (KeyT)(KB[j] - ua + ub);
              }
            } else {
              int r[N > 1 ? N - 1 : 1];
#pragma unroll
              for (int q = 0; q < N - 1; ++q) r[q] = (q < m) ? n[q] : n[q + 1];
#pragma unroll
              for (int q = 0; q < N; ++q) {
                const bool ltq = (q < N - 1) ? (r[q < N - 1 ? q : 0] < tg) : false;
                const bool ltp = (q > 0) ? (r[q > 0 ? q - 1 : 0] < tg) : true;
                nn[q] = ltq ? r[q < N - 1 ? q : 0] : (ltp ? tg : r[q > 0 ? q - 1 : 0]);
              }
              int g[N];
#pragma unroll
              for (int q = 0; q < N - 1; ++q) g[q] = nn[q + 1] - nn[q];
              g[N - 1] = nn[0] + L - nn[N - 1];
#pragma unroll
              for (int j = 0; j < N; ++j) {
                KeyT k = 0;
#pragma unroll
                for (int tt = 1; tt < N; ++tt) k = (KeyT)((k << kb) | (KeyT)g[(j - tt + N) % N]);
                K[j] = k;
              }
            }
            KeyT kmax = K[0];
#pragma unroll
            for (int j = 1; j < N; ++j) kmax = K[j] > kmax ? K[j] : kmax;
            int mx = 0, count = 0;
#pragma unroll
            for (int j = 0; j < N; ++j) {
              const bool tie = (K[j] == kmax);
              count += tie ? 1 : 0;
              mx = (tie && nn[j] > mx) ? nn[j] : mx;
            }
            const bool isrep = (nn[0] == 1) && (K[0] == kmax);
            int c = 0;
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
            const int h = m * DEG + e;
            hc[h] = c;
            hcnt[h] = count;
            hei[h] = isrep ? 0 : (L + 1 - mx) + 3 * (L + 1);
            hj[h] = s_jxy2[btp];
            hbt[h] = btp;
          }
        }
        uint2 wd[H];
#pragma unroll
        for (int h = 0; h < H; ++h) wd[h] = d.bitmap[(uint32_t)(hc[h] >= 0 ? hc[h] : 0) >> 5];
#pragma unroll
        for (int h = 0; h < H; ++h) {
          const uint32_t b = (uint32_t)hc[h] & 31u;
          const bool in = hc[h] >= 0 && ((wd[h].x >> b) & 1u);
          hid[h] = in ? (int)wd[h].y + __popc(wd[h].x & ((1u << b) - 1u)) : -1;
        }
        if (a.rec_id) {
#pragma unroll
          for (int h = 0; h < H; ++h) {
            const uint32_t ell = hei[h] ? (uint32_t)(hei[h] - 3 * (L + 1)) : 0u;
            const uint32_t cd =
                hbt[h] >= 0 ? (1u << 31) | ((uint32_t)hbt[h] << 16) | ((uint32_t)hcnt[h] << 12) | ell : 0u;
            a.rec_id[(size_t)h * a.rec_rows + rrow] = hid[h];
            a.rec_code[(size_t)h * a.rec_rows + rrow] = cd;
          }
        }
      }
      C y[H];
#pragma unroll
      for (int h = 0; h < H; ++h) y[h] = v1[hid[h] >= 0 ? (int64_t)hid[h] : lr];
      R accx = 0, accy = 0;
#pragma unroll
      for (int h = 0; h < H; ++h) {
        if (hid[h] < 0) continue;
        const R cf = hj[h] * (R)d.rcl[hcnt[h]] * rinv;
        const R pr = (R)d.explist[2 * hei[h]] * cf, pi = (R)d.explist[2 * hei[h] + 1] * cf;
        accx += pr * y[h].x - pi * y[h].y;
        accy += pr * y[h].y + pi * y[h].x;
      }
      C x = v1[lr];
      x.x *= scale;
      x.y *= scale;
      const R dg = (R)a.shift - (R)2 * anti;
      C out;
      out.x = dg * x.x + accx * scale;
      out.y = dg * x.y + accy * scale;
      v0[lr] = out;
      dre = (double)x.x * (double)out.x + (double)x.y * (double)out.y;
      dim_ = (double)x.x * (double)out.y - (double)x.y * (double)out.x;
    }
    tile_reduce(dre, dim_, a.part, t);
  }
}

// w = w - a s u - b s' up per sector (b term from step 1 on); part[tile] = |w|^2
template <class R>
__global__ __launch_bounds__(kMT) void k_update3_multi(void *wv, const void *uv, const void *upv,
                                                       const SecDesc *sec, const int64_t *tile0, const int *active,
                                                       int S, int64_t ntiles, const double *sc, Slots sl,
                                                       const int *stepp, double *part) {
  using C = typename CT<R>::T;
  const int step = *stepp;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int s = find_sector(tile0, 0, S, t);
    if (!active[s]) continue;
    const SecDesc &d = sec[s];
    const int64_t lr = (t - d.tile0) * kMT + threadIdx.x;
    double acc = 0;
    if (lr < d.dim) {
      C *w = (C *)wv + d.voff;
      const C *u = (const C *)uv + d.voff;
      const C *up = (const C *)upv + d.voff;
      const R ca = (R)(sc[(size_t)(sl.a_rec + step) * S + s] * sc[(size_t)(sl.scale + step) * S + s]);
      C x = w[lr];
      const C a1 = u[lr];
      x.x -= ca * a1.x;
      x.y -= ca * a1.y;
      if (step > 0) {
        const R cb = (R)(sc[(size_t)(sl.beta + step - 1) * S + s] * sc[(size_t)(sl.scale + step - 1) * S + s]);
        const C b1 = up[lr];
        x.x -= cb * b1.x;
        x.y -= cb * b1.y;
      }
      w[lr] = x;
      acc = (double)x.x * (double)x.x + (double)x.y * (double)x.y;
    }
    tile_reduce(acc, 0.0, part, t);
  }
}

// |v|^2 per tile (start vectors)
template <class R>
__global__ __launch_bounds__(kMT) void k_norm_multi(const void *vv, const SecDesc *sec, const int64_t *tile0, int S,
                                                    int64_t ntiles, double *part) {
  using C = typename CT<R>::T;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int s = find_sector(tile0, 0, S, t);
    const SecDesc &d = sec[s];
    const int64_t lr = (t - d.tile0) * kMT + threadIdx.x;
    double acc = 0;
    if (lr < d.dim) {
      const C x = ((const C *)vv + d.voff)[lr];
      acc = (double)x.x * (double)x.x + (double)x.y * (double)x.y;
    }
    tile_reduce(acc, 0.0, part, t);
  }
}

// Segmented finalize, one block per sector (active sectors only).
// mode 1: sc[i0] = re, sc[i1] = re + shift; mode 2: b = sqrt(re): sc[i0] = b, sc[i1] = 1/b (0 if b = 0).
// Slots are base + step (+ off1 for i1).
__global__ __launch_bounds__(kMT) void k_finalize_multi_sec(const double *part, const int64_t *tile0,
                                                            const int *active, int S, double *sc, int mode, int b0,
                                                            int b1, int off1, double shift, const int *stepp) {
  const int s = blockIdx.x;
  if (s >= S || (active && !active[s])) return;
  const int step = stepp ? *stepp : 0;
  double re = 0;
  for (int64_t t = tile0[s] + threadIdx.x; t < tile0[s + 1]; t += kMT) re += part[2 * t];
  __shared__ double red[kMT];
  red[threadIdx.x] = re;
  __syncthreads();
  for (int k = kMT / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double v = red[0];
    if (mode == 1) {
      sc[(size_t)(b0 + step) * S + s] = v;
      sc[(size_t)(b1 + step + off1) * S + s] = v + shift;
    } else {
      const double b = sqrt(v > 0 ? v : 0.0);
      sc[(size_t)(b0 + step) * S + s] = b;
      sc[(size_t)(b1 + step + off1) * S + s] = b > 0 ? 1.0 / b : 0.0;
    }
  }
}

__global__ void k_step_inc(int *step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

__global__ void k_set_int(int *p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

using MFn = void (*)(MultiArgs);

template <int N, class R>
MFn pick_multi2(int deg, bool key32) {
  if (deg <= 2) return key32 ? k_hv_multi<N, 2, R, uint32_t> : k_hv_multi<N, 2, R, uint64_t>;
  return key32 ? k_hv_multi<N, 4, R, uint32_t> : k_hv_multi<N, 4, R, uint64_t>;
}

template <int N>
MFn pick_multi(int prec, int deg, bool key32) {
  return prec == QS3X_FP64 ? pick_multi2<N, double>(deg, key32) : pick_multi2<N, float>(deg, key32);
}

MFn multi_kernel(int N, int prec, int deg, bool key32) {
  if (deg > 4) return nullptr;
  switch (N) {
    case 1: return pick_multi<1>(prec, deg, key32);
    case 2: return pick_multi<2>(prec, deg, key32);
    case 3: return pick_multi<3>(prec, deg, key32);
    case 4: return pick_multi<4>(prec, deg, key32);
    case 5: return pick_multi<5>(prec, deg, key32);
    case 6: return pick_multi<6>(prec, deg, key32);
    case 7: return pick_multi<7>(prec, deg, key32);
    case 8: return pick_multi<8>(prec, deg, key32);
    default: return nullptr;
  }
}

// Largest hop cache of a group solve (all N_down groups together).
constexpr size_t kHopCacheMax = (size_t)32 << 30;

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

int qs3x_lanczos_group(qs3x_sector *const *secs, int32_t nsec, int32_t prec, int64_t maxitr, int64_t minitr,
                       int64_t itrint, double conv, uint64_t seed, int32_t use_graph, double *ene, int64_t *iters,
                       double *stats) {
  return guarded5([&] {
    QS3X_REQUIRE(secs && ene && nsec >= 1, "qs3x_lanczos_group: bad argument");
    QS3X_REQUIRE(prec == QS3X_FP64 || prec == QS3X_FP32, "qs3x_lanczos_group: bad precision");
    QS3X_REQUIRE(maxitr >= 1 && itrint >= 1, "qs3x_lanczos_group: maxitr and itrint must be >= 1");
    const double t_start = now_s();
    qs3x_model *m = secs[0]->model;
    const int S = nsec;
    // order: by N_down (one H.v launch per N_down), caller order within
    std::vector<int> order(S);
    for (int i = 0; i < S; ++i) {
      order[i] = i;
      qs3x_sector *s = secs[i];
      QS3X_REQUIRE(s && s->model == m, "qs3x_lanczos_group: sectors of different models");
      QS3X_REQUIRE(!is_u1(s), "qs3x_lanczos_group: not available for U(1)-only sectors");
      QS3X_REQUIRE(s->chain && s->mode == kIndexAnchored && s->d_rclass.p && s->d_bitmap.p && !s->all_k,
                   "qs3x_lanczos_group: chain sectors built on the device only");
      QS3X_REQUIRE(s->nod >= 1 && s->nod <= 8, "qs3x_lanczos_group: N_down must be 1..8");
      QS3X_REQUIRE(s->dim >= 1, "qs3x_lanczos_group: empty sector");
    }
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return secs[x]->nod < secs[y]->nod; });
    hipStream_t st = stream();
    const size_t es = prec == QS3X_FP64 ? 16 : 8;
    std::vector<SecDesc> desc(S);
    std::vector<int64_t> tile0(S + 1, 0);
    int64_t rows = 0;
    for (int j = 0; j < S; ++j) {
      qs3x_sector *s = secs[order[j]];
      SecDesc &d = desc[j];
      d.cfg = s->d_cfg.p;
      d.bitmap = s->d_bitmap.p;
      d.explist = s->d_explist.p;
      d.dim = s->dim;
      d.voff = rows;
      d.tile0 = tile0[j];
      for (int c = 0; c < kMaxCls; ++c) d.rcl[c] = c < (int)s->h_rclass.size() ? s->h_rclass[c] : 0.0;
      const int64_t nt = (s->dim + kMT - 1) / kMT;
      tile0[j + 1] = tile0[j] + nt;
      rows += nt * kMT;
    }
    const int64_t ntiles = tile0[S];
    // the steps a sector can take: min(maxitr, dim)
    std::vector<int64_t> cap(S);
    int64_t mitr = 1;
    for (int j = 0; j < S; ++j) {
      // dim steps (the reference's bound) in FP64; FP32 may run to 2 dim + 8 (see qs3x_lanczos)
      const int64_t dj = secs[order[j]]->dim;
      cap[j] = std::min<int64_t>(maxitr, prec == QS3X_FP32 ? 2 * dj + 8 : dj);
      mitr = std::max(mitr, cap[j]);
    }
    DevBuf<SecDesc> d_desc(S);
    d_desc.upload(desc.data(), S);
    DevBuf<int64_t> d_tile0(S + 1);
    d_tile0.upload(tile0.data(), S + 1);
    std::vector<int> act(S, 1);
    DevBuf<int> d_act(S), d_step(1);
    d_act.upload(act.data(), S);
    DevBuf<unsigned char> U[3];
    for (auto &u : U) {
      u.alloc(es * (size_t)rows);
    }
    // slots: arec [mitr] | aout [mitr] | beta [mitr] | scale [mitr + 1] | tmp
    const Slots sl{0, (int)mitr, (int)(2 * mitr), (int)(3 * mitr), (int)(4 * mitr + 1)};
    DevBuf<double> sc((size_t)(4 * mitr + 3) * S);
    DevBuf<double> part((size_t)2 * std::max<int64_t>(ntiles, 1));
    // start vectors (counter-based per sector, keyed by the sector's own rows)
    for (int j = 0; j < S; ++j) vec_random(prec, U[0].p + es * desc[j].voff, desc[j].dim, seed, st);
    const int gU = (int)std::min<int64_t>(ntiles, (int64_t)cus() * 8);
    if (prec == QS3X_FP64)
      hipLaunchKernelGGL(k_norm_multi<double>, dim3(gU), dim3(kMT), 0, st, U[0].p, d_desc.p, d_tile0.p, S, ntiles,
                         part.p);
    else
      hipLaunchKernelGGL(k_norm_multi<float>, dim3(gU), dim3(kMT), 0, st, U[0].p, d_desc.p, d_tile0.p, S, ntiles,
                         part.p);
    hipLaunchKernelGGL(k_finalize_multi_sec, dim3(S), dim3(kMT), 0, st, part.p, d_tile0.p, nullptr, S, sc.p, 2,
                       sl.tmp, sl.scale, 0, 0.0, nullptr);
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, st, d_step.p, 0);
    QS3X_HIP(hipGetLastError());
    // the N_down groups
    struct Grp {
      int N, s0, s1;
      MFn f;
      size_t lds;
      int grid;
      int32_t *rid;
      uint32_t *rcode;
      int64_t rrows;
    };
    std::vector<Grp> grps;
    int kb = 1;
    while ((1 << kb) <= m->NOS) ++kb;
    for (int j = 0; j < S;) {
      const int N = secs[order[j]]->nod;
      int k = j;
      while (k < S && secs[order[k]]->nod == N) ++k;
      Grp g{N, j, k, nullptr, 0, 1, nullptr, nullptr, 0};
      g.f = multi_kernel(N, prec, m->max_deg, (int64_t)(N - 1) * kb <= 32);
      QS3X_REQUIRE(g.f, "qs3x_lanczos_group: no kernel for this N_down / bond degree");
      const size_t rs = prec == QS3X_FP64 ? 8 : 4;
      const size_t nb = N > 1 ? (size_t)(N - 1) * (m->NOS - 1) : 0;
      g.lds = ((nb * 4 + 15) & ~(size_t)15) + 2 * rs * m->bt_jxy2.size() + 4 * (size_t)(m->NOS + 2) + 4 * m->adj.size();
      // up to what a CU holds (160 KB on gfx950, less a margin); beyond 64 KB
      // the kernel must be told before the occupancy query and the launches
      QS3X_REQUIRE(g.lds <= 156 * 1024, "qs3x_lanczos_group: LDS tables too large");
      if (g.lds > 64 * 1024)
        QS3X_HIP(hipFuncSetAttribute((const void *)g.f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)g.lds));
      int nbpc = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbpc, g.f, kMT, g.lds) != hipSuccess || nbpc < 1) nbpc = 1;
      g.grid = (int)std::max<int64_t>(1, std::min<int64_t>(tile0[k] - tile0[j], (int64_t)cus() * nbpc));
      g.rrows = (tile0[k] - tile0[j]) * kMT;
      grps.push_back(g);
      j = k;
    }
    // Option group_hop_cache: every step after the first reads the hops the
    // first one found (8 bytes per hop slot and row) instead of searching
    // for representatives again, when all groups' tables fit in the cap.
    std::vector<DevBuf<int32_t>> rids;
    std::vector<DevBuf<uint32_t>> rcodes;
    size_t hop_cache_bytes = 0;
    {
      const int deg_t = m->max_deg <= 2 ? 2 : 4;  // the kernels' hop slots per spin
      size_t cache = 0;
      for (const Grp &g : grps) cache += (size_t)8 * g.N * deg_t * (size_t)g.rrows;
      // the code's fields: ell < 4096, bond type < 32768
      const bool fits = m->NOS < 4096 && m->bt_jxy2.size() < 32768;
      if (options().group_hop_cache && fits && cache <= kHopCacheMax) {
        for (Grp &g : grps) {
          const size_t n = (size_t)g.N * deg_t * (size_t)g.rrows;
          rids.emplace_back(n);
          rcodes.emplace_back(n);
          g.rid = rids.back().p;
          g.rcode = rcodes.back().p;
        }
        hop_cache_bytes = 0;
        for (const auto &r : rids) hop_cache_bytes += 8 * r.n;
      }
    }
    const double shift = prec == QS3X_FP32 ? m->ezz : 0.0;
    // one Lanczos step (slots from the device step counter)
    auto enqueue_step = [&](int cur, int prv, int nxt) {
      for (const Grp &g : grps) {
        MultiArgs a{};
        a.sec = d_desc.p;
        a.tile0 = d_tile0.p;
        a.active = d_act.p;
        a.s0 = g.s0;
        a.s1 = g.s1;
        a.t0 = tile0[g.s0];
        a.t1 = tile0[g.s1];
        a.L = m->NOS;
        a.bits = secs[order[g.s0]]->bits;
        a.kbits = kb;
        a.kdim = secs[order[g.s0]]->nbinom_k;
        a.binom = secs[order[g.s0]]->d_binom.p;
        a.adj_ptr = m->d_adj_ptr.p;
        a.adj = m->d_adj.p;
        a.nadj = (int)m->adj.size();
        a.bt = m->d_bt.p;
        a.nbt = (int)m->bt_jxy2.size();
        a.shift = prec == QS3X_FP64 ? m->ezz : 0.0;
        a.v1 = U[cur].p;
        a.v0 = U[nxt].p;
        a.sc = sc.p;
        a.S = S;
        a.scale_base = sl.scale;
        a.step = d_step.p;
        a.part = part.p;
        a.rec_id = g.rid;
        a.rec_code = g.rcode;
        a.rec_rows = g.rrows;
        hipLaunchKernelGGL(g.f, dim3(g.grid), dim3(kMT), g.lds, st, a);
      }
      hipLaunchKernelGGL(k_finalize_multi_sec, dim3(S), dim3(kMT), 0, st, part.p, d_tile0.p, d_act.p, S, sc.p, 1,
                         sl.a_rec, sl.a_out, 0, shift, d_step.p);
      if (prec == QS3X_FP64)
        hipLaunchKernelGGL(k_update3_multi<double>, dim3(gU), dim3(kMT), 0, st, U[nxt].p, U[cur].p, U[prv].p,
                           d_desc.p, d_tile0.p, d_act.p, S, ntiles, sc.p, sl, d_step.p, part.p);
      else
        hipLaunchKernelGGL(k_update3_multi<float>, dim3(gU), dim3(kMT), 0, st, U[nxt].p, U[cur].p, U[prv].p,
                           d_desc.p, d_tile0.p, d_act.p, S, ntiles, sc.p, sl, d_step.p, part.p);
      hipLaunchKernelGGL(k_finalize_multi_sec, dim3(S), dim3(kMT), 0, st, part.p, d_tile0.p, d_act.p, S, sc.p, 2,
                         sl.beta, sl.scale, 1, 0.0, d_step.p);
      hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, st, d_step.p);
    };
    // graphs for the three rotations of the vector slots
    hipGraphExec_t gx[3] = {nullptr, nullptr, nullptr};
    hipGraph_t gg[3] = {nullptr, nullptr, nullptr};
    if (use_graph) {
      for (int r = 0; r < 3; ++r) {
        QS3X_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        enqueue_step(r, (r + 2) % 3, (r + 1) % 3);
        QS3X_HIP(hipStreamEndCapture(st, &gg[r]));
        QS3X_HIP(hipGraphInstantiate(&gx[r], gg[r], nullptr, nullptr, 0));
      }
    }
    QS3X_HIP(hipStreamSynchronize(st));
    const double t_setup = now_s() - t_start;
    // host side: per-sector alpha/beta, convergence (lanczos.f90:63-120)
    std::vector<std::vector<double>> ha(S), hb(S);
    std::vector<double> E(S, 0.0), Eprev(S, 0.0);
    std::vector<int64_t> ite(S, 0);
    std::vector<double> buf;
    int64_t done_steps = 0;
    int nactive = S;
    int64_t launches = 0;
    while (nactive > 0 && done_steps < mitr) {
      const int64_t nrun = std::min<int64_t>(itrint, mitr - done_steps);
      for (int64_t q = 0; q < nrun; ++q) {
        const int i = (int)(done_steps + q);
        if (use_graph) QS3X_HIP(hipGraphLaunch(gx[i % 3], st));
        else enqueue_step(i % 3, (i + 2) % 3, (i + 1) % 3);
        ++launches;
      }
      QS3X_HIP(hipGetLastError());
      const int64_t first = done_steps;
      done_steps += nrun;
      // alpha_out and beta rows of the new steps
      buf.assign(sc.n, 0.0);
      sc.download(buf.data(), sc.n);
      bool changed = false;
      for (int j = 0; j < S; ++j) {
        if (!act[j]) continue;
        for (int64_t i = first; i < done_steps && (int64_t)ha[j].size() < cap[j]; ++i) {
          ha[j].push_back(buf[(size_t)(sl.a_out + i) * S + j]);
          hb[j].push_back(buf[(size_t)(sl.beta + i) * S + j]);
        }
        ite[j] = (int64_t)ha[j].size();
        const int64_t it = ite[j];
        bool stop = it >= cap[j];
        if (it == 1 && cap[j] == 1) {
          E[j] = ha[j][0];
        } else if (it >= 1) {
          E[j] = tridiag_lowest(ha[j].data(), hb[j].data(), (int)it);
          const bool breakdown = !(hb[j][it - 1] > 1e-13 * std::fabs(E[j]) + 1e-300);
          if (it >= minitr && (std::fabs(Eprev[j] / E[j] - 1.0) < conv || breakdown)) stop = true;
          Eprev[j] = E[j];
        }
        if (stop) {
          act[j] = 0;
          --nactive;
          changed = true;
        }
      }
      if (changed) d_act.upload(act.data(), S);
    }
    QS3X_HIP(hipStreamSynchronize(st));
    const double t_total = now_s() - t_start;
    for (int r = 0; r < 3; ++r) {
      if (gx[r]) (void)hipGraphExecDestroy(gx[r]);
      if (gg[r]) (void)hipGraphDestroy(gg[r]);
    }
    for (int j = 0; j < S; ++j) {
      ene[order[j]] = E[j];
      if (iters) iters[order[j]] = ite[j];
    }
    if (stats) {
      stats[0] = t_total;
      stats[1] = t_setup;
      stats[2] = (double)done_steps;
      stats[3] = (double)launches;
      stats[4] = (double)rows;
      stats[5] = (double)grps.size();
      stats[6] = (double)hop_cache_bytes;  // bytes of the hop cache (0: none)
      stats[7] = 0.0;
    }
  });
}

}  // extern "C"
