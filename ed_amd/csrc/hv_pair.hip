// hv_pair.hip -- the pair form of the gap-lattice stencil H.v in FP32
// (nearest-neighbour chains with uniform bonds, N_down = 4; the lattice, line
// table and task schedule of hv_nn.hip, the thread-per-column sweep of
// hv_col.h).  Each lane holds two adjacent window positions, so every vector
// load and store moves 16 bytes per lane: 8-byte lanes cap a streaming pass
// near 4.8 TB/s on MI355X, 16-byte lanes near 5.8 TB/s on the same pattern.
//
// Window line: 128 positions, lane l holding positions 2 l and 2 l + 1;
// position p stands for n2 = a0 - 2 + p, so lane 0 holds the left halo (p = 1;
// p = 0 is unused), lane 63 the right halo (p = 126; p = 127 unused), and
// lanes 1 .. 62 the tile's 124 values a0 .. a0 + 123.  A row's hops read:
//   own pair, slabs c - 1, c, c + 1: registers (yc0, x, yc1; the other
//     element of the pair is one of ya0 / ya1);
//   lines j - 1, j + 1 of slab c: one 16-byte LDS read each (yb0, yb1 of
//     both rows);
//   the anchor hops (a - 1, b - 1, c - 1) and (a + 1, b + 1, c + 1): one
//     16-byte LDS read each, at positions 2 l - 1 and 2 l + 1;
//   the neighbouring lanes' elements 2 l - 1 and 2 l + 2: one 8-byte read each.
// Every sum is formed in the order of the other stencil kernels, so the three
// forms agree bit for bit.
//
// Global access goes through buffer descriptors over the whole vector
// (byte offsets below 2 GB: dim < 2^28 rows); an element outside its line is
// loaded from the neighbouring row and set to 0, and one outside the vector
// reads 0 in the hardware's range check.  A pair is stored with one 16-byte
// store when both of its rows belong to this tile and line, otherwise its one
// such row with an 8-byte store.  The 16-byte stores are issued by every lane
// (with an offset past the vector where there is nothing to write); the 8-byte
// ones only by wavefronts that hold a line's first or last row.
// Rows that are not clear are written too and then overwritten by the
// slow-row kernel, which runs after this one on the same stream.
#include "hv_col.h"

namespace qs3x {

using namespace hvk;

namespace {

typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4v ld4(__amdgpu_buffer_rsrc_t r, unsigned off) {
  const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
  f4v o;
  o.x = __uint_as_float(v.x);
  o.y = __uint_as_float(v.y);
  o.z = __uint_as_float(v.z);
  o.w = __uint_as_float(v.w);
  return o;
}
__device__ __forceinline__ void st4(__amdgpu_buffer_rsrc_t r, unsigned off, f4v x) {
  u32x4 v;
  v.x = __float_as_uint(x.x);
  v.y = __float_as_uint(x.y);
  v.z = __float_as_uint(x.z);
  v.w = __float_as_uint(x.w);
  __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, 2);
}

constexpr int kPW = 128;  // window positions per line
constexpr int kPA = 124;  // n2 values per tile

// byte offset of row r in a vector of complex floats, or past every vector
__device__ __forceinline__ unsigned row_off(int r, bool valid) {
  return (valid && r >= 0) ? (unsigned)r * 8u : kOffNone;
}

// D: slab windows in flight (2 or 4, a divisor of the ring length).
template <bool REALPH, bool FUSED, int BT, int D>
__global__ __launch_bounds__(kThreads) void k_hv_stencil_pair(HvArgs a, StencilArgs sa) {
  static_assert(D == 2 || D == 4, "windows in flight");
  using C = float2;
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int B = BT, NL = B + 2;
  constexpr int kF = (NL + 3) / 4;  // window lines per wavefront: j = wid + 4 f
  C *s_ring = (C *)smem;
  int4 *s_tl = (int4 *)(smem + (size_t)kRing * NL * kPW * sizeof(C));
  const int dim = (int)a.dim;  // < 2^28 (the launcher's condition)
  // loads: rows 0 .. dim - 2 (row dim - 1 separately, see fetch)
  const int nrec_ld = (int)((dim - 1) * 8);
  const int nrec_st = (int)(dim * 8);
  const __amdgpu_buffer_rsrc_t r1 = __builtin_amdgcn_make_buffer_rsrc((void *)a.v1, 0, nrec_ld, kBufFlags);
  FusedCoef<float> fc{1.f, 0.f, 0.f, false};
  if constexpr (FUSED) fc = fused_coef<float>(a);
  const bool hasv = FUSED && fc.hasv;
  const __amdgpu_buffer_rsrc_t rw =
      __builtin_amdgcn_make_buffer_rsrc((void *)(hasv ? a.vprev : a.v1), 0, hasv ? nrec_ld : 0, kBufFlags);
  const __amdgpu_buffer_rsrc_t ro = __builtin_amdgcn_make_buffer_rsrc(a.v0, 0, nrec_st, kBufFlags);
  const __amdgpu_buffer_rsrc_t rx =
      __builtin_amdgcn_make_buffer_rsrc(FUSED ? a.xout : a.v0, 0, FUSED ? nrec_st : 0, kBufFlags);
  const C *v1 = (const C *)a.v1;
  const C *vp = (const C *)a.vprev;
  const float scale = (!FUSED && a.scale) ? (float)(*a.scale) : 1.f;
  const float jxy2 = (float)a.bt[0], jz4 = (float)a.bt[1];
  const int L = a.L;
  const int ei_ar = (L - 1) + 3 * (L + 1), ei_al = 1 + 3 * (L + 1);
  float ph_ar_r = 1.f, ph_ar_i = 0.f, ph_al_r = 1.f, ph_al_i = 0.f;
  if constexpr (!REALPH) {
    ph_ar_r = (float)a.explist[2 * ei_ar];
    ph_ar_i = (float)a.explist[2 * ei_ar + 1];
    ph_al_r = (float)a.explist[2 * ei_al];
    ph_al_i = (float)a.explist[2 * ei_al + 1];
  }
  const int lane = threadIdx.x & 63, wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  double dre = 0.0, dim_ = 0.0;
  double acc4[4] = {0.0, 0.0, 0.0, 0.0};  // FUSED: Re<v,Hv>, |P|^2, Re<v,P>, |v|^2
  const int W1 = L + 1;
  f4v fp[D][kF], fw[D][kF];
  f4v xr[kRing][kF], wv[kRing][kF];
  // line entry of window line j of slab s (s_rel = s - c0 + 1); the
  // descriptors span the whole vector, so the entry only enters lane offsets
  // and stays in vector registers (scalar ones would spill)
  auto entry = [&](int s_rel, int j) __attribute__((always_inline)) -> int4 { return s_tl[s_rel * NL + j]; };
  // slab s (s_rel = s - c0 + 1) into a set, with per line the flags that
  // the store applies once the loads have arrived: bit 0 / 1 the pair's first
  // / second element lies in its line; bit 2 that one element is row 0 (bit
  // 4) or row dim - 1, read separately (bit 3: the pair's second element).  Row 0 in the pair
  // (-1, 0) reads 0 in the range check, and the descriptor's range ends before
  // row dim - 1, so that no pair reads past the vector whether the range check
  // takes the access as a whole or each of its dwords.
  int fin[D][kF];
  auto fetch = [&](f4v (&p)[kF], f4v (&w)[kF], int (&fl)[kF], int a0, int s_rel, bool live)
                   __attribute__((always_inline)) {
    const int av0 = a0 - 2 + 2 * lane;  // n2 of the lane's first position
#pragma unroll
    for (int f = 0; f < kF; ++f) {
      const int j = wid + 4 * f;
      const int4 e = entry(s_rel, j < NL ? j : 0);
      const bool lv = live && j < NL;  // uniform
      const bool in0 = lv && av0 >= e.y && av0 < e.z, in1 = lv && av0 + 1 >= e.y && av0 + 1 < e.z;
      const int r0 = e.x - e.y + av0;  // row of the first position
      const bool lo = r0 == -1 && in1, hi0 = r0 == dim - 1 && in0, hi1 = r0 == dim - 2 && in1;
      fl[f] = (in0 ? 1 : 0) | (in1 ? 2 : 0) | ((lo || hi0 || hi1) ? 4 : 0) | ((lo || hi1) ? 8 : 0) | (lo ? 16 : 0);
      const unsigned off = row_off(r0, in0 || in1);
      p[f] = ld4(r1, off);
      if constexpr (FUSED) w[f] = ld4(rw, off);
    }
  };
  auto store = [&](const f4v (&p)[kF], const f4v (&w)[kF], const int (&fl)[kF], int sl) __attribute__((always_inline)) {
#pragma unroll
    for (int f = 0; f < kF; ++f) {
      const int j = wid + 4 * f;
      f4v x = p[f], y = {0.f, 0.f, 0.f, 0.f};
      if constexpr (FUSED) y = w[f];
      if (fl[f] & 4) {  // row 0 or dim - 1: one lane of the sector
        const int rs = (fl[f] & 16) ? 0 : dim - 1;
        const C t = v1[rs];
        const C u = hasv ? vp[rs] : C{0.f, 0.f};
        if (fl[f] & 8) {
          x.z = t.x, x.w = t.y, y.z = u.x, y.w = u.y;
        } else {
          x.x = t.x, x.y = t.y, y.x = u.x, y.y = u.y;
        }
      }
      if (!(fl[f] & 1)) x.x = x.y = y.x = y.y = 0.f;
      if (!(fl[f] & 2)) x.z = x.w = y.z = y.w = 0.f;
      f4v v = x;
      if constexpr (FUSED) {
        const C v0 = fused_v<float>(fc, make_float2(x.x, x.y), make_float2(y.x, y.y));
        const C v1e = fused_v<float>(fc, make_float2(x.z, x.w), make_float2(y.z, y.w));
        v.x = v0.x, v.y = v0.y, v.z = v1e.x, v.w = v1e.y;
      }
      if (j < NL) *(f4v *)(s_ring + ((size_t)sl * NL + j) * kPW + 2 * lane) = v;
      xr[sl][f] = v;
      if constexpr (FUSED) wv[sl][f] = y;
    }
  };
  // one row: element of the pair (hops already gathered); returns H v, P
  auto row = [&](C x, C yc1, C yal, C ya1, C yb0, C ya0, C yar, C yb1, C yc0, int nanti, C vo, C &out, C &pn)
                 __attribute__((always_inline)) {
    float sx = (yc1.x + ya1.x) + (yb0.x + ya0.x) + (yb1.x + yc0.x);
    float sy = (yc1.y + ya1.y) + (yb0.y + ya0.y) + (yb1.y + yc0.y);
    const float alx = yal.x, aly = yal.y, arx = yar.x, ary = yar.y;
    if constexpr (REALPH) {
      sx += arx + alx;
      sy += ary + aly;
    } else {
      sx += ph_ar_r * arx - ph_ar_i * ary + ph_al_r * alx - ph_al_i * aly;
      sy += ph_ar_r * ary + ph_ar_i * arx + ph_al_r * aly + ph_al_i * alx;
    }
    const float dg = 0.f - 2.f * jz4 * (float)nanti;
    out.x = dg * x.x + jxy2 * scale * sx;
    out.y = dg * x.y + jxy2 * scale * sy;
    if constexpr (FUSED) {
      pn.x = out.x - fc.c3 * vo.x;
      pn.y = out.y - fc.c3 * vo.y;
    }
  };
  auto compute = [&](int a0, int b0, int c, int s_rel, int sm, int s0, int sp, bool live)
                     __attribute__((always_inline)) {
    const C *r0p = s_ring + (size_t)s0 * NL * kPW, *rmp = s_ring + (size_t)sm * NL * kPW,
            *rpp = s_ring + (size_t)sp * NL * kPW;
    const int w3 = L + 1 - c - 3;
    const int av0 = a0 - 2 + 2 * lane;
#pragma unroll
    for (int m = 0; m < kF; ++m) {
      const int j = wid + 4 * m;
      const bool jin = live && j >= 1 && j <= B;  // uniform
      const int4 e = entry(jin ? s_rel : 1, jin ? j : 1);
      const int b = b0 - 1 + j;
      const int g3 = c - b;
      const bool tile = lane >= 1 && lane <= 62;  // positions 2 .. 125
      bool wr[2], ok[2];
      int nanti[2];
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int av = av0 + k;
        const int g1 = av - 1, g2 = b - av;
        wr[k] = jin && tile && av >= e.y && av < e.z;
        ok[k] = wr[k] && max(g1, g2) <= w3 && g3 <= w3;
        nanti[k] = 2 + 2 * ((g1 > 1) + (g2 > 1) + (g3 > 1));
      }
      f4v o = {0.f, 0.f, 0.f, 0.f}, xo4 = {0.f, 0.f, 0.f, 0.f};
      if (jin) {  // uniform: LDS reads and arithmetic, no global access
        const int ll = min(max(lane, 1), 62);  // the halo lanes read inside the window
        const int x0 = j * kPW + 2 * ll;       // the pair's first position in a slot
        const f4v xs = xr[s0][m], xc1 = xr[sp][m], xc0 = xr[sm][m];
        const f4v yb0 = *(const f4v *)(r0p + x0 - kPW), yb1 = *(const f4v *)(r0p + x0 + kPW);
        const C yar0 = rmp[x0 - kPW - 1], yar1 = rmp[x0 - kPW];
        const C yal0 = rpp[x0 + kPW + 1], yal1 = rpp[x0 + kPW + 2];
        const C ya0 = r0p[x0 - 1], ya1 = r0p[x0 + 2];
        C x0c = make_float2(xs.x * scale, xs.y * scale), x1c = make_float2(xs.z * scale, xs.w * scale);
        C out0, out1, pn0, pn1;
        pn0 = pn1 = make_float2(0.f, 0.f);
        C vo0 = make_float2(0.f, 0.f), vo1 = make_float2(0.f, 0.f);
        if constexpr (FUSED) {
          vo0 = make_float2(wv[s0][m].x, wv[s0][m].y);
          vo1 = make_float2(wv[s0][m].z, wv[s0][m].w);
        }
        // first row: (a + 1) is the pair's second element; second row: (a - 1) its first
        row(x0c, make_float2(xc1.x, xc1.y), yal0, make_float2(xs.z, xs.w), make_float2(yb0.x, yb0.y), ya0, yar0,
            make_float2(yb1.x, yb1.y), make_float2(xc0.x, xc0.y), nanti[0], vo0, out0, pn0);
        row(x1c, make_float2(xc1.z, xc1.w), yal1, ya1, make_float2(yb0.z, yb0.w), make_float2(xs.x, xs.y), yar1,
            make_float2(yb1.z, yb1.w), make_float2(xc0.z, xc0.w), nanti[1], vo1, out1, pn1);
        if constexpr (FUSED) {
          o.x = pn0.x, o.y = pn0.y, o.z = pn1.x, o.w = pn1.y;
          xo4.x = x0c.x, xo4.y = x0c.y, xo4.z = x1c.x, xo4.w = x1c.y;
        } else {
          o.x = out0.x, o.y = out0.y, o.z = out1.x, o.w = out1.y;
        }
        const C xx[2] = {x0c, x1c}, oo[2] = {out0, out1}, pp[2] = {pn0, pn1};
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          if (!ok[k]) continue;
          const C x = xx[k], out = oo[k];
          if constexpr (FUSED) {
            const C pn = pp[k];
            acc4[0] += (double)x.x * (double)out.x + (double)x.y * (double)out.y;
            acc4[1] += (double)pn.x * (double)pn.x + (double)pn.y * (double)pn.y;
            acc4[2] += (double)x.x * (double)pn.x + (double)x.y * (double)pn.y;
            acc4[3] += (double)x.x * (double)x.x + (double)x.y * (double)x.y;
          } else if (a.dot_partials) {
            dre += (double)x.x * (double)out.x + (double)x.y * (double)out.y;
            dim_ += (double)x.x * (double)out.y - (double)x.y * (double)out.x;
          }
        }
      }
      // 16-byte store where both rows are written (issued by every lane),
      // else one 8-byte store of the single row
      const int rr = e.x - e.y + av0;
      const bool both = wr[0] && wr[1], one = wr[0] != wr[1];
      const unsigned off4 = row_off(rr, both);
      const unsigned off2 = row_off(rr + (wr[1] ? 1 : 0), one);
      const C o2 = wr[1] ? make_float2(o.z, o.w) : make_float2(o.x, o.y);
      st4(ro, off4, o);
      if constexpr (FUSED) st4(rx, off4, xo4);
      if (one) {  // a line's first or last row in the tile: few lanes, most wavefronts none
        buf_st(ro, off2, o2);
        if constexpr (FUSED) {
          const C x2 = wr[1] ? make_float2(xo4.z, xo4.w) : make_float2(xo4.x, xo4.y);
          buf_st(rx, off2, x2);
        }
      }
    }
  };

  for (int t = blockIdx.x; t < sa.ntasks; t += gridDim.x) {
    const int4 task = sa.tasks[t];
    const int a0 = task.x, b0 = task.y, c0 = task.z, c1 = task.w & 0xffff;
    if (c1 <= c0) continue;
    const int nent = (c1 - c0 + 2) * NL;
    for (int q = threadIdx.x; q < nent; q += kThreads) {
      const int cs = c0 - 1 + q / NL, b = b0 - 1 + q % NL;
      int4 e = make_int4(0, 0, 0, 0);
      if (cs >= 0 && cs <= L && b >= 0 && b <= L) e = sa.lines[(size_t)cs * W1 + b];
      s_tl[q] = e;
    }
    __syncthreads();
    {
      f4v zp[kF], zw[kF];
      int zf[kF];
      fetch(fp[0], fw[0], fin[0], a0, 0, true);
      fetch(fp[1], fw[1], fin[1], a0, 1, true);
      fetch(zp, zw, zf, a0, 2, true);
      store(fp[0], fw[0], fin[0], 0);
      store(fp[1], fw[1], fin[1], 1);
      store(zp, zw, zf, 2);
    }
    // slabs c0 + 2 .. c0 + D in flight, slab s in set (s - c0) mod D
    const int srel_max = c1 - c0 + 1;
#pragma unroll
    for (int k = 2; k <= D; ++k) fetch(fp[k % D], fw[k % D], fin[k % D], a0, min(k + 1, srel_max), c0 + k <= c1);
    __syncthreads();
    for (int cb = c0; cb < c1; cb += kRing) {
#pragma unroll
      for (int u = 0; u < kRing; ++u) {
        const int c = cb + u;
        const int s_rel = c - c0 + 1;
        fetch(fp[(u + 1) % D], fw[(u + 1) % D], fin[(u + 1) % D], a0, min(s_rel + D + 1, srel_max), c + D + 1 <= c1);
        compute(a0, b0, c, min(s_rel, srel_max - 1), u % kRing, (u + 1) % kRing, (u + 2) % kRing, c < c1);
        store(fp[(u + 2) % D], fw[(u + 2) % D], fin[(u + 2) % D], (u + 3) % kRing);
        __syncthreads();
      }
    }
  }
  if constexpr (FUSED) {
    if (a.dot_partials) block_reduce_store4(acc4, a.dot_partials);
  } else {
    if (a.dot_partials) block_reduce_store(dre, dim_, a.dot_partials);
  }
}

template <int BT, int D>
SFn pair_b(bool realph, bool fused) {
  if (fused) return realph ? k_hv_stencil_pair<true, true, BT, D> : k_hv_stencil_pair<false, true, BT, D>;
  return realph ? k_hv_stencil_pair<true, false, BT, D> : k_hv_stencil_pair<false, false, BT, D>;
}

}  // namespace

// FP32, lines per tile 2, 4, 6, 8 or 10, two slab windows in flight, or 6,
// 8 or 10 with four (other combinations: null).
SFn stencil_pair_kernel(bool realph, bool fused, int B, int D) {
  if (D == 4) {
    switch (B) {
      case 6: return pair_b<6, 4>(realph, fused);
      case 8: return pair_b<8, 4>(realph, fused);
      case 10: return pair_b<10, 4>(realph, fused);
      default: return nullptr;
    }
  }
  switch (B) {
    case 2: return pair_b<2, 2>(realph, fused);
    case 4: return pair_b<4, 2>(realph, fused);
    case 6: return pair_b<6, 2>(realph, fused);
    case 8: return pair_b<8, 2>(realph, fused);
    case 10: return pair_b<10, 2>(realph, fused);
    default: return nullptr;
  }
}

}  // namespace qs3x
