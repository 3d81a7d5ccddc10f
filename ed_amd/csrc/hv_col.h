// hv_col.h -- the column form of the gap-lattice stencil H.v (nearest-
// neighbour chains with uniform bonds, N_down = 4; the lattice, line table
// and task schedule are those of the slab-window form in hv_nn.hip, with
// tiles 64 P - 2 values of n2 wide).  Instantiated in hv_col.hip (P = 1, one
// window position per lane) and hv_col2.hip (P = 2).
#pragma once
#include <hip/hip_runtime.h>

#include "blas.h"
#include "hv_common.h"
#include "nt.h"
#include "qs3x_internal.h"

namespace qs3x {
namespace hvk {

// Buffer access to one window line: a descriptor whose base is the line's
// first row (n2 = e.y) and whose range is the line's rows, so a position
// outside the line (below e.y the unsigned offset wraps) reads 0 and drops its
// write in the hardware's range check.  Every lane then issues every load and
// store: the number of vector-memory operations between a window's loads and
// the wait for them is the same on every path, so the wait (vmcnt, which
// counts loads and stores in issue order) leaves the next window's loads in
// flight.  Offsets are 32-bit byte offsets within a line.
constexpr unsigned kBufFlags = 0x00020000u;  // raw buffer, 32-bit dwords (gfx9 descriptor word 3)
constexpr unsigned kOffNone = 0x80000000u;   // an offset beyond every line
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
template <class C>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t line_rsrc(const C *base, int row, int nrows) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)(base + row), 0, nrows > 0 ? nrows * (int)sizeof(C) : 0, kBufFlags);
}
__device__ __forceinline__ float2 buf_ld(__amdgpu_buffer_rsrc_t r, unsigned off, float2) {
  const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
  return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
}
__device__ __forceinline__ double2 buf_ld(__amdgpu_buffer_rsrc_t r, unsigned off, double2) {
  const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
  return make_double2(__hiloint2double((int)v.y, (int)v.x), __hiloint2double((int)v.w, (int)v.z));
}
// nontemporal (slc) stores
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, unsigned off, float2 x) {
  u32x2 v;
  v.x = __float_as_uint(x.x);
  v.y = __float_as_uint(x.y);
  __builtin_amdgcn_raw_buffer_store_b64(v, r, off, 0, 2);
}
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, unsigned off, double2 x) {
  u32x4 v;
  v.x = (unsigned)__double2loint(x.x);
  v.y = (unsigned)__double2hiint(x.x);
  v.z = (unsigned)__double2loint(x.y);
  v.w = (unsigned)__double2hiint(x.y);
  __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, 2);
}

// Column form of the gap-lattice stencil (same lattice, tile schedule and
// LDS ring as hv_nn.hip's k_hv_stencil4, and every sum formed in the same
// order, so the two agree bit for bit).  A thread keeps P window positions
// of one line -- lane + 64 q, q < P -- for the whole sweep of a task:
// window lines j = wid + 4 f (f < kF), of which 1 .. B are computed; with
// B + 2 a multiple of 4 (B = 2, 6, 10, 14) no wavefront has an empty line
// slot.  A tile is 64 P - 2 values of n2 by B lines, so a window line is one
// contiguous span of 64 P vector elements.  What the slab-window kernel
// recomputes per row -- the row's place in the tile (a division by 62), its
// line entry, the clear test's gaps and three ring addresses -- is fixed per
// thread here:
//   - the own value and its neighbours in c (slabs c - 1, c + 1) are the
//     thread's own window points, kept in registers xr[slot][f][q] as they
//     are stored to LDS; only the six hops within a slab or to the anchor
//     neighbours read LDS, at constant offsets from one address;
//   - the slab loop is unrolled by the ring length, so every ring slot and
//     register set has a compile-time index;
//   - a line entry is uniform over the wavefront (scalar registers), and so
//     are the line's buffer descriptors; the clear test is max(g1, g2) <=
//     W - 3 per position and g3 <= W - 3 per line;
//   - FUSED: the own v' of a slab is the w fetched for the window, kept
//     (wv[slot][f][q]) until the slab is computed, instead of a second load;
//   - every global access is a buffer access (above) issued by every lane.
// Window positions 0 and 64 P - 1 (a0 - 1 and a0 + 64 P - 2) are the halo
// and compute nothing.  Window points outside a line's range of a hold 0 (the
// absent hops of the file comment); rows that are not clear are left to the
// slow-row kernel.
template <class R, bool REALPH, bool FUSED, int BT, int D, int P>
__global__ __launch_bounds__(kThreads) void k_hv_stencil_col(HvArgs a, StencilArgs sa) {
  static_assert(D == 2 || D == 4, "windows in flight: a divisor of the ring length");
  static_assert(P == 1 || P == 2, "positions per lane");
  using C = typename CT<R>::T;
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int B = BT, NL = B + 2;
  constexpr int SW = kSegW * P;     // window positions per line
  constexpr int SA = SW - 2;        // n2 values per tile
  constexpr int kF = (NL + 3) / 4;  // window lines per wavefront: j = wid + 4 f
  constexpr unsigned ES = sizeof(C);
  C *s_ring = (C *)smem;
  int4 *s_tl = (int4 *)(smem + (size_t)kRing * NL * SW * sizeof(C));
  const C *v1 = (const C *)a.v1;
  const C *vp = (const C *)a.vprev;
  const C *v0 = (const C *)a.v0;
  const C *xo = (const C *)a.xout;
  FusedCoef<R> fc{(R)1, (R)0, (R)0, false};
  if constexpr (FUSED) fc = fused_coef<R>(a);
  const R scale = (!FUSED && a.scale) ? (R)(*a.scale) : (R)1;
  const R jxy2 = (R)a.bt[0], jz4 = (R)a.bt[1];
  const int L = a.L;
  const int ei_ar = (L - 1) + 3 * (L + 1), ei_al = 1 + 3 * (L + 1);
  R ph_ar_r = (R)1, ph_ar_i = (R)0, ph_al_r = (R)1, ph_al_i = (R)0;
  if constexpr (!REALPH) {
    ph_ar_r = (R)a.explist[2 * ei_ar];
    ph_ar_i = (R)a.explist[2 * ei_ar + 1];
    ph_al_r = (R)a.explist[2 * ei_al];
    ph_al_i = (R)a.explist[2 * ei_al + 1];
  }
  // wid in a scalar register: the line of a lane's window points, and with it
  // every buffer descriptor, is uniform over the wavefront
  const int lane = threadIdx.x & 63, wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  double dre = 0.0, dim_ = 0.0;
  double acc4[4] = {0.0, 0.0, 0.0, 0.0};  // FUSED: Re<v,Hv>, |P|^2, Re<v,P>, |v|^2
  const int W1 = L + 1;
  // registers: window sets in flight, own window points per ring slot
  C fp[D][kF][P], fw[D][kF][P];
  C xr[kRing][kF][P], wv[kRing][kF][P];
  // line entry (first row, lowest n2, highest n2 + 1) of window line j of
  // slab s (s_rel = s - c0 + 1), uniform over the wavefront
  auto entry = [&](int s_rel, int j) __attribute__((always_inline)) -> int4 {
    const int4 e = s_tl[s_rel * NL + j];
    return make_int4(__builtin_amdgcn_readfirstlane(e.x), __builtin_amdgcn_readfirstlane(e.y),
                     __builtin_amdgcn_readfirstlane(e.z), 0);
  };
  const bool hasv = FUSED && fc.hasv;
  // slab s into a set (lines of slabs past c1, live false, and line slots
  // past the window read as empty)
  auto fetch = [&](C (&p)[kF][P], C (&w)[kF][P], int a0, int s_rel, bool live = true)
                   __attribute__((always_inline)) {
#pragma unroll
    for (int f = 0; f < kF; ++f) {
      const int j = wid + 4 * f;
      const int4 e = entry(s_rel, j < NL ? j : 0);
      const int n = (live && j < NL) ? e.z - e.y : 0;
      const __amdgpu_buffer_rsrc_t r1 = line_rsrc(v1, e.x, n);
      const __amdgpu_buffer_rsrc_t rw = line_rsrc(hasv ? vp : v1, e.x, hasv ? n : 0);
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const unsigned off = (unsigned)(a0 - 1 + lane + kSegW * q - e.y) * ES;
        p[f][q] = buf_ld(r1, off, C{});
        if constexpr (FUSED) w[f][q] = buf_ld(rw, off, C{});
      }
    }
  };
  // set into ring slot sl (and the own registers of that slot)
  auto store = [&](const C (&p)[kF][P], const C (&w)[kF][P], int sl) __attribute__((always_inline)) {
#pragma unroll
    for (int f = 0; f < kF; ++f) {
      const int j = wid + 4 * f;
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const C v = FUSED ? fused_v<R>(fc, p[f][q], w[f][q]) : p[f][q];
        if (j < NL) s_ring[((size_t)sl * NL + j) * SW + lane + kSegW * q] = v;
        xr[sl][f][q] = v;
        if constexpr (FUSED) wv[sl][f][q] = w[f][q];
      }
    }
  };
  // slab c (s_rel = c - c0 + 1) from slots sm (c - 1), s0 (c), sp (c + 1);
  // live false: slabs past c1, nothing written
  auto compute = [&](int a0, int b0, int c, int s_rel, int sm, int s0, int sp, bool live)
                     __attribute__((always_inline)) {
    const C *r0 = s_ring + (size_t)s0 * NL * SW, *rm = s_ring + (size_t)sm * NL * SW,
            *rp = s_ring + (size_t)sp * NL * SW;
    const int w3 = L + 1 - c - 3;
#pragma unroll
    for (int m = 0; m < kF; ++m) {
      const int j = wid + 4 * m;
      const bool jin = live && j >= 1 && j <= B;  // uniform
      const int4 e = entry(jin ? s_rel : 1, jin ? j : 1);
      const int n = jin ? e.z - e.y : 0;
      const __amdgpu_buffer_rsrc_t ro = line_rsrc(v0, e.x, n);
      const __amdgpu_buffer_rsrc_t rx = line_rsrc(FUSED ? xo : v0, e.x, FUSED ? n : 0);
      const int b = b0 - 1 + j;
      const int g3 = c - b;
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int pos = lane + kSegW * q;
        const int av = a0 - 1 + pos;
        const int g1 = av - 1, g2 = b - av;
        const bool ok = jin && pos >= 1 && pos <= SA && av >= e.y && av < e.z && max(g1, g2) <= w3 && g3 <= w3;
        C out, x, pn;
        out.x = out.y = x.x = x.y = pn.x = pn.y = (R)0;
        if (jin) {  // uniform: the ring reads and arithmetic, no global access
          const int pc = min(max(pos, 1), SA);  // halo lanes read inside the window
          const int x0 = min(max(j, 1), B) * SW + pc;
          x = xr[s0][m][q];
          const C yc1 = xr[sp][m][q];     // (a, b, c + 1)
          const C yal = rp[x0 + SW + 1];  // anchor to site L: (a + 1, b + 1, c + 1)
          const C ya1 = r0[x0 + 1];       // (a + 1, b, c)
          const C yb0 = r0[x0 - SW];      // (a, b - 1, c)
          const C ya0 = r0[x0 - 1];       // (a - 1, b, c)
          const C yar = rm[x0 - SW - 1];  // anchor to site 2: (a - 1, b - 1, c - 1)
          const C yb1 = r0[x0 + SW];      // (a, b + 1, c)
          const C yc0 = xr[sm][m][q];     // (a, b, c - 1)
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
          out.x = dg * x.x + jxy2 * scale * sx;
          out.y = dg * x.y + jxy2 * scale * sy;
          if constexpr (FUSED) {
            const C vo = wv[s0][m][q];
            pn.x = out.x - fc.c3 * vo.x;
            pn.y = out.y - fc.c3 * vo.y;
          }
          if (ok) {
            if constexpr (FUSED) {
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
        const unsigned off = ok ? (unsigned)(av - e.y) * ES : kOffNone;
        if constexpr (FUSED) {
          buf_st(rx, off, x);
          buf_st(ro, off, pn);
        } else {
          buf_st(ro, off, out);
        }
      }
    }
  };

  for (int t = blockIdx.x; t < sa.ntasks; t += gridDim.x) {
    const int4 task = sa.tasks[t];
    const int a0 = task.x, b0 = task.y, c0 = task.z, c1 = task.w & 0xffff;
    if (c1 <= c0) continue;
    // (the previous task's last slab ended with a barrier)
    // line entries of slabs c0 - 1 .. c1 (entry (cs, j) at s_tl[(cs - c0 + 1) NL + j])
    const int nent = (c1 - c0 + 2) * NL;
    for (int q = threadIdx.x; q < nent; q += kThreads) {
      const int cs = c0 - 1 + q / NL, b = b0 - 1 + q % NL;
      int4 e = make_int4(0, 0, 0, 0);
      if (cs >= 0 && cs <= L && b >= 0 && b <= L) e = sa.lines[(size_t)cs * W1 + b];
      s_tl[q] = e;
    }
    __syncthreads();
    // ring slot of slab s: (s - c0 + 1) mod 4; prologue: slabs c0 - 1, c0,
    // c0 + 1 into slots 0, 1, 2
    {
      C zp[kF][P], zw[kF][P];
      fetch(fp[0], fw[0], a0, 0);
      fetch(fp[1], fw[1], a0, 1);
      fetch(zp, zw, a0, 2);
      store(fp[0], fw[0], 0);
      store(fp[1], fw[1], 1);
      store(zp, zw, 2);
    }
    // Slabs c0 + 2 .. c0 + D in flight, slab s in set (s - c0) mod D.  Every
    // step fetches (slabs past c1 as empty windows), so the memory operations
    // between a set's fetch and its store are the same on every path.
    const int srel_max = c1 - c0 + 1;
#pragma unroll
    for (int k = 2; k <= D; ++k) fetch(fp[k % D], fw[k % D], a0, min(k + 1, srel_max), c0 + k <= c1);
    __syncthreads();
    // slab c = c0 + 4 k + u: slab c + D + 1 goes in flight in set (u + 1) %
    // D, slab c is computed from slots u, u + 1, u + 2 (mod 4), and slab
    // c + 2 (set (u + 2) % D) is stored into slot u + 3.  All four steps of a
    // group run (past c1 they fetch and store empty windows and write nothing).
    for (int cb = c0; cb < c1; cb += kRing) {
#pragma unroll
      for (int u = 0; u < kRing; ++u) {
        const int c = cb + u;
        const int s_rel = c - c0 + 1;
        fetch(fp[(u + 1) % D], fw[(u + 1) % D], a0, min(s_rel + D + 1, srel_max), c + D + 1 <= c1);
        compute(a0, b0, c, min(s_rel, srel_max - 1), u % kRing, (u + 1) % kRing, (u + 2) % kRing, c < c1);
        store(fp[(u + 2) % D], fw[(u + 2) % D], (u + 3) % kRing);
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

}  // namespace hvk
}  // namespace qs3x
