// rotate.hip -- thick-restart Ritz rotation as an FP64 MFMA tall-skinny GEMM.
//
// Replaces the zgemm + zcopy of eigen_solver.f90:147-149:
//   Q = Psi(:, 1:NOM) C(1:NOM, 1:NOK);  Psi(:, 1:NOK) = Q.
// The projected eigenvectors C are real (see eig.cpp), so a complex basis
// vector of length D is a column of 2D reals and the rotation is a real
// (2D x NOM) x (NOM x NOK) GEMM.  Each wavefront owns a 16-row tile of the
// output and accumulates it with v_mfma_f64_16x16x4f64 over K = NOM in steps
// of 4; the basis scales s_j are folded into C while it is staged in LDS.
// The update is in place: a wavefront reads all NOM columns of its rows before
// it writes the NOK outputs of the same rows, and tiles are disjoint.
//
// Operand layout of v_mfma_f64_16x16x4f64 (64 lanes): A (16 x 4) lane l holds
// A[l % 16][l / 16]; B (4 x 16) lane l holds B[l / 16][l % 16]; D (16 x 16)
// lane l holds D[l / 16 + 4 e][l % 16] in its e-th result register, e = 0..3
// (checked against a host GEMM by tests/test_gpu_solvers.py).
#include <algorithm>
#include <type_traits>

#include "blas.h"
#include "qs3x_internal.h"

namespace qs3x {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kRotThreads = 256;
constexpr int kMaxNT = 4;  // up to 64 output columns per pass

template <class R, int NT>
__global__ __launch_bounds__(kRotThreads) void k_rotate_mfma(R *base, int64_t M /* 2*dim reals */,
                                                             const double *scales, int nom, int nok,
                                                             const double *C /* nom x nok row-major */) {
  extern __shared__ double cs[];  // [nomp][NT*16]
  const int nomp = (nom + 3) & ~3;
  const int ncol = NT * 16;
  for (int q = threadIdx.x; q < nomp * ncol; q += blockDim.x) {
    int k = q / ncol, n = q % ncol;
    cs[q] = (k < nom && n < nok) ? C[(size_t)k * nok + n] * (scales ? scales[k] : 1.0) : 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int waves_per_block = blockDim.x >> 6;
  const int64_t ntiles = (M + 15) / 16;
  const int li = lane & 15, lk = lane >> 4;
  for (int64_t tile = (int64_t)blockIdx.x * waves_per_block + wave; tile < ntiles;
       tile += (int64_t)gridDim.x * waves_per_block) {
    const int64_t row0 = tile * 16;
    d4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = d4{0.0, 0.0, 0.0, 0.0};
    const int64_t r = row0 + li;
    for (int k0 = 0; k0 < nomp; k0 += 4) {
      const int k = k0 + lk;
      double a = 0.0;
      if (k < nom && r < M) a = (double)base[(size_t)k * M + r];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const double b = cs[k * ncol + t * 16 + li];
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int col = t * 16 + li;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t row = row0 + lk + 4 * e;
        if (row < M && col < nok) base[(size_t)col * M + row] = (R)acc[t][e];
      }
    }
  }
}

// Same product with 16-byte loads: lane (li, lk) reads SUB = 16 / sizeof(R)
// consecutive rows row0 + SUB li .. + SUB - 1 of basis vector k + lk, and
// each of the SUB values is the A operand of its own MFMA (the tile's row i =
// li stands for physical row row0 + SUB li + s).  A wavefront covers 16 SUB
// rows and issues the loads of U K-steps (4U basis vectors) before their
// MFMAs, so that U 16-byte loads per lane are in flight.  C (scaled) is staged
// compactly as [nom][nok] (B = 0 outside), which keeps the LDS small enough
// for several blocks per CU.  Needs 16-byte aligned slots (M * sizeof(R) a
// multiple of 16).
template <class R, int NT, int U>
__global__ __launch_bounds__(kRotThreads) void k_rotate_mfma_v(R *base, int64_t M, const double *scales, int nom,
                                                               int nok, const double *C) {
  constexpr int SUB = 16 / (int)sizeof(R);
  using V4 = typename std::conditional<sizeof(R) == 4, float4, double2>::type;
  extern __shared__ double cs[];  // [nom][nok]
  for (int q = threadIdx.x; q < nom * nok; q += blockDim.x) cs[q] = C[q] * (scales ? scales[q / nok] : 1.0);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int waves_per_block = blockDim.x >> 6;
  const int64_t rows_per_tile = 16 * SUB;
  const int64_t ntiles = (M + rows_per_tile - 1) / rows_per_tile;
  const int li = lane & 15, lk = lane >> 4;
  for (int64_t tile = (int64_t)blockIdx.x * waves_per_block + wave; tile < ntiles;
       tile += (int64_t)gridDim.x * waves_per_block) {
    const int64_t row0 = tile * rows_per_tile;
    const int64_t r = row0 + SUB * li;
    const bool rin = r < M;
    d4 acc[SUB][NT];
#pragma unroll
    for (int q = 0; q < SUB; ++q)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[q][t] = d4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < nom; k0 += 4 * U) {
      V4 a[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 + 4 * u + lk;
        if (k < nom && rin) a[u] = *(const V4 *)(base + (size_t)k * M + r);
        else a[u] = V4{};
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 + 4 * u + lk;
        double av[SUB];
        if constexpr (SUB == 4) {
          av[0] = a[u].x;
          av[1] = a[u].y;
          av[2] = a[u].z;
          av[3] = a[u].w;
        } else {
          av[0] = a[u].x;
          av[1] = a[u].y;
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const int col = t * 16 + li;
          const double b = (k < nom && col < nok) ? cs[k * nok + col] : 0.0;
#pragma unroll
          for (int q = 0; q < SUB; ++q)
            acc[q][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], b, acc[q][t], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int col = t * 16 + li;
#pragma unroll
      for (int q = 0; q < SUB; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int64_t row = row0 + SUB * (lk + 4 * e) + q;
          if (row < M && col < nok) base[(size_t)col * M + row] = (R)acc[q][t][e];
        }
    }
  }
}

}  // namespace

size_t ritz_rotate_lds(int prec, int64_t dim, int nom, int nok, bool base16) {
  const size_t rsz = prec == QS3X_FP64 ? sizeof(double) : sizeof(float);
  const bool vec = base16 && ((2 * dim * (int64_t)rsz) % 16 == 0) && options().rot_vec;
  if (vec) return sizeof(double) * (size_t)nom * nok;
  return sizeof(double) * (size_t)((nom + 3) & ~3) * (size_t)((nok + 15) / 16) * 16;
}

void ritz_rotate(int prec, void *base, int64_t dim, const double *scales_dev, int nom, int nok,
                 const double *C_dev, hipStream_t st) {
  QS3X_REQUIRE(nok >= 1 && nok <= 16 * kMaxNT, "ritz_rotate: NOK must be in 1..64");
  const int NT = (nok + 15) / 16;
  const int64_t M = 2 * dim;
  const int64_t ntiles = (M + 15) / 16;
  const int waves = kRotThreads / 64;
  int grid = (int)std::min<int64_t>((ntiles + waves - 1) / waves, (int64_t)vec_grid(dim) * 2);
  grid = std::max(grid, 1);
  size_t lds = sizeof(double) * (size_t)((nom + 3) & ~3) * NT * 16;
  const size_t rsz = prec == QS3X_FP64 ? sizeof(double) : sizeof(float);
  const bool vec = ((uintptr_t)base % 16 == 0) && ((M * rsz) % 16 == 0) && options().rot_vec;
  if (vec) {
    lds = sizeof(double) * (size_t)nom * nok;
    const int64_t vt = (M + 16 * (16 / (int64_t)rsz) - 1) / (16 * (16 / (int64_t)rsz));
    grid = (int)std::max<int64_t>(1, std::min<int64_t>((vt + waves - 1) / waves, (int64_t)vec_grid(dim) * 2));
  }
  // up to what a CU holds (160 KB on gfx950, less a margin); beyond 64 KB the
  // kernel must be told before its launch
  QS3X_REQUIRE(lds <= 156 * 1024, "ritz_rotate: NOM x NOK too large for LDS staging");
#define ROT(R, NTV)                                                                                           \
  do {                                                                                                        \
    if (lds > 64 * 1024)                                                                                      \
      QS3X_HIP(hipFuncSetAttribute(vec ? (const void *)&k_rotate_mfma_v<R, NTV, 4>                            \
                                       : (const void *)&k_rotate_mfma<R, NTV>,                                \
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));                    \
    if (vec)                                                                                                  \
      hipLaunchKernelGGL((k_rotate_mfma_v<R, NTV, 4>), dim3(grid), dim3(kRotThreads), lds, st, (R *)base, M, \
                         scales_dev, nom, nok, C_dev);                                                        \
    else                                                                                                      \
      hipLaunchKernelGGL((k_rotate_mfma<R, NTV>), dim3(grid), dim3(kRotThreads), lds, st, (R *)base, M,        \
                         scales_dev, nom, nok, C_dev);                                                        \
  } while (0)
  if (prec == QS3X_FP64) {
    switch (NT) {
      case 1: ROT(double, 1); break;
      case 2: ROT(double, 2); break;
      case 3: ROT(double, 3); break;
      default: ROT(double, 4); break;
    }
  } else {
    switch (NT) {
      case 1: ROT(float, 1); break;
      case 2: ROT(float, 2); break;
      case 3: ROT(float, 3); break;
      default: ROT(float, 4); break;
    }
  }
#undef ROT
  QS3X_HIP(hipGetLastError());
}

}  // namespace qs3x

extern "C" {
// Test hook (tests/test_gpu_solvers.py): in-place rotation of a device basis
// (prec, (nom+1) slots of dim complex values at base) by host C (nom x nok,
// row-major) with unit scales.
int qs3x_test_ritz_rotate(int32_t prec, void *base, int64_t dim, int32_t nom, int32_t nok, const double *C) {
  try {
    qs3x::DevBuf<double> c((size_t)nom * nok);
    c.upload(C, (size_t)nom * nok);
    qs3x::ritz_rotate(prec, base, dim, nullptr, nom, nok, c.p, qs3x::stream());
    QS3X_HIP(hipStreamSynchronize(qs3x::stream()));
    return 0;
  } catch (const qs3x::Error &e) {
    qs3x::set_error(e.msg);
    return 1;
  }
}
}
