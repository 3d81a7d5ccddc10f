// blas.h -- launchers of the fused vector kernels (blas.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace qs3x {

constexpr int kVecThreads = 256;
constexpr int kMaxGroup = 8;    // basis vectors per pass of the multi-dot kernel
constexpr int kMaxBasis = 1024;  // largest Krylov basis (NOM + 1)
constexpr int kMaxBlocks = 4096;  // upper bound of every reduction grid (vec_grid, H.v)

int vec_grid(int64_t n);
void vec_random(int prec, void *v, int64_t n, uint64_t seed, hipStream_t st);
// dst = s * src with s = sc[is] (is >= 0) or s0; precisions may differ.
void vec_copy_scale(int pd, void *dst, int ps, const void *src, int64_t n, const double *sc, int is, double s0,
                    hipStream_t st);
// block partials of <x, y> = sum conj(x) y; returns the block count.
int vec_dot(int px, const void *x, int py, const void *y, int64_t n, double *part, hipStream_t st);
// see k_finalize modes in blas.hip
void finalize(const double *part, int nb, double *sc, int mode, int i0, int i1, double shift, hipStream_t st);
// alpha and beta of a fused Lanczos step from its 4 partial sums per block
// (k_finalize_fused in blas.hip)
void finalize_fused(const double *part, int nb, double *sc, int ia, int io, int ib, int is, double shift,
                    hipStream_t st);
int vec_update3(int prec, void *w, const void *u, const void *up, int64_t n, const double *sc, int ia, int isa,
                int ib, int isb, double *part, hipStream_t st);
void vec_axpy_acc(int prec, double *psi, const void *u, int64_t n, double z, const double *sc, int is,
                  hipStream_t st);
// V: device array of m vector pointers; scales[sidx ? sidx[j] : j] scales V[j]
void vec_mdot(int prec, const void *const *V, const double *scales, const int *sidx, int m, const void *w, int64_t n,
              double *part, double *h, hipStream_t st);
int vec_maxpy(int prec, const void *const *V, const double *scales, const int *sidx, int m, const double *h,
              void *w, int64_t n, double *part, hipStream_t st);

// Tunables (qs3x_set_option): thick-restart reorthogonalisation and the FP32 phase.
struct Options {
  int tr_reorth = 0;          // 0: every previous basis vector; 1: locked Ritz vectors + last two;
                              // 2: the three-term recurrence in the first cycle (with tr_check_every), then as 1
  int tr_cgs_passes = 2;      // classical Gram-Schmidt passes per step
  double fp32_floor = 1e-6;   // |delta theta| that ends the FP32 phase of the mixed thick restart
  int fp64_refine = 4;         // FP32-mixed thick restart: block Krylov steps of the final FP64 refinement (0: none)
  // FP32-mixed thick restart with one wanted pair: instead of the block
  // Krylov refinement, at most this many plain FP64 Lanczos steps from the
  // Rayleigh-Ritz vector, until the residual bound certifies the value (0: off)
  int fp64_lanczos = 0;
  int rot_vec = 1;            // Ritz rotation: 16-byte loads (aligned slots) instead of 4/8-byte loads
  int tr_check_every = 0;     // FP32 thick-restart stage: convergence check every n steps inside cycles (0: at restarts)
  // H.v kernel choice: 0 = fastest available (nearest-neighbour chain, then
  // batched-load chain, then per-hop chain, then general lattice); 1 = from
  // the batched-load chain kernel on; 2 = from the per-hop chain kernel on;
  // 3 = general lattice kernel (translation scan).
  int hv_kernel = 0;
  // L2 domains (XCDs) for the H.v tile schedule; 1 = plain round-robin.
  int hv_xcd = 0;  // 0: per precision (see hv_xcd_for)
  // Nearest-neighbour chain kernels with hv_xcd > 1: 0 gives each domain one
  // contiguous range of tiles; c > 0 deals chunks of 2^c tiles to the domains
  // in turn, so every domain works through the whole row range (the rows that
  // need the key search cluster at the end of the row order, and one range
  // per domain leaves them to the last one).  -1: 5.
  int hv_chunk = -1;
  // Gap-lattice stencil H.v (nearest-neighbour chains, N_down = 4, uniform
  // bonds): 1 = use it where it applies (default), 0 = never.
  int hv_stencil = 1;
  int hv_stencil_lines = 0;   // n3-lines per tile of the stencil kernel (0: per precision; even, 2 .. 14)
  int hv_stencil_slabs = 0;   // n4-slabs per stencil task (0: per precision)
  int hv_stencil_depth = 0;   // slab windows in flight per block (0: 2; 3 slab-window form, 4 pair form)
  // stencil kernel form: 0 = slab window (threads take the rows of each slab
  // in turn), 1 = column (a thread keeps its window positions for a task's
  // sweep), 2 = pair (column form with two adjacent positions per lane and
  // 16-byte accesses; FP32, the column form in FP64)
  int hv_stencil_kernel = 2;
  int hv_stencil_span = 0;  // column form: window positions per lane, 1, 2 or 4 (0: per precision)
  // single-vector stencil task order: 0 tile-major (each XCD a run of tiles,
  // every tile swept over its slabs), 1 n4-major (chunks of slabs in turn;
  // 2-5 % faster at L = 1000)
  int hv_stencil_order = 1;
  // stencil: key-search rows on a side stream, overlapping the stencil launch
  // (1), or after it on the caller's stream (0, the default: the overlap
  // measured no gain at L = 1000 -- the stencil fills the CUs -- and slowed
  // the stencil by 4 %)
  int hv_stencil_side = 0;
  // Batched 64-q DSF block H.v: the gap-lattice stencil for the clear rows of
  // N_down = 4 nearest-neighbour chains (1, default) or the per-row block
  // kernel for every row (0); its lines per tile (even, 2..14; 0: 12) and
  // slabs per task (0: 32).
  // Plain Lanczos on stencil sectors: the fused step (H.v and the three-term
  // update in one pass over the two previous vectors) in FP64 only (1,
  // default), in both precisions (2), or never (0: the H.v and update
  // kernels).  At L = 1000, N_down = 4 the fused step is faster in FP64
  // (0.736 against 0.784 ms per step) and slower in FP32 (0.464 against
  // 0.421 ms: the FP32 stencil is bound by its VALU work and memory latency,
  // and the fused window reads two vectors where the FP32 update streams
  // them at 6.6 TB/s).
  int lanczos_fused = 1;
  int block_stencil = 1;
  int block_stencil_lines = 0;
  int block_stencil_slabs = 0;
  int block_stencil_threads = 0;  // threads per block of the block stencil (0: 512; 256 or 512)
  int block_autotune = 0;         // 1: the DSF batch times the block stencil tile shapes and keeps the fastest
  int group_hop_cache = 1;        // 1: the group solve's later steps read the hops its first step found
  int block_fused = 1;            // 1: the DSF batch on block-stencil sectors runs the fused step (H.v + update)
  // dense eigensolver of qs3x_full_diag: 0 = QR iteration on the tridiagonal
  // form (rocSOLVER zheev, the reference's LAPACK zheev algorithm; default),
  // 1 = divide and conquer (zheevd; faster for large dim, but it has returned
  // NaN eigenvectors with info 0 on exactly degenerate spectra)
  int dense_solver = 0;
  int verbose = 0;            // 1: one stderr line per restart / per 100 Lanczos steps
};
Options &options();
// Tile schedule of the H.v kernels: option hv_xcd, or by default the 8 XCDs
// of MI355X, each working through interleaved chunks of 2^5 tiles (option
// hv_chunk) in the nearest-neighbour kernels.  At L = 1000, N_down = 4 this
// beats plain round robin and one range per XCD in both precisions (FP32
// 0.392 / 0.398 / 0.445 ms, FP64 0.590 / 0.680 / 0.644 ms).
inline int hv_xcd_for(int prec) {
  (void)prec;
  return options().hv_xcd > 0 ? options().hv_xcd : 8;
}
inline int hv_chunk_log2() { return options().hv_chunk >= 0 ? options().hv_chunk : 5; }

// host eigensolvers (eig.cpp)
double tridiag_eigenvalue(const double *d, const double *e, int n, int k);
double tridiag_lowest(const double *d, const double *e, int n);
void tridiag_vector(const double *d, const double *e, int n, double lam, double *z);
void sym_eig(int n, double *A, double *w, double *V);
// Thick-restart projection: k locked values w (couplings c to row k), then a
// tridiagonal tail a[0..m-k), b[0..m-k-1).
struct Arrow {
  int k, m;
  const double *w, *c, *a, *b;
};
double arrow_eigenvalue(const Arrow &A, int idx);
void arrow_lowest_pairs(const Arrow &A, int nv, double *lam, double *Z);

}  // namespace qs3x
