// qs3x_internal.h -- shared definitions for libqs3x (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/qs3x.h"

namespace qs3x {

// ----------------------------------------------------------------- errors
void set_error(const std::string &msg);
struct Error {
  std::string msg;
};
#define QS3X_HIP(call)                                                                 \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw ::qs3x::Error{std::string(#call) + ": " + hipGetErrorString(e_)};          \
  } while (0)
#define QS3X_REQUIRE(cond, msg)                                                        \
  do {                                                                                 \
    if (!(cond)) throw ::qs3x::Error{std::string(msg)};                                \
  } while (0)

hipStream_t stream();  // the library stream (default priority)
int device();

// Device buffer with RAII.
template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) QS3X_HIP(hipMalloc(&p, sizeof(T) * count));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    release();
    p = o.p; n = o.n; o.p = nullptr; o.n = 0;
    return *this;
  }
  void upload(const T *h, size_t count) {
    QS3X_HIP(hipMemcpyAsync(p, h, sizeof(T) * count, hipMemcpyHostToDevice, stream()));
  }
  void download(T *h, size_t count) const {
    QS3X_HIP(hipMemcpyAsync(h, p, sizeof(T) * count, hipMemcpyDeviceToHost, stream()));
    QS3X_HIP(hipStreamSynchronize(stream()));
  }
};

constexpr int kMaxNod = 16;      // largest N_down any kernel handles
constexpr int kMaxChainNod = 12; // compile-time specialisations of the chain kernels

}  // namespace qs3x

// ------------------------------------------------------------------ model
struct qs3x_model {
  int32_t NOS, LX, LY, LZ, NOxxz;
  std::vector<int32_t> p1, p2;       // 1-based, p1 <= p2 (reference swap)
  std::vector<double> jxy, jz;       // raw couplings in bond-file order
  double ezz = 0.0;                  // sum_b Jz_b / 4
  bool chain = false;                // LY == LZ == 1
  // Bond adjacency (self-bonds dropped: never antiparallel).  adj_ptr has
  // NOS + 2 entries indexed by site (1..NOS); adj[k] = other | (btype << 16).
  std::vector<int32_t> adj_ptr, adj;
  std::vector<double> bt_jxy2, bt_jz4;  // per distinct (Jxy, Jz): Jxy/2, Jz/4
  int32_t max_deg = 0;
  // Nearest-neighbour chain (LY = LZ = 1, NOS >= 3, every site bonded exactly
  // to s - 1 and s + 1 cyclically, once each): nn_bt[s] = bond type of
  // (s, s + 1 mod NOS) for s = 1..NOS (entry 0 unused).
  bool nn_chain = false;
  std::vector<uint16_t> nn_bt;
  qs3x::DevBuf<uint16_t> d_nn_bt;
  // device copies
  qs3x::DevBuf<int32_t> d_adj_ptr, d_adj;
  qs3x::DevBuf<double> d_bt;  // [2*nbt]: jxy2, jz4 interleaved
};

// Sector: rows are the representatives in ascending rank.
//
// Packed configuration word (d_cfg): down-spin sites n_0 < ... < n_{N-1}
// (1-based) in consecutive `bits`-wide fields from bit 0, and the norm class
// in bits 56..63 (chain sectors: the stabiliser order; general sectors: 0).
enum IndexMode : int32_t {
  kIndexAnchored = 0,  // chain: dense table over configurations containing site 1
  kIndexFullRank = 1,  // dense table over all C(NOS, N) ranks
  kIndexSearch = 2,    // binary search in list_s
  kIndexU1 = 3         // U(1)-only block: row = colex rank - 1, no lists (source_only_u1)
};

struct qs3x_sector {
  qs3x_model *model = nullptr;
  int32_t nod = 0;
  double rk[3] = {0, 0, 0};
  int64_t dim = 0;
  int32_t bits = 0;        // bits per site field in d_cfg
  int32_t mode = 0;        // IndexMode
  bool real_phase = true;  // every explist entry the operator can use is exactly 1
  bool all_k = false;      // chain sector of every translation representative (batched multi-k work)
  bool chain = false;      // chain fast path usable (LY = LZ = 1 and built on device)
  int64_t index_len = 0;
  int32_t nbinom_n = 0, nbinom_k = 0;  // binomial table C(n, k), n < nbinom_n, k < nbinom_k
  qs3x::DevBuf<int64_t> d_list_s;
  qs3x::DevBuf<double> d_list_r;
  qs3x::DevBuf<uint64_t> d_cfg;
  qs3x::DevBuf<int32_t> d_index;
  // Anchored mode: one word per 32 anchored ranks, x = presence bits, y = row
  // of the word's first present rank (rows ascend with rank).
  qs3x::DevBuf<uint2> d_bitmap;
  qs3x::DevBuf<double> d_explist;   // complex, (LX+1)(LY+1)(LZ+1)
  qs3x::DevBuf<double> d_rclass;    // chain: r = sqrt(ro) per stabiliser order (0..nod)
  qs3x::DevBuf<int64_t> d_binom;    // row-major [n][k]
  std::vector<double> h_rclass;
  std::vector<double> h_explist;
  int64_t nnz = -1;
  // scratch for reductions
  qs3x::DevBuf<double> d_partials;
  // H.v launch configuration chosen by qs3x_hv_autotune, per precision
  // (0: FP64, 1: FP32); unset entries fall back to the options.
  struct HvTune {
    bool set = false;
    int kernel = 0;  // first kernel tried, as option hv_kernel
    int xcd = 1;     // L2 domains of the tile schedule, as option hv_xcd
    int grid_q = 4;  // quarters of the resident grid launched (4: all of it)
    int stencil = 1;  // kernel 0 on a stencil sector: 1 the stencil, 0 the nearest-neighbour kernel
    int st_lines = 0, st_slabs = 0, st_depth = 0;  // stencil tile height, task slabs, windows in flight (0: options)
    int st_col = -1;  // stencil kernel form: 0 slab window, 1 column, 2 pair (-1: option hv_stencil_kernel)
    int st_span = 0;  // column form: window positions per lane (0: option hv_stencil_span)
    int chunk = -1;   // nearest-neighbour kernels: log2 of the interleaved XCD chunk (-1: option hv_chunk)
    int pairs = -1;   // FP32 nearest-neighbour kernel: 0 one row per lane, -1 the default rule (two where it applies)
  } tune[2];
  // the last H.v launch: kernel family (0 nearest-neighbour chain, 1
  // batched-load chain, 2 per-hop chain, 3 general lattice, 4 U(1)-only) and
  // its dynamic LDS bytes (qs3x_hv_last_launch)
  int last_family = -1;
  int64_t last_lds = 0;
  // Gap-lattice stencil H.v (hv_nn.hip): line table over (n3, n4) of the
  // anchored rows (int4: first row, lowest n2, highest n2 + 1, 0) and, per
  // precision, the task list of its launch plan.
  struct StencilPlan {
    bool tried = false, ok = false;
    int B = 0, slabs = 0, P = 1, form = 0, order = 0, ntasks = 0, grid = 0, grid_fused = 0;  // P: window positions per lane
    const void *grid_for = nullptr, *grid_fused_for = nullptr;  // the kernels the grids were sized for
    int64_t nlines = 0;  // non-empty (n3, n4) lines
    size_t lds = 0;
    qs3x::DevBuf<int4> tasks;
  } stencil[2];
  qs3x::DevBuf<int4> d_lines;
  qs3x::DevBuf<int32_t> d_binom32;  // C(x, q) at [(q - 1)(L - 1) + x], q = 1..3
  qs3x::DevBuf<int32_t> d_slow;     // rows the stencil kernel skips (not clear), ascending
  int64_t nslow = 0;
};

namespace qs3x {

// Host helpers (sector.hip)
int64_t binom_host(int64_t n, int64_t k);  // exact C(n,k) with the reference's recurrence
void build_sector_device(qs3x_sector *s);
// U(1)-only block of all C(NOS, N) configurations (source_only_u1/main.f90:23-29)
void build_sector_u1(qs3x_sector *s);
// Entry points that need the translation-symmetric lists reject U(1)-only
// sectors, which support H.v, the Lanczos solvers and the DSF recurrence.
inline bool is_u1(const qs3x_sector *s) { return s && s->mode == kIndexU1; }
void sector_from_lists(qs3x_sector *s, const int64_t *list_s, const double *list_r,
                       const double *explist);
void compute_explist(const qs3x_model *m, const double rk[3], std::vector<double> &ex);
// Chain norm ro of a configuration with stabiliser order `count` at wave
// vector rk (the reference's phase sum, state_lists.f90:84-110).
double chain_ro(int L, int count, const double rk[3]);

// H.v launch (hv.hip). v0 = H (scale * v1) [- shift for FP32]; if dot_partials
// is non-null each block writes sum_i conj(scale*v1_i) * v0_i (complex,
// 2 doubles) to dot_partials[2*block] and nblocks_out receives the count.
void launch_hv(qs3x_sector *s, int prec, void *v0, const void *v1, const double *scale_dev,
               double *dot_partials, int *nblocks_out, hipStream_t st);
int64_t count_nnz(qs3x_sector *s);
// Fused Lanczos step on stencil sectors (hv.hip): v = c1 pprev - c2 vprev
// (c1 = fsc[fis], c2 = fsc[fia] fsc[fis]; vprev null at the first step, then
// v = c1 pprev) is written to xout and P = (H - e_zz) v - fsc[fib] vprev to
// pout, with 4 partial sums per block in part4 (may be null): Re<v, Hv>,
// |P|^2, Re<v, P>, |v|^2 (finalize_fused turns them into alpha and beta).
// Returns false, launching nothing, where the stencil does not apply or
// option lanczos_fused excludes the precision.
bool hv_fused_available(const qs3x_sector *s, int prec);
bool launch_hv_fused(qs3x_sector *s, int prec, void *pout, void *xout, const void *pprev, const void *vprev,
                     const double *fsc, int fis, int fia, int fib, double *part4, int *nblocks_out, hipStream_t st);

// Gap-lattice stencil (hv_nn.hip): whether a sector qualifies (anchored
// nearest-neighbour chain with uniform bonds and N_down = 4); its line table,
// binomials and non-clear row list (built once per sector; with lines_out the
// line table is also copied to the host); and the XCD-dealt task list of a
// sweep (see the definition).
bool stencil_applies(const qs3x_sector *s);
void stencil_tables(qs3x_sector *s, std::vector<int4> *lines_out);
std::vector<int4> stencil_tasks(const std::vector<int4> &lines, int L, int segA, int B, int slabs, int groups);

}  // namespace qs3x

// ------------------------------------------------------------ device utils
namespace qs3x {
namespace dev {

__device__ __forceinline__ int unpack_site(uint64_t cfg, int q, int bits) {
  return (int)((cfg >> (q * bits)) & ((1ull << bits) - 1ull));
}
__device__ __forceinline__ int cfg_class(uint64_t cfg) { return (int)(cfg >> 56); }

}  // namespace dev
}  // namespace qs3x
