// api.hip -- C ABI entry points of libqs3x (model, sectors, H.v, timing).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <utility>

#include "blas.h"
#include "qs3x_internal.h"

namespace qs3x {

static thread_local std::string g_err;
static int g_device = 0;
static hipStream_t g_stream = nullptr;

void set_error(const std::string &msg) { g_err = msg; }

Options &options() {
  static Options o;
  return o;
}

int device() { return g_device; }

hipStream_t stream() {
  if (!g_stream) {
    QS3X_HIP(hipSetDevice(g_device));
    QS3X_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  }
  return g_stream;
}

template <class F>
static int guarded(F &&f) {
  try {
    f();
    return 0;
  } catch (const Error &e) {
    set_error(e.msg);
    return 1;
  } catch (const std::exception &e) {
    set_error(std::string("qs3x: ") + e.what());
    return 2;
  } catch (...) {
    set_error("qs3x: unknown error");
    return 3;
  }
}

}  // namespace qs3x

using namespace qs3x;

extern "C" {

const char *qs3x_last_error(void) { return g_err.c_str(); }

int qs3x_set_device(int32_t dev) {
  return guarded([&] {
    int n = 0;
    QS3X_HIP(hipGetDeviceCount(&n));
    QS3X_REQUIRE(dev >= 0 && dev < n, "qs3x_set_device: no such device");
    if (g_stream && dev != g_device) {
      (void)hipStreamDestroy(g_stream);
      g_stream = nullptr;
    }
    g_device = dev;
    QS3X_HIP(hipSetDevice(dev));
  });
}

int qs3x_set_option(const char *name, double value) {
  return guarded([&] {
    QS3X_REQUIRE(name, "qs3x_set_option: NULL name");
    const std::string n(name);
    if (n == "tr_reorth") {
      QS3X_REQUIRE(value >= 0 && value <= 2, "tr_reorth must be 0, 1 or 2");
      options().tr_reorth = (int)value;
    } else if (n == "tr_cgs_passes") {
      QS3X_REQUIRE(value >= 1 && value <= 3, "tr_cgs_passes must be 1..3");
      options().tr_cgs_passes = (int)value;
    } else if (n == "verbose") {
      QS3X_REQUIRE(value >= 0 && value <= 2, "verbose must be 0..2");
      options().verbose = (int)value;
    } else if (n == "fp64_refine") {
      QS3X_REQUIRE(value >= 0 && value <= 32, "fp64_refine must be 0..32");
      options().fp64_refine = (int)value;
    } else if (n == "fp64_lanczos") {
      QS3X_REQUIRE(value >= 0 && value <= 100000, "fp64_lanczos must be 0..100000");
      options().fp64_lanczos = (int)value;
    } else if (n == "rot_vec") {
      QS3X_REQUIRE(value == 0 || value == 1, "rot_vec must be 0 or 1");
      options().rot_vec = (int)value;
    } else if (n == "tr_check_every") {
      QS3X_REQUIRE(value >= 0 && value <= 1000, "tr_check_every must be 0..1000");
      options().tr_check_every = (int)value;
    } else if (n == "hv_xcd") {
      QS3X_REQUIRE(value >= 0 && value <= 64, "hv_xcd must be 0..64");
      options().hv_xcd = (int)value;
    } else if (n == "hv_chunk") {
      QS3X_REQUIRE(value >= -1 && value <= 24, "hv_chunk must be -1..24");
      options().hv_chunk = (int)value;
    } else if (n == "hv_kernel") {
      QS3X_REQUIRE(value >= 0 && value <= 3, "hv_kernel must be 0..3");
      options().hv_kernel = (int)value;
    } else if (n == "hv_stencil") {
      QS3X_REQUIRE(value == 0 || value == 1, "hv_stencil must be 0 or 1");
      options().hv_stencil = (int)value;
    } else if (n == "hv_stencil_lines") {
      QS3X_REQUIRE(value == 0 || (value >= 2 && value <= 14 && (int)value % 2 == 0),
                   "hv_stencil_lines must be 0 (per precision) or even, 2 .. 14 (10 .. 14: column form with one "
                   "position per lane; the slab-window form takes at most 8)");
      options().hv_stencil_lines = (int)value;
    } else if (n == "hv_stencil_slabs") {
      QS3X_REQUIRE(value >= 0 && value <= 1024, "hv_stencil_slabs must be 0..1024");
      options().hv_stencil_slabs = (int)value;
    } else if (n == "hv_stencil_depth") {
      QS3X_REQUIRE(value == 0 || value == 2 || value == 3 || value == 4,
                   "hv_stencil_depth must be 0 (per precision), 2, 3 (slab-window form) or 4 (pair form; the "
                   "slab-window form takes 3, the column form 2)");
      options().hv_stencil_depth = (int)value;
    } else if (n == "hv_stencil_kernel") {
      QS3X_REQUIRE(value == 0 || value == 1 || value == 2,
                   "hv_stencil_kernel must be 0 (slab window), 1 (column) or 2 (pair: FP32, else column)");
      options().hv_stencil_kernel = (int)value;
    } else if (n == "hv_stencil_span") {
      QS3X_REQUIRE(value == 0 || value == 1 || value == 2,
                   "hv_stencil_span must be 0 (per precision), 1 or 2 (2: FP32, up to 8 lines per tile)");
      options().hv_stencil_span = (int)value;
    } else if (n == "hv_stencil_order") {
      QS3X_REQUIRE(value == 0 || value == 1, "hv_stencil_order must be 0 (tile-major) or 1 (n4-major)");
      options().hv_stencil_order = (int)value;
    } else if (n == "hv_stencil_side") {
      QS3X_REQUIRE(value == 0 || value == 1, "hv_stencil_side must be 0 or 1");
      options().hv_stencil_side = (int)value;
    } else if (n == "lanczos_fused") {
      QS3X_REQUIRE(value == 0 || value == 1 || value == 2, "lanczos_fused must be 0, 1 (FP64) or 2 (both)");
      options().lanczos_fused = (int)value;
    } else if (n == "block_stencil") {
      QS3X_REQUIRE(value == 0 || value == 1, "block_stencil must be 0 or 1");
      options().block_stencil = (int)value;
    } else if (n == "block_stencil_lines") {
      const int v = (int)value;
      QS3X_REQUIRE(v == value && v >= 0 && v <= 14 && v % 2 == 0,
                   "block_stencil_lines must be 0 (default) or even, 2..14");
      options().block_stencil_lines = v;
    } else if (n == "block_stencil_threads") {
      QS3X_REQUIRE(value == 0 || value == 256 || value == 512, "block_stencil_threads must be 0, 256 or 512");
      options().block_stencil_threads = (int)value;
    } else if (n == "group_hop_cache") {
      QS3X_REQUIRE(value == 0 || value == 1, "group_hop_cache must be 0 or 1");
      options().group_hop_cache = (int)value;
    } else if (n == "block_fused") {
      QS3X_REQUIRE(value == 0 || value == 1, "block_fused must be 0 or 1");
      options().block_fused = (int)value;
    } else if (n == "block_autotune") {
      QS3X_REQUIRE(value == 0 || value == 1, "block_autotune must be 0 or 1");
      options().block_autotune = (int)value;
    } else if (n == "block_stencil_slabs") {
      QS3X_REQUIRE(value >= 0 && value <= 1024, "block_stencil_slabs must be 0..1024");
      options().block_stencil_slabs = (int)value;
    } else if (n == "dense_solver") {
      QS3X_REQUIRE(value == 0 || value == 1, "dense_solver must be 0 or 1");
      options().dense_solver = (int)value;
    } else if (n == "fp32_floor") {
      QS3X_REQUIRE(value > 0, "fp32_floor must be > 0");
      options().fp32_floor = value;
    } else {
      throw Error{"qs3x_set_option: unknown option " + n};
    }
  });
}

}  // extern "C"

// Options from the environment, applied once before the first model is
// created: QS3X_OPTIONS="name=value,name=value" with the names of
// qs3x_set_option (so that a driver that cannot call it, such as QS3's own
// Fortran main, can still choose e.g. the FP32 thick-restart settings).
static void apply_env_options() {
  static bool done = false;
  if (done) return;
  done = true;
  const char *env = std::getenv("QS3X_OPTIONS");
  if (!env) return;
  std::string all(env);
  size_t pos = 0;
  while (pos < all.size()) {
    size_t end = all.find(',', pos);
    if (end == std::string::npos) end = all.size();
    const std::string item = all.substr(pos, end - pos);
    pos = end + 1;
    if (item.empty()) continue;
    const size_t eq = item.find('=');
    QS3X_REQUIRE(eq != std::string::npos, "QS3X_OPTIONS: expected name=value, got '" + item + "'");
    char *stop = nullptr;
    const std::string val = item.substr(eq + 1);
    const double v = std::strtod(val.c_str(), &stop);
    QS3X_REQUIRE(stop && *stop == '\0' && !val.empty(), "QS3X_OPTIONS: bad value in '" + item + "'");
    if (qs3x_set_option(item.substr(0, eq).c_str(), v) != 0)
      throw Error{std::string("QS3X_OPTIONS: ") + qs3x_last_error()};
  }
}

extern "C" {

int qs3x_model_create(int32_t NOS, int32_t LX, int32_t LY, int32_t LZ, int32_t NOxxz,
                      const int32_t *p_xxz, const double *jxy, const double *jz,
                      qs3x_model **out) {
  return guarded([&] {
    apply_env_options();
    QS3X_REQUIRE(out, "qs3x_model_create: out is NULL");
    QS3X_REQUIRE(NOS > 0 && LX > 0 && LY > 0 && LZ > 0 && (int64_t)LX * LY * LZ == NOS,
                 "qs3x_model_create: NOS must equal LX*LY*LZ");
    QS3X_REQUIRE(NOS < 65536, "qs3x_model_create: NOS must be < 65536");
    QS3X_REQUIRE(NOxxz >= 0 && (NOxxz == 0 || (p_xxz && jxy && jz)), "qs3x_model_create: bonds");
    QS3X_HIP(hipSetDevice(g_device));
    auto *m = new qs3x_model();
    m->NOS = NOS; m->LX = LX; m->LY = LY; m->LZ = LZ; m->NOxxz = NOxxz;
    m->chain = (LY == 1 && LZ == 1);
    m->p1.resize(NOxxz); m->p2.resize(NOxxz);
    m->jxy.assign(jxy, jxy + NOxxz);
    m->jz.assign(jz, jz + NOxxz);
    // read_ip (input_param.f90:92-103): swap so that p1 < p2; Jz/4 summed for
    // the ferromagnetic energy (eigen_solver.f90:341-347).
    double ezz = 0.0;
    std::map<std::pair<double, double>, int> types;
    std::vector<std::vector<int32_t>> adj(NOS + 1);
    for (int b = 0; b < NOxxz; ++b) {
      int a = p_xxz[2 * b], c = p_xxz[2 * b + 1];
      if (!(a >= 1 && a <= NOS && c >= 1 && c <= NOS)) {
        delete m;
        throw Error{"qs3x_model_create: bond site out of range"};
      }
      if (a > c) std::swap(a, c);
      m->p1[b] = a; m->p2[b] = c;
      ezz += jz[b] / 4.0;
      auto key = std::make_pair(jxy[b], jz[b]);
      auto it = types.find(key);
      int bt;
      if (it == types.end()) {
        bt = (int)m->bt_jxy2.size();
        types[key] = bt;
        m->bt_jxy2.push_back(jxy[b] * 0.5);
        m->bt_jz4.push_back(jz[b] * 0.25);
      } else {
        bt = it->second;
      }
      if (bt >= 32768) {
        delete m;
        throw Error{"qs3x_model_create: more than 32767 distinct (Jxy, Jz) pairs"};
      }
      if (a == c) continue;  // a self-bond is never antiparallel: diagonal only
      adj[a].push_back(c | (bt << 16));
      adj[c].push_back(a | (bt << 16));
    }
    m->ezz = ezz;
    m->adj_ptr.assign(NOS + 2, 0);
    for (int s = 1; s <= NOS; ++s) {
      m->adj_ptr[s + 1] = m->adj_ptr[s] + (int)adj[s].size();
      m->max_deg = std::max<int>(m->max_deg, (int)adj[s].size());
      for (int32_t e : adj[s]) m->adj.push_back(e);
    }
    m->adj_ptr[0] = 0;
    if (m->chain && NOS >= 3) {
      bool nn = true;
      m->nn_bt.assign(NOS + 1, 0);
      for (int s = 1; s <= NOS && nn; ++s) {
        const int right = s % NOS + 1, left = (s + NOS - 2) % NOS + 1;
        if (adj[s].size() != 2) {
          nn = false;
          break;
        }
        const int o0 = adj[s][0] & 0xffff, o1 = adj[s][1] & 0xffff;
        if (!((o0 == right && o1 == left) || (o0 == left && o1 == right))) nn = false;
        m->nn_bt[s] = (uint16_t)((o0 == right ? adj[s][0] : adj[s][1]) >> 16);
      }
      m->nn_chain = nn;
      if (nn) {
        m->d_nn_bt.alloc(NOS + 1);
        m->d_nn_bt.upload(m->nn_bt.data(), NOS + 1);
      }
    }
    m->d_adj_ptr.alloc(m->adj_ptr.size());
    m->d_adj_ptr.upload(m->adj_ptr.data(), m->adj_ptr.size());
    m->d_adj.alloc(std::max<size_t>(1, m->adj.size()));
    if (!m->adj.empty()) m->d_adj.upload(m->adj.data(), m->adj.size());
    std::vector<double> bt(2 * std::max<size_t>(1, m->bt_jxy2.size()), 0.0);
    for (size_t q = 0; q < m->bt_jxy2.size(); ++q) {
      bt[2 * q] = m->bt_jxy2[q];
      bt[2 * q + 1] = m->bt_jz4[q];
    }
    m->d_bt.alloc(bt.size());
    m->d_bt.upload(bt.data(), bt.size());
    QS3X_HIP(hipStreamSynchronize(stream()));
    *out = m;
  });
}

int qs3x_model_destroy(qs3x_model *m) {
  return guarded([&] { delete m; });
}

int qs3x_fm_energy(const qs3x_model *m, double *e) {
  return guarded([&] {
    QS3X_REQUIRE(m && e, "qs3x_fm_energy: NULL argument");
    // calcu_FM_energy sums Jint(:,3) then divides by 4
    double s = 0.0;
    for (double z : m->jz) s += z;
    *e = s / 4.0;
  });
}

double qs3x_wavevector(int32_t K, int32_t L) {
  const double PI = std::acos(-1.0);
  return 2.0 * PI * (double)K / (double)L;
}

// A wave vector the lattice's translations can carry: rk_a L_a / (2 pi) an
// integer in every direction.  The reference builds a target sector at
// rk + q without this check (get_expectation_values.f90:430-447), so an
// incommensurate q (input_ex3's QZ = pi/2 on LZ = 10) gives norms and phases
// of no translation eigenspace; this build refuses it.
static void require_commensurate(const qs3x_model *m, const double *rk, const char *who) {
  const double PI = std::acos(-1.0);
  const int Ls[3] = {m->LX, m->LY, m->LZ};
  const char *ax = "xyz";
  for (int a = 0; a < 3; ++a) {
    const double K = rk[a] * Ls[a] / (2.0 * PI);
    if (std::fabs(K - std::nearbyint(K)) > 1e-8 * std::max(1.0, std::fabs(K))) {
      char buf[256];
      std::snprintf(buf, sizeof buf,
                    "%s: wave vector rk%c = %.17g is not a multiple of 2 pi / L%c (L%c = %d, K = %.12g); "
                    "snap q to 2 pi K / L with an integer K",
                    who, ax[a], rk[a], ax[a] + ('A' - 'a'), ax[a] + ('A' - 'a'), Ls[a], K);
      throw Error{buf};
    }
  }
}

int qs3x_sector_build(qs3x_model *m, int32_t nod, double rkx, double rky, double rkz,
                      qs3x_sector **out, int64_t *dim) {
  return guarded([&] {
    QS3X_REQUIRE(m && out, "qs3x_sector_build: NULL argument");
    const double rkv[3] = {rkx, rky, rkz};
    require_commensurate(m, rkv, "qs3x_sector_build");
    auto *s = new qs3x_sector();
    s->model = m;
    s->nod = nod;
    s->rk[0] = rkx; s->rk[1] = rky; s->rk[2] = rkz;
    try {
      build_sector_device(s);
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
    if (dim) *dim = s->dim;
  });
}

int qs3x_sector_build_u1(qs3x_model *m, int32_t nod, qs3x_sector **out, int64_t *dim) {
  return guarded([&] {
    QS3X_REQUIRE(m && out, "qs3x_sector_build_u1: NULL argument");
    auto *s = new qs3x_sector();
    s->model = m;
    s->nod = nod;
    try {
      build_sector_u1(s);
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
    if (dim) *dim = s->dim;
  });
}

int qs3x_sector_from_lists(qs3x_model *m, int32_t nod, int64_t dim, const int64_t *list_s,
                           const double *list_r, const double *explist, qs3x_sector **out) {
  return guarded([&] {
    QS3X_REQUIRE(m && out && (dim == 0 || (list_s && list_r)), "qs3x_sector_from_lists: NULL argument");
    auto *s = new qs3x_sector();
    s->model = m;
    s->nod = nod;
    s->dim = dim;
    try {
      sector_from_lists(s, list_s, list_r, explist);
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
  });
}

int qs3x_sector_destroy(qs3x_sector *s) {
  return guarded([&] { delete s; });
}

int qs3x_sector_dim(const qs3x_sector *s, int64_t *dim) {
  return guarded([&] {
    QS3X_REQUIRE(s && dim, "qs3x_sector_dim: NULL argument");
    *dim = s->dim;
  });
}

int qs3x_sector_lists(const qs3x_sector *s, int64_t *list_s, double *list_r) {
  return guarded([&] {
    QS3X_REQUIRE(s, "qs3x_sector_lists: NULL sector");
    if (is_u1(s)) {
      // the U(1)-only block: list_s(i) = i, norms 1 (rows are the ranks)
      for (int64_t i = 0; list_s && i < s->dim; ++i) list_s[i] = i + 1;
      for (int64_t i = 0; list_r && i < s->dim; ++i) list_r[i] = 1.0;
      return;
    }
    if (list_s && s->dim) s->d_list_s.download(list_s, s->dim);
    if (list_r && s->dim) s->d_list_r.download(list_r, s->dim);
  });
}

int qs3x_sector_nnz(qs3x_sector *s, int64_t *nnz) {
  return guarded([&] {
    QS3X_REQUIRE(s && nnz, "qs3x_sector_nnz: NULL argument");
    if (s->nnz < 0) s->nnz = count_nnz(s);
    *nnz = s->nnz;
  });
}

int qs3x_ham_to_vec_wave_vector(qs3x_sector *s, double *v0, const double *v1) {
  return guarded([&] {
    QS3X_REQUIRE(s && v0 && v1, "qs3x_ham_to_vec_wave_vector: NULL argument");
    const size_t n = (size_t)2 * s->dim;
    DevBuf<double> a(std::max<size_t>(n, 2)), b(std::max<size_t>(n, 2));
    if (n) a.upload(v1, n);
    launch_hv(s, QS3X_FP64, b.p, a.p, nullptr, nullptr, nullptr, stream());
    if (n) b.download(v0, n);
  });
}

int qs3x_hv_dev(qs3x_sector *s, int32_t prec, void *v0_dev, const void *v1_dev, void *st) {
  return guarded([&] {
    QS3X_REQUIRE(s && v0_dev && v1_dev, "qs3x_hv_dev: NULL argument");
    QS3X_REQUIRE(prec == QS3X_FP64 || prec == QS3X_FP32, "qs3x_hv_dev: bad precision");
    launch_hv(s, prec, v0_dev, v1_dev, nullptr, nullptr, nullptr, (hipStream_t)st);
  });
}

int qs3x_hv_last_launch(const qs3x_sector *s, int32_t *family, int64_t *lds_bytes) {
  return guarded([&] {
    QS3X_REQUIRE(s && family && lds_bytes, "qs3x_hv_last_launch: NULL argument");
    *family = s->last_family;
    *lds_bytes = s->last_lds;
  });
}

int qs3x_hv_stencil_info(const qs3x_sector *s, int32_t prec, int64_t *info) {
  return guarded([&] {
    QS3X_REQUIRE(s && info, "qs3x_hv_stencil_info: NULL argument");
    QS3X_REQUIRE(prec == QS3X_FP64 || prec == QS3X_FP32, "qs3x_hv_stencil_info: bad precision");
    const qs3x_sector::StencilPlan &pl = s->stencil[prec == QS3X_FP64 ? 0 : 1];
    QS3X_REQUIRE(pl.tried && pl.ok, "qs3x_hv_stencil_info: no stencil launch on this sector and precision yet");
    info[0] = s->nslow;
    info[1] = pl.nlines;
    info[2] = pl.ntasks;
    info[3] = pl.B;
    info[4] = pl.slabs;
    info[5] = pl.grid;
    info[6] = pl.form;
    info[7] = pl.P;
  });
}

int qs3x_time_hv(qs3x_sector *s, int32_t prec, int32_t n_rep, double *ms) {
  return guarded([&] {
    QS3X_REQUIRE(s && ms && n_rep > 0, "qs3x_time_hv: bad argument");
    const size_t es = (prec == QS3X_FP64) ? 16 : 8;
    DevBuf<unsigned char> a(std::max<size_t>(1, es * s->dim)), b(std::max<size_t>(1, es * s->dim));
    QS3X_HIP(hipMemsetAsync(a.p, 0, es * s->dim, stream()));
    hipEvent_t e0, e1;
    QS3X_HIP(hipEventCreate(&e0));
    QS3X_HIP(hipEventCreate(&e1));
    launch_hv(s, prec, b.p, a.p, nullptr, nullptr, nullptr, stream());
    QS3X_HIP(hipEventRecord(e0, stream()));
    for (int r = 0; r < n_rep; ++r) launch_hv(s, prec, b.p, a.p, nullptr, nullptr, nullptr, stream());
    QS3X_HIP(hipEventRecord(e1, stream()));
    QS3X_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    QS3X_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = (double)t / n_rep;
  });
}

int qs3x_hv_autotune(qs3x_sector *s, int32_t prec, int32_t n_rep, int32_t *best, double *best_ms,
                     int32_t max_results, int32_t *n_results, double *results) {
  return guarded([&] {
    QS3X_REQUIRE(s && n_rep > 0, "qs3x_hv_autotune: bad argument");
    QS3X_REQUIRE(prec == QS3X_FP64 || prec == QS3X_FP32, "qs3x_hv_autotune: bad precision");
    qs3x_sector::HvTune &tn = s->tune[prec == QS3X_FP64 ? 0 : 1];
    tn.set = false;
    if (s->dim == 0) {
      if (n_results) *n_results = 0;
      return;
    }
    const size_t es = (prec == QS3X_FP64) ? 16 : 8;
    DevBuf<unsigned char> a(es * s->dim), b(es * s->dim);
    hipStream_t st = stream();
    vec_random(prec, a.p, s->dim, 12345, st);
    hipEvent_t e0, e1;
    QS3X_HIP(hipEventCreate(&e0));
    QS3X_HIP(hipEventCreate(&e1));
    auto timed = [&](int reps) {
      QS3X_HIP(hipEventRecord(e0, st));
      for (int r = 0; r < reps; ++r) launch_hv(s, prec, b.p, a.p, nullptr, nullptr, nullptr, st);
      QS3X_HIP(hipEventRecord(e1, st));
      QS3X_HIP(hipEventSynchronize(e1));
      float t = 0.f;
      QS3X_HIP(hipEventElapsedTime(&t, e0, e1));
      return (double)t / reps;
    };
    const int xd = hv_xcd_for(prec);  // the default schedule first
    const int xalt = xd == 1 ? 8 : 1;
    double best_t = 1e300;
    qs3x_sector::HvTune best_c;
    int best_family = -1;
    int nres = 0;
    // One candidate: a launch that is warm (its first launch builds the
    // kernel's plan) is timed once; over 3x the best so far it is recorded
    // negated and not repeated, else timed over n_rep launches.  A later
    // candidate must win by 2 %, so that noise does not move the choice off
    // the earlier (default) configurations.  Returns the family that ran.
    auto run = [&](const qs3x_sector::HvTune &c, bool warm) {
      tn = c;
      if (!warm) launch_hv(s, prec, b.p, a.p, nullptr, nullptr, nullptr, st);  // plans, LDS attributes
      const int fam = s->last_family;
      const double first = timed(1);
      double ms;
      if (first > 3.0 * best_t) {
        ms = -first;
      } else {
        ms = timed(n_rep);
        if (ms < 0.98 * best_t) {
          best_t = ms;
          best_c = c;
          best_family = fam;
        }
      }
      if (results && nres < max_results) {
        double *r = results + 12 * (size_t)nres;
        const double row[12] = {(double)fam, (double)c.kernel, (double)c.stencil, (double)c.st_lines,
                                (double)c.st_slabs, (double)c.st_depth, (double)c.pairs, (double)c.chunk,
                                (double)c.xcd, (double)c.grid_q, (double)c.st_col, ms};
        for (int k = 0; k < 12; ++k) r[k] = row[k];
      }
      ++nres;
      return fam;
    };
    auto base = [&](int kernel, int stencil) {
      qs3x_sector::HvTune c;
      c.set = true;
      c.kernel = kernel;
      c.stencil = stencil;
      c.xcd = xd;
      c.grid_q = 4;
      return c;
    };
    // Stage 1: the kernel families, each in its default configuration.  The
    // launcher falls through to the next family where one does not apply; a
    // family seen before is not timed again.
    bool seen[8] = {false, false, false, false, false, false, false, false};
    for (int k = -1; k <= 3; ++k) {
      const qs3x_sector::HvTune c = base(std::max(k, 0), k < 0 ? 1 : (k == 0 ? 0 : 1));
      tn = c;
      launch_hv(s, prec, b.p, a.p, nullptr, nullptr, nullptr, st);
      const int fam = s->last_family;
      if (fam < 0 || fam >= 8 || seen[fam]) continue;
      seen[fam] = true;
      run(c, true);  // the launch above was its warm-up
    }
    // Stage 2: the winning family's own parameters.
    const qs3x_sector::HvTune w = best_c;
    if (best_family == 5) {
      // stencil: kernel form (the precision's default, pair in FP32 and
      // column in FP64, then the slab-window form), tile height, task slabs,
      // slab windows in flight
      const int fdef = prec == QS3X_FP64 ? 1 : 2;
      for (int form : {fdef, 0})
        for (int B : {4, 6, 8, 10, 14}) {
          if ((form == 0 && B > 8) || (form == 2 && B > 10) || (form != 0 && B < 6)) continue;
          for (int sl : {8, 16, 32})
            for (int d : {2, 3, 4}) {
              if ((form == 0 && d == 4) || (form == 1 && d != 2) || (form == 2 && d == 3)) continue;
              qs3x_sector::HvTune c = w;
              c.st_col = form;
              c.st_lines = B;
              c.st_slabs = sl;
              c.st_depth = d;
              run(c, false);
            }
        }
    } else if (best_family == 0) {
      // nearest-neighbour kernel: rows per lane (FP32), interleaved chunk,
      // schedule, grid
      for (int pr : {-1, 0}) {
        if (pr == 0 && prec == QS3X_FP64) continue;
        for (int ch : {0, 3, 5, 7})
          for (int x : {xd, xalt})
            for (int q : {4, 2}) {
              qs3x_sector::HvTune c = w;
              c.pairs = pr;
              c.chunk = ch;
              c.xcd = x;
              c.grid_q = q;
              run(c, false);
            }
      }
    } else if (best_family >= 0) {
      for (int x : {xd, xalt})
        for (int q : {4, 3, 2, 1}) {
          qs3x_sector::HvTune c = w;
          c.xcd = x;
          c.grid_q = q;
          run(c, false);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    tn = best_c;
    if (best) {
      const int v[11] = {best_family, best_c.kernel, best_c.stencil, best_c.st_lines, best_c.st_slabs,
                         best_c.st_depth, best_c.pairs, best_c.chunk, best_c.xcd, best_c.grid_q, best_c.st_col};
      for (int k = 0; k < 11; ++k) best[k] = v[k];
    }
    if (best_ms) *best_ms = best_t;
    if (n_results) *n_results = std::min(nres, max_results);
  });
}

int qs3x_hv_tune_clear(qs3x_sector *s) {
  return guarded([&] {
    QS3X_REQUIRE(s, "qs3x_hv_tune_clear: NULL sector");
    s->tune[0].set = s->tune[1].set = false;
  });
}

}  // extern "C"
