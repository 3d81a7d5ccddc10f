// solvers.hip -- device-resident Lanczos solvers (plain and thick restart).
//
// All Krylov vectors stay in HBM; per step the host only enqueues kernels.
// Lanczos scalars live in a device array (see blas.hip); the host reads them
// back at convergence checks and restarts, where it solves the small
// projected eigenproblem (eig.cpp).
//
// FP32 mode: Krylov vectors are complex float and the operator applied is
// H' = H - e_zz (e_zz = sum_b Jz_b/4, the ferromagnetic energy), which keeps
// the FP32 diagonal O(number of antiparallel bonds) instead of O(L).  Krylov
// spaces are shift invariant, so only the alphas move: alpha = alpha' + e_zz,
// added in FP64.  Dots, norms, reorthogonalisation coefficients and the Ritz
// rotation are FP64.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "blas.h"
#include "qs3x_internal.h"

using namespace qs3x;

namespace qs3x {
size_t ritz_rotate_lds(int prec, int64_t dim, int nom, int nok, bool base16);
void ritz_rotate(int prec, void *base, int64_t dim, const double *scales_dev, int nom, int nok,
                 const double *C_dev, hipStream_t st);
}

namespace {

template <class F>
int guarded2(F &&f) {
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

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t elem_bytes(int prec) { return prec == QS3X_FP64 ? 16 : 8; }

// Convergence trace of the last solve: one row per convergence check (plain
// Lanczos) or restart (thick restart): [seconds since the solve started,
// H.v applications so far, lowest Ritz value (FP64, unshifted), precision].
struct TracePoint {
  double t, hv, e, prec;
};
std::vector<TracePoint> &trace() {
  static std::vector<TracePoint> tr;
  return tr;
}
void trace_point(double t0, int64_t hv, double e, int prec) {
  trace().push_back({now_s() - t0, (double)hv, e, prec == QS3X_FP64 ? 64.0 : 32.0});
}

// Acceptance of the last FP32-mixed thick restart's lowest pair: spectral
// gap estimate, FP64 residual, bound r^2 / gap, acceptance (1 certified, 2
// agreement only, 0 not accepted; -1: no mixed solve yet); then the last
// thick restart's seconds allocating and freeing Krylov bases.
double *certificate() {
  static double c[6] = {0.0, 0.0, 0.0, -1.0, 0.0, 0.0};
  return c;
}

// Loads a start vector (host complex double) or generates the counter-based one.
void load_start(int prec, void *dst, int64_t dim, const double *start, uint64_t seed, hipStream_t st) {
  if (start) {
    DevBuf<double> tmp((size_t)2 * dim);
    tmp.upload(start, (size_t)2 * dim);
    vec_copy_scale(prec, dst, QS3X_FP64, tmp.p, dim, nullptr, -1, 1.0, st);
    QS3X_HIP(hipStreamSynchronize(st));
  } else {
    vec_random(prec, dst, dim, seed, st);
  }
}

}  // namespace

extern "C" {

int qs3x_solver_certificate(double *out) {
  return guarded2([&] {
    QS3X_REQUIRE(out, "qs3x_solver_certificate: NULL argument");
    for (int k = 0; k < 6; ++k) out[k] = certificate()[k];
  });
}

int qs3x_solver_trace(double *out, int64_t cap, int64_t *n) {
  return guarded2([&] {
    QS3X_REQUIRE(n, "qs3x_solver_trace: NULL count");
    const auto &tr = trace();
    *n = (int64_t)tr.size();
    if (out)
      for (int64_t i = 0; i < std::min<int64_t>(cap, (int64_t)tr.size()); ++i) {
        out[4 * i] = tr[i].t;
        out[4 * i + 1] = tr[i].hv;
        out[4 * i + 2] = tr[i].e;
        out[4 * i + 3] = tr[i].prec;
      }
  });
}

// ------------------------------------------------------------ plain Lanczos
// lancz_eigen_val (lanczos.f90:50-152).
int qs3x_lanczos(qs3x_sector *s, int32_t prec, int64_t maxitr, int64_t minitr, int64_t itrint, double conv,
                 const double *start, uint64_t seed, double *ene, double *psi, int64_t *iters) {
  return guarded2([&] {
    QS3X_REQUIRE(s && ene, "qs3x_lanczos: NULL argument");
    QS3X_REQUIRE(prec == QS3X_FP64 || prec == QS3X_FP32, "qs3x_lanczos: bad precision");
    QS3X_REQUIRE(maxitr >= 1 && itrint >= 1, "qs3x_lanczos: maxitr and itrint must be >= 1");
    hipStream_t st = stream();
    const int64_t dim = s->dim;
    QS3X_REQUIRE(dim >= 1, "qs3x_lanczos: empty sector");
    // The reference stops at dim steps (exact arithmetic spans the sector).
    // FP32 vectors lose orthogonality long before, so the FP32 recurrence may
    // run past dim (up to 2 dim + 8 steps): ghosts of converged Ritz values
    // never lie below the lowest eigenvalue, and the change test still stops it.
    const int64_t mitr = std::min<int64_t>(maxitr, prec == QS3X_FP32 ? 2 * dim + 8 : dim);
    const size_t es = elem_bytes(prec);
    // Fused steps (stencil sectors): one kernel pass forms the Lanczos vector
    // v_i = (P_{i-1} - alpha_{i-1} v_{i-1}) / beta_{i-1} from the previous
    // step's P and v while applying H, and writes v_i and P_i = (H - e_zz) v_i
    // - beta_{i-1} v_{i-1}; alpha_i and beta_i come from its partial sums.
    // Buffers: U[0], U[1] hold P (U[1]: the start vector), U[2], U[3] hold v.
    // Otherwise the H.v and update kernels run on three buffers U[0..2].
    const bool fused = hv_fused_available(s, prec);
    DevBuf<unsigned char> U[4];
    for (int k = 0; k < (fused ? 4 : 3); ++k) U[k].alloc(es * dim);
    // scalar slots
    const int64_t A_REC = 0, A_OUT = mitr, BETA = 2 * mitr, SCALE = 3 * mitr + 1;
    DevBuf<double> sc((size_t)4 * mitr + 4);
    DevBuf<double> part((size_t)4 * kMaxBlocks);
    const double shift = (prec == QS3X_FP32 || fused) ? s->model->ezz : 0.0;
    trace().clear();
    const double t_start = now_s();
    void *u0 = fused ? U[1].p : U[0].p;
    load_start(prec, u0, dim, start, seed, st);
    int nb = vec_dot(prec, u0, prec, u0, dim, part.p, st);
    finalize(part.p, nb, sc.p, 2, (int)(SCALE + mitr + 1), (int)SCALE, 0.0, st);  // scale[0] = 1/|u0|
    // one step: alpha_i, beta_i and scale_{i+1} = 1/beta_i into sc
    auto step = [&](int64_t i, double *pt) {
      if (fused) {
        void *pprev = U[(i + 1) % 2].p, *pout = U[i % 2].p;
        void *vprev = i > 0 ? U[2 + (i + 1) % 2].p : nullptr, *xout = U[2 + i % 2].p;
        QS3X_REQUIRE(launch_hv_fused(s, prec, pout, xout, pprev, vprev, sc.p, (int)(SCALE + i), (int)(A_REC + i - 1),
                                     i > 0 ? (int)(BETA + i - 1) : -1, pt, &nb, st),
                     "qs3x_lanczos: fused step unavailable");
        if (pt) finalize_fused(pt, nb, sc.p, (int)(A_REC + i), (int)(A_OUT + i), (int)(BETA + i), (int)(SCALE + i + 1),
                               shift, st);
        return;
      }
      void *cur = U[i % 3].p, *prev = U[(i + 2) % 3].p, *w = U[(i + 1) % 3].p;
      launch_hv(s, prec, w, cur, sc.p + SCALE + i, pt, &nb, st);
      if (pt) finalize(pt, nb, sc.p, 1, (int)(A_REC + i), (int)(A_OUT + i), shift, st);
      nb = vec_update3(prec, w, cur, prev, dim, sc.p, (int)(A_REC + i), (int)(SCALE + i),
                       i > 0 ? (int)(BETA + i - 1) : -1, (int)(SCALE + i - 1), pt, st);
      if (pt) finalize(pt, nb, sc.p, 2, (int)(BETA + i), (int)(SCALE + i + 1), 0.0, st);
    };
    std::vector<double> ha, hb;
    double E = 0.0, tmp = 0.0;
    int64_t ite = 0;
    bool done = false;
    // The convergence test (lanczos.f90:90-110) runs one check interval
    // behind the device: at a check point the alpha/beta so far are copied
    // to pinned host memory and an event recorded, the next steps are
    // launched, and the previous check is evaluated once its event has
    // completed.  The device never waits for the host's tridiagonal solve;
    // it runs at most itrint steps past the step that met the test, and the
    // result (E, iteration count, eigenvector pass) is that of the step that
    // met it, as in the reference.
    struct Pending {
      int64_t ite = 0;
      double *host = nullptr;
      hipEvent_t ev = nullptr;
    } pend[2];
    for (auto &p : pend) {
      QS3X_HIP(hipHostMalloc((void **)&p.host, sizeof(double) * 2 * (size_t)mitr, hipHostMallocDefault));
      QS3X_HIP(hipEventCreateWithFlags(&p.ev, hipEventDisableTiming));
    }
    struct PendingFree {
      Pending *p;
      ~PendingFree() {
        for (int k = 0; k < 2; ++k) {
          if (p[k].host) (void)hipHostFree(p[k].host);
          if (p[k].ev) (void)hipEventDestroy(p[k].ev);
        }
      }
    } pend_free{pend};
    int npend = 0, slot = 0;
    int64_t ite_done = 0;  // step count of the check that ended the solve
    auto evaluate = [&](Pending &p) {
      QS3X_HIP(hipEventSynchronize(p.ev));
      const int64_t n = p.ite;
      ha.assign(p.host, p.host + n);
      hb.assign(p.host + mitr, p.host + mitr + n);
      if (n == 1) {  // dim == 1 (lanczos.f90:95-101)
        E = ha[0];
        done = true;
        ite_done = n;
        return;
      }
      E = tridiag_lowest(ha.data(), hb.data(), (int)n);
      const double cv = std::fabs(tmp / E - 1.0);
      const bool breakdown = !(hb[n - 1] > 1e-13 * std::fabs(E) + 1e-300);
      if (n == maxitr || n == mitr || cv < conv || breakdown) {
        done = true;
        ite_done = n;
      }
      trace_point(t_start, n, E, prec);
      if (options().verbose && (done || options().verbose > 1 || n % 100 == 0))
        std::fprintf(stderr, "[qs3x] lanczos step %lld E=%.15g change %.3e\n", (long long)n, E, cv);
      tmp = E;
    };
    for (int64_t i = 0; i < mitr && !done; ++i) {
      step(i, part.p);
      ite = i + 1;
      if (ite == mitr || (ite >= minitr && ite % itrint == 0)) {
        Pending &p = pend[slot];
        p.ite = ite;
        QS3X_HIP(hipMemcpyAsync(p.host, sc.p + A_OUT, sizeof(double) * ite, hipMemcpyDeviceToHost, st));
        QS3X_HIP(hipMemcpyAsync(p.host + mitr, sc.p + BETA, sizeof(double) * ite, hipMemcpyDeviceToHost, st));
        QS3X_HIP(hipEventRecord(p.ev, st));
        slot ^= 1;
        if (++npend == 2) {  // evaluate the previous check while this one is in flight
          evaluate(pend[slot]);
          --npend;
        }
        if (ite == mitr && !done) {  // last step: evaluate it now
          evaluate(pend[slot ^ 1]);
          npend = 0;
        }
      }
    }
    if (!done && npend == 1) evaluate(pend[slot ^ 1]);
    if (ite_done > 0) ite = ite_done;
    // coefficients of the accepted step count (the extra steps past it are dropped)
    ha.resize(ite);
    hb.resize(ite);
    ene[0] = E;
    ene[1] = 99.0;  // the reference's sentinel for ene(2) (lanczos.f90:23-24)
    if (iters) *iters = ite;
    if (psi) {
      DevBuf<double> p64((size_t)2 * dim);
      QS3X_HIP(hipMemsetAsync(p64.p, 0, 16 * dim, st));
      std::vector<double> z(ite);
      if (ite == 1) z[0] = 1.0;
      else tridiag_vector(ha.data(), hb.data(), (int)ite, E, z.data());
      load_start(prec, u0, dim, start, seed, st);
      if (fused) {
        // the second pass repeats the steps with the stored alpha and beta;
        // step i writes v_i itself (normalised) into U[2 + i % 2]
        for (int64_t i = 0; i < ite; ++i) {
          step(i, nullptr);
          vec_axpy_acc(prec, p64.p, U[2 + i % 2].p, dim, z[i], sc.p, -1, st);
        }
      } else {
        vec_axpy_acc(prec, p64.p, U[0].p, dim, z[0], sc.p, (int)SCALE, st);
        for (int64_t i = 0; i + 1 < ite; ++i) {
          step(i, nullptr);
          vec_axpy_acc(prec, p64.p, U[(i + 1) % 3].p, dim, z[i + 1], sc.p, (int)(SCALE + i + 1), st);
        }
      }
      nb = vec_dot(QS3X_FP64, p64.p, QS3X_FP64, p64.p, dim, part.p, st);
      finalize(part.p, nb, sc.p, 2, 0, 1, 0.0, st);
      vec_copy_scale(QS3X_FP64, p64.p, QS3X_FP64, p64.p, dim, sc.p, 1, 1.0, st);
      p64.download(psi, (size_t)2 * dim);
    }
    QS3X_HIP(hipStreamSynchronize(st));
  });
}

}  // extern "C"

// -------------------------------------------------------- thick restart
namespace {

struct TR {
  qs3x_sector *s;
  int prec;
  int NOK, NOM;
  int64_t dim;
  double shift;
  hipStream_t st;
  DevBuf<unsigned char> basis;  // (NOM+1) x dim, slot j at j*dim
  DevBuf<double> sc;            // [0,NOM): alpha' ; [NOM,2NOM): beta ; [2NOM, 3NOM+1): scales ; tmp
  DevBuf<double> part, h, Cdev;
  // per-step reorthogonalisation sets: tab[(phase * NOM + i) * (NOM + 1) + j]
  // (phase 0: first cycle, 1: restart cycles), cnt[phase * NOM + i] entries
  DevBuf<const void *> ptab;
  DevBuf<int> stab;
  std::vector<int> cnt;
  int64_t hv_count = 0;
  int64_t hv_base = 0;   // H.v applications of earlier phases (for the trace)
  double t_origin = 0;   // start of the solve (for the trace)
  double alloc_s = 0;    // seconds spent allocating the basis in init()
  int cgs_passes = 2;
  int reorth = 0;

  int A() const { return 0; }
  int Bt() const { return NOM; }
  int S() const { return 2 * NOM; }
  int TMP() const { return 3 * NOM + 2; }

  void *slot(int j) { return basis.p + (size_t)j * dim * elem_bytes(prec); }

  // reorth_ / cgs_ < 0: take tr_reorth / tr_cgs_passes from the options
  void init(qs3x_sector *s_, int prec_, int nok, int nom, int reorth_ = -1, int cgs_ = -1) {
    s = s_;
    prec = prec_;
    NOK = nok;
    NOM = nom;
    dim = s->dim;
    st = stream();
    shift = (prec == QS3X_FP32) ? s->model->ezz : 0.0;
    {
      const double ta = now_s();
      basis.alloc((size_t)(NOM + 1) * dim * elem_bytes(prec));
      alloc_s = now_s() - ta;
      if (options().verbose)
        std::fprintf(stderr, "[qs3x] thick-restart basis: %d x %lld %s elements, %.1f GB, allocated in %.3f s\n",
                     NOM + 1, (long long)dim, prec == QS3X_FP32 ? "complex-float" : "complex-double",
                     1e-9 * (double)(NOM + 1) * (double)dim * (double)elem_bytes(prec), now_s() - ta);
    }
    sc.alloc((size_t)3 * NOM + 8);
    const int nb = vec_grid(dim);
    part.alloc((size_t)2 * std::max(nb, 1) * (NOM + 2));
    h.alloc((size_t)2 * (NOM + 2));
    Cdev.alloc((size_t)NOM * NOM);
    cgs_passes = std::max(1, cgs_ < 0 ? options().tr_cgs_passes : cgs_);
    reorth = reorth_ < 0 ? options().tr_reorth : reorth_;
    t_origin = now_s();
    // Step i writes slot i+1 and orthogonalises it against: every slot 0..i
    // (reorth 0, the reference's full reorthogonalisation), or the locked Ritz
    // vectors 0..NOK-1 (restart cycles) plus slots i-1, i (reorth 1).
    const int W = NOM + 1;
    std::vector<const void *> pt((size_t)2 * NOM * W, nullptr);
    std::vector<int> sx((size_t)2 * NOM * W, 0);
    cnt.assign((size_t)2 * NOM, 0);
    for (int phase = 0; phase < 2; ++phase)
      for (int i = 0; i < NOM; ++i) {
        std::vector<int> set;
        if (reorth == 0) {
          for (int j = 0; j <= i; ++j) set.push_back(j);
        } else {
          if (phase == 1)
            for (int j = 0; j < NOK && j <= i; ++j) set.push_back(j);
          for (int j = std::max(0, i - 1); j <= i; ++j)
            if (std::find(set.begin(), set.end(), j) == set.end()) set.push_back(j);
        }
        cnt[(size_t)phase * NOM + i] = (int)set.size();
        for (size_t q = 0; q < set.size(); ++q) {
          pt[((size_t)phase * NOM + i) * W + q] = slot(set[q]);
          sx[((size_t)phase * NOM + i) * W + q] = set[q];
        }
      }
    ptab.alloc(pt.size());
    ptab.upload(pt.data(), pt.size());
    stab.alloc(sx.size());
    stab.upload(sx.data(), sx.size());
    QS3X_HIP(hipStreamSynchronize(st));
  }

  // slot j := v / |v| (scale slot set to 1/|v|, data unchanged)
  void set_norm_scale(int j) {
    int nb = vec_dot(prec, slot(j), prec, slot(j), dim, part.p, st);
    finalize(part.p, nb, sc.p, 2, TMP(), S() + j, 0.0, st);
  }

  // One Lanczos step from slot i into slot i+1, then classical Gram-Schmidt
  // (cgs_passes times, FP64 coefficients) against the step's slot set.  In
  // restart cycles (phase 1) at least two passes: the new vector must stay
  // orthogonal to the converged locked Ritz vectors, which one pass does not
  // keep over many restarts (measured: FP64 at L = 1000, NOM 128 ended far
  // outside the spectrum with one pass, at the ground state with two).
  void step(int i, int phase) {
    int nb = 0;
    launch_hv(s, prec, slot(i + 1), slot(i), sc.p + S() + i, part.p, &nb, st);
    ++hv_count;
    finalize(part.p, nb, sc.p, 1, A() + i, TMP(), shift, st);
    const size_t row = (size_t)phase * NOM + i;
    const int m = cnt[row];
    const void *const *P = (const void *const *)(ptab.p + row * (NOM + 1));
    const int *X = stab.p + row * (NOM + 1);
    const int passes = phase == 1 ? std::max(2, cgs_passes) : cgs_passes;
    for (int pass = 0; pass < passes; ++pass) {
      vec_mdot(prec, P, sc.p + S(), X, m, slot(i + 1), dim, part.p, h.p, st);
      nb = vec_maxpy(prec, P, sc.p + S(), X, m, h.p, slot(i + 1), dim, part.p, st);
    }
    finalize(part.p, nb, sc.p, 2, Bt() + i, S() + i + 1, 0.0, st);
  }

  // One Lanczos step from slot i into slot i+1 by the three-term recurrence
  // (no Gram-Schmidt): w = H v_i - alpha_i v_i - beta_{i-1} v_{i-1}, with
  // alpha from the fused dot of the H.v and beta_{i-1} from the previous step.
  void step3(int i) {
    int nb = 0;
    launch_hv(s, prec, slot(i + 1), slot(i), sc.p + S() + i, part.p, &nb, st);
    ++hv_count;
    finalize(part.p, nb, sc.p, 1, A() + i, TMP(), shift, st);
    nb = vec_update3(prec, slot(i + 1), slot(i), i > 0 ? slot(i - 1) : slot(i), dim, sc.p, A() + i, S() + i,
                     i > 0 ? Bt() + i - 1 : -1, i > 0 ? S() + i - 1 : -1, part.p, st);
    finalize(part.p, nb, sc.p, 2, Bt() + i, S() + i + 1, 0.0, st);
  }

  // alpha'[i0..i1), beta[i0..i1) to host
  void read_ab(int i0, int i1, std::vector<double> &a, std::vector<double> &b) {
    std::vector<double> all(sc.n);
    sc.download(all.data(), all.size());
    a.assign(all.begin() + A() + i0, all.begin() + A() + i1);
    b.assign(all.begin() + Bt() + i0, all.begin() + Bt() + i1);
  }

  // slots 0..NOK-1 := V[0..m-1] Z (Z row-major m x NOK on host, m <= NOM)
  void rotate(const std::vector<double> &Z, int m) {
    Cdev.upload(Z.data(), (size_t)m * NOK);
    ritz_rotate(prec, basis.p, dim, sc.p + S(), m, NOK, Cdev.p, st);
    std::vector<double> ones(NOK, 1.0);
    QS3X_HIP(hipMemcpyAsync(sc.p + S(), ones.data(), sizeof(double) * NOK, hipMemcpyHostToDevice, st));
    QS3X_HIP(hipStreamSynchronize(st));
  }

  void copy_slot(int dst, int src) {
    QS3X_HIP(hipMemcpyAsync(slot(dst), slot(src), dim * elem_bytes(prec), hipMemcpyDeviceToDevice, st));
    QS3X_HIP(hipMemcpyAsync(sc.p + S() + dst, sc.p + S() + src, sizeof(double), hipMemcpyDeviceToDevice, st));
  }
};

// Thick-restart cycles (eigen_solver.f90:91-216).  On entry slot 0 holds the
// start vector (any norm).  Returns the number of restarts; theta receives the
// NOM Ritz values (shifted back by e_zz), slots 0..NOK-1 the Ritz vectors.
struct TRResult {
  int restarts = 0;
  bool converged = false;
  double max_change = 0.0;
};

// abs_floor > 0 (FP32 phase): converged when the nwant lowest Ritz values
// moved less than abs_floor (or tol relative) over two consecutive restarts;
// abs_floor = 0: the reference's test over all NOK values (eigen_solver.f90:150-162).
// Lowest nok eigenpairs of the projection (k locked values w with couplings
// cb to row k, tridiagonal tail a, b; m rows): Jacobi on the dense matrix for
// small m (the path of the reference's zheevr call), the structured Sturm /
// inverse-iteration solver of eig.cpp otherwise.  Z is row-major m x nok.
void projected_pairs(int k, int m, const std::vector<double> &w, const std::vector<double> &cb,
                     const std::vector<double> &a, const std::vector<double> &b, int nok, std::vector<double> &lam,
                     std::vector<double> &Z) {
  lam.assign(nok, 0.0);
  Z.assign((size_t)m * nok, 0.0);
  if (m <= 96) {
    std::vector<double> T((size_t)m * m, 0.0), C((size_t)m * m), ww(m);
    for (int j = 0; j < k; ++j) {
      T[(size_t)j * m + j] = w[j];
      T[(size_t)j * m + k] = T[(size_t)k * m + j] = cb[j];
    }
    for (int i = k; i < m; ++i) {
      T[(size_t)i * m + i] = a[i - k];
      if (i + 1 < m) T[(size_t)i * m + i + 1] = T[(size_t)(i + 1) * m + i] = b[i - k];
    }
    sym_eig(m, T.data(), ww.data(), C.data());
    for (int j = 0; j < nok; ++j) {
      lam[j] = ww[j];
      for (int i = 0; i < m; ++i) Z[(size_t)i * nok + j] = C[(size_t)i * m + j];
    }
    return;
  }
  Arrow A{k, m, k ? w.data() : nullptr, k ? cb.data() : nullptr, a.data(), b.data()};
  arrow_lowest_pairs(A, nok, lam.data(), Z.data());
}

TRResult tr_cycles(TR &t, int64_t maxitr, double tol, double abs_floor, std::vector<double> &theta,
                   std::vector<double> &resid, int nwant = 0) {
  const int NOK = t.NOK, NOM = t.NOM;
  const int ncheck = (abs_floor > 0 && nwant > 0) ? std::min(nwant, NOK) : NOK;
  TRResult r;
  t.set_norm_scale(0);
  for (int i = 0; i < NOM; ++i) t.step(i, 0);
  std::vector<double> a, b;
  t.read_ab(0, NOM, a, b);
  std::vector<double> w, C, none;
  double beta_last = b[NOM - 1];
  projected_pairs(0, NOM, none, none, a, b, NOK, w, C);
  std::vector<double> prev(a.begin(), a.begin() + NOK);  // calpha(1:NOK) = alpha (eigen_solver.f90:293-294)
  int flag = 0;
  int64_t tritr = 0;
  for (tritr = 1; tritr <= maxitr; ++tritr) {
    t.rotate(C, NOM);
    double mc = 0.0;
    for (int j = 0; j < ncheck; ++j) {
      const double e = w[j] + t.shift, pe = prev[j] + t.shift;
      double ch = std::fabs(pe / e - 1.0);
      if (abs_floor > 0) ch = std::min(ch, std::fabs(pe - e) / abs_floor * tol);
      mc = std::max(mc, ch);
    }
    r.max_change = mc;
    trace_point(t.t_origin, t.hv_base + t.hv_count, w[0] + t.shift, t.prec);
    if (options().verbose)
      std::fprintf(stderr, "[qs3x] thick restart %lld (%s): theta0=%.15g change %.3e, %lld H.v\n",
                   (long long)tritr, t.prec == QS3X_FP32 ? "fp32" : "fp64", w[0] + t.shift, mc,
                   (long long)t.hv_count);
    flag = (mc < tol) ? flag + 1 : 0;
    if (flag == 2) {
      r.converged = true;
      break;
    }
    if (tritr == maxitr) break;
    t.copy_slot(NOK, NOM);
    for (int j = 0; j < NOK; ++j) prev[j] = w[j];
    std::vector<double> cb(NOK), wl(w.begin(), w.begin() + NOK);
    for (int j = 0; j < NOK; ++j) cb[j] = beta_last * C[(size_t)(NOM - 1) * NOK + j];
    for (int i = NOK; i < NOM; ++i) t.step(i, 1);
    t.read_ab(NOK, NOM, a, b);
    beta_last = b[NOM - 1 - NOK];
    projected_pairs(NOK, NOM, wl, cb, a, b, NOK, w, C);
  }
  r.restarts = (int)std::min<int64_t>(tritr, maxitr);
  theta.resize(NOK);
  for (int j = 0; j < NOK; ++j) theta[j] = w[j] + t.shift;
  resid.resize(NOK);
  for (int j = 0; j < NOK; ++j) resid[j] = std::fabs(beta_last * C[(size_t)(NOM - 1) * NOK + j]);
  return r;
}

// Thick restart with convergence checks inside the cycles (an extension of
// eigen_solver.f90:91-216, used by the FP32 stage of the mixed solver when
// option tr_check_every > 0): every itrint steps the nwant lowest eigenvalues
// of the projection so far (structured Sturm bisection, O(m) per count) are
// compared with the previous check; when they moved less than tol (relative)
// or abs_floor (absolute; 0 disables) and the extrapolated remaining change
// is below tol too, the cycle ends there.  Then (or at NOM)
// the NOK lowest Ritz pairs are rotated into slots 0..NOK-1 as in tr_cycles.
// In the first cycle, with tr_reorth >= 1, each step is the three-term
// recurrence (no Gram-Schmidt); restart cycles keep the locked + local CGS of
// step().
TRResult tr_cycles_checked(TR &t, int64_t maxitr, double tol, double abs_floor, int itrint,
                           std::vector<double> &theta, std::vector<double> &resid, int nwant) {
  const int NOK = t.NOK, NOM = t.NOM;
  const int nw = std::max(1, std::min(nwant, NOK));
  TRResult r;
  t.set_norm_scale(0);
  std::vector<double> wl, cb, a, b, lam, Z, prev;
  double beta_last = 0.0, mc_prev = -1.0;
  // The checks run one interval behind the device, as in qs3x_lanczos: at a
  // check point the coefficients so far are copied to pinned host memory and
  // an event recorded; the previous check is evaluated once its event has
  // completed, while the next steps run.  A cycle that met the test at step
  // m ends at m (m_end), although up to itrint more steps were launched: the
  // restart only uses slots 0..m and the coefficients before m.
  struct PendCheck {
    int m = 0;
    int64_t hv = 0;
    double *host = nullptr;
    hipEvent_t ev = nullptr;
  } pc[2];
  for (auto &p : pc) {
    QS3X_HIP(hipHostMalloc((void **)&p.host, sizeof(double) * 2 * (size_t)NOM, hipHostMallocDefault));
    QS3X_HIP(hipEventCreateWithFlags(&p.ev, hipEventDisableTiming));
  }
  struct PendFree {
    PendCheck *p;
    ~PendFree() {
      for (int k = 0; k < 2; ++k) {
        if (p[k].host) (void)hipHostFree(p[k].host);
        if (p[k].ev) (void)hipEventDestroy(p[k].ev);
      }
    }
  } pend_free{pc};
  int k0 = 0;
  int64_t cycle = 0;
  for (cycle = 1; cycle <= maxitr; ++cycle) {
    const int i0 = k0;
    int m_end = NOM;
    bool conv = false;
    int npend = 0, slot = 0;
    // evaluates a recorded check; true when the cycle may end at its step
    auto evaluate = [&](PendCheck &p) -> bool {
      QS3X_HIP(hipEventSynchronize(p.ev));
      const int m = p.m;
      a.assign(p.host, p.host + (m - i0));
      b.assign(p.host + NOM, p.host + NOM + (m - i0));
      Arrow A{k0, m, k0 ? wl.data() : nullptr, k0 ? cb.data() : nullptr, a.data(), b.data()};
      std::vector<double> ev(nw);
      for (int j = 0; j < nw; ++j) ev[j] = arrow_eigenvalue(A, j);
      const bool first_check = prev.empty();
      double mc = first_check ? 1.0 : 0.0;
      for (int j = 0; j < nw && !prev.empty(); ++j) {
        const double e = ev[j] + t.shift, pe = prev[j] + t.shift;
        double ch = std::fabs(pe / e - 1.0);
        if (abs_floor > 0) ch = std::min(ch, std::fabs(pe - e) / abs_floor * tol);
        mc = std::max(mc, ch);
      }
      prev = ev;
      trace_point(t.t_origin, p.hv, ev[0] + t.shift, t.prec);
      if (options().verbose > 1)
        std::fprintf(stderr, "[qs3x] cycle %lld step %d: theta0=%.15g change %.3e\n", (long long)cycle, m,
                     ev[0] + t.shift, mc);
      r.max_change = mc;
      // Stop when the change is below tol and so is the remaining error
      // estimated from the last two changes (geometric tail
      // change * rho / (1 - rho), rho = change / previous change), or when
      // the change is below tol / 10 (fast convergence or the Ritz values
      // have reached the working precision's resolution).
      bool stop = false;
      if (mc < tol) {
        if (mc < 0.1 * tol) stop = true;
        else if (mc_prev > 0.0) {
          const double rho = mc / mc_prev;
          stop = rho < 1.0 && mc * rho / (1.0 - rho) < tol;
        }
      }
      mc_prev = first_check ? -1.0 : mc;
      return stop;
    };
    for (int i = i0; i < NOM; ++i) {
      // tr_reorth >= 1: the first cycle is the three-term recurrence; restart
      // cycles orthogonalise every step against the locked Ritz vectors and
      // the last two Lanczos vectors.  (Three-term steps in restart cycles
      // lose orthogonality to the converged locked vectors over many
      // restarts, and the arrowhead projection, which assumes it, then gives
      // Ritz values below the spectrum: measured at L = 1000, NOM 128.)
      if (k0 == 0 && t.reorth >= 1) t.step3(i);
      else t.step(i, k0 == 0 ? 0 : 1);
      const int m = i + 1;
      if ((m - i0) % itrint != 0 || m < NOK + 1) continue;
      PendCheck &p = pc[slot];
      p.m = m;
      p.hv = t.hv_base + t.hv_count;
      QS3X_HIP(hipMemcpyAsync(p.host, t.sc.p + t.A() + i0, sizeof(double) * (m - i0), hipMemcpyDeviceToHost, t.st));
      QS3X_HIP(hipMemcpyAsync(p.host + NOM, t.sc.p + t.Bt() + i0, sizeof(double) * (m - i0), hipMemcpyDeviceToHost,
                              t.st));
      QS3X_HIP(hipEventRecord(p.ev, t.st));
      slot ^= 1;
      if (++npend == 2) {  // the previous check, while this one is in flight
        --npend;
        if (evaluate(pc[slot])) {
          conv = true;
          m_end = pc[slot].m;
          npend = 0;
          break;
        }
      }
    }
    if (!conv && npend == 1 && evaluate(pc[slot ^ 1])) {
      conv = true;
      m_end = pc[slot ^ 1].m;
    }
    QS3X_HIP(hipStreamSynchronize(t.st));
    t.read_ab(i0, m_end, a, b);
    beta_last = b[m_end - 1 - i0];
    projected_pairs(k0, m_end, wl, cb, a, b, NOK, lam, Z);
    t.rotate(Z, m_end);
    if (options().verbose)
      std::fprintf(stderr, "[qs3x] cycle %lld (%s): %d steps, theta0=%.15g, %lld H.v%s\n", (long long)cycle,
                   t.prec == QS3X_FP32 ? "fp32" : "fp64", m_end - i0, lam[0] + t.shift, (long long)t.hv_count,
                   conv ? ", converged" : "");
    if (conv) {
      r.converged = true;
      break;
    }
    if (cycle == maxitr) break;
    // restart: the residual vector becomes slot NOK, the locked Ritz values
    // couple to it through beta_last * z_{m-1, j}
    t.copy_slot(NOK, m_end);
    wl.assign(lam.begin(), lam.begin() + NOK);
    cb.resize(NOK);
    for (int j = 0; j < NOK; ++j) cb[j] = beta_last * Z[(size_t)(m_end - 1) * NOK + j];
    k0 = NOK;
    prev.clear();
    mc_prev = -1.0;
  }
  r.restarts = (int)std::min<int64_t>(cycle, maxitr);
  theta.resize(NOK);
  for (int j = 0; j < NOK; ++j) theta[j] = lam[j] + t.shift;
  const int m_last = (int)(Z.size() / NOK);
  resid.resize(NOK);
  for (int j = 0; j < NOK; ++j) resid[j] = std::fabs(beta_last * Z[(size_t)(m_last - 1) * NOK + j]);
  return r;
}

// FP64 Lanczos refinement of one Ritz vector (option fp64_lanczos): the
// plain FP64 recurrence of lanczos.f90:50-92 (fused steps where they apply)
// started from y0 (device, complex double), without storing its basis.
// Every 5 steps from step 10 the tridiagonal's lowest eigenvalue theta and
// the residual norm of its Ritz vector, r = beta_m |s_m| (beta_m the last
// step's beta, s_m the last component of the normalised tridiagonal
// eigenvector: an identity of the recurrence, no vector needed), give the
// Kato-Temple bound r^2 / gap; the run stops when it is <= conv |theta| or
// after maxsteps.  With out (device, complex double) the Ritz vector of the
// accepted step is rebuilt by a second pass of the same steps.
struct Lz64Result {
  double theta = 0.0, resid = 0.0;
  int64_t steps = 0;
  bool certified = false;
};

Lz64Result lanczos_refine64(qs3x_sector *s, const void *y0, int64_t maxsteps, double gap, double conv,
                            double *out, hipStream_t st, double t_origin, int64_t hv_base) {
  const int prec = QS3X_FP64;
  const int64_t dim = s->dim;
  const int64_t mitr = std::max<int64_t>(std::min<int64_t>(maxsteps, dim), 1);
  const bool fused = hv_fused_available(s, prec);
  DevBuf<unsigned char> U[4];
  for (int k = 0; k < (fused ? 4 : 3); ++k) U[k].alloc(16 * (size_t)dim);
  const int64_t A_REC = 0, A_OUT = mitr, BETA = 2 * mitr, SCALE = 3 * mitr + 1;
  DevBuf<double> sc((size_t)4 * mitr + 4);
  DevBuf<double> part((size_t)4 * kMaxBlocks);
  const double shift = fused ? s->model->ezz : 0.0;
  void *u0 = fused ? U[1].p : U[0].p;
  auto start = [&] {
    vec_copy_scale(prec, u0, QS3X_FP64, y0, dim, nullptr, -1, 1.0, st);
    const int nb = vec_dot(prec, u0, prec, u0, dim, part.p, st);
    finalize(part.p, nb, sc.p, 2, (int)(SCALE + mitr + 1), (int)SCALE, 0.0, st);
  };
  int nb = 0;
  auto step = [&](int64_t i, double *pt) {
    if (fused) {
      void *pprev = U[(i + 1) % 2].p, *pout = U[i % 2].p;
      void *vprev = i > 0 ? U[2 + (i + 1) % 2].p : nullptr, *xout = U[2 + i % 2].p;
      QS3X_REQUIRE(launch_hv_fused(s, prec, pout, xout, pprev, vprev, sc.p, (int)(SCALE + i), (int)(A_REC + i - 1),
                                   i > 0 ? (int)(BETA + i - 1) : -1, pt, &nb, st),
                   "lanczos_refine64: fused step unavailable");
      if (pt) finalize_fused(pt, nb, sc.p, (int)(A_REC + i), (int)(A_OUT + i), (int)(BETA + i), (int)(SCALE + i + 1),
                             shift, st);
      return;
    }
    void *cur = U[i % 3].p, *prev = U[(i + 2) % 3].p, *w = U[(i + 1) % 3].p;
    launch_hv(s, prec, w, cur, sc.p + SCALE + i, pt, &nb, st);
    if (pt) finalize(pt, nb, sc.p, 1, (int)(A_REC + i), (int)(A_OUT + i), shift, st);
    nb = vec_update3(prec, w, cur, prev, dim, sc.p, (int)(A_REC + i), (int)(SCALE + i),
                     i > 0 ? (int)(BETA + i - 1) : -1, (int)(SCALE + i - 1), pt, st);
    if (pt) finalize(pt, nb, sc.p, 2, (int)(BETA + i), (int)(SCALE + i + 1), 0.0, st);
  };
  Lz64Result r;
  std::vector<double> ha, hb, z;
  start();
  for (int64_t i = 0; i < mitr; ++i) {
    step(i, part.p);
    const int64_t n = i + 1;
    if (!(n == mitr || (n >= 10 && n % 5 == 0))) continue;
    ha.resize(n);
    hb.resize(n);
    QS3X_HIP(hipMemcpyAsync(ha.data(), sc.p + A_OUT, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    QS3X_HIP(hipMemcpyAsync(hb.data(), sc.p + BETA, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    QS3X_HIP(hipStreamSynchronize(st));
    const double E = n == 1 ? ha[0] : tridiag_lowest(ha.data(), hb.data(), (int)n);
    z.assign(n, 0.0);
    if (n == 1) z[0] = 1.0;
    else tridiag_vector(ha.data(), hb.data(), (int)n, E, z.data());
    double zn = 0.0;
    for (double x : z) zn += x * x;
    const double res = hb[n - 1] * std::fabs(z[n - 1]) / std::sqrt(std::max(zn, 1e-300));
    r.theta = E;
    r.resid = res;
    r.steps = n;
    trace_point(t_origin, hv_base + n, E, QS3X_FP64);
    if (options().verbose)
      std::fprintf(stderr, "[qs3x] fp64 lanczos refinement step %lld: theta %.15g residual %.3e bound %.3e\n",
                   (long long)n, E, res, res * res / std::max(gap, 1e-300));
    if (gap > 0.0 && res * res / gap <= conv * std::fabs(E)) {
      r.certified = true;
      break;
    }
  }
  if (out) {
    // second pass: psi = sum_i z_i v_i over the accepted steps
    const int64_t n = r.steps;
    z.assign(n, 0.0);
    if (n == 1) z[0] = 1.0;
    else tridiag_vector(ha.data(), hb.data(), (int)n, r.theta, z.data());
    QS3X_HIP(hipMemsetAsync(out, 0, 16 * (size_t)dim, st));
    start();
    if (fused) {
      for (int64_t i = 0; i < n; ++i) {
        step(i, nullptr);
        vec_axpy_acc(prec, out, U[2 + i % 2].p, dim, z[i], sc.p, -1, st);
      }
    } else {
      vec_axpy_acc(prec, out, U[0].p, dim, z[0], sc.p, (int)SCALE, st);
      for (int64_t i = 0; i + 1 < n; ++i) {
        step(i, nullptr);
        vec_axpy_acc(prec, out, U[(i + 1) % 3].p, dim, z[i + 1], sc.p, (int)(SCALE + i + 1), st);
      }
    }
    const int nb2 = vec_dot(QS3X_FP64, out, QS3X_FP64, out, dim, part.p, st);
    finalize(part.p, nb2, sc.p, 2, 0, 1, 0.0, st);
    vec_copy_scale(QS3X_FP64, out, QS3X_FP64, out, dim, sc.p, 1, 1.0, st);
  }
  QS3X_HIP(hipStreamSynchronize(st));
  return r;
}

// FP64 Rayleigh-Ritz on k vectors: slots 0..k-1 of t (FP64) hold the
// candidate vectors (any scale).  They are orthonormalised (CGS2), W = H Y is
// formed in slots k..2k-1, G = Y^H W is diagonalised on the host (as the real
// symmetric 2k x 2k form of the Hermitian matrix), and for the nwant lowest
// pairs the residual norms |W c - theta Y c| are computed.  On return slot
// 2k + j (j < nwant) holds the refined Ritz vector y_j = Y c_j (scale 1).
// Requires 2k + nwant <= NOM + 1.
struct RRResult {
  std::vector<double> theta, resid;
};

RRResult rayleigh_ritz(TR &t, int k, int nwant) {
  hipStream_t st = t.st;
  const int64_t dim = t.dim;
  QS3X_REQUIRE(2 * k + nwant <= t.NOM + 1, "rayleigh_ritz: basis too small");
  std::vector<const void *> hp(2 * k + 1);
  for (int j = 0; j < 2 * k; ++j) hp[j] = t.slot(j);
  DevBuf<const void *> P(2 * k);
  P.upload(hp.data(), 2 * k);
  std::vector<int> hs(2 * k);
  for (int j = 0; j < 2 * k; ++j) hs[j] = j;
  DevBuf<int> X(2 * k);
  X.upload(hs.data(), 2 * k);
  const void *const *PY = (const void *const *)P.p;
  const void *const *PW = (const void *const *)(P.p + k);
  // orthonormalise Y (CGS2), scales in sc[S()+j]
  int nb = 0;
  for (int j = 0; j < k; ++j) {
    if (j > 0)
      for (int pass = 0; pass < 2; ++pass) {
        vec_mdot(QS3X_FP64, PY, t.sc.p + t.S(), X.p, j, t.slot(j), dim, t.part.p, t.h.p, st);
        vec_maxpy(QS3X_FP64, PY, t.sc.p + t.S(), X.p, j, t.h.p, t.slot(j), dim, nullptr, st);
      }
    std::vector<double> one(1, 1.0);
    QS3X_HIP(hipMemcpyAsync(t.sc.p + t.S() + j, one.data(), sizeof(double), hipMemcpyHostToDevice, st));
    nb = vec_dot(QS3X_FP64, t.slot(j), QS3X_FP64, t.slot(j), dim, t.part.p, st);
    finalize(t.part.p, nb, t.sc.p, 2, t.TMP(), t.S() + j, 0.0, st);
    QS3X_HIP(hipStreamSynchronize(st));
  }
  // W_j = H y_j (y_j = s_j Y_j), W scale 1
  for (int j = 0; j < k; ++j) {
    launch_hv(t.s, QS3X_FP64, t.slot(k + j), t.slot(j), t.sc.p + t.S() + j, nullptr, nullptr, st);
    ++t.hv_count;
  }
  {
    std::vector<double> ones(k, 1.0);
    QS3X_HIP(hipMemcpyAsync(t.sc.p + t.S() + k, ones.data(), sizeof(double) * k, hipMemcpyHostToDevice, st));
  }
  // G = Y^H W (complex k x k), column j from one multi-dot
  std::vector<double> G((size_t)2 * k * k), hcol(2 * k);
  for (int j = 0; j < k; ++j) {
    vec_mdot(QS3X_FP64, PY, t.sc.p + t.S(), X.p, k, t.slot(k + j), dim, t.part.p, t.h.p, st);
    QS3X_HIP(hipMemcpyAsync(hcol.data(), t.h.p, sizeof(double) * 2 * k, hipMemcpyDeviceToHost, st));
    QS3X_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < k; ++i) {
      G[2 * ((size_t)i * k + j)] = hcol[2 * i];
      G[2 * ((size_t)i * k + j) + 1] = hcol[2 * i + 1];
    }
  }
  // Hermitian part as the real symmetric [[Re, -Im], [Im, Re]]
  const int n2 = 2 * k;
  std::vector<double> A((size_t)n2 * n2), wv(n2), V((size_t)n2 * n2);
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) {
      const double re = 0.5 * (G[2 * ((size_t)i * k + j)] + G[2 * ((size_t)j * k + i)]);
      const double im = 0.5 * (G[2 * ((size_t)i * k + j) + 1] - G[2 * ((size_t)j * k + i) + 1]);
      A[(size_t)i * n2 + j] = re;
      A[(size_t)(i + k) * n2 + (j + k)] = re;
      A[(size_t)i * n2 + (j + k)] = -im;
      A[(size_t)(i + k) * n2 + j] = im;
    }
  sym_eig(n2, A.data(), wv.data(), V.data());
  RRResult r;
  r.theta.resize(k);
  std::vector<std::vector<double>> cvec(k, std::vector<double>(2 * k));
  for (int j = 0; j < k; ++j) {
    // eigenvalues come in equal pairs; column 2j gives c = u + i v
    r.theta[j] = wv[2 * j];
    double nrm = 0.0;
    for (int i = 0; i < k; ++i) {
      const double u = V[(size_t)i * n2 + 2 * j], v = V[(size_t)(i + k) * n2 + 2 * j];
      cvec[j][2 * i] = u;
      cvec[j][2 * i + 1] = v;
      nrm += u * u + v * v;
    }
    nrm = std::sqrt(nrm);
    for (double &x : cvec[j]) x /= nrm;
  }
  // residuals and Ritz vectors of the nwant lowest pairs
  r.resid.assign(nwant, 0.0);
  const void *const *PYW = (const void *const *)P.p;  // Y then W
  DevBuf<double> hh((size_t)4 * k);
  for (int j = 0; j < nwant; ++j) {
    void *z = t.slot(2 * k + j);
    // z = sum_i c_ij W_i - theta_j sum_i c_ij Y_i  (maxpy: z -= sum h_i s_i u_i)
    std::vector<double> h((size_t)4 * k);
    for (int i = 0; i < k; ++i) {
      h[2 * i] = r.theta[j] * cvec[j][2 * i];
      h[2 * i + 1] = r.theta[j] * cvec[j][2 * i + 1];
      h[2 * (k + i)] = -cvec[j][2 * i];
      h[2 * (k + i) + 1] = -cvec[j][2 * i + 1];
    }
    hh.upload(h.data(), h.size());
    QS3X_HIP(hipMemsetAsync(z, 0, 16 * dim, st));
    nb = vec_maxpy(QS3X_FP64, PYW, t.sc.p + t.S(), X.p, 2 * k, hh.p, z, dim, t.part.p, st);
    finalize(t.part.p, nb, t.sc.p, 3, t.TMP(), -1, 0.0, st);
    std::vector<double> sc1(1);
    QS3X_HIP(hipMemcpyAsync(sc1.data(), t.sc.p + t.TMP(), sizeof(double), hipMemcpyDeviceToHost, st));
    QS3X_HIP(hipStreamSynchronize(st));
    r.resid[j] = sc1[0];
    // y_j = sum_i c_ij Y_i
    for (int i = 0; i < k; ++i) {
      h[2 * i] = -cvec[j][2 * i];
      h[2 * i + 1] = -cvec[j][2 * i + 1];
    }
    hh.upload(h.data(), 2 * k);
    QS3X_HIP(hipMemsetAsync(z, 0, 16 * dim, st));
    vec_maxpy(QS3X_FP64, PY, t.sc.p + t.S(), X.p, k, hh.p, z, dim, nullptr, st);
    std::vector<double> one(1, 1.0);
    QS3X_HIP(hipMemcpyAsync(t.sc.p + t.S() + 2 * k + j, one.data(), sizeof(double), hipMemcpyHostToDevice, st));
    QS3X_HIP(hipStreamSynchronize(st));
  }
  return r;
}

// Final FP64 refinement of p accepted Ritz vectors (slots src0 .. src0+p-1,
// src0 >= p): Rayleigh-Ritz on the block Krylov space span{Y, HY, ..., H^S Y}.
// Each new block vector is H applied to the one p slots back, orthonormalised
// at once (CGS2) against every earlier slot, so the basis stays well
// conditioned.  The space ends early when a new vector lies in the span of the
// earlier ones (norm after CGS2 below 1e-10 of its norm before).  Slots 0..k-1
// hold the basis; the refined vectors land in slots 2k + j (see rayleigh_ritz).
struct RefineResult {
  RRResult rr;
  int k;
};

RefineResult refine_block_krylov(TR &t, int src0, int p, int S) {
  hipStream_t st = t.st;
  const int64_t dim = t.dim;
  int kmax = (S + 1) * p;
  while (kmax > p && (2 * kmax + p > t.NOM + 1 || kmax > dim)) kmax -= p;
  QS3X_REQUIRE(src0 >= p && 2 * kmax + p <= t.NOM + 1, "refine_block_krylov: basis too small");
  for (int j = 0; j < p; ++j) t.copy_slot(j, src0 + j);
  std::vector<const void *> hp(kmax);
  std::vector<int> hs(kmax);
  for (int j = 0; j < kmax; ++j) {
    hp[j] = t.slot(j);
    hs[j] = j;
  }
  DevBuf<const void *> P(kmax);
  P.upload(hp.data(), kmax);
  DevBuf<int> X(kmax);
  X.upload(hs.data(), kmax);
  const void *const *PV = (const void *const *)P.p;
  int k = kmax;
  const std::vector<double> one(1, 1.0);
  for (int i = p; i < kmax; ++i) {
    launch_hv(t.s, QS3X_FP64, t.slot(i), t.slot(i - p), t.sc.p + t.S() + i - p, nullptr, nullptr, st);
    ++t.hv_count;
    int nb = vec_dot(QS3X_FP64, t.slot(i), QS3X_FP64, t.slot(i), dim, t.part.p, st);
    finalize(t.part.p, nb, t.sc.p, 3, t.TMP() + 1, -1, 0.0, st);
    for (int pass = 0; pass < 2; ++pass) {
      vec_mdot(QS3X_FP64, PV, t.sc.p + t.S(), X.p, i, t.slot(i), dim, t.part.p, t.h.p, st);
      vec_maxpy(QS3X_FP64, PV, t.sc.p + t.S(), X.p, i, t.h.p, t.slot(i), dim, nullptr, st);
    }
    QS3X_HIP(hipMemcpyAsync(t.sc.p + t.S() + i, one.data(), sizeof(double), hipMemcpyHostToDevice, st));
    nb = vec_dot(QS3X_FP64, t.slot(i), QS3X_FP64, t.slot(i), dim, t.part.p, st);
    finalize(t.part.p, nb, t.sc.p, 2, t.TMP(), t.S() + i, 0.0, st);
    double nrm[2];
    QS3X_HIP(hipMemcpyAsync(nrm, t.sc.p + t.TMP(), 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    QS3X_HIP(hipStreamSynchronize(st));
    if (!(nrm[0] > 1e-10 * nrm[1])) {
      k = i - (i % p);  // keep whole blocks
      break;
    }
  }
  RefineResult out;
  out.k = std::max(k, p);
  out.rr = rayleigh_ritz(t, out.k, p);
  return out;
}

}  // namespace

extern "C" {

int qs3x_trlanczos(qs3x_sector *s, int32_t prec, int32_t NOE, int32_t NOK, int32_t NOM, int64_t maxitr,
                   double conv, const double *start, uint64_t seed, double *ene, double *psi, double *stats) {
  return guarded2([&] {
    QS3X_REQUIRE(s && ene, "qs3x_trlanczos: NULL argument");
    QS3X_REQUIRE(prec == QS3X_FP64 || prec == QS3X_FP32, "qs3x_trlanczos: bad precision");
    // eigen_solver.f90:17-22
    QS3X_REQUIRE(NOE >= 1 && NOE <= NOK, "ne must be eaqual or smaller than NOK!");
    if (NOM < 2 * NOK) NOM = 2 * NOK;
    const int64_t dim = s->dim;
    QS3X_REQUIRE(!(dim < NOE || dim < NOK || dim < NOM), "qs3x_trlanczos: dim < NOM (stop 999)");
    QS3X_REQUIRE(NOM + 1 <= kMaxBasis, "qs3x_trlanczos: NOM too large");
    // the Ritz rotation stages NOM x NOK coefficients in LDS: refuse here,
    // before the basis is allocated, rather than at the first restart
    if (NOK < NOM) {
      const size_t lds_rot = ritz_rotate_lds(prec, dim, NOM, NOK, true);
      QS3X_REQUIRE(lds_rot <= 156 * 1024,
                   "qs3x_trlanczos: NOM x NOK too large for the Ritz rotation's LDS staging (" +
                       std::to_string(lds_rot) + " bytes of coefficients, at most 156 KB)");
    }
    hipStream_t st = stream();
    // Solve time and the trace exclude the Krylov basis allocations (their
    // cost depends on what the process freed before: 0 to about 4.5 s for
    // 160-250 GB); they are reported separately in stats[7].
    double t0 = now_s(), t_alloc = 0.0, t_free = 0.0;
    std::vector<double> theta, resid;
    double t32 = 0.0;
    int64_t hv32 = 0, hv64 = 0;
    int restarts = 0;
    std::vector<double> startv;  // FP64 start for the refinement (host not needed: device copy)
    TR t64;
    bool refined = false;
    double accept_code = 0.0;
    int rr_j0 = 0;
    trace().clear();
    if (prec == QS3X_FP32) {
      TR t32s;
      t32s.init(s, QS3X_FP32, NOK, NOM);
      t0 += t32s.alloc_s;
      t_alloc += t32s.alloc_s;
      t32s.t_origin = t0;
      load_start(QS3X_FP32, t32s.slot(0), dim, start, seed, st);
      // FP32 phase: until the NOE lowest Ritz values stop moving at the FP32
      // resolution of the shifted spectrum
      const double floor32 = options().fp32_floor;
      const int every = options().tr_check_every;
      TRResult r = every > 0
                       ? tr_cycles_checked(t32s, maxitr, std::max(conv, 1e-30), 0.0, every, theta, resid, NOE)
                       : tr_cycles(t32s, maxitr, std::max(conv, 1e-30), floor32, theta, resid, NOE);
      restarts += r.restarts;
      hv32 = t32s.hv_count;
      QS3X_HIP(hipStreamSynchronize(st));
      t32 = now_s() - t0;
      // FP64 Rayleigh-Ritz on the NOK FP32 Ritz vectors.  The FP64 stage
      // always reorthogonalises fully (CGS2): started next to an eigenvector,
      // local orthogonalisation loses the projection property in one cycle.
      // The FP64 stage first holds only the Rayleigh-Ritz block and the
      // block-Krylov refinement (2 NOK + NOE and (2 s + 3) NOE vectors), so
      // that the FP32 basis can take nearly all of HBM; an FP64 thick restart
      // of at most 64 vectors replaces it only if the values are not accepted,
      // after the FP32 basis is freed.
      const int S_ref = std::max(0, options().fp64_refine);
      const int NOM64_rr = std::max(2 * NOK + NOE, (2 * S_ref + 3) * NOE) - 1;
      const int NOM64 = std::max(std::min(NOM, 64), 2 * NOK + NOE);
      t64.init(s, QS3X_FP64, NOK, NOM64_rr, 0, 2);
      t0 += t64.alloc_s;
      t_alloc += t64.alloc_s;
      t64.t_origin = t0;
      t64.hv_base = hv32;
      for (int j = 0; j < NOK; ++j)
        vec_copy_scale(QS3X_FP64, t64.slot(j), QS3X_FP32, t32s.slot(j), dim, nullptr, -1, 1.0, st);
      QS3X_HIP(hipStreamSynchronize(st));
      RRResult rr = rayleigh_ritz(t64, NOK, NOE);
      trace_point(t0, hv32 + t64.hv_count, rr.theta[0], QS3X_FP64);
      // Spectral gaps of the wanted pairs, from the FP32 projection's NOK
      // lowest Ritz values: distance from theta_j to the nearest other Ritz
      // value outside its cluster (which also holds any ghost copies of a
      // converged value in a recurrence without reorthogonalisation).  A pair is certified when r^2 / gap <= conv |theta| with r
      // its FP64 residual (the Kato-Temple bound, gap estimated from above by
      // the next Ritz value).  Values the bound does not certify are still
      // accepted when the FP32 phase met its convergence test and the FP64
      // Rayleigh quotient agrees with the FP32 Ritz value to conv |theta|
      // (stats[6] = 2: accepted, not certified).
      // Values closer than 1e-8 max(1, |theta_j|) form one cluster (degenerate
      // levels, split only by FP32 rounding); the gap of a cluster member is
      // to the nearest value outside it, as in a block (cluster) bound.
      std::vector<double> gap32(NOE, 0.0);
      for (int j = 0; j < NOE; ++j) {
        const double tol_c = 1e-8 * std::max(1.0, std::fabs(theta[j]));
        double g = 1e300;
        for (int i = 0; i < (int)theta.size(); ++i) {
          const double d = std::fabs(theta[i] - theta[j]);
          if (i != j && d > tol_c) g = std::min(g, d);
        }
        gap32[j] = g < 1e300 ? g : 0.0;
      }
      auto bound_ok = [&](const RRResult &q, int j) {
        const double bound = q.resid[j] * q.resid[j] / std::max(gap32[j], 1e-300);
        return std::make_pair(gap32[j] > 0.0 && bound <= conv * std::fabs(q.theta[j]), bound);
      };
      bool ok = true, certified = true;
      for (int j = 0; j < NOE; ++j) {
        const auto b = bound_ok(rr, j);
        const bool agree = r.converged && std::fabs(rr.theta[j] - theta[j]) <= conv * std::fabs(rr.theta[j]);
        if (options().verbose)
          std::fprintf(stderr, "[qs3x] pair %d: fp32 %.15g fp64 RR %.15g residual %.3e gap %.3e bound %.3e\n", j,
                       theta[j], rr.theta[j], rr.resid[j], gap32[j], b.second);
        if (!b.first) certified = false;
        if (!(b.first || agree)) ok = false;
      }
      resid = rr.resid;
      theta = rr.theta;
      rr_j0 = 2 * NOK;
      const int S = options().fp64_refine;
      const int m_lz = options().fp64_lanczos;
      if (m_lz > 0 && NOE == 1) {
        // FP64 Lanczos from the Rayleigh-Ritz vector until the residual bound
        // certifies the value (its basis is not stored; the vector, if
        // wanted, comes from a second pass into slot rr_j0)
        DevBuf<double> yout(psi ? (size_t)2 * dim : 0);
        Lz64Result lz = lanczos_refine64(s, t64.slot(rr_j0), m_lz, gap32[0], conv, psi ? yout.p : nullptr, st, t0,
                                         hv32 + t64.hv_count);
        t64.hv_count += lz.steps * (psi ? 2 : 1);
        if (psi)
          QS3X_HIP(hipMemcpyAsync(t64.slot(rr_j0), yout.p, 16 * (size_t)dim, hipMemcpyDeviceToDevice, st));
        QS3X_HIP(hipStreamSynchronize(st));
        theta[0] = lz.theta;
        resid[0] = lz.resid;
        if (options().verbose)
          std::fprintf(stderr, "[qs3x] fp64 lanczos refinement: %.15g residual %.3e bound %.3e after %lld steps%s\n",
                       lz.theta, lz.resid, lz.resid * lz.resid / std::max(gap32[0], 1e-300), (long long)lz.steps,
                       lz.certified ? " (certified)" : "");
        certified = lz.certified;
        ok = ok || lz.certified;
      } else if (S > 0) {
        // final FP64 residual correction of the Ritz vectors: Rayleigh-Ritz
        // on their block Krylov space.  Values not accepted above are
        // accepted when the refined pairs pass the residual bound.
        RefineResult rf = refine_block_krylov(t64, rr_j0, NOE, S);
        bool ok2 = true;
        for (int j = 0; j < NOE; ++j) {
          theta[j] = rf.rr.theta[j];
          resid[j] = rf.rr.resid[j];
          const auto b = bound_ok(rf.rr, j);
          if (!b.first) ok2 = false;
          if (options().verbose)
            std::fprintf(stderr, "[qs3x] refined pair %d: %.15g residual %.3e gap %.3e bound %.3e (block Krylov, %d vectors)\n",
                         j, theta[j], resid[j], gap32[j], b.second, rf.k);
        }
        certified = certified || ok2;
        ok = ok || ok2;
        rr_j0 = 2 * rf.k;
        trace_point(t0, hv32 + t64.hv_count, theta[0], QS3X_FP64);
      }
      certificate()[0] = gap32[0];
      certificate()[1] = resid[0];
      certificate()[2] = resid[0] * resid[0] / std::max(gap32[0], 1e-300);
      if (ok) {
        refined = true;
        accept_code = certified ? 1.0 : 2.0;
      } else {
        // fall back to FP64 thick restart from the refined vectors: keep their
        // sum, free the FP32 basis, give the FP64 stage its full basis
        DevBuf<double> acc((size_t)2 * dim);
        QS3X_HIP(hipMemsetAsync(acc.p, 0, 16 * dim, st));
        for (int j = 0; j < NOE; ++j)
          vec_axpy_acc(QS3X_FP64, acc.p, t64.slot(rr_j0 + j), dim, 1.0, nullptr, -1, st);
        QS3X_HIP(hipStreamSynchronize(st));
        {
          const double tf = now_s();
          t32s.basis.release();
          const double dt_free = now_s() - tf;
          t0 += dt_free;
          t_free += dt_free;
        }
        const int64_t hv_rr = t64.hv_count;
        t64.init(s, QS3X_FP64, NOK, NOM64, 0, 2);
        t0 += t64.alloc_s;
        t_alloc += t64.alloc_s;
        t64.t_origin = t0;
        t64.hv_base = hv32;
        t64.hv_count = hv_rr;
        vec_copy_scale(QS3X_FP64, t64.slot(0), QS3X_FP64, acc.p, dim, nullptr, -1, 1.0, st);
        QS3X_HIP(hipStreamSynchronize(st));
        rr_j0 = 0;  // the FP64 thick restart leaves its Ritz vectors in slots 0..NOK-1
      }
      // freeing a basis of hundreds of GB takes up to a second; like the
      // allocation it is kept out of the solve time (stats[7])
      const double tf = now_s();
      t32s.basis.release();
      const double dt_free = now_s() - tf;
      t0 += dt_free;
      t_free += dt_free;
      t64.t_origin = t0;
    } else {
      t64.init(s, QS3X_FP64, NOK, NOM);
      t0 += t64.alloc_s;
      t_alloc += t64.alloc_s;
      t64.t_origin = t0;
      load_start(QS3X_FP64, t64.slot(0), dim, start, seed, st);
    }
    if (!refined) {
      // FP64 solve (or the fallback of the mixed one): with tr_check_every > 0
      // the FP64 solve checks inside its cycles as the FP32 stage does; the
      // fallback keeps the reference's restart-level test
      const int every = options().tr_check_every;
      TRResult r = (prec == QS3X_FP64 && every > 0)
                       ? tr_cycles_checked(t64, maxitr, std::max(conv, 1e-30), 0.0, every, theta, resid, NOE)
                       : tr_cycles(t64, maxitr, conv, 0.0, theta, resid);
      restarts += r.restarts;
    }
    hv64 = t64.hv_count;
    for (int j = 0; j < NOE; ++j) ene[j] = theta[j];
    if (psi) {
      for (int j = 0; j < NOE; ++j) {
        // Ritz vectors are stored normalised to FP64 accuracy after rotation;
        // normalise explicitly for the host copy.
        t64.set_norm_scale(rr_j0 + j);
        DevBuf<double> tmp((size_t)2 * dim);
        vec_copy_scale(QS3X_FP64, tmp.p, QS3X_FP64, t64.slot(rr_j0 + j), dim, t64.sc.p, t64.S() + rr_j0 + j, 1.0, st);
        tmp.download(psi + (size_t)2 * dim * j, (size_t)2 * dim);
      }
    }
    QS3X_HIP(hipStreamSynchronize(st));
    if (stats) {
      double mr = 0.0;
      for (int j = 0; j < NOE && j < (int)resid.size(); ++j) mr = std::max(mr, resid[j]);
      stats[0] = restarts;
      stats[1] = (double)hv32;
      stats[2] = (double)hv64;
      stats[3] = now_s() - t0;
      stats[4] = t32;
      stats[5] = mr;
      // FP32 Ritz vectors accepted after the FP64 Rayleigh-Ritz step: 1
      // certified by the residual bound, 2 by agreement only, 0 not accepted
      stats[6] = refined ? accept_code : 0.0;
      if (prec == QS3X_FP32) certificate()[3] = stats[6];
      certificate()[4] = t_alloc;
      certificate()[5] = t_free;
      stats[7] = t_alloc;  // seconds allocating Krylov bases (not in stats[3] or the trace)
    }
  });
}

// DSF_analysis (source_DSFLan/main.f90:52-84).  Host arithmetic: O(levels x w).
int qs3x_continued_fraction(int32_t nlev, const double *alpha, const double *beta2, double norm1, double E0,
                            double eta, int64_t nomega, const double *omega, int32_t reference_quirk,
                            double *dsf) {
  return guarded2([&] {
    QS3X_REQUIRE(nlev >= 1 && alpha && beta2 && omega && dsf, "qs3x_continued_fraction: bad argument");
    const double pi = std::acos(-1.0);
    const int M = nlev;
    for (int64_t o = 0; o < nomega; ++o) {
      const double w = omega[o];
      // deepest level M: beta(M+1) = 0 (the file ends after M lines)
      double kr = w + E0 - alpha[M - 1];
      double bunbo = kr * kr + eta * eta;
      double tr = beta2[M - 1] * kr / bunbo;
      double ti = reference_quirk ? -beta2[M - 1] * eta : -beta2[M - 1] * eta / bunbo;
      for (int n = M - 1; n >= 2; --n) {
        kr = w + E0 - alpha[n - 1] - tr;
        double ki = eta - ti;
        bunbo = kr * kr + ki * ki;
        tr = beta2[n - 1] * kr / bunbo;
        ti = -beta2[n - 1] * ki / bunbo;
      }
      kr = w + E0 - alpha[0] - tr;
      double ki = eta - ti;
      bunbo = kr * kr + ki * ki;
      ti = -norm1 * ki / bunbo;
      dsf[o] = -ti / pi;
    }
  });
}

}  // extern "C"
