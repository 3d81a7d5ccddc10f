// eig.cpp -- small dense FP64 eigensolvers on the host for the Lanczos
// projections.  They stand where the reference calls LAPACK:
//   dstevx range 'I' (lanczos.f90:104-105)      -> tridiag_lowest + tridiag_vector
//   dstevx range 'A' (eigen_solver.f90:128-129) -> sym_eig on the tridiagonal
//   zheevr           (eigen_solver.f90:205-207) -> sym_eig on the arrowhead matrix
// The thick-restart projection is real symmetric: the Lanczos alpha are real
// (<v|H|v> with H Hermitian), the betas are norms, and the arrowhead entries
// beta_NOM * c(NOM, j) are products of reals, so a real solver is exact here.
//
// Algorithms (textbook, written from the mathematics):
//   * Sturm-sequence bisection: the number of eigenvalues of T below x equals
//     the number of negative pivots of the LDL^T factorisation of T - x I.
//   * inverse iteration with a partially pivoted tridiagonal LU solve.
//   * cyclic Jacobi rotations for dense symmetric matrices.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

#include "blas.h"

namespace qs3x {

// number of eigenvalues of the symmetric tridiagonal (d[0..n-1], off-diagonal
// e[0..n-2]) strictly below x
static int sturm_count(const double *d, const double *e, int n, double x, double pivmin) {
  int c = 0;
  double q = d[0] - x;
  if (std::fabs(q) < pivmin) q = -pivmin;
  if (q < 0) ++c;
  for (int i = 1; i < n; ++i) {
    q = d[i] - x - (e[i - 1] * e[i - 1]) / q;
    if (std::fabs(q) < pivmin) q = -pivmin;
    if (q < 0) ++c;
  }
  return c;
}

// k-th smallest eigenvalue (k = 0 lowest) by bisection to full precision.
double tridiag_eigenvalue(const double *d, const double *e, int n, int k) {
  if (n == 1) return d[0];
  double lo = d[0], hi = d[0], emax = 0.0;
  for (int i = 0; i < n; ++i) {
    double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i < n - 1 ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
    if (i < n - 1) emax = std::max(emax, e[i] * e[i]);
  }
  const double eps = std::numeric_limits<double>::epsilon();
  const double pivmin = std::numeric_limits<double>::min() * std::max(1.0, emax);
  const double tnorm = std::max(std::fabs(lo), std::fabs(hi));
  lo -= 2 * eps * tnorm + pivmin;
  hi += 2 * eps * tnorm + pivmin;
  for (int it = 0; it < 200; ++it) {
    double mid = 0.5 * (lo + hi);
    if (hi - lo <= 2 * eps * std::max(std::fabs(lo), std::fabs(hi)) + pivmin || mid == lo || mid == hi) break;
    if (sturm_count(d, e, n, mid, pivmin) > k) hi = mid;
    else lo = mid;
  }
  return 0.5 * (lo + hi);
}

double tridiag_lowest(const double *d, const double *e, int n) { return tridiag_eigenvalue(d, e, n, 0); }

// Eigenvector of the tridiagonal for eigenvalue lam: inverse iteration on
// (T - lam I) with Gaussian elimination with partial pivoting (tridiagonal
// band, one extra super-diagonal for the row swaps).
void tridiag_vector(const double *d, const double *e, int n, double lam, double *z) {
  if (n == 1) {
    z[0] = 1.0;
    return;
  }
  double tnorm = 0.0;
  for (int i = 0; i < n; ++i)
    tnorm = std::max(tnorm, std::fabs(d[i]) + (i ? std::fabs(e[i - 1]) : 0.0) + (i < n - 1 ? std::fabs(e[i]) : 0.0));
  const double eps = std::numeric_limits<double>::epsilon();
  const double tiny = eps * std::max(tnorm, 1e-300);
  // LU of T - lam I with partial pivoting (row i holds dl[i-1], d[i], du[i];
  // a row swap creates the second super-diagonal du2).
  std::vector<double> dl(n, 0.0), dg(n), du(n, 0.0), du2(n, 0.0);
  std::vector<char> sw(n, 0);
  for (int i = 0; i < n; ++i) dg[i] = d[i] - lam;
  for (int i = 0; i < n - 1; ++i) dl[i] = du[i] = e[i];
  for (int i = 0; i < n - 1; ++i) {
    if (std::fabs(dg[i]) >= std::fabs(dl[i])) {
      if (std::fabs(dg[i]) < tiny) dg[i] = (dg[i] >= 0 ? tiny : -tiny);
      double f = dl[i] / dg[i];
      dl[i] = f;
      dg[i + 1] -= f * du[i];
    } else {
      double f = dg[i] / dl[i];
      dg[i] = dl[i];
      dl[i] = f;
      double t = du[i];
      du[i] = dg[i + 1];
      dg[i + 1] = t - f * dg[i + 1];
      if (i < n - 2) {
        du2[i] = du[i + 1];
        du[i + 1] = -f * du[i + 1];
      }
      sw[i] = 1;
    }
  }
  if (std::fabs(dg[n - 1]) < tiny) dg[n - 1] = tiny;
  std::vector<double> x(n);
  for (int i = 0; i < n; ++i) x[i] = 1.0 + 0.01 * std::sin(1.0 + i);
  for (int it = 0; it < 4; ++it) {
    for (int i = 0; i < n - 1; ++i) {
      if (!sw[i]) {
        x[i + 1] -= dl[i] * x[i];
      } else {
        double t = x[i];
        x[i] = x[i + 1];
        x[i + 1] = t - dl[i] * x[i];
      }
    }
    x[n - 1] /= dg[n - 1];
    if (n >= 2) x[n - 2] = (x[n - 2] - du[n - 2] * x[n - 1]) / dg[n - 2];
    for (int i = n - 3; i >= 0; --i) x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / dg[i];
    double nr = 0.0;
    for (int i = 0; i < n; ++i) nr += x[i] * x[i];
    nr = std::sqrt(nr);
    for (int i = 0; i < n; ++i) x[i] /= nr;
  }
  // sign convention: first nonzero component positive
  int s = 0;
  while (s < n - 1 && std::fabs(x[s]) < 1e-300) ++s;
  double sg = x[s] < 0 ? -1.0 : 1.0;
  for (int i = 0; i < n; ++i) z[i] = sg * x[i];
}

// Dense real symmetric eigensolver (cyclic Jacobi).  A row-major n x n (is
// overwritten); w ascending eigenvalues; V row-major, column j = eigenvector j.
void sym_eig(int n, double *A, double *w, double *V) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) V[i * n + j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double v = A[i * n + j] * A[i * n + j];
        tot += v;
        if (i != j) off += v;
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = A[p * n + q];
        if (apq == 0.0) continue;
        double app = A[p * n + p], aqq = A[q * n + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = c * vkp - s * vkq;
          V[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> ord(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int x, int y) { return A[x * n + x] < A[y * n + y]; });
  std::vector<double> V2((size_t)n * n);
  for (int j = 0; j < n; ++j) {
    w[j] = A[ord[j] * n + ord[j]];
    // sign convention: largest-magnitude component positive
    int imax = 0;
    for (int i = 1; i < n; ++i)
      if (std::fabs(V[i * n + ord[j]]) > std::fabs(V[imax * n + ord[j]])) imax = i;
    double sg = V[imax * n + ord[j]] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) V2[i * n + j] = sg * V[i * n + ord[j]];
  }
  std::copy(V2.begin(), V2.end(), V);
}

// ------------------------------------------------ arrowhead + tridiagonal
// The thick-restart projection of size m: rows 0..k-1 are the locked Ritz
// values w_j coupled only to row k (c_j); rows k..m-1 are tridiagonal
// (a[i-k] diagonal, b[i-k] coupling of i and i+1).  k = 0 is a tridiagonal.
//
// Inertia: eliminating the tail bottom-up (i = m-1 .. k+1), then the locked
// rows, then row k, is an LDL^T factorisation of a symmetric permutation of
// A - x I, so its negative pivots count the eigenvalues below x:
//   d_{m-1} = a_{m-1} - x,  d_i = a_i - x - b_i^2 / d_{i+1},
//   p_j = w_j - x,          d_k = a_k - x - b_k^2 / d_{k+1} - sum_j c_j^2 / p_j.
static int arrow_count(const Arrow &A, double x, double pivmin) {
  const int k = A.k, m = A.m;
  auto guard = [&](double q) { return std::fabs(q) < pivmin ? -pivmin : q; };
  int c = 0;
  double dn = 0.0;  // pivot of row i + 1
  for (int i = m - 1; i >= k; --i) {
    double q = A.a[i - k] - x;
    if (i + 1 < m) q -= A.b[i - k] * A.b[i - k] / dn;
    if (i == k)
      for (int j = 0; j < k; ++j) q -= A.c[j] * A.c[j] / guard(A.w[j] - x);
    q = guard(q);
    if (q < 0) ++c;
    dn = q;
  }
  for (int j = 0; j < k; ++j)
    if (guard(A.w[j] - x) < 0) ++c;
  return c;
}

static void arrow_bounds(const Arrow &A, double &lo, double &hi, double &pivmin) {
  const int k = A.k, m = A.m;
  lo = std::numeric_limits<double>::max();
  hi = -lo;
  double smax = 0.0;
  double ck = 0.0;
  for (int j = 0; j < k; ++j) {
    const double r = std::fabs(A.c[j]);
    lo = std::min(lo, A.w[j] - r);
    hi = std::max(hi, A.w[j] + r);
    ck += r;
    smax = std::max(smax, A.c[j] * A.c[j]);
  }
  for (int i = k; i < m; ++i) {
    double r = (i > k ? std::fabs(A.b[i - k - 1]) : 0.0) + (i + 1 < m ? std::fabs(A.b[i - k]) : 0.0);
    if (i == k) r += ck;
    lo = std::min(lo, A.a[i - k] - r);
    hi = std::max(hi, A.a[i - k] + r);
    if (i + 1 < m) smax = std::max(smax, A.b[i - k] * A.b[i - k]);
  }
  const double eps = std::numeric_limits<double>::epsilon();
  pivmin = std::numeric_limits<double>::min() * std::max(1.0, smax);
  const double tn = std::max(std::fabs(lo), std::fabs(hi));
  lo -= 2 * eps * tn + pivmin;
  hi += 2 * eps * tn + pivmin;
}

double arrow_eigenvalue(const Arrow &A, int idx) {
  double lo, hi, pivmin;
  arrow_bounds(A, lo, hi, pivmin);
  const double eps = std::numeric_limits<double>::epsilon();
  for (int it = 0; it < 300; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (hi - lo <= 2 * eps * std::max(std::fabs(lo), std::fabs(hi)) + pivmin || mid == lo || mid == hi) break;
    if (arrow_count(A, mid, pivmin) > idx) hi = mid;
    else lo = mid;
  }
  return 0.5 * (lo + hi);
}

// Solve (A - lam I) x = r in place by the elimination order of arrow_count.
static void arrow_solve(const Arrow &A, double lam, double tiny, std::vector<double> &x) {
  const int k = A.k, m = A.m;
  auto guard = [&](double q) { return std::fabs(q) < tiny ? (q >= 0 ? tiny : -tiny) : q; };
  std::vector<double> d(m), g(m);
  for (int i = m - 1; i >= k; --i) {
    double q = A.a[i - k] - lam, r = x[i];
    if (i + 1 < m) {
      q -= A.b[i - k] * A.b[i - k] / d[i + 1];
      r -= A.b[i - k] * g[i + 1] / d[i + 1];
    }
    if (i == k)
      for (int j = 0; j < k; ++j) {
        const double p = guard(A.w[j] - lam);
        q -= A.c[j] * A.c[j] / p;
        r -= A.c[j] * x[j] / p;
      }
    d[i] = guard(q);
    g[i] = r;
  }
  // back-substitution: row k, then the locked rows, then the tail downwards
  x[k] = g[k] / d[k];
  for (int j = 0; j < k; ++j) x[j] = (x[j] - A.c[j] * x[k]) / guard(A.w[j] - lam);
  for (int i = k + 1; i < m; ++i) x[i] = (g[i] - A.b[i - k - 1] * x[i - 1]) / d[i];
}

// Lowest nv eigenpairs: lam[0..nv) ascending, Z row-major m x nv (column j =
// eigenvector j, unit norm, largest-magnitude component positive).  Inverse
// iteration; vectors of (nearly) equal eigenvalues are orthogonalised.
void arrow_lowest_pairs(const Arrow &A, int nv, double *lam, double *Z) {
  const int m = A.m;
  double lo, hi, pivmin;
  arrow_bounds(A, lo, hi, pivmin);
  const double anorm = std::max(std::fabs(lo), std::fabs(hi));
  const double eps = std::numeric_limits<double>::epsilon();
  const double tiny = eps * std::max(anorm, 1e-300);
  std::vector<std::vector<double>> V;
  for (int v = 0; v < nv; ++v) {
    lam[v] = arrow_eigenvalue(A, v);
    std::vector<double> x(m);
    for (int i = 0; i < m; ++i) x[i] = 1.0 + 0.37 * std::sin(1.3 * (i + 1) + 0.7 * (v + 1));
    for (int it = 0; it < 5; ++it) {
      arrow_solve(A, lam[v], tiny, x);
      // orthogonalise against the vectors of close eigenvalues (twice)
      for (int pass = 0; pass < 2; ++pass)
        for (int u = 0; u < v; ++u) {
          if (std::fabs(lam[u] - lam[v]) > 1e-3 * std::max(anorm, 1e-300)) continue;
          double dot = 0.0;
          for (int i = 0; i < m; ++i) dot += V[u][i] * x[i];
          for (int i = 0; i < m; ++i) x[i] -= dot * V[u][i];
        }
      double nr = 0.0;
      for (int i = 0; i < m; ++i) nr += x[i] * x[i];
      nr = std::sqrt(nr);
      if (!(nr > 0)) {
        for (int i = 0; i < m; ++i) x[i] = (i == v % m) ? 1.0 : 0.0;
        continue;
      }
      for (int i = 0; i < m; ++i) x[i] /= nr;
    }
    int imax = 0;
    for (int i = 1; i < m; ++i)
      if (std::fabs(x[i]) > std::fabs(x[imax])) imax = i;
    const double sg = x[imax] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < m; ++i) x[i] *= sg;
    for (int i = 0; i < m; ++i) Z[(size_t)i * nv + v] = x[i];
    V.push_back(std::move(x));
  }
}

}  // namespace qs3x

extern "C" {
// Test hook: lowest nv eigenpairs of the arrowhead + tridiagonal matrix (k
// locked values w, couplings c; tail diagonal a (m - k), off-diagonal b (m - k - 1)).
void qs3x_test_arrow_pairs(int k, int m, const double *w, const double *c, const double *a, const double *b,
                           int nv, double *lam, double *Z) {
  qs3x::Arrow A{k, m, w, c, a, b};
  qs3x::arrow_lowest_pairs(A, nv, lam, Z);
}
// Exposed for tests of the host eigensolvers (not part of include/qs3x.h's
// solver contract; used by tests/test_cpu_eig.py).
double qs3x_test_tridiag_eigenvalue(const double *d, const double *e, int n, int k) {
  return qs3x::tridiag_eigenvalue(d, e, n, k);
}
void qs3x_test_tridiag_vector(const double *d, const double *e, int n, double lam, double *z) {
  qs3x::tridiag_vector(d, e, n, lam, z);
}
void qs3x_test_sym_eig(int n, double *A, double *w, double *V) { qs3x::sym_eig(n, A, w, V); }
}
