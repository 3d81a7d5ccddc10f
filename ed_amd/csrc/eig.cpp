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

}  // namespace qs3x

extern "C" {
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
