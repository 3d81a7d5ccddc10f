// dense.hip -- the dense Hamiltonian of a sector and its full diagonalisation
// (the ALG = 3 path).
//
// make_full_hamiltonian (ham2vec_v3.f90:5-35) fills Ham(i, id) with the
// pull-form hop terms of row i and the diagonal, so column j of the matrix is
// H e_j.  Here column j is one FP64 H.v launch of the sector's own operator on
// the unit vector e_j: the dense matrix is the matrix-free operator itself
// (same kernel family, same norms and phases), with no second copy of the hop
// code.  The columns land in a device matrix whose leading dimension is
// padded to 4 elements (64 B), so every column starts where the H.v kernels'
// vector stores expect.
//
// fulldia_routines_with_momentum (eigen_solver.f90:263-339) calls LAPACK
// zheev('V', 'U', ...): the eigenvalues ascending and, in the columns of A,
// orthonormal eigenvectors, reading only the upper triangle.  rocSOLVER's
// zheev (Householder tridiagonalisation, then implicit QR on the tridiagonal
// form: the same algorithm) does this on the device, also from the upper
// triangle, so the pull-form entries H(i, j) with i <= j are the ones used, as
// in the reference.  Option dense_solver 1 takes zheevd (divide and conquer)
// instead, faster for large dim; it has returned NaN eigenvectors with info 0
// on the exactly degenerate spectrum of a Heisenberg ferromagnet, so every
// result is checked to be finite before it is returned.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "blas.h"
#include "qs3x_internal.h"

namespace qs3x {
namespace {

// e = e_j, given e = e_{j-1} (or zero for j = 0): one thread.
__global__ void k_unit_step(double2 *e, int64_t j) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (j > 0) e[j - 1] = make_double2(0.0, 0.0);
    e[j] = make_double2(1.0, 0.0);
  }
}

int64_t padded_ld(int64_t dim) { return (dim + 3) / 4 * 4; }

// Counts the non-finite values of a (n values, or n columns of m with leading
// dimension lda when lda > 0) into *bad.
__global__ void k_count_nonfinite(const double *a, int64_t m, int64_t n, int64_t lda, unsigned long long *bad) {
  unsigned long long c = 0;
  const int64_t total = lda > 0 ? 2 * m * n : n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t at = lda > 0 ? (t / (2 * m)) * 2 * lda + t % (2 * m) : t;
    c += !isfinite(a[at]);
  }
  for (int d = 32; d >= 1; d >>= 1) c += __shfl_down(c, d, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(bad, c);
}

unsigned long long count_nonfinite(const double *a, int64_t m, int64_t n, int64_t lda, hipStream_t st) {
  DevBuf<unsigned long long> bad(1);
  QS3X_HIP(hipMemsetAsync(bad.p, 0, sizeof(unsigned long long), st));
  const int64_t total = lda > 0 ? 2 * m * n : n;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 4096));
  k_count_nonfinite<<<grid, 256, 0, st>>>(a, m, n, lda, bad.p);
  QS3X_HIP(hipGetLastError());
  unsigned long long h = 0;
  bad.download(&h, 1);
  return h;
}

void check_sector(const qs3x_sector *s, const char *who) {
  QS3X_REQUIRE(!s->all_k, std::string(who) + ": the all-k chain sector is not an operator sector");
  QS3X_REQUIRE(s->dim > 0, std::string(who) + ": empty sector");
  QS3X_REQUIRE(s->dim <= INT32_MAX, std::string(who) + ": dim exceeds the dense solver's 32-bit size");
}

// Column j of the device matrix a (leading dimension lda) = H e_j.
void assemble(qs3x_sector *s, double2 *a, int64_t lda, hipStream_t st) {
  DevBuf<double2> e((size_t)s->dim);
  QS3X_HIP(hipMemsetAsync(e.p, 0, sizeof(double2) * (size_t)s->dim, st));
  for (int64_t j = 0; j < s->dim; ++j) {
    k_unit_step<<<1, 64, 0, st>>>(e.p, j);
    launch_hv(s, QS3X_FP64, a + j * lda, e.p, nullptr, nullptr, nullptr, st);
  }
  QS3X_HIP(hipGetLastError());
}

void copy_out(double *host, int64_t ld, const double2 *a, int64_t lda, int64_t dim, hipStream_t st) {
  QS3X_HIP(hipMemcpy2DAsync(host, sizeof(double2) * (size_t)ld, a, sizeof(double2) * (size_t)lda,
                            sizeof(double2) * (size_t)dim, (size_t)dim, hipMemcpyDeviceToHost, st));
  QS3X_HIP(hipStreamSynchronize(st));
}

struct Handle {
  rocblas_handle h = nullptr;
  explicit Handle(hipStream_t st) {
    QS3X_REQUIRE(rocblas_create_handle(&h) == rocblas_status_success, "rocblas_create_handle failed");
    QS3X_REQUIRE(rocblas_set_stream(h, st) == rocblas_status_success, "rocblas_set_stream failed");
  }
  ~Handle() {
    if (h) (void)rocblas_destroy_handle(h);
  }
  Handle(const Handle &) = delete;
  Handle &operator=(const Handle &) = delete;
};

template <class F>
int guarded_dense(F &&f) {
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

}  // namespace
}  // namespace qs3x

using namespace qs3x;

extern "C" int qs3x_make_full_hamiltonian(qs3x_sector *s, double *ham, int64_t ld) {
  return guarded_dense([&] {
    QS3X_REQUIRE(s && ham, "qs3x_make_full_hamiltonian: NULL argument");
    check_sector(s, "qs3x_make_full_hamiltonian");
    QS3X_REQUIRE(ld >= s->dim, "qs3x_make_full_hamiltonian: ld < dim");
    const int64_t lda = padded_ld(s->dim);
    DevBuf<double2> a((size_t)(lda * s->dim));
    assemble(s, a.p, lda, stream());
    copy_out(ham, ld, a.p, lda, s->dim, stream());
  });
}

extern "C" int qs3x_full_diag(qs3x_sector *s, double *ene, double *psi, int64_t ld) {
  return guarded_dense([&] {
    QS3X_REQUIRE(s && ene, "qs3x_full_diag: NULL argument");
    check_sector(s, "qs3x_full_diag");
    QS3X_REQUIRE(!psi || ld >= s->dim, "qs3x_full_diag: ld < dim");
    const int64_t n = s->dim, lda = padded_ld(n);
    hipStream_t st = stream();
    DevBuf<double2> a((size_t)(lda * n));
    DevBuf<double> w((size_t)n), offd((size_t)n);
    DevBuf<rocblas_int> info(1);
    assemble(s, a.p, lda, st);
    const bool dc = options().dense_solver == 1;
    const char *name = dc ? "zheevd" : "zheev";
    {
      Handle h(st);
      const rocblas_evect ev = psi ? rocblas_evect_original : rocblas_evect_none;
      const rocblas_status rs =
          dc ? rocsolver_zheevd(h.h, ev, rocblas_fill_upper, (rocblas_int)n, (rocblas_double_complex *)a.p,
                                (rocblas_int)lda, w.p, offd.p, info.p)
             : rocsolver_zheev(h.h, ev, rocblas_fill_upper, (rocblas_int)n, (rocblas_double_complex *)a.p,
                               (rocblas_int)lda, w.p, offd.p, info.p);
      QS3X_REQUIRE(rs == rocblas_status_success,
                   std::string("qs3x_full_diag: rocsolver_") + name + " returned " + rocblas_status_to_string(rs));
      QS3X_HIP(hipStreamSynchronize(st));
    }
    rocblas_int inf = 0;
    info.download(&inf, 1);
    QS3X_REQUIRE(inf == 0, std::string("qs3x_full_diag: ") + name + " did not converge (info = " +
                               std::to_string(inf) + ")");
    QS3X_REQUIRE(count_nonfinite(w.p, 0, n, 0, st) == 0,
                 std::string("qs3x_full_diag: ") + name + " returned non-finite eigenvalues");
    if (psi)
      QS3X_REQUIRE(count_nonfinite((const double *)a.p, n, n, lda, st) == 0,
                   std::string("qs3x_full_diag: ") + name + " returned non-finite eigenvectors");
    w.download(ene, (size_t)n);
    if (psi) copy_out(psi, ld, a.p, lda, n, st);
  });
}
