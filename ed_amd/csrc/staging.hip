// staging.hip -- host-vector H . v that allocates no device memory.
//
// qs3x_ham_to_vec_wave_vector (api.hip) stages v1 and v0 through two device
// buffers it allocates on every call.  Callers that must not allocate (a
// Fortran driver calling ham_to_vec_wave_vector once per Lanczos step in a
// loop that forbids hipMalloc, harnesses that pre-size all device memory)
// use qs3x_ham_to_vec_wave_vector_noalloc instead: the same H . v kernels,
// with both vectors staged through one file-scope device array.
//
// The array is a per-device static of kStagingBytes; each GPU the library
// runs on holds its own copy until the process ends.  Host code never uses
// the array's name as an address: on a device's first call its address is
// taken with hipGetSymbolAddress (that device current) and kept per device.
// Every call on one GPU shares that GPU's array, so calls that overlap on one
// GPU (from several host threads) race; each call overwrites what it reads.
#include <mutex>

#include "qs3x_internal.h"

namespace {

using namespace qs3x;

// v1 and v0 as complex doubles: 2 * 16 bytes per row, 2,097,152 rows at most
constexpr size_t kStagingBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kStagingBytes];

std::mutex g_mu;
void *g_scratch_dev[kMaxDevices] = {};

// Device address of this device's staging array (taken once per device).
unsigned char *staging_ptr(int dev) {
  QS3X_REQUIRE(dev >= 0 && dev < kMaxDevices, "qs3x_ham_to_vec_wave_vector_noalloc: device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch_dev[dev]) {
    QS3X_HIP(hipSetDevice(dev));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    QS3X_REQUIRE(e == hipSuccess && p != nullptr,
                 std::string("qs3x_ham_to_vec_wave_vector_noalloc: hipGetSymbolAddress failed: ") +
                     hipGetErrorString(e));
    g_scratch_dev[dev] = p;
  }
  return (unsigned char *)g_scratch_dev[dev];
}

template <class F>
int guarded_staging(F &&f) {
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

extern "C" {

int qs3x_ham_to_vec_wave_vector_noalloc(qs3x_sector *s, double *v0, const double *v1) {
  return guarded_staging([&] {
    QS3X_REQUIRE(s && v0 && v1, "qs3x_ham_to_vec_wave_vector_noalloc: NULL argument");
    const size_t bytes = (size_t)16 * (size_t)s->dim;
    const size_t off = (bytes + 255) & ~(size_t)255;  // v0 follows v1, 256-byte aligned
    QS3X_REQUIRE(off + bytes <= kStagingBytes,
                 "qs3x_ham_to_vec_wave_vector_noalloc: sector dimension " + std::to_string(s->dim) +
                     " exceeds the static staging limit of " + std::to_string(kStagingBytes / 32) +
                     " rows (use qs3x_ham_to_vec_wave_vector)");
    if (s->dim == 0) return;
    unsigned char *base = staging_ptr(device());
    hipStream_t st = stream();
    QS3X_HIP(hipMemcpyAsync(base, v1, bytes, hipMemcpyHostToDevice, st));
    launch_hv(s, QS3X_FP64, base + off, base, nullptr, nullptr, nullptr, st);
    QS3X_HIP(hipMemcpyAsync(v0, base + off, bytes, hipMemcpyDeviceToHost, st));
    QS3X_HIP(hipStreamSynchronize(st));
  });
}

}  // extern "C"
