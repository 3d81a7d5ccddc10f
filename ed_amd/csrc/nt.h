// nt.h -- streaming loads and stores with the nontemporal hint, for complex
// vectors that a pass reads or writes once and that are far larger than the
// caches (measured on a 41.4M-element Lanczos update: 5.1 -> 5.6-5.8 TB/s).
#pragma once
#include <hip/hip_runtime.h>

namespace qs3x {

typedef float nt_f2 __attribute__((ext_vector_type(2)));
typedef double nt_d2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float2 ld_nt(const float2 *p) {
  const nt_f2 v = __builtin_nontemporal_load((const nt_f2 *)p);
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ double2 ld_nt(const double2 *p) {
  const nt_d2 v = __builtin_nontemporal_load((const nt_d2 *)p);
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st_nt(float2 *p, float2 x) {
  nt_f2 v;
  v.x = x.x;
  v.y = x.y;
  __builtin_nontemporal_store(v, (nt_f2 *)p);
}
__device__ __forceinline__ void st_nt(double2 *p, double2 x) {
  nt_d2 v;
  v.x = x.x;
  v.y = x.y;
  __builtin_nontemporal_store(v, (nt_d2 *)p);
}

}  // namespace qs3x
