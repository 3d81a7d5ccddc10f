// hv_col.hip -- instantiations of the column-form stencil kernel (hv_col.h)
// with one window position per lane, and the kernel selector.
#include "hv_col.h"

namespace qs3x {

using namespace hvk;

namespace {
template <int BT>
SFn col_b1(int prec, bool realph, bool fused) {
  if (fused) {
    if (prec == QS3X_FP64)
      return realph ? k_hv_stencil_col<double, true, true, BT, 2, 1> : k_hv_stencil_col<double, false, true, BT, 2, 1>;
    return realph ? k_hv_stencil_col<float, true, true, BT, 2, 1> : k_hv_stencil_col<float, false, true, BT, 2, 1>;
  }
  if (prec == QS3X_FP64)
    return realph ? k_hv_stencil_col<double, true, false, BT, 2, 1> : k_hv_stencil_col<double, false, false, BT, 2, 1>;
  return realph ? k_hv_stencil_col<float, true, false, BT, 2, 1> : k_hv_stencil_col<float, false, false, BT, 2, 1>;
}
}  // namespace

SFn stencil_col_kernel_p2(int prec, bool realph, bool fused, int B);  // hv_col2.hip

// Lines per tile: 2, 4, 6, 8, 10, 12 or 14 with one position per lane; 2,
// 4, 6 or 8 with two (other combinations: null).
SFn stencil_col_kernel(int prec, bool realph, bool fused, int B, int P) {
  if (P == 2) return stencil_col_kernel_p2(prec, realph, fused, B);
  switch (B) {
    case 2: return col_b1<2>(prec, realph, fused);
    case 4: return col_b1<4>(prec, realph, fused);
    case 6: return col_b1<6>(prec, realph, fused);
    case 8: return col_b1<8>(prec, realph, fused);
    case 10: return col_b1<10>(prec, realph, fused);
    case 12: return col_b1<12>(prec, realph, fused);
    case 14: return col_b1<14>(prec, realph, fused);
    default: return nullptr;
  }
}

}  // namespace qs3x
