// hv_col2.hip -- instantiations of the column-form stencil kernel (hv_col.h)
// with two window positions per lane (FP32).
#include "hv_col.h"

namespace qs3x {

using namespace hvk;

namespace {
// FP32 only: an FP64 window line of 64 positions is already 1 KB, and two
// positions per lane in FP64 exceed 256 registers in the fused step.
template <int BT>
SFn col_b2(int prec, bool realph, bool fused) {
  if (prec == QS3X_FP64) return nullptr;
  if (fused) return realph ? k_hv_stencil_col<float, true, true, BT, 2, 2> : k_hv_stencil_col<float, false, true, BT, 2, 2>;
  return realph ? k_hv_stencil_col<float, true, false, BT, 2, 2> : k_hv_stencil_col<float, false, false, BT, 2, 2>;
}
}  // namespace

SFn stencil_col_kernel_p2(int prec, bool realph, bool fused, int B) {
  switch (B) {
    case 2: return col_b2<2>(prec, realph, fused);
    case 4: return col_b2<4>(prec, realph, fused);
    case 6: return col_b2<6>(prec, realph, fused);
    case 8: return col_b2<8>(prec, realph, fused);
    default: return nullptr;
  }
}

}  // namespace qs3x
