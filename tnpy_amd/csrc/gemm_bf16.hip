// Kernel instantiations and launcher of the GEMM in arithmetic BF16X3 (three bf16 MFMAs per product).
#include "gemm_kernel.h"

namespace tnx {

void gemm_launch_bf16(int ci, const GemmDesc& d, const GemmArgs& g, int splits, bool typed,
                    hipStream_t s) {
  if (!typed) return gemm_launch_index<2, void, void>(ci, d, g, splits, s);
  if (d.ta == DT::F64 && d.tb == DT::F64) return gemm_launch_index<2, double, double>(ci, d, g, splits, s);
  if (d.ta == DT::F64) return gemm_launch_index<2, double, float>(ci, d, g, splits, s);
  if (d.tb == DT::F64) return gemm_launch_index<2, float, double>(ci, d, g, splits, s);
  return gemm_launch_index<2, float, float>(ci, d, g, splits, s);
}

}  // namespace tnx
