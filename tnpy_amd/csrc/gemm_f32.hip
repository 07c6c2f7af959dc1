// Kernel instantiations and launcher of the GEMM in arithmetic F32 (FP32 MFMA).
#include "gemm_kernel.h"

namespace tnx {

void gemm_launch_f32(int ci, const GemmDesc& d, const GemmArgs& g, int splits, bool typed,
                    hipStream_t s) {
  if (!typed) return gemm_launch_index<1, void, void>(ci, d, g, splits, s);
  if (d.ta == DT::F64 && d.tb == DT::F64) return gemm_launch_index<1, double, double>(ci, d, g, splits, s);
  if (d.ta == DT::F64) return gemm_launch_index<1, double, float>(ci, d, g, splits, s);
  if (d.tb == DT::F64) return gemm_launch_index<1, float, double>(ci, d, g, splits, s);
  return gemm_launch_index<1, float, float>(ci, d, g, splits, s);
}

}  // namespace tnx
