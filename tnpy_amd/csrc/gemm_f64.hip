// Kernel instantiations and launcher of the GEMM in arithmetic F64 (FP64 MFMA).
#include "gemm_kernel.h"

namespace tnx {

void gemm_launch_f64(int ci, const GemmDesc& d, const GemmArgs& g, int splits, bool typed,
                    hipStream_t s) {
  if (!typed) return gemm_launch_index<0, void, void>(ci, d, g, splits, s);
  if (d.ta == DT::F64 && d.tb == DT::F64) return gemm_launch_index<0, double, double>(ci, d, g, splits, s);
  // (F64, F32): the only other pair gemm_typed_ok admits for F64
  return gemm_launch_index<0, double, float>(ci, d, g, splits, s);
}

}  // namespace tnx
