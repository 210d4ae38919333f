// The batch pair kernels that carry two tile paths (32- and 64-wide virtual
// blocks, chosen per fragment): the autotuner's mixed candidate for fp64 and
// fp32.  Kept apart from batch.hip, whose one-tile kernels instantiate the
// same pair_tile specializations (see batch_kernel.h).
#include "batch_kernel.h"

namespace rimp2 {

void launch_batch_mixed(int mode, hipStream_t s, const BatchParams& p, int ncu) {
  switch (mode) {
    case RIMP2_MODE_FP64: return launch_batch_t<PolF64, 32, PolF64, 64>(s, p, ncu);
    case RIMP2_MODE_FP32: return launch_batch_t<PolF32R16, 32, PolF32, 64>(s, p, ncu);
  }
  fail(RIMP2_ERR_ARG, "no mixed-tile batch pair kernel for mode %d", mode);
}

}  // namespace rimp2
