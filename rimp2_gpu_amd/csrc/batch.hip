// Batched DEC fragments (filled in below the pair kernel work).
#include "rimp2_common.h"

namespace rimp2 {
struct BatchState {};
void batch_destroy(BatchState* b) { delete b; }
}  // namespace rimp2

using namespace rimp2;
extern "C" {
int rimp2_batch(void* ctx, int64_t, const rimp2_frag_desc*, int, double*, double*) {
  set_error(ctx, "rimp2_batch: not built yet");
  return RIMP2_ERR_STATE;
}
int rimp2_batch_prepare(void* ctx, int64_t, const rimp2_frag_desc*, int) {
  set_error(ctx, "rimp2_batch_prepare: not built yet");
  return RIMP2_ERR_STATE;
}
int rimp2_batch_run(void* ctx, double*) { set_error(ctx, "not built yet"); return RIMP2_ERR_STATE; }
int rimp2_batch_result(void* ctx, double*, double*) { set_error(ctx, "not built yet"); return RIMP2_ERR_STATE; }
int rimp2_batch_tile(void* ctx, int64_t*) { set_error(ctx, "not built yet"); return RIMP2_ERR_STATE; }
}
