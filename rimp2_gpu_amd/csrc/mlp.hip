// Pair-screening surrogate (filled in after the batch path).
#include "rimp2_common.h"

namespace rimp2 {
struct MlpState {};
void mlp_destroy(MlpState* m) { delete m; }
}  // namespace rimp2

using namespace rimp2;
extern "C" {
int rimp2_pair_features(void* ctx, int64_t, int64_t, const double*, const double*, float*) { set_error(ctx, "not built yet"); return RIMP2_ERR_STATE; }
int rimp2_mlp_init(void* ctx, uint64_t) { set_error(ctx, "not built yet"); return RIMP2_ERR_STATE; }
int rimp2_mlp_train_step(void* ctx, int64_t, const float*, const float*, float, float*) { set_error(ctx, "not built yet"); return RIMP2_ERR_STATE; }
int rimp2_mlp_predict(void* ctx, int64_t, const float*, float*) { set_error(ctx, "not built yet"); return RIMP2_ERR_STATE; }
int rimp2_mlp_get_params(void* ctx, float*, int64_t) { set_error(ctx, "not built yet"); return RIMP2_ERR_STATE; }
int rimp2_mlp_set_params(void* ctx, const float*, int64_t) { set_error(ctx, "not built yet"); return RIMP2_ERR_STATE; }
int64_t rimp2_mlp_num_params(void) { return 0; }
int rimp2_mlp_grad(void* ctx, int64_t, const float*, const float*, float*, float*) { set_error(ctx, "not built yet"); return RIMP2_ERR_STATE; }
int rimp2_screened_energy(void* ctx, int64_t, int64_t, const double*, const double*, int, double, double*, double*, int64_t*) { set_error(ctx, "not built yet"); return RIMP2_ERR_STATE; }
}
