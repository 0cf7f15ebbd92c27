// fx_fp32.hip -- FP32-mixed path (placeholder entry points until the fused
// kernel lands).
#include "fx_params.h"

int fx32_init_state(fx_domain) { return fx_fail(FX_ERR_UNSUPPORTED, "FP32-mixed path not built yet"); }
int fx32_launch_step(fx_domain) { return fx_fail(FX_ERR_UNSUPPORTED, "FP32-mixed path not built yet"); }
int fx32_launch_cfl_initial(fx_domain) { return fx_fail(FX_ERR_UNSUPPORTED, "FP32-mixed path not built yet"); }
int fx32_get_field(fx_domain, int, void *) { return fx_fail(FX_ERR_UNSUPPORTED, "FP32-mixed path not built yet"); }
int fx32_set_field(fx_domain, int, const void *) { return fx_fail(FX_ERR_UNSUPPORTED, "FP32-mixed path not built yet"); }
int fx32_launch_volume(fx_domain, double *) { return fx_fail(FX_ERR_UNSUPPORTED, "FP32-mixed path not built yet"); }
