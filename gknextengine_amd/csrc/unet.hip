// unet.hip - learned denoiser (placeholder until the MFMA kernels land).
#include <hip/hip_runtime.h>

#include "gk_ctx.hpp"

namespace gk {
int unet_denoise_frame(Ctx&) { return GK_ERR_STATE; }
}  // namespace gk

struct gk_unet {
    int unused;
};

extern "C" {
int gk_unet_create(gk_ctx*, uint32_t, uint64_t, gk_unet** out) { if (out) *out = nullptr; return GK_ERR_STATE; }
int gk_unet_destroy(gk_unet*) { return GK_ERR_STATE; }
int gk_unet_num_params(gk_unet*, uint64_t*) { return GK_ERR_STATE; }
int gk_unet_params(gk_unet*, float**) { return GK_ERR_STATE; }
int gk_unet_save(gk_unet*, const char*) { return GK_ERR_STATE; }
int gk_unet_load(gk_unet*, const char*) { return GK_ERR_STATE; }
int gk_unet_train_step(gk_unet*, const float*, const float*, uint32_t, uint32_t, uint32_t, float, float*) { return GK_ERR_STATE; }
int gk_unet_infer(gk_unet*, const float*, float*, uint32_t, uint32_t) { return GK_ERR_STATE; }
}
