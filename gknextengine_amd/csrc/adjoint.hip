// adjoint.hip - path-replay backpropagation (placeholder until the kernels land).
#include <hip/hip_runtime.h>

#include "gk_ctx.hpp"

extern "C" {
int gk_adjoint(gk_ctx*, const float*, float*) { return GK_ERR_STATE; }
int gk_render_radiance(gk_ctx*, float*) { return GK_ERR_STATE; }
}
