// Wavefront path integrator (filled in next; entry points report that the
// integrator is unavailable until then).
#include "internal.hpp"

extern "C" {

int ilm_scene_create(ilm_ctx*, ilm_mesh*, const float*, const int32_t*, const int32_t*, const float*, int,
                     const float*, int, ilm_scene** out) {
    if (out) *out = nullptr;
    return ilm::fail(ILM_ERR_STATE, "path integrator not available in this build");
}

int ilm_scene_destroy(ilm_scene*) { return ILM_OK; }

int ilm_render(ilm_scene*, const ilm_render_params*, double*, ilm_render_stats*) {
    return ilm::fail(ILM_ERR_STATE, "path integrator not available in this build");
}

}  // extern "C"
