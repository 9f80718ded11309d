// integrator.hip -- placeholder until the wavefront integrator lands.
#include "api_internal.hpp"
extern "C" int bc_render(bc_ctx* c, const bc_render_desc*, int, int, float*, uint32_t*, uint64_t*, const volatile int32_t*) {
    if (c) c->err = "render not available";
    return BC_ERR_ARG;
}
extern "C" int bc_render_device(bc_ctx* c, const bc_render_desc*, int, int, float*, uint32_t*, uint64_t*, uint64_t*) {
    if (c) c->err = "render not available";
    return BC_ERR_ARG;
}
void bc_render_release(bc_ctx*) {}
