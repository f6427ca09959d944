// Wavefront engine (stages: generate, extend, shade, shadow).
#include "ngi_render.h"

namespace ngi {

int Session::render_wavefront(void*, uint64_t) { return fail(NGI_ERR_UNSUPPORTED, "wavefront engine is not available yet"); }
void Session::release_engine() {}

}  // namespace ngi
