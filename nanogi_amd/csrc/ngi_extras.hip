// Adjoint and batched entry points.
#include "ngi_render.h"

using namespace ngi;

extern "C" {

int ngi_render_batch(ngi_scene* const*, const ngi_render_params*, int, double* const*, double*, ngi_stats*) {
    return fail(NGI_ERR_UNSUPPORTED, "batched rendering is not available yet");
}

}  // extern "C"
