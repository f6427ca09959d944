// Adjoint and batched entry points.
#include "ngi_render.h"

using namespace ngi;

extern "C" {

int ngi_render_adjoint(ngi_scene*, const ngi_render_params*, const double*, double*, double*, ngi_stats*) {
    return fail(NGI_ERR_UNSUPPORTED, "adjoint is not available yet");
}

int ngi_render_batch(ngi_scene* const*, const ngi_render_params*, int, double* const*, double*, ngi_stats*) {
    return fail(NGI_ERR_UNSUPPORTED, "batched rendering is not available yet");
}

}  // extern "C"
