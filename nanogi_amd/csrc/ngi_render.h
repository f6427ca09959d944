// Render session shared by the megakernel and wavefront engines.
#pragma once

#include <vector>

#include "ngi_path.h"

namespace ngi {

enum { K_MEGA = 0, K_EXTEND = 1, K_SHADOW = 2, K_SHADE = 3, K_OTHER = 4 };

struct TimedLaunch {
    int kind;
    int64_t items;
    hipEvent_t a, b;
};

struct WavefrontState;   // ngi_wavefront.hip

void launch_megakernel(const RenderArgs& a, int precision, bool direct, int width, bool count, hipStream_t st);
int launch_persistent(const RenderArgs& a, unsigned long long* next, int precision, bool direct, int width, bool count, hipStream_t st);
// Light tracers (ngi_light.hip): lt (direct = false) and ltdirect.
int launch_light(const RenderArgs& a, unsigned long long* next, int precision, bool direct, int width, bool count, hipStream_t st);

inline bool is_light_tracer(int renderer) { return renderer == NGI_RENDERER_LT || renderer == NGI_RENDERER_LTDIRECT; }
// Bidirectional path tracer (ngi_bdpt.hip): FP32 shading; (re)allocates the
// per-lane subpath records in *mem as needed.
int launch_bdpt(const RenderArgs& a, unsigned long long* next, void** mem, size_t* mem_bytes, bool film64, int width, bool count,
                hipStream_t st);
// Renderers that always run on the persistent engine.
inline bool persistent_only(int renderer) { return is_light_tracer(renderer) || renderer == NGI_RENDERER_BDPT; }

struct Session {
    ngi_scene* scene = nullptr;
    ngi_render_params p;
    bool film64 = true;
    uint32_t npix = 0;
    void* film = nullptr;        // FilmT[npix*3], unnormalised sums
    void* counters = nullptr;    // unsigned long long[CNT_N]
    void* m1 = nullptr;          // double[npix*3] sum of normalised batch films
    void* m2 = nullptr;          // double[npix*3] sum of squares
    void* batch = nullptr;       // FilmT[npix*3] current batch film
    void* next_path = nullptr;   // unsigned long long path counter of the persistent engine
    void* bdpt_mem = nullptr;    // bdpt subpath records
    size_t bdpt_bytes = 0;
    uint64_t paths_done = 0;
    uint32_t samples_done = 0;
    int batches = 0;
    ngi_stats stats;
    std::vector<TimedLaunch> pending;
    WavefrontState* wf = nullptr;

    hipStream_t stream() const { return scene->dev->stream; }
    int init(ngi_scene* sc, const ngi_render_params& pp);
    int reset();
    void destroy();
    int zero(void* p, size_t n_elems, bool f64);
    RenderArgs args(void* target_film) const;
    int render_into(void* target, int spp);
    int render(int spp);
    int render_batch(int spp);
    int sync();
    int get_film(double* out, double* var);
    int get_stats(ngi_stats* out);
    int timed_begin(int kind, int64_t items, TimedLaunch& t);
    int timed_end(TimedLaunch& t);
    // Wavefront engine (ngi_wavefront.hip).
    int render_wavefront(void* target, uint64_t total_paths);
    void release_engine();
};

}  // namespace ngi
