// Render session and the launch entry points of its engines (megakernel,
// wavefront, persistent path tracer, light tracers, bdpt).
#pragma once

#include <algorithm>
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

// Scenes below this many triangles take path ids in chunks (RenderArgs::path_chunk).
#ifndef NGI_CHUNKED_FEED_TRIS
#define NGI_CHUNKED_FEED_TRIS 16384
#endif

// Concurrency of a persistent launch.  want > 0 caps the paths in flight
// (rounded up to whole blocks); 0 keeps every block the device can hold
// resident; want < 0 launches nothing and only reports that full size.
struct Lanes {
    int64_t want = 0;
    int64_t* full = nullptr;   // optional: full size in lanes
};

// Grid of a persistent kernel under `ln`; *grid = 0 for a query (want < 0).
inline int persistent_grid(const void* kern, int block, size_t lds, uint64_t paths, Lanes ln, unsigned* grid) {
    int dev = 0, cus = 0, per_cu = 0;
    NGI_HIP_CHECK(hipGetDevice(&dev));
    NGI_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    NGI_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, lds));
    const uint64_t full = (uint64_t)std::max(1, cus) * (uint64_t)std::max(1, per_cu);
    if (ln.full) *ln.full = (int64_t)(full * (uint64_t)block);
    *grid = 0;
    if (ln.want < 0) return NGI_OK;
    uint64_t blocks = full;
    if (ln.want > 0) blocks = std::min<uint64_t>(full, std::max<uint64_t>(1, ((uint64_t)ln.want + block - 1) / block));
    const uint64_t need = (paths + block - 1) / block;
    *grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(need, blocks));
    return NGI_OK;
}

// Device scratch owned by a session and grown on demand (per-lane bdpt
// subpath records, per-wavefront shadow queues).  Growing waits for the
// session's stream first: an earlier launch may still be using the buffer.
struct LaneMem {
    void** ptr;
    size_t* bytes;
    hipStream_t stream;
    bool fixed = false;   // caller-owned buffer of *bytes bytes: never reallocated
    int reserve(size_t n) const {
        if (n == 0) return fail(NGI_ERR_INVALID_ARG, "empty per-lane record buffer");
        if (*bytes >= n) return NGI_OK;
        if (fixed) return fail(NGI_ERR_INVALID_ARG, "launch needs %zu bytes of per-lane records, the fixed buffer holds %zu", n, *bytes);
        if (*ptr) {
            NGI_HIP_CHECK(hipStreamSynchronize(stream));
            NGI_HIP_CHECK(hipFree(*ptr));
        }
        *ptr = nullptr;
        *bytes = 0;
        if (hipMalloc(ptr, n) != hipSuccess) {
            *ptr = nullptr;
            (void)hipGetLastError();   // a later call's launch check must not see this refusal
            return fail(NGI_ERR_OOM, "cannot allocate %zu bytes of per-lane records", n);
        }
        *bytes = n;
        return NGI_OK;
    }
};

void launch_megakernel(const RenderArgs& a, int precision, bool direct, int width, bool count, hipStream_t st);
int launch_persistent(const RenderArgs& a, unsigned long long* next, LaneMem mem, int precision, bool direct, int width, bool count,
                      hipStream_t st, Lanes ln = Lanes());
// Light tracers (ngi_light.hip): lt (direct = false) and ltdirect.
int launch_light(const RenderArgs& a, unsigned long long* next, LaneMem mem, int precision, bool direct, int width, bool count,
                 hipStream_t st, Lanes ln = Lanes());

inline bool is_light_tracer(int renderer) { return renderer == NGI_RENDERER_LT || renderer == NGI_RENDERER_LTDIRECT; }
// Bidirectional path tracer (ngi_bdpt.hip): FP32 shading; (re)allocates the
// per-lane subpath records in *mem as needed.
int launch_bdpt(const RenderArgs& a, unsigned long long* next, LaneMem mem, bool film64, int width, bool count, hipStream_t st,
                Lanes ln = Lanes());
// Renderers that always run on the persistent engine.
inline bool persistent_only(int renderer) { return is_light_tracer(renderer) || renderer == NGI_RENDERER_BDPT; }

// Validation shared by every render entry point (renderer, sizes, sampling,
// precision, engine, roulette, vertex cap, sensor).
int check_render_params(const ngi_scene* sc, const ngi_render_params& p);

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
    void* lane_mem = nullptr;    // bdpt subpath records or ptdirect shadow queues
    size_t lane_bytes = 0;
    uint64_t paths_done = 0;
    uint32_t samples_done = 0;
    int batches = 0;
    ngi_stats stats;
    std::vector<TimedLaunch> pending;
    WavefrontState* wf = nullptr;
    uint32_t wf_cap_max = 1u << 25;   // default wavefront queues: largest capacity not refused by the allocator

    hipStream_t stream() const { return scene->dev->stream; }   // valid after ensure_device()
    int ensure_device();
    int init(ngi_scene* sc, const ngi_render_params& pp);
    int reset();
    void destroy();
    int zero(void* p, size_t n_elems, bool f64);
    RenderArgs args(void* target_film) const;
    int render_into(void* target, int spp);
    // Launch the persistent kernel of this renderer under `ln`.
    int launch_pk_engine(const RenderArgs& a, Lanes ln);
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
