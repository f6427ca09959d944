// Rendering without device allocations (no counterpart in the reference).
//
// ngi_render_device_noalloc renders into a caller-owned device film with the
// megakernel or the persistent kernels, using for everything else -- the
// ray counters, the persistent kernels' path counter and the per-wavefront
// shadow-ray queues -- one file-scope device array, `scratch`, instead of
// session buffers.  It calls no hipMalloc / hipFree: for callers that must
// not allocate (code under stream capture, serving loops, harnesses that
// forbid device allocation).
//
// Layout of `scratch` (every GPU holds its own copy until the process ends):
//   [0, 256)             CNT_N 64-bit ray/sample counters, then the 64-bit
//                        path counter of the persistent kernels
//   [256, 256 + queues)  shadow-ray rings: NGI_NOALLOC_WAVES wavefronts x
//                        NGI_SHQ records x 48 B (ptdirect / ltdirect, FP32
//                        shading); FP64 shading's 64-B records fill it with
//                        3,840 wavefronts, and its launches take as many
//                        blocks as have rings (960 of 1,024 at 4 per CU)
// Every call clears the counters with a fill kernel on the scene's stream,
// so a call sees nothing of the previous one; calls that overlap on one GPU
// (other streams, other threads) share the array and race.
#include <cstring>
#include <mutex>

#include "ngi_render.h"

namespace ngi {

// Wavefronts whose shadow rings fit in the scratch array.  MI355X holds
// 256 CUs x 4 SIMDs x 5 waves = 5,120 waves of the persistent kernels at
// full occupancy; 5,120 rings of 6 KiB are 30 MiB.
#define NGI_NOALLOC_WAVES 5120u
#define NGI_NOALLOC_HEAD 256u
#define NGI_NOALLOC_QUEUE_BYTES ((size_t)3 * NGI_NOALLOC_WAVES * NGI_SHQ * sizeof(float4))
#define NGI_NOALLOC_BYTES (NGI_NOALLOC_HEAD + NGI_NOALLOC_QUEUE_BYTES)

static_assert((CNT_N + 1) * sizeof(unsigned long long) <= NGI_NOALLOC_HEAD, "counters must fit the scratch head");

static __device__ __align__(256) unsigned char scratch[NGI_NOALLOC_BYTES];

// Device address of `scratch` on each device, looked up once per device
// with that device current.  The array's name in host code is a host-side
// stand-in, never a device address.
#define NGI_NOALLOC_MAX_DEVICES 64
static void* g_scratch_ptr[NGI_NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

static int scratch_address(void** out) {
    int dev = 0;
    NGI_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NGI_NOALLOC_MAX_DEVICES) return fail(NGI_ERR_DEVICE, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(NGI_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return NGI_OK;
}

__global__ void noalloc_clear(unsigned long long* head, int n) {
    if ((int)threadIdx.x < n) head[threadIdx.x] = 0ull;
}

template <typename FilmT>
__global__ void noalloc_fill(FilmT* film, size_t n, FilmT v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) film[i] = v;
}

template <typename FilmT>
__global__ void noalloc_scale(FilmT* film, size_t n, double s) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        film[i] = (FilmT)((double)film[i] * s);
}

static unsigned fill_grid(size_t n) { return (unsigned)std::min<size_t>(4096, std::max<size_t>(1, (n + 255) / 256)); }

static int render_noalloc(ngi_scene* scene, const ngi_render_params& p, void* d_film, ngi_stats* stats) {
    int st = check_render_params(scene, p);
    if (st) return st;
    if (p.renderer == NGI_RENDERER_BDPT)
        return fail(NGI_ERR_UNSUPPORTED, "bdpt keeps per-lane subpath records too large for the fixed scratch; use ngi_render");
    const bool persistent = p.engine == NGI_ENGINE_PERSISTENT || is_light_tracer(p.renderer);
    if (!persistent && p.engine != NGI_ENGINE_MEGAKERNEL)
        return fail(NGI_ERR_UNSUPPORTED, "the wavefront engine keeps per-path queues; use the persistent or megakernel engine");
    if ((int64_t)p.paths_in_flight > (int64_t)NGI_NOALLOC_WAVES * 64)
        return fail(NGI_ERR_INVALID_ARG, "paths_in_flight %d exceeds the %u lanes whose shadow queues fit the scratch", p.paths_in_flight,
                    NGI_NOALLOC_WAVES * 64u);
    if (!scene->dev || scene->dev->params_version != scene->params_version || !scene->built)
        return fail(NGI_ERR_INVALID_ARG, "the scene is not resident on the device with its current parameters: render or trace "
                                         "it once with an allocating entry point first");
    void* base = nullptr;
    if ((st = scratch_address(&base))) return st;
    hipStream_t str = scene->dev->stream;
    unsigned long long* head = reinterpret_cast<unsigned long long*>(base);
    unsigned long long* next = head + CNT_N;
    void* queues = reinterpret_cast<unsigned char*>(base) + NGI_NOALLOC_HEAD;
    size_t queue_bytes = NGI_NOALLOC_QUEUE_BYTES;

    const bool film64 = p.precision != NGI_PRECISION_FP32_ALL;
    const uint32_t npix = (uint32_t)p.width * (uint32_t)p.height;
    const size_t nf = (size_t)npix * 3;
    const uint64_t total = p.sampling == NGI_SAMPLING_PER_PIXEL ? (uint64_t)npix * (uint64_t)p.spp : (uint64_t)p.num_samples;
    if (p.sampling == NGI_SAMPLING_PER_PIXEL && p.spp <= 0) return fail(NGI_ERR_INVALID_ARG, "spp must be positive");

    noalloc_clear<<<1, 64, 0, str>>>(head, CNT_N + 1);
    if (film64) noalloc_fill<double><<<fill_grid(nf), 256, 0, str>>>(reinterpret_cast<double*>(d_film), nf, 0.0);
    else noalloc_fill<float><<<fill_grid(nf), 256, 0, str>>>(reinterpret_cast<float*>(d_film), nf, 0.0f);
    NGI_HIP_CHECK(hipGetLastError());

    // Same arguments as a session's render of these parameters.
    Session S;
    S.scene = scene;
    S.p = p;
    S.npix = npix;
    S.counters = head;
    const int width = scene->dev->view.bvh_width;
    if (persistent) {
        RenderArgs a = S.args(d_film);
        a.sample_base = 0;
        a.path_begin = 0;
        a.path_count = total;
        LaneMem mem{&queues, &queue_bytes, str};
        mem.fixed = true;
        Lanes ln;
        ln.want = p.paths_in_flight > 0 ? p.paths_in_flight : (int64_t)NGI_NOALLOC_WAVES * 64;
        if (is_light_tracer(p.renderer))
            st = launch_light(a, next, mem, p.precision, p.renderer == NGI_RENDERER_LTDIRECT, width, p.collect_counters != 0, str, ln);
        else
            st = launch_persistent(a, next, mem, p.precision, p.renderer == NGI_RENDERER_PTDIRECT, width, p.collect_counters != 0, str, ln);
        if (st) return st;
    } else {
        // Launches of at most 2^22 paths (whole rows of samples), as a session's.
        const uint64_t chunk = p.sampling == NGI_SAMPLING_PER_PIXEL ? (uint64_t)npix * std::max<uint64_t>(1, (1u << 22) / npix) : (1u << 22);
        for (uint64_t done = 0; done < total; done += std::min<uint64_t>(chunk, total - done)) {
            RenderArgs a = S.args(d_film);
            a.path_begin = done;
            a.sample_base = 0;
            a.path_count = std::min<uint64_t>(chunk, total - done);
            launch_megakernel(a, p.precision, p.renderer == NGI_RENDERER_PTDIRECT, width, p.collect_counters != 0, str);
            NGI_HIP_CHECK(hipGetLastError());
        }
    }
    // W*H/N scaling (src/nanogi.cpp:434-437), on the device.
    const double scale = total ? (double)npix / (double)total : 0.0;
    if (film64) noalloc_scale<double><<<fill_grid(nf), 256, 0, str>>>(reinterpret_cast<double*>(d_film), nf, scale);
    else noalloc_scale<float><<<fill_grid(nf), 256, 0, str>>>(reinterpret_cast<float*>(d_film), nf, scale);
    NGI_HIP_CHECK(hipGetLastError());
    if (stats) {
        unsigned long long c[CNT_N];
        NGI_HIP_CHECK(hipMemcpyAsync(c, head, sizeof(c), hipMemcpyDeviceToHost, str));
        NGI_HIP_CHECK(hipStreamSynchronize(str));
        memset(stats, 0, sizeof(*stats));
        stats->samples = (int64_t)c[CNT_SAMPLES];
        stats->extension_rays = (int64_t)c[CNT_EXT_RAYS];
        stats->shadow_rays = (int64_t)c[CNT_SHADOW_RAYS];
        stats->node_visits = (double)c[CNT_EXT_NODES];
        stats->tri_tests = (double)c[CNT_EXT_TRIS];
        stats->shadow_node_visits = (double)c[CNT_SH_NODES];
        stats->shadow_tri_tests = (double)c[CNT_SH_TRIS];
    }
    return NGI_OK;
}

}  // namespace ngi

extern "C" int ngi_render_device_noalloc(ngi_scene* scene, const ngi_render_params* p, void* d_film, ngi_stats* stats) {
    ngi::clear_error();
    if (!scene || !p || !d_film) return ngi::fail(NGI_ERR_INVALID_ARG, "null argument");
    return ngi::render_noalloc(scene, *p, d_film, stats);
}
