// Light tracers (lt, ltdirect) on the persistent engine: lanes take the
// next light-path id when their path ends (from the wavefront's chunk of
// ids), as the path tracers' persistent kernel does.  The vertex loop is in ngi_light.h.
#include <algorithm>
#include <type_traits>

#include "ngi_light.h"
#include "ngi_render.h"

namespace ngi {

#define NGI_LT_BLOCK 256
// Minimum waves per SIMD for the light kernels' register allocation: 5
// (96 VGPRs, some spilled) measured fastest for ltdirect against 3 (the
// unconstrained 141 VGPRs) and 4: +22% on the sphere field, +19% on the nave,
// -4% on the 36-triangle Cornell box.
#ifndef NGI_LT_WAVES
#define NGI_LT_WAVES 5
#endif
// ltdirect's variant for small scenes (chunked path feed) gets 4, without
// spills: 1,072 against 1,056 M samples/s on the Cornell box (lt loses 3%
// with 4 and keeps 5).
#ifndef NGI_LT_WAVES_SMALL
#define NGI_LT_WAVES_SMALL 4
#endif

// ltdirect in FP32 sends its sensor connections through the wavefront's
// shadow queue and traces them 64 at a time.
template <typename R, typename FilmT, bool DIRECT, int WIDTH, bool COUNT, bool CHUNKED>
__global__ void __launch_bounds__(NGI_LT_BLOCK, CHUNKED && DIRECT ? NGI_LT_WAVES_SMALL : NGI_LT_WAVES) light_kernel(RenderArgs a, unsigned long long* next, ShadowQueue q) {
    constexpr bool QUEUE = DIRECT && std::is_same<R, float>::value;
    extern __shared__ uint32_t stk[];   // a.sc.stack_entries x block (wide BVH only)
    uint32_t cnt[CNT_N] = {0, 0, 0, 0, 0, 0, 0};
    const Cam<R> cam = load_cam<R>(a.sc.cam);
    const DevPrim<R>* prims = prims_of<R>(a.sc);
    const int max_nv = clamp_vertices(a.max_nv);
    const uint32_t lane = __lane_id();
    const uint32_t wave = blockIdx.x * (NGI_LT_BLOCK / 64) + (threadIdx.x >> 6);
    uint32_t qhead = 0, qcount = 0;   // wave-uniform shadow queue state (< NGI_SHQ)
    LightState<R> ls;
    PathFeed feed;
    bool active = false;
    for (;;) {
        const bool exhausted = feed.done();
        if (!exhausted) {
            const unsigned long long need = __ballot(!active);
            if (need) {
                const unsigned long long id = feed.take<CHUNKED>(next, a.path_count, need, a.path_chunk);
                if (!active && id != ~0ull) {
                    cnt[CNT_SAMPLES]++;
                    active = light_start<R>(a, a.path_begin + id, ls);
                }
            }
        }
        const bool any = __ballot(active) != 0;
        if (QUEUE) shq_flush<FilmT, WIDTH, COUNT>(a, q, wave, qhead, qcount, !any && exhausted, stk + threadIdx.x, NGI_LT_BLOCK, cnt);
        if (!any) {
            if (exhausted) break;
            continue;
        }
        bool pushed = false;
        if (active)
            active = light_vertex<R, FilmT, DIRECT, WIDTH, COUNT, QUEUE>(a, cam, prims, ls, stk + threadIdx.x, NGI_LT_BLOCK, max_nv, cnt, &q,
                                                                         wave, qhead + qcount, &pushed);
        if (QUEUE) qcount += (uint32_t)__popcll(__ballot(pushed));   // at most 63 + 64 < NGI_SHQ
    }
    if (a.counters) {
        for (int k = 0; k < CNT_N; ++k) {
            if (!COUNT && (k == CNT_EXT_NODES || k == CNT_EXT_TRIS || k == CNT_SH_NODES || k == CNT_SH_TRIS)) continue;
            unsigned long long x = cnt[k];
            for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
            if ((threadIdx.x & 63) == 0 && x) atomicAdd(a.counters + k, x);
        }
    }
}

template <typename R, typename FilmT, bool DIRECT, int WIDTH, bool COUNT>
static int launch_lt(const RenderArgs& a, unsigned long long* next, LaneMem mem, Lanes ln, hipStream_t st) {
    const size_t lds = WIDTH == 8 ? (size_t)a.sc.stack_entries * NGI_LT_BLOCK * 4 : 0;
    auto kern = a.path_chunk ? light_kernel<R, FilmT, DIRECT, WIDTH, COUNT, true> : light_kernel<R, FilmT, DIRECT, WIDTH, COUNT, false>;
    unsigned grid = 0;
    int s = persistent_grid(reinterpret_cast<const void*>(kern), NGI_LT_BLOCK, lds, a.path_count, ln, &grid);
    if (s || !grid) return s;
    ShadowQueue q;
    q.waves = grid * (NGI_LT_BLOCK / 64);
    q.rec = nullptr;
    if (DIRECT && std::is_same<R, float>::value) {
        if ((s = mem.reserve((size_t)3 * q.waves * NGI_SHQ * sizeof(float4)))) return s;
        q.rec = reinterpret_cast<float4*>(*mem.ptr);
    }
    kern<<<grid, NGI_LT_BLOCK, lds, st>>>(a, next, q);
    NGI_HIP_CHECK(hipGetLastError());
    return NGI_OK;
}

template <typename R, typename FilmT, bool COUNT>
static int launch_lt3(const RenderArgs& a, unsigned long long* next, LaneMem mem, bool direct, int width, Lanes ln, hipStream_t st) {
    if (direct)
        return width == 8 ? launch_lt<R, FilmT, true, 8, COUNT>(a, next, mem, ln, st) : launch_lt<R, FilmT, true, 2, COUNT>(a, next, mem, ln, st);
    return width == 8 ? launch_lt<R, FilmT, false, 8, COUNT>(a, next, mem, ln, st) : launch_lt<R, FilmT, false, 2, COUNT>(a, next, mem, ln, st);
}

int launch_light(const RenderArgs& a, unsigned long long* next, LaneMem mem, int precision, bool direct, int width, bool count,
                 hipStream_t st, Lanes ln) {
    if (precision == NGI_PRECISION_FP64_ALL) return launch_lt3<double, double, false>(a, next, mem, direct, width, ln, st);
    if (precision == NGI_PRECISION_FP32_ALL) return launch_lt3<float, float, false>(a, next, mem, direct, width, ln, st);
    if (count) return launch_lt3<float, double, true>(a, next, mem, direct, width, ln, st);
    return launch_lt3<float, double, false>(a, next, mem, direct, width, ln, st);
}

}  // namespace ngi
