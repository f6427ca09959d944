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

// The light kernel for FP32 shading over the wide BVH with resumable
// extension traversals, as resume_kernel for the path tracers: a wavefront
// stops traversing once at most NGI_LT_KEEP of its lanes are still inside
// the tree; the others complete their vertices, take new light paths and
// start their next rays while the waiting lanes keep their traversal state,
// direction and roulette number in LDS.  ltdirect's sensor connections go
// through the shadow rings: drained with refills from rings of
// NGI_SHQ_LONG records on large scenes, in rounds of 64 on small ones.
// Measured (ltdirect, M samples/s, sphere field / nave): keep 4: 1,952 / 716,
// 8: 2,061 / 742, 16: 2,098 / 743, 24: 1,924 / 660.  FP64 shading (R =
// double) parks its direction and roulette number in FP64, queues its
// sensor connections with FP64 contributions on large scenes and makes them
// in place on small ones, at NGI_LT_WAVES_FP64 waves per SIMD (ltdirect,
// sphere field / nave: 4 waves 1,498 / 610, 3 waves 1,465 / 564, 5 waves
// 1,210 / 563).
#ifndef NGI_LT_KEEP
#define NGI_LT_KEEP 16
#endif
#ifndef NGI_LT_WAVES_FP64
#define NGI_LT_WAVES_FP64 4
#endif
// Light paths are refilled once NGI_LT_PATH_REFILL lanes wait (exact feed),
// as the path tracer's lanes are.  Measured (ltdirect, M samples/s, nave
// FP32 / FP64, sphere field FP32): at once 737 / 612 / 2,095; 8 waiting
// 766 / 620 / 2,098; 16: 770 / 627 / 2,095; 24: 747 / 609 / 2,105.
#ifndef NGI_LT_PATH_REFILL
#define NGI_LT_PATH_REFILL 16
#endif
template <typename R, bool DIRECT, bool CHUNKED>
constexpr unsigned light_resume_waves() {
    return std::is_same<R, double>::value ? NGI_LT_WAVES_FP64 : (CHUNKED && DIRECT ? NGI_LT_WAVES_SMALL : NGI_LT_WAVES);
}
template <typename R, typename FilmT, bool DIRECT, bool COUNT, bool CHUNKED>
__global__ void __launch_bounds__(NGI_LT_BLOCK, (light_resume_waves<R, DIRECT, CHUNKED>())) light_resume_kernel(RenderArgs a, unsigned long long* next, ShadowQueue q) {
    constexpr bool QUEUE = DIRECT && (std::is_same<R, float>::value || !CHUNKED);
    // LDS: extension stacks, shadow stacks (a.sc.stack_entries words per
    // lane each), then the parked lane state.
    extern __shared__ uint32_t stk[];
    uint32_t* ext_stk = stk + threadIdx.x;
    uint32_t* sh_stk = stk + (size_t)a.sc.stack_entries * NGI_LT_BLOCK + threadIdx.x;
    uint32_t* pk = stk + (size_t)2 * a.sc.stack_entries * NGI_LT_BLOCK + threadIdx.x;
    uint32_t cnt[CNT_N] = {0, 0, 0, 0, 0, 0, 0};
    TraceCounters tc{0, 0};
    const Cam<R> cam = load_cam<R>(a.sc.cam);
    const DevPrim<R>* prims = prims_of<R>(a.sc);
    const int max_nv = clamp_vertices(a.max_nv);
    const uint32_t wave = blockIdx.x * (NGI_LT_BLOCK / 64) + (threadIdx.x >> 6);
    const bool geometry = a.sc.num_tris != 0;
    uint32_t qhead = 0, qcount = 0;   // wave-uniform: oldest slot, queued rays
    LightState<R> ls;
    PathFeed feed;
    bool active = false;   // the lane holds a light path
    bool busy = false;     // ... whose extension ray is being traced
    for (;;) {
        // Vertices whose extension rays have finished.
        if (active && !busy)
            active = light_end<R>(a, prims, ls, parked_dir_r<NGI_LT_BLOCK, R>(pk), parked_urr_r<NGI_LT_BLOCK, R>(pk), parked_hit<NGI_LT_BLOCK>(pk));
        if (!feed.done()) {
            const unsigned long long need = __ballot(!active);
            if (need && ((uint32_t)__popcll(need) >= (CHUNKED ? 1u : (uint32_t)NGI_LT_PATH_REFILL) || need == ~0ull)) {
                const unsigned long long id = feed.take<CHUNKED>(next, a.path_count, need, a.path_chunk);
                if (!active && id != ~0ull) {
                    cnt[CNT_SAMPLES]++;
                    active = light_start<R>(a, a.path_begin + id, ls);
                }
            }
        }
        // Next vertex of every light path not inside the tree.
        bool pushed = false;
        if (active && !busy) {
            V3<R> wo;
            R urr;
            if (light_begin<R, FilmT, DIRECT, 8, COUNT, QUEUE>(a, cam, prims, ls, sh_stk, NGI_LT_BLOCK, max_nv, cnt, &q, wave, qhead + qcount,
                                                               &pushed, wo, urr)) {
                cnt[CNT_EXT_RAYS]++;
                Trav8 tv;
                trav8_begin(FLT_MAX, tv);
                busy = geometry;
                park_trav<NGI_LT_BLOCK>(pk, tv);
                park_dir<NGI_LT_BLOCK>(pk, wo, urr);
            } else {
                active = false;
            }
        }
        const bool any = __ballot(active) != 0;
        if (QUEUE) {
            qcount += (uint32_t)__popcll(__ballot(pushed));   // at most q.cap - 65 + 64 (63 + 64 chunked)
            if (CHUNKED) shq_flush<FilmT, 8, COUNT, R>(a, q, wave, qhead, qcount, !any && feed.done(), sh_stk, NGI_LT_BLOCK, cnt);
            else shq_drain<FilmT, COUNT, R>(a, q, wave, qhead, qcount, !any && feed.done(), sh_stk, NGI_LT_BLOCK, cnt);
        }
        if (!any) {
            if (feed.done()) break;
            continue;
        }
        if (__ballot(busy)) {
            // Lanes without a ray load and store back what they parked last.
            Trav8 tv = parked_trav<NGI_LT_BLOCK>(pk);
            const V3<R> d = parked_dir_r<NGI_LT_BLOCK, R>(pk);
            const RayPre r = make_pre(TraceRay{(float)ls.s.p.x, (float)ls.s.p.y, (float)ls.s.p.z, (float)d.x, (float)d.y, (float)d.z, NGI_EPS_F,
                                               FLT_MAX});
            const int keep = __builtin_amdgcn_readfirstlane(feed.done() ? 0 : NGI_LT_KEEP);
            trav8_run<COUNT>(a.sc, r, NGI_EPS_F, tv, busy, keep, tc, ext_stk, NGI_LT_BLOCK);
            park_trav<NGI_LT_BLOCK>(pk, tv);
        }
    }
    if (COUNT) cnt[CNT_EXT_NODES] += tc.nodes, cnt[CNT_EXT_TRIS] += tc.tris;
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
    // Resumable traversals while their LDS leaves room for
    // NGI_RESUME_MIN_BLOCKS blocks per CU.
    const size_t resume_lds = ((size_t)2 * a.sc.stack_entries + ParkWords<R>::value) * NGI_LT_BLOCK * 4;
    const unsigned blocks_per_cu = a.path_chunk ? light_resume_waves<R, DIRECT, true>() : light_resume_waves<R, DIRECT, false>();
    const bool resume = WIDTH == 8 && resume_lds * resume_min_blocks(blocks_per_cu) <= NGI_CU_LDS_BYTES;
    const size_t lds = WIDTH == 8 ? (resume ? resume_lds : (size_t)a.sc.stack_entries * NGI_LT_BLOCK * 4) : 0;
    auto kern = resume ? (a.path_chunk ? light_resume_kernel<R, FilmT, DIRECT, COUNT, true> : light_resume_kernel<R, FilmT, DIRECT, COUNT, false>)
                       : (a.path_chunk ? light_kernel<R, FilmT, DIRECT, WIDTH, COUNT, true> : light_kernel<R, FilmT, DIRECT, WIDTH, COUNT, false>);
    unsigned grid = 0;
    int s = persistent_grid(reinterpret_cast<const void*>(kern), NGI_LT_BLOCK, lds, a.path_count, ln, &grid);
    if (s || !grid) return s;
    ShadowQueue q;
    q.waves = grid * (NGI_LT_BLOCK / 64);
    q.rec = nullptr;
    if (DIRECT && (std::is_same<R, float>::value || (resume && !a.path_chunk))) {
        // Longer rings for the refilling drains; a caller-owned buffer gets
        // the longest ring it holds, and as many blocks as it has rings for.
        const size_t rec = (size_t)ShqWords<R>::value * sizeof(float4);
        if (resume && !a.path_chunk) {
            q.cap = NGI_SHQ_LONG;
            while (mem.fixed && q.cap > NGI_SHQ && rec * q.waves * q.cap > *mem.bytes) q.cap >>= 1;
        }
        if (mem.fixed && rec * q.waves * q.cap > *mem.bytes) {
            grid = (unsigned)std::min<size_t>(grid, *mem.bytes / (rec * q.cap * (NGI_LT_BLOCK / 64)));
            if (!grid) return fail(NGI_ERR_INVALID_ARG, "the fixed buffer of %zu bytes holds no block's shadow rings", *mem.bytes);
            q.waves = grid * (NGI_LT_BLOCK / 64);
        }
        if ((s = mem.reserve(rec * q.waves * q.cap))) return s;
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
