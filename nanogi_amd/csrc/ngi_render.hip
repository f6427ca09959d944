// Rendering entry points: the one-path-per-lane megakernel and the render
// session (film, statistics, batch moments).  The wavefront engine lives in
// ngi_wavefront.hip and shares the session.
//
// Replaces Renderer::Render / RenderProcess (src/nanogi.cpp:182-440): the TBB
// loop over samples becomes launches over path ids; per-thread FP64 films
// become one FP64 film updated with hardware FP64 atomics; the final
// W*H/N scaling (src/nanogi.cpp:434-437) is applied on read-back.
#include <chrono>
#include <cstring>
#include <type_traits>
#include <vector>

#include "ngi_path.h"
#include "ngi_render.h"

namespace ngi {

#define NGI_MK_BLOCK 256
// Minimum waves per SIMD the register allocator must leave room for in the
// persistent kernel: 5 (at most 102 VGPRs) measured fastest on the 1.8M
// triangle ptdirect workload (960 vs 846 M samples/s unconstrained at 147
// VGPRs, 933 at 4 waves, 948 at 6; with the shadow queue 1,447 M at 5, 1,362
// at 4, 1,426 at 6).  ptdirect's variant for small scenes (chunked path
// feed) gets 4, without spills: on the 36-triangle Cornell box 1,200 M
// samples/s against 1,075 M with 5 (pt loses 1% with 4 and keeps 5).  With
// resumable traversals 4 waves (no spills) still lose on every scene: glossy
// sphere fields 0.4-1.5%, the nave 1.9%, a diffuse room 0.1%.
#ifndef NGI_PK_WAVES
#define NGI_PK_WAVES 5
#endif
#ifndef NGI_PK_WAVES_SMALL
#define NGI_PK_WAVES_SMALL 4
#endif

template <typename T>
__device__ __forceinline__ void wave_add(unsigned long long* p, T v) {
    unsigned long long x = (unsigned long long)v;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63) == 0 && x) atomicAdd(p, x);
}

template <typename T>
__device__ __forceinline__ void flush_counters(unsigned long long* counters, const T* cnt, bool count) {
    if (!counters) return;
    for (int k = 0; k < CNT_N; ++k)
        if (count || (k != CNT_EXT_NODES && k != CNT_EXT_TRIS && k != CNT_SH_NODES && k != CNT_SH_TRIS)) wave_add(counters + k, cnt[k]);
}

// One lane = one path sample (ProcessSample_PT / ProcessSample_PTDirect).
template <typename R, typename FilmT, bool DIRECT, int WIDTH, bool COUNT>
__global__ void __launch_bounds__(NGI_MK_BLOCK) megakernel(RenderArgs a) {
    extern __shared__ uint32_t stk[];   // a.sc.stack_entries x block (wide BVH only)
    const uint64_t local = (uint64_t)blockIdx.x * NGI_MK_BLOCK + threadIdx.x;
    uint32_t cnt[CNT_N] = {0, 0, 0, 0, 0, 0, 0};
    if (local < a.path_count) {
        const Cam<R> cam = load_cam<R>(a.sc.cam);
        const DevPrim<R>* prims = prims_of<R>(a.sc);
        PathState<R> ps;
        path_start<R>(a, a.path_begin + local, cam, ps);
        cnt[CNT_SAMPLES] = 1;
        const int max_nv = clamp_vertices(a.max_nv);
        while (path_vertex<R, FilmT, DIRECT, WIDTH, COUNT>(a, cam, prims, ps, stk + threadIdx.x, NGI_MK_BLOCK, max_nv, cnt)) {
        }
    }
    flush_counters(a.counters, cnt, COUNT);
}

// Persistent lanes with path regeneration: a lane whose path ended takes the
// next path id (from its wavefront's chunk of ids), so every iteration
// traces with all lanes busy regardless of how uneven path lengths are.
// With ptdirect in FP32 the NEE shadow rays go through the wavefront's
// shadow queue (q, a ring per wavefront) and are traced 64 at a time.
template <typename R, typename FilmT, bool DIRECT, int WIDTH, bool COUNT, bool CHUNKED>
__global__ void __launch_bounds__(NGI_MK_BLOCK, CHUNKED && DIRECT ? NGI_PK_WAVES_SMALL : NGI_PK_WAVES) persistent_kernel(RenderArgs a, unsigned long long* next, ShadowQueue q) {
    constexpr bool QUEUE = DIRECT && std::is_same<R, float>::value;
    extern __shared__ uint32_t stk[];
    uint32_t cnt[CNT_N] = {0, 0, 0, 0, 0, 0, 0};
    const Cam<R> cam = load_cam<R>(a.sc.cam);
    const DevPrim<R>* prims = prims_of<R>(a.sc);
    const int max_nv = clamp_vertices(a.max_nv);
    const uint32_t lane = __lane_id();
    const uint32_t wave = blockIdx.x * (NGI_MK_BLOCK / 64) + (threadIdx.x >> 6);
    uint32_t qhead = 0, qcount = 0;   // wave-uniform: oldest slot, queued rays (< NGI_SHQ)
    PathState<R> ps;
    PathFeed feed;
    bool active = false;
    for (;;) {
        if (!feed.done()) {
            const unsigned long long need = __ballot(!active);
            if (need) {
                const unsigned long long id = feed.take<CHUNKED>(next, a.path_count, need, a.path_chunk);
                if (!active && id != ~0ull) {
                    path_start<R>(a, a.path_begin + id, cam, ps);
                    cnt[CNT_SAMPLES]++;
                    active = true;
                }
            }
        }
        const bool any = __ballot(active) != 0;
        if (QUEUE) shq_flush<FilmT, WIDTH, COUNT>(a, q, wave, qhead, qcount, !any, stk + threadIdx.x, NGI_MK_BLOCK, cnt);
        if (!any) break;
        bool pushed = false;
        if (active)
            active = path_vertex<R, FilmT, DIRECT, WIDTH, COUNT, QUEUE>(a, cam, prims, ps, stk + threadIdx.x, NGI_MK_BLOCK, max_nv, cnt,
                                                                        &q, wave, qhead + qcount, &pushed);
        if (QUEUE) qcount += (uint32_t)__popcll(__ballot(pushed));   // at most 63 + 64 < NGI_SHQ
    }
    flush_counters(a.counters, cnt, COUNT);
}

// The persistent kernel for FP32 shading over the wide BVH, with extension
// traversals that outlive a wavefront's shading step.  A wavefront stops
// traversing as soon as at most NGI_PK_KEEP of its lanes are still inside the
// tree (instead of waiting for its slowest ray); the lanes whose rays have
// finished complete their vertices, take new paths and start their next rays,
// while the others keep their traversal state and carry on in the next
// traversal step alongside the new rays.  Between traversal steps a lane's
// traversal state, direction and roulette number are parked in LDS (the
// shading code needs the registers), next to its traversal stack; shadow rays
// use a second LDS stack per lane.  Every path draws the same random numbers
// and makes the same node visits and triangle tests as in persistent_kernel;
// the film receives its additions in another order.
//
// Waiting for every lane kept 31% of the lanes busy per traversal iteration
// on the 1.8M-triangle sphere field; stopping at 8 lanes left, 63%, with 1.13
// traversal steps per ray.  Measured (M samples/s, sphere field / nave,
// ptdirect, 64 spp per launch), shadow rays in rounds of 64: waiting for all
// 1,455 / 597, keep 4: 1,612 / 655, 8: 1,639 / 674, 12: 1,638 / 682, 16:
// 1,616 / 684, 24: 1,543 / 662 (more keep, more shading steps with fewer
// lanes); with the drained 512-record shadow rings: keep 8: 1,753 / 709,
// 12: 1,738 / 717, 16: 1,704 / 715; with the light guide tables and the
// primitive type in the shading record: keep 6: 1,783 / 709, 8: 1,789 /
// 717, 10: 1,782 / 723, 12: 1,769 / 725 (refills once 12 / 16 / 24 lanes of
// a drain are idle: 1,782 / 1,789 / 1,788 on the sphere field).  The best
// keep follows the scene, not the tree depth: keep 8 against 12 (ptdirect,
// 1024², 32 spp) on the glossy sphere field at 4 sizes, 29k-1.8M triangles
// (wide BVH 6-9 levels), 2,207 / 2,108 / 2,059 / 1,776 against 2,139 / 2,052
// / 2,004 / 1,755 M; on the diffuse nave 716 against 721 M and a diffuse
// 4k-triangle room 1,421 against 1,468 M (both 6 levels).  A glossy or
// specular vertex costs about twice a diffuse one to shade, so scenes
// without such surfaces keep NGI_PK_KEEP_DIFFUSE: nave 720 -> 728 M, room
// 1,421 -> 1,473 M, FP64 nave 615 -> 629 M, the glossy scenes unchanged (12
// lanes).  With paths refilled 16 at a time (nave / room): diffuse keep 10:
// 739 / 1,447, 12: 743 / 1,470, 16: 748 / 1,496, 20: 749 / 1,505, 24: 746 /
// 1,504; glossy keep 6 / 8 / 10 on the sphere field: 1,861 / 1,863 / 1,867.
//
// FP64 shading (R = double) parks its direction and roulette number in FP64
// and queues its shadow rays with FP64 contributions (four float4s per
// record), at NGI_PK_WAVES_FP64 waves per SIMD; on small scenes (chunked
// path feed) it traces them in place.  Measured (M samples/s, FP64 ptdirect,
// sphere field / nave / Cornell box): the kernel that runs every traversal
// to its end 1,148 / 506 / 1,132; resumable with shadow rays in place 1,140
// / 537 / 1,140; with the drained rings (rounds of 64 on the Cornell box)
// 1,421 / 622 / 1,090.  Register budgets (sphere field / nave, ptdirect):
// 4 waves/SIMD (128 VGPRs) 1,422 / 623, 3 waves 1,300 / 541, 5 waves 1,194
// / 586.
#ifndef NGI_PK_KEEP
#define NGI_PK_KEEP 8
#endif
// Lanes whose paths have ended take new ones (an atomic on the path counter,
// then the camera vertex) only once NGI_PK_PATH_REFILL of them wait, so most
// shading steps skip the refill; the chunked feed (small scenes) takes ids
// from a wave's own chunk and refills at once.  Measured (M samples/s,
// ptdirect, 1024², 32 spp; sphere field 29k / 90k / 336k / 1.8M triangles,
// nave, a 4k-triangle room on the chunked feed): refill at once 2,209 /
// 2,104 / 2,057 / 1,778 / 722 / 1,470; once 4 wait 2,306 / 2,188 / 2,134 /
// 1,804 / 724 / 1,466; 8: 2,384 / 2,252 / 2,193 / 1,846 / 734 / 1,465; 12:
// 2,416 / 2,270 / 2,211 / 1,865 / 743 / 1,462; 16: 2,424 / 2,280 / 2,221 /
// 1,866 / 743 / 1,457; 24: 2,437 / 2,283 / 2,227 / 1,865 / 726 / 1,438;
// 32: 2,434 / 2,281 / 2,231 / 1,861 / 712 / 1,402.
#ifndef NGI_PK_PATH_REFILL
#define NGI_PK_PATH_REFILL 16
#endif
#ifndef NGI_PK_KEEP_DIFFUSE
#define NGI_PK_KEEP_DIFFUSE 16
#endif
#ifndef NGI_PK_WAVES_FP64
#define NGI_PK_WAVES_FP64 4
#endif
template <typename R, bool DIRECT, bool CHUNKED>
constexpr unsigned resume_waves() {
    return std::is_same<R, double>::value ? NGI_PK_WAVES_FP64 : (CHUNKED && DIRECT ? NGI_PK_WAVES_SMALL : NGI_PK_WAVES);
}
template <typename R, typename FilmT, bool DIRECT, bool COUNT, bool CHUNKED>
__global__ void __launch_bounds__(NGI_MK_BLOCK, (resume_waves<R, DIRECT, CHUNKED>())) resume_kernel(RenderArgs a, unsigned long long* next, ShadowQueue q) {
    constexpr bool QUEUE = DIRECT && (std::is_same<R, float>::value || !CHUNKED);
    // LDS: extension stacks, shadow stacks (a.sc.stack_entries words per lane
    // each), then the parked lane state (ParkWords<R> per lane).
    extern __shared__ uint32_t stk[];
    uint32_t* ext_stk = stk + threadIdx.x;
    uint32_t* sh_stk = stk + (size_t)a.sc.stack_entries * NGI_MK_BLOCK + threadIdx.x;
    uint32_t* pk = stk + (size_t)2 * a.sc.stack_entries * NGI_MK_BLOCK + threadIdx.x;
    uint32_t cnt[CNT_N] = {0, 0, 0, 0, 0, 0, 0};
    TraceCounters tc{0, 0};
    const Cam<R> cam = load_cam<R>(a.sc.cam);
    const DevPrim<R>* prims = prims_of<R>(a.sc);
    const int max_nv = clamp_vertices(a.max_nv);
    const uint32_t wave = blockIdx.x * (NGI_MK_BLOCK / 64) + (threadIdx.x >> 6);
    const bool geometry = a.sc.num_tris != 0;
    uint32_t qhead = 0, qcount = 0;   // wave-uniform: oldest slot, queued rays (< NGI_SHQ)
    PathState<R> ps;
    PathFeed feed;
    bool active = false;   // the lane holds a path
    bool busy = false;     // ... whose extension ray is being traced
    for (;;) {
        // Vertices whose extension rays have finished.
        if (active && !busy)
            active = vertex_end<R, FilmT, DIRECT>(a, prims, ps, parked_dir_r<NGI_MK_BLOCK, R>(pk), parked_urr_r<NGI_MK_BLOCK, R>(pk),
                                                  parked_hit<NGI_MK_BLOCK>(pk));
        if (!feed.done()) {
            const unsigned long long need = __ballot(!active);
            if (need && ((uint32_t)__popcll(need) >= (CHUNKED ? 1u : (uint32_t)NGI_PK_PATH_REFILL) || need == ~0ull)) {
                const unsigned long long id = feed.take<CHUNKED>(next, a.path_count, need, a.path_chunk);
                if (!active && id != ~0ull) {
                    path_start<R>(a, a.path_begin + id, cam, ps);
                    cnt[CNT_SAMPLES]++;
                    active = true;
                }
            }
        }
        // Next vertex of every path that is not inside the tree.
        bool pushed = false;
        if (active && !busy) {
            V3<R> wo;
            R urr;
            if (vertex_begin<R, FilmT, DIRECT, 8, COUNT, QUEUE>(a, cam, prims, ps, sh_stk, NGI_MK_BLOCK, max_nv, cnt, &q, wave, qhead + qcount,
                                                                 &pushed, wo, urr)) {
                cnt[CNT_EXT_RAYS]++;
                Trav8 tv;
                trav8_begin(FLT_MAX, tv);
                busy = geometry;
                park_trav<NGI_MK_BLOCK>(pk, tv);
                park_dir<NGI_MK_BLOCK>(pk, wo, urr);
            } else {
                active = false;
            }
        }
        const bool any = __ballot(active) != 0;
        if (QUEUE) {
            qcount += (uint32_t)__popcll(__ballot(pushed));   // at most q.cap - 65 + 64 (63 + 64 chunked)
            if (CHUNKED) shq_flush<FilmT, 8, COUNT, R>(a, q, wave, qhead, qcount, !any && feed.done(), sh_stk, NGI_MK_BLOCK, cnt);
            else shq_drain<FilmT, COUNT, R>(a, q, wave, qhead, qcount, !any && feed.done(), sh_stk, NGI_MK_BLOCK, cnt);
        }
        if (!any) {
            if (feed.done()) break;
            continue;
        }
        if (__ballot(busy)) {
            // Lanes without a ray load and store back what they parked last.
            Trav8 tv = parked_trav<NGI_MK_BLOCK>(pk);
            const RayPre r = make_pre(extension_ray(ps, parked_dir_r<NGI_MK_BLOCK, R>(pk)));
            const int keep = __builtin_amdgcn_readfirstlane(feed.done() ? 0 : (a.sc.any_glossy ? NGI_PK_KEEP : NGI_PK_KEEP_DIFFUSE));
            trav8_run<COUNT>(a.sc, r, NGI_EPS_F, tv, busy, keep, tc, ext_stk, NGI_MK_BLOCK);
            park_trav<NGI_MK_BLOCK>(pk, tv);
        }
    }
    if (COUNT) cnt[CNT_EXT_NODES] += tc.nodes, cnt[CNT_EXT_TRIS] += tc.tris;
    flush_counters(a.counters, cnt, COUNT);
}

template <typename R, typename FilmT, bool DIRECT, int WIDTH, bool COUNT>
static void launch_mk(const RenderArgs& a, hipStream_t st) {
    const unsigned grid = (unsigned)((a.path_count + NGI_MK_BLOCK - 1) / NGI_MK_BLOCK);
    const size_t lds = WIDTH == 8 ? (size_t)a.sc.stack_entries * NGI_MK_BLOCK * 4 : 0;
    megakernel<R, FilmT, DIRECT, WIDTH, COUNT><<<grid, NGI_MK_BLOCK, lds, st>>>(a);
}

template <typename R, typename FilmT, bool COUNT>
static void launch_mk3(const RenderArgs& a, bool direct, int width, hipStream_t st) {
    if (direct) {
        if (width == 8) launch_mk<R, FilmT, true, 8, COUNT>(a, st);
        else launch_mk<R, FilmT, true, 2, COUNT>(a, st);
    } else {
        if (width == 8) launch_mk<R, FilmT, false, 8, COUNT>(a, st);
        else launch_mk<R, FilmT, false, 2, COUNT>(a, st);
    }
}

template <typename R, typename FilmT, bool DIRECT, int WIDTH, bool COUNT>
static int launch_pk(const RenderArgs& a, unsigned long long* next, LaneMem mem, Lanes ln, hipStream_t st) {
    // Resumable traversals need a second LDS stack and the parked state per
    // lane: used while that leaves room for NGI_RESUME_MIN_BLOCKS blocks per
    // CU (on the 1.8M-triangle sphere field the wide BVH is 9 levels deep:
    // 31 KiB per block, the full 5 blocks per CU; deeper trees run at 4).
    const size_t resume_lds = ((size_t)2 * a.sc.stack_entries + ParkWords<R>::value) * NGI_MK_BLOCK * 4;
    const unsigned blocks_per_cu = a.path_chunk ? resume_waves<R, DIRECT, true>() : resume_waves<R, DIRECT, false>();
    const bool resume = WIDTH == 8 && resume_lds * resume_min_blocks(blocks_per_cu) <= NGI_CU_LDS_BYTES;
    const size_t lds = WIDTH == 8 ? (resume ? resume_lds : (size_t)a.sc.stack_entries * NGI_MK_BLOCK * 4) : 0;
    auto kern = resume ? (a.path_chunk ? resume_kernel<R, FilmT, DIRECT, COUNT, true> : resume_kernel<R, FilmT, DIRECT, COUNT, false>)
                       : (a.path_chunk ? persistent_kernel<R, FilmT, DIRECT, WIDTH, COUNT, true>
                                       : persistent_kernel<R, FilmT, DIRECT, WIDTH, COUNT, false>);
    unsigned grid = 0;
    int s = persistent_grid(reinterpret_cast<const void*>(kern), NGI_MK_BLOCK, lds, a.path_count, ln, &grid);
    if (s || !grid) return s;
    ShadowQueue q;
    q.waves = grid * (NGI_MK_BLOCK / 64);
    q.rec = nullptr;
    if (DIRECT && (std::is_same<R, float>::value || (resume && !a.path_chunk))) {
        // Longer rings for the refilling drains; a caller-owned buffer gets
        // the longest ring it holds, and as many blocks as it has rings for
        // (the persistent grid takes every path with fewer blocks too).
        const size_t rec = (size_t)ShqWords<R>::value * sizeof(float4);
        if (resume && !a.path_chunk) {
            q.cap = NGI_SHQ_LONG;
            while (mem.fixed && q.cap > NGI_SHQ && rec * q.waves * q.cap > *mem.bytes) q.cap >>= 1;
        }
        if (mem.fixed && rec * q.waves * q.cap > *mem.bytes) {
            grid = (unsigned)std::min<size_t>(grid, *mem.bytes / (rec * q.cap * (NGI_MK_BLOCK / 64)));
            if (!grid) return fail(NGI_ERR_INVALID_ARG, "the fixed buffer of %zu bytes holds no block's shadow rings", *mem.bytes);
            q.waves = grid * (NGI_MK_BLOCK / 64);
        }
        if ((s = mem.reserve(rec * q.waves * q.cap))) return s;
        q.rec = reinterpret_cast<float4*>(*mem.ptr);
    }
    kern<<<grid, NGI_MK_BLOCK, lds, st>>>(a, next, q);
    NGI_HIP_CHECK(hipGetLastError());
    return NGI_OK;
}

template <typename R, typename FilmT, bool COUNT>
static int launch_pk3(const RenderArgs& a, unsigned long long* next, LaneMem mem, bool direct, int width, Lanes ln, hipStream_t st) {
    if (direct)
        return width == 8 ? launch_pk<R, FilmT, true, 8, COUNT>(a, next, mem, ln, st) : launch_pk<R, FilmT, true, 2, COUNT>(a, next, mem, ln, st);
    return width == 8 ? launch_pk<R, FilmT, false, 8, COUNT>(a, next, mem, ln, st) : launch_pk<R, FilmT, false, 2, COUNT>(a, next, mem, ln, st);
}

int launch_persistent(const RenderArgs& a, unsigned long long* next, LaneMem mem, int precision, bool direct, int width, bool count,
                      hipStream_t st, Lanes ln) {
    if (precision == NGI_PRECISION_FP64_ALL) return launch_pk3<double, double, false>(a, next, mem, direct, width, ln, st);
    if (precision == NGI_PRECISION_FP32_ALL) return launch_pk3<float, float, false>(a, next, mem, direct, width, ln, st);
    if (count) return launch_pk3<float, double, true>(a, next, mem, direct, width, ln, st);
    return launch_pk3<float, double, false>(a, next, mem, direct, width, ln, st);
}

void launch_megakernel(const RenderArgs& a, int precision, bool direct, int width, bool count, hipStream_t st) {
    if (precision == NGI_PRECISION_FP64_ALL) launch_mk3<double, double, false>(a, direct, width, st);
    else if (precision == NGI_PRECISION_FP32_ALL) launch_mk3<float, float, false>(a, direct, width, st);
    else if (count) launch_mk3<float, double, true>(a, direct, width, st);
    else launch_mk3<float, double, false>(a, direct, width, st);
}

// ---------------------------------------------------------------------------
// Film moments: m1 += batch, m2 += batch^2 (batch films normalised).
// ---------------------------------------------------------------------------
template <typename FilmT>
__global__ void fold_batch(const FilmT* __restrict__ batch, double scale, double* __restrict__ m1, double* __restrict__ m2,
                           FilmT* __restrict__ film, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double b = (double)batch[i];
    const double v = b * scale;
    m1[i] += v;
    m2[i] += v * v;
    film[i] += (FilmT)b;
}

template <typename FilmT>
__global__ void zero_kernel(FilmT* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = FilmT(0);
}

// A rebuilt scene (ngi_scene_build) has released its device copy and its
// stream; the session's own buffers survive, and the scene is re-uploaded
// before the session enqueues anything again.
int Session::ensure_device() {
    if (scene->dev) return NGI_OK;
    return device_prepare(*scene);
}

int Session::zero(void* p, size_t n_elems, bool f64) {
    if (n_elems == 0) return NGI_OK;
    int st = ensure_device();
    if (st) return st;
    const unsigned grid = (unsigned)((n_elems + 255) / 256);
    if (f64) zero_kernel<double><<<grid, 256, 0, stream()>>>(reinterpret_cast<double*>(p), n_elems);
    else zero_kernel<float><<<grid, 256, 0, stream()>>>(reinterpret_cast<float*>(p), n_elems);
    NGI_HIP_CHECK(hipGetLastError());
    return NGI_OK;
}

int check_render_params(const ngi_scene* sc, const ngi_render_params& p) {
    if (p.renderer != NGI_RENDERER_PT && p.renderer != NGI_RENDERER_PTDIRECT && !is_light_tracer(p.renderer) &&
        p.renderer != NGI_RENDERER_BDPT)
        return fail(NGI_ERR_UNSUPPORTED, "renderer %d is not supported (pt, ptdirect, lt, ltdirect and bdpt are)", p.renderer);
    if (p.width <= 0 || p.height <= 0 || (int64_t)p.width * p.height > (1ll << 28)) return fail(NGI_ERR_INVALID_ARG, "bad image size");
    if (p.sampling != NGI_SAMPLING_PER_PIXEL && p.sampling != NGI_SAMPLING_REFERENCE_UNIFORM) return fail(NGI_ERR_INVALID_ARG, "bad sampling mode");
    if (p.sampling == NGI_SAMPLING_REFERENCE_UNIFORM && p.num_samples <= 0) return fail(NGI_ERR_INVALID_ARG, "num_samples must be positive");
    if (p.precision < 0 || p.precision > 2) return fail(NGI_ERR_INVALID_ARG, "bad precision");
    if (p.engine != NGI_ENGINE_MEGAKERNEL && p.engine != NGI_ENGINE_WAVEFRONT && p.engine != NGI_ENGINE_PERSISTENT)
        return fail(NGI_ERR_INVALID_ARG, "bad engine");
    if (p.renderer == NGI_RENDERER_BDPT && p.precision == NGI_PRECISION_FP64_ALL)
        return fail(NGI_ERR_UNSUPPORTED, "bdpt shades in FP32 (FP64 film with FP32_FP64ACC)");
    if (p.rr_depth >= 0 && !(p.rr_prob > 0.0 && p.rr_prob <= 1.0)) return fail(NGI_ERR_INVALID_ARG, "rr_prob must be in (0, 1]");
    if (p.max_num_vertices < -1) return fail(NGI_ERR_INVALID_ARG, "bad max_num_vertices");
    if (sc->sensor_prim < 0) return fail(NGI_ERR_INVALID_SCENE, "scene has no sensor (E) primitive");
    if (sc->prims[sc->sensor_prim].sensor_type != NGI_SENSOR_PINHOLE) return fail(NGI_ERR_UNSUPPORTED, "only pinhole sensors are supported");
    return NGI_OK;
}

int Session::init(ngi_scene* sc, const ngi_render_params& pp) {
    scene = sc;
    p = pp;
    int st = check_render_params(sc, p);
    if (st) return st;
    st = device_prepare(*sc);
    if (st) return st;
    film64 = p.precision != NGI_PRECISION_FP32_ALL;
    npix = (uint32_t)p.width * (uint32_t)p.height;
    const size_t fb = (size_t)npix * 3 * (film64 ? 8 : 4);
    NGI_HIP_CHECK(hipMalloc(&film, fb));
    NGI_HIP_CHECK(hipMalloc(&counters, CNT_N * sizeof(unsigned long long)));
    return reset();
}

int Session::reset() {
    paths_done = 0;
    samples_done = 0;
    batches = 0;
    memset(&stats, 0, sizeof(stats));
    int st = zero(film, (size_t)npix * 3, film64);
    if (st) return st;
    if (m1) {
        if ((st = zero(m1, (size_t)npix * 3, true))) return st;
        if ((st = zero(m2, (size_t)npix * 3, true))) return st;
    }
    if ((st = zero(counters, CNT_N, true))) return st;   // 64-bit words
    return sync();
}

void Session::destroy() {
    if (scene && scene->dev) (void)hipStreamSynchronize(stream());
    for (void* q : {film, counters, m1, m2, batch, next_path, lane_mem}) if (q) (void)hipFree(q);
    film = counters = m1 = m2 = batch = next_path = lane_mem = nullptr;
    lane_bytes = 0;
    release_engine();
    for (auto& e : pending) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    pending.clear();
}

int Session::timed_begin(int kind, int64_t items, TimedLaunch& t) {
    t.kind = kind;
    t.items = items;
    t.a = t.b = nullptr;
    if (!p.time_kernels) return NGI_OK;
    NGI_HIP_CHECK(hipEventCreate(&t.a));
    NGI_HIP_CHECK(hipEventCreate(&t.b));
    NGI_HIP_CHECK(hipEventRecord(t.a, stream()));
    return NGI_OK;
}

int Session::timed_end(TimedLaunch& t) {
    if (!t.a) return NGI_OK;
    NGI_HIP_CHECK(hipEventRecord(t.b, stream()));
    pending.push_back(t);
    return NGI_OK;
}

int Session::sync() {
    int st = ensure_device();
    if (st) return st;
    NGI_HIP_CHECK(hipStreamSynchronize(stream()));
    for (auto& t : pending) {
        float ms = 0;
        NGI_HIP_CHECK(hipEventElapsedTime(&ms, t.a, t.b));
        switch (t.kind) {
            case K_EXTEND: stats.ms_extend += ms, stats.extend_launches++, stats.extend_rays_timed += t.items; break;
            case K_SHADOW: stats.ms_shadow += ms, stats.shadow_launches++; break;
            case K_SHADE: stats.ms_shade += ms; break;
            case K_MEGA: stats.ms_shade += ms; break;
            default: break;
        }
        if (t.kind == K_EXTEND || t.kind == K_SHADOW) stats.ms_trace += ms, stats.trace_launches++;
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    pending.clear();
    return NGI_OK;
}

RenderArgs Session::args(void* target_film) const {
    RenderArgs a;
    memset(&a, 0, sizeof(a));
    a.sc = scene->dev->view;
    a.film = target_film;
    a.w = p.width, a.h = p.height;
    a.npix = npix;
    a.sampling = p.sampling;
    a.max_nv = p.max_num_vertices;
    a.rr_depth = p.rr_depth;
    a.rr_prob_f = (float)p.rr_prob;
    a.rr_prob_d = p.rr_prob;
    a.seed = p.seed;
    a.stream_id = p.stream_id;
    a.counters = reinterpret_cast<unsigned long long*>(counters);
    // Persistent engines: on small scenes an iteration is short and one
    // counter atomic per refill serialises the grid (ptdirect / ltdirect on the
    // 36-triangle Cornell box: +27% / +40% with chunks of 128); on large ones
    // exact refills keep each wavefront's paths spread over the frontier
    // (chunks cost 5-10% on the 77k-triangle nave and the sphere field).
    a.path_chunk = scene->dev->view.num_tris < NGI_CHUNKED_FEED_TRIS ? 128u : 0u;
    return a;
}

// Enqueue `spp` samples per pixel (or passes) into `target`.
int Session::render_into(void* target, int spp) {
    if (spp <= 0) return fail(NGI_ERR_INVALID_ARG, "spp must be positive");
    int st = device_prepare(*scene);   // refresh parameters if they changed
    if (st) return st;
    const auto t0 = std::chrono::steady_clock::now();
    uint64_t total;
    if (p.sampling == NGI_SAMPLING_PER_PIXEL) total = (uint64_t)npix * (uint64_t)spp;
    else total = (uint64_t)p.num_samples * (uint64_t)spp;
    if (p.engine == NGI_ENGINE_WAVEFRONT && !persistent_only(p.renderer)) {
        st = render_wavefront(target, total);
    } else if (p.engine == NGI_ENGINE_PERSISTENT || persistent_only(p.renderer)) {
        // The light tracers and bdpt always run on the persistent engine.
        if (!next_path) NGI_HIP_CHECK(hipMalloc(&next_path, sizeof(unsigned long long)));
        if ((st = zero(next_path, 1, true))) return st;
        RenderArgs a = args(target);
        a.sample_base = samples_done;
        a.path_begin = p.sampling == NGI_SAMPLING_PER_PIXEL ? 0 : paths_done;
        a.path_count = total;
        TimedLaunch tl;
        if ((st = timed_begin(K_MEGA, (int64_t)total, tl))) return st;
        Lanes ln;
        ln.want = std::max(0, p.paths_in_flight);
        st = launch_pk_engine(a, ln);
        if (!st) st = timed_end(tl);
    } else {
        // Launches of at most 2^22 paths (whole pixels rows of samples).
        const uint64_t chunk = p.sampling == NGI_SAMPLING_PER_PIXEL ? (uint64_t)npix * std::max<uint64_t>(1, (1u << 22) / npix) : (1u << 22);
        for (uint64_t done = 0; done < total && !st;) {
            const uint64_t n = std::min<uint64_t>(chunk, total - done);
            RenderArgs a = args(target);
            if (p.sampling == NGI_SAMPLING_PER_PIXEL) {
                a.path_begin = done;   // id -> (pixel = id % npix, sample = base + id / npix)
                a.sample_base = samples_done;
            } else {
                a.path_begin = paths_done + done;
            }
            a.path_count = n;
            TimedLaunch tl;
            if ((st = timed_begin(K_MEGA, (int64_t)n, tl))) break;
            launch_megakernel(a, p.precision, p.renderer == NGI_RENDERER_PTDIRECT, scene->dev->view.bvh_width, p.collect_counters != 0, stream());
            NGI_HIP_CHECK(hipGetLastError());
            if ((st = timed_end(tl))) break;
            done += n;
        }
    }
    if (st) return st;
    paths_done += total;
    samples_done += (uint32_t)spp;
    stats.ms_total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return NGI_OK;
}

int Session::launch_pk_engine(const RenderArgs& a, Lanes ln) {
    unsigned long long* next = reinterpret_cast<unsigned long long*>(next_path);
    const int width = scene->dev->view.bvh_width;
    const bool count = p.collect_counters != 0;
    LaneMem mem{&lane_mem, &lane_bytes, stream()};
    if (p.renderer == NGI_RENDERER_BDPT) return launch_bdpt(a, next, mem, film64, width, count, stream(), ln);
    if (is_light_tracer(p.renderer))
        return launch_light(a, next, mem, p.precision, p.renderer == NGI_RENDERER_LTDIRECT, width, count, stream(), ln);
    return launch_persistent(a, next, mem, p.precision, p.renderer == NGI_RENDERER_PTDIRECT, width, count, stream(), ln);
}

int Session::render(int spp) { return render_into(film, spp); }

int Session::render_batch(int spp) {
    const size_t n = (size_t)npix * 3;
    if (!m1) {
        NGI_HIP_CHECK(hipMalloc(&m1, n * 8));
        NGI_HIP_CHECK(hipMalloc(&m2, n * 8));
        NGI_HIP_CHECK(hipMalloc(&batch, n * (film64 ? 8 : 4)));
        int st = zero(m1, n, true);
        if (!st) st = zero(m2, n, true);
        if (st) return st;
    }
    int st = zero(batch, n, film64);
    if (st) return st;
    const uint64_t before = paths_done;
    if ((st = render_into(batch, spp))) return st;
    const double scale = (double)npix / (double)(paths_done - before);
    const unsigned grid = (unsigned)((n + 255) / 256);
    if (film64) fold_batch<double><<<grid, 256, 0, stream()>>>(reinterpret_cast<double*>(batch), scale, reinterpret_cast<double*>(m1), reinterpret_cast<double*>(m2), reinterpret_cast<double*>(film), n);
    else fold_batch<float><<<grid, 256, 0, stream()>>>(reinterpret_cast<float*>(batch), scale, reinterpret_cast<double*>(m1), reinterpret_cast<double*>(m2), reinterpret_cast<float*>(film), n);
    NGI_HIP_CHECK(hipGetLastError());
    batches++;
    return NGI_OK;
}

int Session::get_film(double* out, double* var) {
    int st = sync();
    if (st) return st;
    const size_t n = (size_t)npix * 3;
    const double scale = paths_done ? (double)npix / (double)paths_done : 0.0;
    if (out) {
        if (film64) {
            NGI_HIP_CHECK(hipMemcpy(out, film, n * 8, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i) out[i] *= scale;
        } else {
            std::vector<float> tmp(n);
            NGI_HIP_CHECK(hipMemcpy(tmp.data(), film, n * 4, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i) out[i] = (double)tmp[i] * scale;
        }
    }
    if (var) {
        if (!m1 || batches < 2) return fail(NGI_ERR_INVALID_ARG, "variance needs at least two batches (ngi_session_render_batch)");
        std::vector<double> a(n), b(n);
        NGI_HIP_CHECK(hipMemcpy(a.data(), m1, n * 8, hipMemcpyDeviceToHost));
        NGI_HIP_CHECK(hipMemcpy(b.data(), m2, n * 8, hipMemcpyDeviceToHost));
        const double k = (double)batches;
        for (size_t i = 0; i < n; ++i) {
            const double mean = a[i] / k;
            const double sv = (b[i] - k * mean * mean) / (k - 1.0);   // sample variance of batch means
            var[i] = (sv > 0 ? sv : 0.0) / k;
        }
    }
    return NGI_OK;
}

int Session::get_stats(ngi_stats* out) {
    int st = sync();
    if (st) return st;
    unsigned long long c[CNT_N];
    NGI_HIP_CHECK(hipMemcpy(c, counters, sizeof(c), hipMemcpyDeviceToHost));
    *out = stats;
    out->samples = (int64_t)c[CNT_SAMPLES];
    out->extension_rays = (int64_t)c[CNT_EXT_RAYS];
    out->shadow_rays = (int64_t)c[CNT_SHADOW_RAYS];
    out->node_visits = (double)c[CNT_EXT_NODES];
    out->tri_tests = (double)c[CNT_EXT_TRIS];
    out->shadow_node_visits = (double)c[CNT_SH_NODES];
    out->shadow_tri_tests = (double)c[CNT_SH_TRIS];
    return NGI_OK;
}

}  // namespace ngi

using namespace ngi;

struct ngi_session : ngi::Session {};

extern "C" {

int ngi_session_create(ngi_scene* scene, const ngi_render_params* p, ngi_session** out) {
    clear_error();
    if (!scene || !p || !out) return fail(NGI_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    ngi_session* s = new ngi_session();
    int st = s->init(scene, *p);
    if (st) {
        s->destroy();
        delete s;
        return st;
    }
    *out = s;
    return NGI_OK;
}

int ngi_session_render(ngi_session* s, int spp) {
    clear_error();
    if (!s) return fail(NGI_ERR_INVALID_ARG, "null session");
    return s->render(spp);
}

int ngi_session_render_batch(ngi_session* s, int spp) {
    clear_error();
    if (!s) return fail(NGI_ERR_INVALID_ARG, "null session");
    return s->render_batch(spp);
}

int ngi_session_sync(ngi_session* s) {
    clear_error();
    if (!s) return fail(NGI_ERR_INVALID_ARG, "null session");
    return s->sync();
}

int ngi_session_get_film(ngi_session* s, double* film, double* var) {
    clear_error();
    if (!s) return fail(NGI_ERR_INVALID_ARG, "null session");
    return s->get_film(film, var);
}

int ngi_session_get_stats(ngi_session* s, ngi_stats* out) {
    clear_error();
    if (!s || !out) return fail(NGI_ERR_INVALID_ARG, "null argument");
    return s->get_stats(out);
}

int ngi_session_reset(ngi_session* s) {
    clear_error();
    if (!s) return fail(NGI_ERR_INVALID_ARG, "null session");
    return s->reset();
}

void ngi_session_destroy(ngi_session* s) {
    if (!s) return;
    s->destroy();
    delete s;
}

int ngi_autotune(ngi_scene* scene, ngi_render_params* p, int64_t tune_paths, ngi_tune_trial* trials, int max_trials, int* num_trials) {
    clear_error();
    if (!scene || !p) return fail(NGI_ERR_INVALID_ARG, "null argument");
    if (tune_paths < 0 || max_trials < 0 || (max_trials > 0 && !trials)) return fail(NGI_ERR_INVALID_ARG, "bad tuning arguments");
    if (num_trials) *num_trials = 0;
    if (tune_paths == 0) tune_paths = (int64_t)1 << 24;
    ngi_render_params q = *p;
    q.collect_counters = 0;
    q.time_kernels = 0;
    q.engine = NGI_ENGINE_PERSISTENT;
    q.paths_in_flight = 0;
    int passes = 1;
    if (q.sampling == NGI_SAMPLING_PER_PIXEL) {
        const int64_t npix = std::max<int64_t>(1, (int64_t)q.width * q.height);
        passes = (int)std::min<int64_t>(std::max<int64_t>(1, tune_paths / npix), 1 << 16);
    } else {
        q.num_samples = tune_paths;
    }
    Session s;
    int st = s.init(scene, q);
    int64_t full = 0;
    if (!st) {
        RenderArgs a = s.args(s.film);
        a.path_count = 1;
        Lanes query;
        query.want = -1;
        query.full = &full;
        st = s.launch_pk_engine(a, query);
    }
    struct Cand {
        int engine;
        int64_t lanes;
        double ms;
    };
    std::vector<Cand> cands;
    for (int64_t l : {(int64_t)0, full * 3 / 4, full / 2}) cands.push_back({NGI_ENGINE_PERSISTENT, l, 0.0});
    if (!persistent_only(q.renderer)) {
        if (q.precision != NGI_PRECISION_FP64_ALL)
            for (int64_t c : {(int64_t)1 << 22, (int64_t)1 << 23, (int64_t)1 << 24, (int64_t)1 << 25})
                cands.push_back({NGI_ENGINE_WAVEFRONT, c, 0.0});
        cands.push_back({NGI_ENGINE_MEGAKERNEL, 0, 0.0});
    }
    // Each candidate: one untimed warm-up render, then the best of two.
    for (size_t i = 0; i < cands.size() && !st; ++i) {
        s.release_engine();
        s.p.engine = cands[i].engine;
        s.p.paths_in_flight = (int)cands[i].lanes;
        double best = 0.0;
        for (int rep = 0; rep < 3 && !st; ++rep) {
            const auto t0 = std::chrono::steady_clock::now();
            st = s.render(passes);
            if (!st) st = s.sync();
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (rep == 1 || (rep > 1 && ms < best)) best = ms;
        }
        cands[i].ms = best;
    }
    const std::string err = st ? std::string(ngi_last_error()) : std::string();
    s.destroy();
    if (st) return fail(st, "autotune: %s", err.c_str());
    size_t bi = 0;
    for (size_t i = 1; i < cands.size(); ++i)
        if (cands[i].ms < cands[bi].ms) bi = i;
    p->engine = cands[bi].engine;
    p->paths_in_flight = (int)cands[bi].lanes;
    const double paths = (double)passes * (double)(q.sampling == NGI_SAMPLING_PER_PIXEL ? (int64_t)q.width * q.height : q.num_samples);
    const int n = (int)std::min<size_t>(cands.size(), (size_t)max_trials);
    for (int i = 0; i < n; ++i) {
        trials[i].engine = cands[i].engine;
        trials[i].paths_in_flight = (int)cands[i].lanes;
        trials[i].ms = cands[i].ms;
        trials[i].samples_per_s = cands[i].ms > 0 ? paths / (cands[i].ms * 1e-3) : 0.0;
    }
    if (num_trials) *num_trials = (int)cands.size();
    return NGI_OK;
}

int ngi_render(ngi_scene* scene, const ngi_render_params* p, double* film, ngi_stats* stats) {
    clear_error();
    if (!scene || !p || !film) return fail(NGI_ERR_INVALID_ARG, "null argument");
    ngi_session* s = nullptr;
    int st = ngi_session_create(scene, p, &s);
    if (st) return st;
    const int passes = p->sampling == NGI_SAMPLING_PER_PIXEL ? p->spp : 1;
    st = s->render(passes);
    if (!st) st = s->get_film(film, nullptr);
    if (!st && stats) st = s->get_stats(stats);
    const std::string err = st ? std::string(ngi_last_error()) : std::string();
    ngi_session_destroy(s);
    if (st) fail(st, "%s", err.c_str());
    return st;
}

}  // extern "C"
