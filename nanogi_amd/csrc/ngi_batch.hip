// Batched rendering of many independent scenes in one persistent launch
// (no reference counterpart: nanogi renders one scene per process).
//
// Every scene's paths are cut into chunks numbered scene after scene.  Waves
// of the persistent kernel take chunks from one counter; their lanes take
// the chunk's paths as soon as their own path ends and run the same vertex
// loop as the single-scene engines (path_vertex) on that scene's BVH, film
// and parameters.  Small and large scenes therefore share the GPU instead
// of leaving it idle between per-scene launches.  The wave that counts a
// scene's last path records a device wall-clock time stamp, which gives
// per-scene completion times.
//
// A batch owns its device memory (all films in one buffer, per-scene
// arguments, counters) and replays one captured hipGraph per frame: a fill
// kernel that clears films and counters, then the persistent kernel.  The
// block size and the number of resident blocks per CU are picked by timing
// candidates on a proportional sample of every scene's paths; the chunk size
// follows from the paths per resident wave.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "ngi_render.h"

namespace ngi {

// Minimum waves per SIMD for the batch kernel's register allocation: 4
// (at most 128 VGPRs) measured fastest on 256 small scenes (1302 vs 1217 M
// samples/s unconstrained, 1242 at 5, 1157 at 6).
#ifndef NGI_BATCH_WAVES
#define NGI_BATCH_WAVES 4
#endif

struct BatchDev {
    const RenderArgs* __restrict__ args;           // per scene
    const unsigned long long* __restrict__ offset; // per scene + 1: prefix sums of the path counts
    const unsigned long long* __restrict__ chunk;  // per scene + 1: prefix sums of the chunk counts
    unsigned long long* done;                      // per scene: completed paths
    unsigned long long* t_done;                    // per scene: wall clock at completion
    unsigned long long* misc;                      // MISC_* words
    unsigned long long num_chunks;
    uint32_t chunk_paths;                          // paths per chunk (the last chunk of a scene may be shorter)
    int num_scenes;
    uint32_t stack;                                // LDS traversal-stack entries per lane (the deepest scene's)
    ShadowQueue q;                                 // ptdirect: per-wave shadow-ray rings
};

// misc words: wall clock at start (atomic min), chunk counter, ray counters.
enum { MISC_T_START = 0, MISC_CHUNK = 1, MISC_EXT = 2, MISC_SHADOW = 3, MISC_N = 4 };

__device__ __forceinline__ void wave_sum_add(unsigned long long* p, uint32_t v) {
    unsigned long long x = v;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63) == 0 && x) atomicAdd(p, x);
}

// Last entry of the prefix array p[0..n] that is <= x (entries repeat for
// scenes without paths; the last repeat is the scene that owns x).
__device__ __forceinline__ int owner(const unsigned long long* __restrict__ p, int n, unsigned long long x) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (p[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Scheduling is per wavefront: a wave takes a chunk of consecutive paths of
// one scene (one atomic on the chunk counter), its lanes take the chunk's
// paths as their paths end, and the next chunk is taken when this one is
// handed out.  Chunks are numbered scene after scene, so a wave stays on a
// scene until the scene's chunks are gone; a chunk of another scene waits
// until the wave's lanes have finished, so every lane of a wave always
// works on the same scene and the scene's arguments are wave-uniform.
// Completed paths are counted per wave and added to the scene's counter
// when the wave leaves the scene; the addition that completes the count
// stamps the scene's completion time.  With ptdirect the NEE shadow rays go
// through the wave's shadow queue (traced 64 at a time, like the
// single-scene persistent engine); the queue only ever holds rays of the
// wave's current scene, since it is drained before the wave counts its
// finished paths and leaves the scene.
template <bool DIRECT, int WIDTH, int BLOCK>
__global__ void __launch_bounds__(BLOCK, NGI_BATCH_WAVES) batch_kernel(BatchDev b) {
    extern __shared__ uint32_t stk[];   // max stack_entries over the scenes x BLOCK (wide BVH only)
    if (threadIdx.x == 0) atomicMin(b.misc + MISC_T_START, (unsigned long long)wall_clock64());
    const uint32_t lane = __lane_id();
    const uint32_t wave = blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
    uint32_t qhead = 0, qcount = 0;   // wave-uniform shadow queue state (< NGI_SHQ)
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t cnt[CNT_N] = {0, 0, 0, 0, 0, 0, 0};
    PathState<float> ps;
    bool active = false;
    // Wave-uniform state.
    int si = -1;                              // scene of the wave
    unsigned long long next = 0, end = 0;     // unhanded paths of the current chunk (scene-local ids)
    int pend_si = -1;                         // a chunk of another scene, taken while lanes were busy
    unsigned long long pend_next = 0, pend_end = 0;
    unsigned long long ended = 0;             // paths of scene si this wave finished, not yet counted
    bool chunks_left = true;
    for (;;) {
        if (next >= end && chunks_left && pend_si < 0) {
            unsigned long long k = 0;
            if (lane == 0) k = atomicAdd(b.misc + MISC_CHUNK, 1ull);
            k = __shfl(k, 0);
            if (k >= b.num_chunks) {
                chunks_left = false;
            } else {
                const int cs = owner(b.chunk, b.num_scenes, k);
                const unsigned long long n = b.offset[cs + 1] - b.offset[cs];
                const unsigned long long cb = (k - b.chunk[cs]) * (unsigned long long)b.chunk_paths;
                const unsigned long long ce = cb + b.chunk_paths < n ? cb + b.chunk_paths : n;
                if (si < 0 || cs == si) {
                    si = cs;
                    next = cb, end = ce;
                } else {
                    pend_si = cs;
                    pend_next = cb, pend_end = ce;
                }
            }
        }
        const int su = __builtin_amdgcn_readfirstlane(si < 0 ? 0 : si);
        const RenderArgs& a = b.args[su];
        if (next < end) {
            const unsigned long long need = __ballot(!active);
            if (need) {
                const unsigned long long avail = end - next;
                const unsigned long long rank = (unsigned long long)__popcll(need & below);
                if (!active && rank < avail) {
                    path_start<float>(a, next + rank, load_cam<float>(a.sc.cam), ps);
                    active = true;
                }
                const unsigned long long took = (unsigned long long)__popcll(need);
                next += took < avail ? took : avail;
            }
        }
        const bool any = __ballot(active) != 0;
        // All queued rays are traced before the wave leaves the scene.
        if (DIRECT) shq_flush<double, WIDTH, false>(a, b.q, wave, qhead, qcount, !any, stk + threadIdx.x, BLOCK, cnt);
        if (!any) {
            // The wave has no path in flight: count what it finished on this
            // scene, then move to the waiting chunk, the next chunk, or stop.
            if (si >= 0 && ended) {
                if (lane == 0) {
                    const unsigned long long n = b.offset[su + 1] - b.offset[su];
                    if (atomicAdd(b.done + su, ended) + ended == n) b.t_done[su] = (unsigned long long)wall_clock64();
                }
                ended = 0;
            }
            if (pend_si >= 0) {
                si = pend_si;
                next = pend_next, end = pend_end;
                pend_si = -1;
                continue;
            }
            if (next < end || chunks_left) continue;
            break;
        }
        const bool was = active;
        bool pushed = false;
        if (active)
            active = path_vertex<float, double, DIRECT, WIDTH, false, DIRECT>(a, load_cam<float>(a.sc.cam), prims_of<float>(a.sc), ps,
                                                                              stk + threadIdx.x, BLOCK, clamp_vertices(a.max_nv), cnt,
                                                                              &b.q, wave, qhead + qcount, &pushed);
        if (DIRECT) qcount += (uint32_t)__popcll(__ballot(pushed));   // at most 63 + 64 < NGI_SHQ
        ended += (unsigned long long)__popcll(__ballot(was && !active));
    }
    wave_sum_add(b.misc + MISC_EXT, cnt[CNT_EXT_RAYS]);
    wave_sum_add(b.misc + MISC_SHADOW, cnt[CNT_SHADOW_RAYS]);
}

// The batch kernel with resumable extension traversals (wide BVHs), as the
// single-scene resume_kernel: a wavefront stops traversing once at most
// NGI_BATCH_KEEP of its lanes are still inside the tree, completes the
// others' vertices and starts their next rays; the waiting lanes keep their
// traversal state, direction and roulette number in LDS.  A wave leaves a
// scene only with no path in flight, so no traversal outlives its scene.
// Measured on the 256 rooms (M samples/s): keep 4: 1,791, 8: 1,838,
// 16: 1,866, 24: 1,858, 32: 1,833.
#ifndef NGI_BATCH_KEEP
#define NGI_BATCH_KEEP 16
#endif
template <bool DIRECT, int BLOCK>
__global__ void __launch_bounds__(BLOCK, NGI_BATCH_WAVES) batch_resume_kernel(BatchDev b) {
    // LDS: extension stacks, shadow stacks (b.stack words per lane each),
    // then the parked lane state.
    extern __shared__ uint32_t stk[];
    uint32_t* ext_stk = stk + threadIdx.x;
    uint32_t* sh_stk = stk + (size_t)b.stack * BLOCK + threadIdx.x;
    uint32_t* pk = stk + (size_t)2 * b.stack * BLOCK + threadIdx.x;
    if (threadIdx.x == 0) atomicMin(b.misc + MISC_T_START, (unsigned long long)wall_clock64());
    const uint32_t lane = __lane_id();
    const uint32_t wave = blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
    uint32_t qhead = 0, qcount = 0;   // wave-uniform shadow queue state (< NGI_SHQ)
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t cnt[CNT_N] = {0, 0, 0, 0, 0, 0, 0};
    PathState<float> ps;
    bool active = false;   // the lane holds a path
    bool busy = false;     // ... whose extension ray is being traced
    // Wave-uniform state.
    int si = -1;                              // scene of the wave
    unsigned long long next = 0, end = 0;     // unhanded paths of the current chunk (scene-local ids)
    int pend_si = -1;                         // a chunk of another scene, taken while lanes were busy
    unsigned long long pend_next = 0, pend_end = 0;
    unsigned long long ended = 0;             // paths of scene si this wave finished, not yet counted
    bool chunks_left = true;
    for (;;) {
        if (next >= end && chunks_left && pend_si < 0) {
            unsigned long long k = 0;
            if (lane == 0) k = atomicAdd(b.misc + MISC_CHUNK, 1ull);
            k = __shfl(k, 0);
            if (k >= b.num_chunks) {
                chunks_left = false;
            } else {
                const int cs = owner(b.chunk, b.num_scenes, k);
                const unsigned long long n = b.offset[cs + 1] - b.offset[cs];
                const unsigned long long cb = (k - b.chunk[cs]) * (unsigned long long)b.chunk_paths;
                const unsigned long long ce = cb + b.chunk_paths < n ? cb + b.chunk_paths : n;
                if (si < 0 || cs == si) {
                    si = cs;
                    next = cb, end = ce;
                } else {
                    pend_si = cs;
                    pend_next = cb, pend_end = ce;
                }
            }
        }
        const int su = __builtin_amdgcn_readfirstlane(si < 0 ? 0 : si);
        const RenderArgs& a = b.args[su];
        const DevPrim<float>* prims = prims_of<float>(a.sc);
        // Vertices whose extension rays have finished.
        bool ended_now = false;
        if (active && !busy) {
            active = vertex_end<float, double, DIRECT>(a, prims, ps, parked_dir<BLOCK>(pk), parked_urr<BLOCK>(pk), parked_hit<BLOCK>(pk));
            ended_now = !active;
        }
        if (next < end) {
            const unsigned long long need = __ballot(!active);
            if (need) {
                const unsigned long long avail = end - next;
                const unsigned long long rank = (unsigned long long)__popcll(need & below);
                if (!active && rank < avail) {
                    path_start<float>(a, next + rank, load_cam<float>(a.sc.cam), ps);
                    active = true;
                }
                const unsigned long long took = (unsigned long long)__popcll(need);
                next += took < avail ? took : avail;
            }
        }
        // Next vertex of every path not inside the tree.
        bool pushed = false;
        if (active && !busy) {
            V3<float> wo;
            float urr;
            if (vertex_begin<float, double, DIRECT, 8, false, DIRECT>(a, load_cam<float>(a.sc.cam), prims, ps, sh_stk, BLOCK,
                                                                    clamp_vertices(a.max_nv), cnt, &b.q, wave, qhead + qcount, &pushed, wo,
                                                                    urr)) {
                cnt[CNT_EXT_RAYS]++;
                Trav8 tv;
                trav8_begin(FLT_MAX, tv);
                busy = a.sc.num_tris != 0;
                park_trav<BLOCK>(pk, tv);
                park_dir<BLOCK>(pk, wo, urr);
            } else {
                active = false;
                ended_now = true;
            }
        }
        ended += (unsigned long long)__popcll(__ballot(ended_now));
        const bool any = __ballot(active) != 0;
        // All queued rays are traced before the wave leaves the scene.
        if (DIRECT) {
            qcount += (uint32_t)__popcll(__ballot(pushed));   // at most 63 + 64 < NGI_SHQ
            shq_flush<double, 8, false>(a, b.q, wave, qhead, qcount, !any, sh_stk, BLOCK, cnt);
        }
        if (!any) {
            // The wave has no path in flight: count what it finished on this
            // scene, then move to the waiting chunk, the next chunk, or stop.
            if (si >= 0 && ended) {
                if (lane == 0) {
                    const unsigned long long n = b.offset[su + 1] - b.offset[su];
                    if (atomicAdd(b.done + su, ended) + ended == n) b.t_done[su] = (unsigned long long)wall_clock64();
                }
                ended = 0;
            }
            if (pend_si >= 0) {
                si = pend_si;
                next = pend_next, end = pend_end;
                pend_si = -1;
                continue;
            }
            if (next < end || chunks_left) continue;
            break;
        }
        if (__ballot(busy)) {
            // Lanes without a ray load and store back what they parked last.
            Trav8 tv = parked_trav<BLOCK>(pk);
            const RayPre r = make_pre(extension_ray(ps, parked_dir<BLOCK>(pk)));
            // The last paths of the wave's chunk run their traversals to the end.
            const int keep = __builtin_amdgcn_readfirstlane(next < end ? NGI_BATCH_KEEP : 0);
            TraceCounters tc{0, 0};
            trav8_run<false>(a.sc, r, NGI_EPS_F, tv, busy, keep, tc, ext_stk, BLOCK);
            park_trav<BLOCK>(pk, tv);
        }
    }
    wave_sum_add(b.misc + MISC_EXT, cnt[CNT_EXT_RAYS]);
    wave_sum_add(b.misc + MISC_SHADOW, cnt[CNT_SHADOW_RAYS]);
}

// Clears the films and per-scene counters and resets the chunk counter: the
// first node of the captured graph (a kernel, not a memset node).
__global__ void batch_clear(double* film, size_t film_n, unsigned long long* done, unsigned long long* t_done, int n, unsigned long long* misc) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t i = t; i < film_n; i += stride) film[i] = 0.0;
    for (size_t i = t; i < (size_t)n; i += stride) done[i] = 0ull, t_done[i] = 0ull;
    if (t < MISC_N) misc[t] = t == MISC_T_START ? ~0ull : 0ull;
}

using BatchKernel = void (*)(BatchDev);

template <bool D, int W>
static BatchKernel pick_block(int block, bool resume) {
    switch (block) {
        case 64: return resume && W == 8 ? batch_resume_kernel<D, 64> : batch_kernel<D, W, 64>;
        case 128: return resume && W == 8 ? batch_resume_kernel<D, 128> : batch_kernel<D, W, 128>;
        case 256: return resume && W == 8 ? batch_resume_kernel<D, 256> : batch_kernel<D, W, 256>;
        default: return nullptr;
    }
}

static BatchKernel pick_kernel(bool direct, int width, int block, bool resume) {
    if (direct) return width == 8 ? pick_block<true, 8>(block, resume) : pick_block<true, 2>(block, resume);
    return width == 8 ? pick_block<false, 8>(block, resume) : pick_block<false, 2>(block, resume);
}

struct Batch {
    std::vector<ngi_scene*> scenes;
    std::vector<ngi_render_params> params;
    std::vector<RenderArgs> args;
    std::vector<unsigned long long> offset;     // n + 1
    std::vector<size_t> film_off;               // n + 1, in doubles
    std::vector<const void*> nodes_seen;        // device BVH of each scene at the last upload
    std::vector<uint64_t> version_seen;
    ngi_batch_opts opts;
    bool direct = false;
    int width = 0;
    uint32_t stack = 0;
    int num_cus = 0, rate_khz = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double* film = nullptr;
    RenderArgs* d_args = nullptr;
    unsigned long long *d_off = nullptr, *d_done = nullptr, *d_tdone = nullptr, *d_misc = nullptr, *d_tune_off = nullptr;
    unsigned long long *d_chunk = nullptr, *d_tune_chunk = nullptr;   // chunk prefix sums (n + 1)
    unsigned long long num_chunks = 0;
    float4* shq = nullptr;          // ptdirect shadow queues, sized for the most resident waves of any block size
    size_t shq_waves = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t frame = 0;
    bool rendered = false;
    double tune_ms = 0.0;

    int n() const { return (int)scenes.size(); }
    // Chunk prefix sums of the path counts in `off` for chunks of c paths.
    std::vector<unsigned long long> chunk_table(const std::vector<unsigned long long>& off, uint32_t c) const {
        std::vector<unsigned long long> ch(off.size(), 0);
        for (size_t i = 0; i + 1 < off.size(); ++i) ch[i + 1] = ch[i] + (off[i + 1] - off[i] + c - 1) / c;
        return ch;
    }
    BatchDev dev(const unsigned long long* off, const unsigned long long* chunk, unsigned long long nchunks, uint32_t c) const {
        BatchDev b;
        b.args = d_args;
        b.offset = off;
        b.chunk = chunk;
        b.done = d_done;
        b.t_done = d_tdone;
        b.misc = d_misc;
        b.num_chunks = nchunks;
        b.chunk_paths = c;
        b.num_scenes = n();
        b.stack = stack;
        b.q.rec = shq;
        b.q.waves = 0;
        return b;
    }
    // Resumable traversals (wide BVHs) while their LDS (two stacks and the
    // parked state per lane) leaves NGI_BATCH_WAVES waves per SIMD.
    bool resume() const { return width == 8 && ((size_t)2 * stack + PK_WORDS) * 4 * 256 * NGI_BATCH_WAVES <= NGI_CU_LDS_BYTES; }
    size_t lds(int block) const {
        if (width != 8) return 0;
        return (resume() ? (size_t)2 * stack + PK_WORDS : (size_t)stack) * block * 4;
    }
    int max_per_cu(int block) const {
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, pick_kernel(direct, width, block, resume()), block, lds(block)) != hipSuccess)
            return 0;
        return occ;
    }
    int enqueue_clear() const {
        const size_t fn = film_off.back();
        const int blocks = (int)std::min<size_t>(4096, std::max<size_t>(1, (fn + 255) / 256));
        batch_clear<<<blocks, 256, 0, stream>>>(film, fn, d_done, d_tdone, n(), d_misc);
        NGI_HIP_CHECK(hipGetLastError());
        return NGI_OK;
    }
    int enqueue_render(const unsigned long long* off, const unsigned long long* chunk, unsigned long long nchunks, uint32_t c, int block,
                       int per_cu) const {
        if (!nchunks) return NGI_OK;
        BatchKernel k = pick_kernel(direct, width, block, resume());
        if (!k) return fail(NGI_ERR_INVALID_ARG, "batch block size %d is not one of 64, 128, 256", block);
        const unsigned grid = (unsigned)std::max(1, num_cus * per_cu);
        BatchDev d = dev(off, chunk, nchunks, c);
        d.q.waves = grid * (unsigned)(block / 64);
        if (direct && d.q.waves > shq_waves)
            return fail(NGI_ERR_INVALID_ARG, "batch launch of %u waves exceeds the %zu shadow queues", d.q.waves, shq_waves);
        k<<<grid, block, lds(block), stream>>>(d);
        NGI_HIP_CHECK(hipGetLastError());
        return NGI_OK;
    }
    int enqueue_frame() const {
        return enqueue_render(d_off, d_chunk, num_chunks, (uint32_t)opts.chunk_paths, opts.block_size, opts.blocks_per_cu);
    }
    int refresh();
    int autotune();
    int capture();
    int create(ngi_scene* const* sc, const ngi_render_params* p, int count, const ngi_batch_opts& o);
    int render(uint32_t frame);
    int sync();
    int get_film(int i, double* out);
    int get_times(double* scene_ms, ngi_stats* st);
    void destroy();
};

static int validate(ngi_scene* const* scenes, const ngi_render_params* params, int n) {
    if (n <= 0) return fail(NGI_ERR_INVALID_ARG, "no scenes");
    const bool direct = params[0].renderer == NGI_RENDERER_PTDIRECT;
    for (int i = 0; i < n; ++i) {
        const ngi_render_params& p = params[i];
        const ngi_scene* sc = scenes[i];
        if (!sc) return fail(NGI_ERR_INVALID_ARG, "scene %d is null", i);
        if (p.renderer != NGI_RENDERER_PT && p.renderer != NGI_RENDERER_PTDIRECT)
            return fail(NGI_ERR_UNSUPPORTED, "scene %d: batches render pt or ptdirect only", i);
        if ((p.renderer == NGI_RENDERER_PTDIRECT) != direct)
            return fail(NGI_ERR_UNSUPPORTED, "scene %d: every scene of a batch uses the same renderer", i);
        if (p.precision != NGI_PRECISION_FP32_FP64ACC)
            return fail(NGI_ERR_UNSUPPORTED, "scene %d: batches render with FP32 shading and an FP64 film", i);
        if (p.sampling != NGI_SAMPLING_PER_PIXEL && p.sampling != NGI_SAMPLING_REFERENCE_UNIFORM)
            return fail(NGI_ERR_INVALID_ARG, "scene %d: unknown sampling mode %d", i, p.sampling);
        if (p.width <= 0 || p.height <= 0 || (int64_t)p.width * p.height > (1ll << 28))
            return fail(NGI_ERR_INVALID_ARG, "scene %d: bad image size", i);
        if (p.sampling == NGI_SAMPLING_PER_PIXEL ? p.spp <= 0 : p.num_samples <= 0)
            return fail(NGI_ERR_INVALID_ARG, "scene %d: sample count must be positive", i);
        if (p.max_num_vertices < -1) return fail(NGI_ERR_INVALID_ARG, "scene %d: bad max_num_vertices", i);
        if (p.rr_depth >= 0 && !(p.rr_prob > 0.0 && p.rr_prob <= 1.0))
            return fail(NGI_ERR_INVALID_ARG, "scene %d: Russian roulette probability must be in (0, 1]", i);
        if (sc->sensor_prim < 0) return fail(NGI_ERR_INVALID_SCENE, "scene %d has no sensor (E) primitive", i);
        if (sc->prims[sc->sensor_prim].sensor_type != NGI_SENSOR_PINHOLE)
            return fail(NGI_ERR_UNSUPPORTED, "scene %d: only pinhole sensors are supported", i);
    }
    return NGI_OK;
}

// (Re)uploads every scene and the per-scene arguments.  Scenes whose BVH or
// parameters changed since the last upload are re-read; the traversal stack
// of the captured launch must still cover every scene's BVH depth.
int Batch::refresh() {
    bool changed = false;
    uint32_t need_stack = 0;
    int w = 0;
    for (int i = 0; i < n(); ++i) {
        ngi_scene* sc = scenes[i];
        int st = device_prepare(*sc);
        if (st) return st;
        const DevScene& v = sc->dev->view;
        if (w && v.bvh_width != w) return fail(NGI_ERR_INVALID_ARG, "all scenes of a batch need the same BVH width");
        w = v.bvh_width;
        need_stack = std::max(need_stack, v.stack_entries);
        if (nodes_seen[i] != (const void*)v.nodes || version_seen[i] != sc->params_version) {
            NGI_HIP_CHECK(hipStreamSynchronize(sc->dev->stream));   // uploads run on the scene's stream
            nodes_seen[i] = v.nodes;
            version_seen[i] = sc->params_version;
            changed = true;
        }
    }
    if (d_args && (w != width || need_stack > stack))
        return fail(NGI_ERR_INVALID_ARG, "a scene was rebuilt with another BVH width or a deeper BVH than the batch was created for");
    width = w;
    stack = std::max(stack, need_stack);
    for (int i = 0; i < n(); ++i) {
        RenderArgs& a = args[i];
        a.sc = scenes[i]->dev->view;
        const uint32_t seed = params[i].seed + frame;
        if (a.seed != seed) changed = true;
        a.seed = seed;
    }
    if (changed && d_args) NGI_HIP_CHECK(hipMemcpyAsync(d_args, args.data(), args.size() * sizeof(RenderArgs), hipMemcpyHostToDevice, stream));
    return NGI_OK;
}

template <typename T>
static int dev_alloc(T** ptr, size_t elems) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(elems, 1) * sizeof(T)) != hipSuccess) {
        *ptr = nullptr;
        return fail(NGI_ERR_OOM, "batch: hipMalloc of %zu bytes failed", elems * sizeof(T));
    }
    *ptr = reinterpret_cast<T*>(p);
    return NGI_OK;
}

int Batch::create(ngi_scene* const* sc, const ngi_render_params* p, int count, const ngi_batch_opts& o) {
    int st = validate(sc, p, count);
    if (st) return st;
    opts = o;
    if (opts.block_size && opts.block_size != 64 && opts.block_size != 128 && opts.block_size != 256)
        return fail(NGI_ERR_INVALID_ARG, "batch block size %d is not one of 64, 128, 256", opts.block_size);
    if (opts.blocks_per_cu < 0) return fail(NGI_ERR_INVALID_ARG, "blocks_per_cu must be >= 0");
    if (opts.chunk_paths < 0 || opts.chunk_paths > (1 << 24)) return fail(NGI_ERR_INVALID_ARG, "chunk_paths must be in [0, 2^24]");
    scenes.assign(sc, sc + count);
    params.assign(p, p + count);
    direct = p[0].renderer == NGI_RENDERER_PTDIRECT;
    nodes_seen.assign(count, nullptr);
    version_seen.assign(count, ~0ull);
    args.assign(count, RenderArgs());
    offset.assign(count + 1, 0);
    film_off.assign(count + 1, 0);
    for (int i = 0; i < count; ++i) {
        const ngi_render_params& q = params[i];
        RenderArgs& a = args[i];
        memset(&a, 0, sizeof(a));
        a.w = q.width, a.h = q.height;
        a.npix = (uint32_t)q.width * (uint32_t)q.height;
        a.sampling = q.sampling;
        a.path_count = q.sampling == NGI_SAMPLING_PER_PIXEL ? (uint64_t)a.npix * (uint64_t)q.spp : (uint64_t)q.num_samples;
        a.max_nv = q.max_num_vertices;
        a.rr_depth = q.rr_depth;
        a.rr_prob_f = (float)q.rr_prob;
        a.rr_prob_d = q.rr_prob;
        a.seed = q.seed;
        a.stream_id = q.stream_id;
        offset[i + 1] = offset[i] + a.path_count;
        film_off[i + 1] = film_off[i] + 3 * (size_t)a.npix;
    }
    if ((st = refresh())) return st;
    int dev = 0;
    NGI_HIP_CHECK(hipGetDevice(&dev));
    NGI_HIP_CHECK(hipDeviceGetAttribute(&num_cus, hipDeviceAttributeMultiprocessorCount, dev));
    NGI_HIP_CHECK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev));
    NGI_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    // set_params / build on a scene of the batch wait for this stream before
    // they overwrite or free the scene's device arrays.
    for (ngi_scene* sc : scenes)
        if (std::find(sc->user_streams.begin(), sc->user_streams.end(), (void*)stream) == sc->user_streams.end())
            sc->user_streams.push_back((void*)stream);
    NGI_HIP_CHECK(hipEventCreate(&ev0));
    NGI_HIP_CHECK(hipEventCreate(&ev1));
    if ((st = dev_alloc(&film, film_off.back())) || (st = dev_alloc(&d_args, count)) || (st = dev_alloc(&d_off, count + 1)) ||
        (st = dev_alloc(&d_tune_off, count + 1)) || (st = dev_alloc(&d_chunk, count + 1)) || (st = dev_alloc(&d_tune_chunk, count + 1)) ||
        (st = dev_alloc(&d_done, count)) || (st = dev_alloc(&d_tdone, count)) ||
        (st = dev_alloc(&d_misc, MISC_N)))
        return st;
    if (direct) {
        // One shadow ring per wave the device can hold at any block size.
        for (int block : {64, 128, 256}) shq_waves = std::max(shq_waves, (size_t)num_cus * max_per_cu(block) * (size_t)(block / 64));
        if ((st = dev_alloc(&shq, 3 * shq_waves * NGI_SHQ))) return st;
    }
    for (int i = 0; i < count; ++i) args[i].film = film + film_off[i];
    NGI_HIP_CHECK(hipMemcpyAsync(d_args, args.data(), count * sizeof(RenderArgs), hipMemcpyHostToDevice, stream));
    NGI_HIP_CHECK(hipMemcpyAsync(d_off, offset.data(), (count + 1) * sizeof(unsigned long long), hipMemcpyHostToDevice, stream));
    NGI_HIP_CHECK(hipStreamSynchronize(stream));
    if ((st = autotune())) return st;
    const std::vector<unsigned long long> ch = chunk_table(offset, (uint32_t)opts.chunk_paths);
    num_chunks = ch.back();
    NGI_HIP_CHECK(hipMemcpyAsync(d_chunk, ch.data(), (count + 1) * sizeof(unsigned long long), hipMemcpyHostToDevice, stream));
    NGI_HIP_CHECK(hipStreamSynchronize(stream));
    return opts.use_graph ? capture() : NGI_OK;
}

// Paths per chunk when opts.chunk_paths is 0: about 16 chunks per resident
// wave, a power of two between 256 and 8192.  Larger chunks mean fewer scene
// switches (each drains the wave's shadow queue and takes a chunk with an
// atomic); on 256 rooms (2.18e9 paths, 4096 waves) frames took 1,433 /
// 1,335 / 1,350 / 1,519 ms with chunks of 2,048 / 8,192 / 32,768 / 131,072,
// and 16 chunks per wave keep every wave busy until the frame's tail.
static int chunk_for(unsigned long long paths, unsigned long long waves) {
    const unsigned long long per = paths / std::max<unsigned long long>(1, 16 * waves);
    int c = 256;
    while (c < 8192 && (unsigned long long)(2 * c) <= per) c *= 2;
    return c;
}

// Times every (block size, resident blocks per CU) candidate on a sample of
// about opts.tune_paths paths taken proportionally from every scene and
// keeps the fastest.  Each candidate's chunk size follows chunk_for() on the
// paths it renders (the sample while tuning, the whole frame afterwards), so
// the sample's scenes switch about as often per wave as the frame's.
// Choices fixed in opts are kept.
int Batch::autotune() {
    const auto t0 = std::chrono::steady_clock::now();
    const std::vector<int> blocks = opts.block_size ? std::vector<int>{opts.block_size} : std::vector<int>{64, 128, 256};
    const bool fixed_chunk = opts.chunk_paths != 0;
    const unsigned long long total = offset.back();
    const double f = std::min(1.0, (double)std::max<int64_t>(opts.tune_paths, 1) / (double)std::max<unsigned long long>(total, 1));
    std::vector<unsigned long long> toff(n() + 1, 0);
    for (int i = 0; i < n(); ++i) toff[i + 1] = toff[i] + (unsigned long long)((double)(offset[i + 1] - offset[i]) * f);
    NGI_HIP_CHECK(hipMemcpyAsync(d_tune_off, toff.data(), (n() + 1) * sizeof(unsigned long long), hipMemcpyHostToDevice, stream));
    struct Cand {
        int block, per_cu, chunk;
    };
    const auto waves_of = [&](int bs, int per) { return (unsigned long long)num_cus * (unsigned long long)per * (unsigned long long)(bs / 64); };
    std::vector<Cand> cands;
    for (int bs : blocks) {
        const int occ = max_per_cu(bs);
        if (occ <= 0) continue;
        std::vector<int> pers;
        if (opts.blocks_per_cu) pers.push_back(std::min(opts.blocks_per_cu, occ));
        else
            for (int per : {occ, (3 * occ + 3) / 4, (occ + 1) / 2})
                if (std::find(pers.begin(), pers.end(), std::max(1, per)) == pers.end()) pers.push_back(std::max(1, per));
        for (int per : pers) cands.push_back({bs, per, fixed_chunk ? opts.chunk_paths : chunk_for(toff.back(), waves_of(bs, per))});
    }
    if (cands.empty()) return fail(NGI_ERR_DEVICE, "no batch kernel configuration fits on this device");
    int best = 0;
    if (cands.size() > 1 && toff.back() > 0) {
        float best_ms = 0.0f;
        int loaded_chunk = -1;
        unsigned long long tchunks = 0;
        for (size_t c = 0; c < cands.size(); ++c) {
            if (cands[c].chunk != loaded_chunk) {
                const std::vector<unsigned long long> tch = chunk_table(toff, (uint32_t)cands[c].chunk);
                tchunks = tch.back();
                NGI_HIP_CHECK(hipMemcpyAsync(d_tune_chunk, tch.data(), (n() + 1) * sizeof(unsigned long long), hipMemcpyHostToDevice, stream));
                NGI_HIP_CHECK(hipStreamSynchronize(stream));
                loaded_chunk = cands[c].chunk;
            }
            float ms = 0.0f;
            for (int rep = 0; rep < 2; ++rep) {   // the first run warms caches and loads the code
                int st = enqueue_clear();
                if (st) return st;
                NGI_HIP_CHECK(hipEventRecord(ev0, stream));
                if ((st = enqueue_render(d_tune_off, d_tune_chunk, tchunks, (uint32_t)cands[c].chunk, cands[c].block, cands[c].per_cu))) return st;
                NGI_HIP_CHECK(hipEventRecord(ev1, stream));
                NGI_HIP_CHECK(hipEventSynchronize(ev1));
                NGI_HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
            }
            if (c == 0 || ms < best_ms) best_ms = ms, best = (int)c;
        }
    }
    opts.block_size = cands[best].block;
    opts.blocks_per_cu = cands[best].per_cu;
    if (!fixed_chunk) opts.chunk_paths = chunk_for(total, waves_of(opts.block_size, opts.blocks_per_cu));
    tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return NGI_OK;
}

int Batch::capture() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr, graph = nullptr;
    NGI_HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int st = enqueue_clear();
    if (!st) st = enqueue_frame();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(stream, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return fail(NGI_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    graph = g;
    NGI_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    return NGI_OK;
}

int Batch::render(uint32_t f) {
    frame = f;
    int st = refresh();
    if (st) return st;
    NGI_HIP_CHECK(hipEventRecord(ev0, stream));
    if (exec) {
        NGI_HIP_CHECK(hipGraphLaunch(exec, stream));
    } else {
        if ((st = enqueue_clear())) return st;
        if ((st = enqueue_frame())) return st;
    }
    NGI_HIP_CHECK(hipEventRecord(ev1, stream));
    rendered = true;
    return NGI_OK;
}

int Batch::sync() {
    NGI_HIP_CHECK(hipStreamSynchronize(stream));
    return NGI_OK;
}

int Batch::get_film(int i, double* out) {
    if (i < 0 || i >= n()) return fail(NGI_ERR_INVALID_ARG, "scene index %d out of range", i);
    if (!rendered) return fail(NGI_ERR_INVALID_ARG, "the batch has not been rendered");
    const size_t m = film_off[i + 1] - film_off[i];
    NGI_HIP_CHECK(hipMemcpyAsync(out, film + film_off[i], m * sizeof(double), hipMemcpyDeviceToHost, stream));
    NGI_HIP_CHECK(hipStreamSynchronize(stream));
    // W*H/N scaling (src/nanogi.cpp:434-437).
    const double scale = args[i].path_count ? (double)args[i].npix / (double)args[i].path_count : 0.0;
    for (size_t k = 0; k < m; ++k) out[k] *= scale;
    return NGI_OK;
}

int Batch::get_times(double* scene_ms, ngi_stats* stats) {
    if (!rendered) return fail(NGI_ERR_INVALID_ARG, "the batch has not been rendered");
    NGI_HIP_CHECK(hipStreamSynchronize(stream));
    std::vector<unsigned long long> td(n()), misc(MISC_N);
    NGI_HIP_CHECK(hipMemcpy(td.data(), d_tdone, n() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    NGI_HIP_CHECK(hipMemcpy(misc.data(), d_misc, MISC_N * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    if (scene_ms)
        for (int i = 0; i < n(); ++i)
            scene_ms[i] = (td[i] && td[i] >= misc[MISC_T_START] && rate_khz > 0) ? (double)(td[i] - misc[MISC_T_START]) / (double)rate_khz : 0.0;
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        float ms = 0.0f;
        NGI_HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
        stats->samples = (int64_t)offset.back();
        stats->extension_rays = (int64_t)misc[MISC_EXT];
        stats->shadow_rays = (int64_t)misc[MISC_SHADOW];
        stats->ms_total = ms;
        stats->ms_shade = ms;
        stats->trace_launches = 1;
    }
    return NGI_OK;
}

void Batch::destroy() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (ngi_scene* sc : scenes) {
        auto& u = sc->user_streams;
        u.erase(std::remove(u.begin(), u.end(), (void*)stream), u.end());
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void* q : {(void*)film, (void*)d_args, (void*)d_off, (void*)d_tune_off, (void*)d_chunk, (void*)d_tune_chunk, (void*)d_done,
                    (void*)d_tdone, (void*)d_misc, (void*)shq})
        if (q) (void)hipFree(q);
    shq = nullptr, shq_waves = 0;
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
    exec = nullptr, graph = nullptr, stream = nullptr, ev0 = ev1 = nullptr;
    film = nullptr, d_args = nullptr, d_off = d_tune_off = d_chunk = d_tune_chunk = d_done = d_tdone = d_misc = nullptr;
}

}  // namespace ngi

struct ngi_batch : ngi::Batch {};

using namespace ngi;

extern "C" {

void ngi_batch_opts_default(ngi_batch_opts* o) {
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->use_graph = 1;
    o->chunk_paths = 0;
    o->tune_paths = 1ll << 24;
}

int ngi_batch_create(ngi_scene* const* scenes, const ngi_render_params* params, int num_scenes, const ngi_batch_opts* opts, ngi_batch** out) {
    clear_error();
    if (!scenes || !params || !out || num_scenes <= 0) return fail(NGI_ERR_INVALID_ARG, "null argument or no scenes");
    *out = nullptr;
    ngi_batch_opts o;
    if (opts) o = *opts;
    else ngi_batch_opts_default(&o);
    ngi_batch* b = new ngi_batch();
    const int st = b->create(scenes, params, num_scenes, o);
    if (st) {
        b->destroy();
        delete b;
        return st;
    }
    *out = b;
    return NGI_OK;
}

int ngi_batch_render(ngi_batch* b, uint32_t frame) {
    clear_error();
    if (!b) return fail(NGI_ERR_INVALID_ARG, "null batch");
    return b->render(frame);
}

int ngi_batch_sync(ngi_batch* b) {
    clear_error();
    if (!b) return fail(NGI_ERR_INVALID_ARG, "null batch");
    return b->sync();
}

int ngi_batch_get_film(ngi_batch* b, int scene, double* film_out) {
    clear_error();
    if (!b || !film_out) return fail(NGI_ERR_INVALID_ARG, "null argument");
    return b->get_film(scene, film_out);
}

int ngi_batch_get_times(ngi_batch* b, double* scene_ms, ngi_stats* stats) {
    clear_error();
    if (!b) return fail(NGI_ERR_INVALID_ARG, "null batch");
    return b->get_times(scene_ms, stats);
}

int ngi_batch_get_opts(ngi_batch* b, ngi_batch_opts* chosen, double* tune_ms) {
    clear_error();
    if (!b || !chosen) return fail(NGI_ERR_INVALID_ARG, "null argument");
    *chosen = b->opts;
    if (tune_ms) *tune_ms = b->tune_ms;
    return NGI_OK;
}

void ngi_batch_destroy(ngi_batch* b) {
    if (!b) return;
    b->destroy();
    delete b;
}

int ngi_render_batch(ngi_scene* const* scenes, const ngi_render_params* params, int num_scenes, double* const* films, double* scene_ms,
                     ngi_stats* stats) {
    clear_error();
    if (!scenes || !params || !films || num_scenes <= 0) return fail(NGI_ERR_INVALID_ARG, "null argument or no scenes");
    int st = validate(scenes, params, num_scenes);
    if (st) return st;
    // One launch: no tuning trials and no graph.
    ngi_batch_opts o;
    ngi_batch_opts_default(&o);
    o.block_size = 256;
    o.blocks_per_cu = 1 << 20;   // clamped to the occupancy limit
    o.chunk_paths = 1024;
    o.use_graph = 0;
    ngi_batch* b = nullptr;
    if ((st = ngi_batch_create(scenes, params, num_scenes, &o, &b))) return st;
    const auto t0 = std::chrono::steady_clock::now();
    st = b->render(0);
    if (!st) st = b->sync();
    const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (int i = 0; !st && i < num_scenes; ++i) st = b->get_film(i, films[i]);
    if (!st) st = b->get_times(scene_ms, stats);
    if (!st && stats) stats->ms_total = wall;
    const std::string err = st ? std::string(ngi_last_error()) : std::string();
    ngi_batch_destroy(b);
    if (st) fail(st, "%s", err.c_str());
    return st;
}

}  // extern "C"
