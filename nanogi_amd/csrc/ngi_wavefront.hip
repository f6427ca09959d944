// Wavefront engine: the per-path vertex loop of ProcessSample_PT /
// ProcessSample_PTDirect (src/nanogi.cpp:446-802) split into stages that
// each run over a queue of paths:
//
//   generate : new paths (camera vertex: pinhole sample, camera-vertex NEE)
//   extend   : closest-hit traversal of the extension-ray queue
//   shade    : hit surface, emission (pt), Russian roulette, next vertex:
//              NEE shadow ray + BSDF sample -> next extension-ray queue
//   shadow   : any-hit traversal of the shadow queue, FP64 film splat
//
// Queues are compacted with a wavefront ballot + prefix count and one
// global atomic per wavefront.  Paths live in the queue entries themselves
// (ray + 32-byte state), so terminated paths cost nothing in later stages
// and new paths refill the queue up to its capacity every iteration.
//
// HBM layout per queue entry (16-byte aligned float4 streams):
//   ray   : (o.xyz, tmin) (d.xyz, tmax)                     32 B
//   state : (T.rgb, pixel) (c0, c1, vertex, prim<<5|type)    32 B
//   hit   : (t, u, v, leaf)                                  16 B
//   shadow: (o.xyz, tmax) (d.xyz, pixel) (C.rgb, -)          48 B
#include <algorithm>
#include <cstring>

#include "ngi_render.h"

namespace ngi {

#define NGI_WF_BLOCK 256
#ifndef NGI_TRACE_WAVES
#define NGI_TRACE_WAVES 1    // minimum waves per SIMD requested for the traversal kernels
#endif

struct WavefrontState {
    uint32_t capacity = 0;
    float4* ray[2] = {nullptr, nullptr};
    float4* st[2] = {nullptr, nullptr};
    float4* hit = nullptr;
    float4* shadow = nullptr;
    uint32_t* counts = nullptr;        // device: [0] queue-out count, [1] shadow count
    uint32_t* host_counts = nullptr;   // pinned mirror
};

// Queue slots for two queues at once, aggregated per block: each wavefront
// takes its offset within the block with one LDS atomic per queue, then one
// thread adds the block's totals to the global counters (one global atomic
// per block and queue instead of one per wavefront: every wavefront of the
// grid would otherwise hit the same two addresses).  Every thread of the
// block must call it.  Returns 0xFFFFFFFF for a lane without the predicate.
__device__ __forceinline__ void block_push2(uint32_t* counter_a, bool pa, uint32_t& slot_a, uint32_t* counter_b, bool pb,
                                            uint32_t& slot_b) {
    __shared__ uint32_t tot[2], base[2];
    const uint32_t lane = __lane_id();
    const unsigned long long ma = __ballot(pa), mb = __ballot(pb);
    if (threadIdx.x == 0) tot[0] = tot[1] = 0;
    __syncthreads();
    uint32_t wa = 0, wb = 0;
    if (lane == 0) {
        if (ma) wa = atomicAdd(&tot[0], (uint32_t)__popcll(ma));
        if (mb) wb = atomicAdd(&tot[1], (uint32_t)__popcll(mb));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        base[0] = tot[0] ? atomicAdd(counter_a, tot[0]) : 0u;
        base[1] = tot[1] && counter_b ? atomicAdd(counter_b, tot[1]) : 0u;
    }
    __syncthreads();
    wa = __shfl(wa, 0);
    wb = __shfl(wb, 0);
    const unsigned long long below = (1ull << lane) - 1ull;
    slot_a = pa ? base[0] + wa + (uint32_t)__popcll(ma & below) : 0xFFFFFFFFu;
    slot_b = pb ? base[1] + wb + (uint32_t)__popcll(mb & below) : 0xFFFFFFFFu;
}

__device__ __forceinline__ void wave_count(unsigned long long* p, uint32_t v) {
    unsigned long long x = v;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if (__lane_id() == 0 && x) atomicAdd(p, x);
}

__device__ __forceinline__ void store_shadow(float4* q, uint32_t slot, const TraceRay& r, int pixel, f3 C) {
    q[3 * (size_t)slot + 0] = make_float4(r.ox, r.oy, r.oz, r.tmax);
    q[3 * (size_t)slot + 1] = make_float4(r.dx, r.dy, r.dz, __int_as_float(pixel));
    q[3 * (size_t)slot + 2] = make_float4(C.x, C.y, C.z, 0.0f);
}

__device__ __forceinline__ void store_path(float4* ray, float4* st, uint32_t slot, f3 o, f3 d, f3 T, int pixel, uint32_t c0,
                                           uint32_t c1, int nv, int prim, int type, float tmax = FLT_MAX) {
    ray[2 * (size_t)slot + 0] = make_float4(o.x, o.y, o.z, NGI_EPS_F);
    ray[2 * (size_t)slot + 1] = make_float4(d.x, d.y, d.z, tmax);
    st[2 * (size_t)slot + 0] = make_float4(T.x, T.y, T.z, __int_as_float(pixel));
    st[2 * (size_t)slot + 1] = __builtin_bit_cast(float4, make_uint4(c0, c1, (uint32_t)nv, ((uint32_t)prim << 5) | (uint32_t)type));
}

// New paths: ids path_base .. path_base + n - 1 into queue slots slot_base..
template <bool DIRECT>
__global__ void __launch_bounds__(NGI_WF_BLOCK) wf_generate(RenderArgs a, uint64_t path_base, uint32_t n, uint32_t slot_base,
                                                           float4* ray, float4* st, float4* shadow, uint32_t* counts) {
    const uint32_t i = blockIdx.x * NGI_WF_BLOCK + threadIdx.x;
    // The camera vertex is vertex 1: with a limit of one vertex or none the
    // path ends before it does anything (src/nanogi.cpp:485-489).
    const int max_nv = (a.max_nv < 0 || a.max_nv > NGI_MAX_VERTICES) ? NGI_MAX_VERTICES : a.max_nv;
    const bool active = i < n && max_nv > 1;
    const bool dead = i < n && !active;   // occupies its queue slot with a ray that hits nothing
    const DevScene& sc = a.sc;
    const Cam<float> cam = load_cam<float>(sc.cam);
    uint32_t c0 = 0, c1 = 0;
    int pixel = -1;
    if (active) path_coords(a, path_base + i, c0, c1, pixel);
    bool emit = false;
    TraceRay sr;
    int target = 0;
    f3 C(0.0f);
    if (DIRECT && active) {
        // Camera-vertex NEE (src/nanogi.cpp:656-710 with type == E).
        Surf<float> s;
        s.p = cam.eye;
        s.degenerate = true;
        const U4 rn = rng4(c0, c1, 1u, NGI_STREAM_NEE, a.seed, a.stream_id);
        C = nee_contribution<float>(sc, cam, prims_of<float>(sc), NGI_PRIM_E, sc.sensor_prim, s, f3(0.0f), f3(1.0f), rn, a.w, a.h,
                                    pixel, sr, target);
        emit = !is_zero(C);
    }
    if (DIRECT) {
        uint32_t slot = 0, unused = 0;
        block_push2(counts + 1, emit, slot, nullptr, false, unused);
        if (emit) store_shadow(shadow, slot, sr, target, C);
    }
    if (dead) store_path(ray, st, slot_base + i, cam.eye, f3(0.0f, 0.0f, 1.0f), f3(0.0f), 0, 0, 0, 1, sc.sensor_prim, NGI_PRIM_E, -1.0f);
    if (!active) return;
    const U4 rx = rng4(c0, c1, 1u, NGI_STREAM_EXT, a.seed, a.stream_id);
    const f3 wo = camera_sample<float>(a, cam, u24f(rx.x), u24f(rx.y), pixel);
    store_path(ray, st, slot_base + i, cam.eye, wo, f3(1.0f), pixel, c0, c1, 1, sc.sensor_prim, NGI_PRIM_E);
}

template <int WIDTH, bool COUNT>
__global__ void __launch_bounds__(NGI_WF_BLOCK, NGI_TRACE_WAVES) wf_extend(DevScene sc, const float4* __restrict__ ray, float4* __restrict__ hit, uint32_t n,
                                                         unsigned long long* counters) {
    extern __shared__ uint32_t stk[];   // sc.stack_entries x block (wide BVH only)
    const uint32_t i = blockIdx.x * NGI_WF_BLOCK + threadIdx.x;
    TraceCounters tc{0, 0};
    if (i < n) {
        const float4 a = ray[2 * (size_t)i], b = ray[2 * (size_t)i + 1];
        TraceRay r{a.x, a.y, a.z, b.x, b.y, b.z, a.w, b.w};
        TraceHit h;
        const bool ok = trace<WIDTH, false, COUNT>(sc, r, h, tc, stk + threadIdx.x, NGI_WF_BLOCK);
        hit[i] = make_float4(h.t, h.u, h.v, __int_as_float(ok ? h.leaf : -1));
    }
    if (COUNT) {
        wave_count(counters + CNT_EXT_NODES, tc.nodes);
        wave_count(counters + CNT_EXT_TRIS, tc.tris);
    }
}

template <bool DIRECT, typename FilmT>
__global__ void __launch_bounds__(NGI_WF_BLOCK) wf_shade(RenderArgs a, const float4* __restrict__ ray_in, const float4* __restrict__ st_in,
                                                        const float4* __restrict__ hit, uint32_t n, float4* __restrict__ ray_out,
                                                        float4* __restrict__ st_out, float4* __restrict__ shadow, uint32_t* counts) {
    const uint32_t i = blockIdx.x * NGI_WF_BLOCK + threadIdx.x;
    const DevScene& sc = a.sc;
    bool alive = false, emit = false;
    f3 C(0.0f), wo(0.0f), T(0.0f);
    TraceRay sr;
    int target = 0, pixel = 0, hp = 0, type = 0, nv = 0;
    uint32_t c0 = 0, c1 = 0;
    Surf<float> ns;
    if (i < n) {
        const float4 h4 = hit[i];
        const int leaf = __float_as_int(h4.w);
        if (leaf >= 0) {
            const float4 r0 = ray_in[2 * (size_t)i], r1 = ray_in[2 * (size_t)i + 1];
            const float4 s0 = st_in[2 * (size_t)i];
            const uint4 s1 = __builtin_bit_cast(uint4, st_in[2 * (size_t)i + 1]);
            const f3 o(r0.x, r0.y, r0.z), d(r1.x, r1.y, r1.z);
            T = f3(s0.x, s0.y, s0.z);
            pixel = __float_as_int(s0.w);
            c0 = s1.x, c1 = s1.y, nv = (int)s1.z;
            TraceHit h;
            h.t = h4.x, h.u = h4.y, h.v = h4.z, h.leaf = leaf;
            int htype;
            ns = hit_surface<float>(sc, o, d, h, hp, htype);
            const DevPrim<float>* prims = prims_of<float>(sc);
            if (!DIRECT && (htype & NGI_PRIM_L)) {
                const DevPrim<float>& hm = prims[hp];
                if (hm.light >= 0) {
                    const DevLight<float>& L = lights_of<float>(sc)[hm.light];
                    if (L.type == NGI_LIGHT_AREA && dot(ns.n, -d) > 0.0f) film_add<FilmT>(a.film, pixel, T * f3(hm.Le[0], hm.Le[1], hm.Le[2]));
                }
            }
            const U4 rx = rng4(c0, c1, (uint32_t)nv, NGI_STREAM_EXT, a.seed, a.stream_id);
            alive = true;
            if (a.rr_depth >= 0 && nv >= a.rr_depth) {   // src/nanogi.cpp:583-591
                if (u24f(rx.w) > a.rr_prob_f) alive = false;
                else T = T / a.rr_prob_f;
            }
            type = htype & ~(NGI_PRIM_L | NGI_PRIM_E);
            if (type == 0) alive = false;
            nv++;
            const int max_nv = (a.max_nv < 0 || a.max_nv > NGI_MAX_VERTICES) ? NGI_MAX_VERTICES : a.max_nv;
            if (nv >= max_nv) alive = false;
            const f3 wi = -d;
            if (alive && DIRECT) {
                const U4 rn = rng4(c0, c1, (uint32_t)nv, NGI_STREAM_NEE, a.seed, a.stream_id);
                Cam<float> cam;   // unused at surface vertices
                C = nee_contribution<float>(sc, cam, prims, type, hp, ns, wi, T, rn, a.w, a.h, pixel, sr, target);
                emit = !is_zero(C);
            }
            if (alive) {
                const U4 ry = rng4(c0, c1, (uint32_t)nv, NGI_STREAM_EXT, a.seed, a.stream_id);
                f3 wgt;
                if (bsdf_sample<float>(prims[hp], type, ns, wi, u24f(ry.x), u24f(ry.y), u24f(ry.z), wo, wgt)) T = T * wgt;
                else alive = false;
            }
        }
    }
    uint32_t slot = 0, sslot = 0;
    block_push2(counts + 0, alive, slot, DIRECT ? counts + 1 : nullptr, DIRECT && emit, sslot);
    if (DIRECT && emit) store_shadow(shadow, sslot, sr, target, C);
    if (alive) store_path(ray_out, st_out, slot, ns.p, wo, T, pixel, c0, c1, nv, hp, type);
}

template <int WIDTH, typename FilmT, bool COUNT>
__global__ void __launch_bounds__(NGI_WF_BLOCK, NGI_TRACE_WAVES) wf_shadow(DevScene sc, const float4* __restrict__ q, const uint32_t* counts, uint32_t bound,
                                                         void* film, unsigned long long* counters) {
    extern __shared__ uint32_t stk[];   // sc.stack_entries x block (wide BVH only)
    const uint32_t i = blockIdx.x * NGI_WF_BLOCK + threadIdx.x;
    const uint32_t n = min(counts[1], bound);
    TraceCounters tc{0, 0};
    if (i < n) {
        const float4 a = q[3 * (size_t)i], b = q[3 * (size_t)i + 1], c = q[3 * (size_t)i + 2];
        TraceRay r{a.x, a.y, a.z, b.x, b.y, b.z, NGI_EPS_F, a.w};
        TraceHit h;
        const bool occluded = trace<WIDTH, true, COUNT>(sc, r, h, tc, stk + threadIdx.x, NGI_WF_BLOCK);
        if (!occluded) film_add<FilmT>(film, __float_as_int(b.w), f3(c.x, c.y, c.z));
    }
    if (COUNT) {
        wave_count(counters + CNT_SH_NODES, tc.nodes);
        wave_count(counters + CNT_SH_TRIS, tc.tris);
    }
}

__global__ void wf_finish(uint32_t* counts, unsigned long long* counters, uint32_t ext_rays, uint32_t new_paths) {
    // Book-keeping after an iteration: statistics, then clear the shadow count.
    counters[CNT_EXT_RAYS] += ext_rays;
    counters[CNT_SAMPLES] += new_paths;
    counters[CNT_SHADOW_RAYS] += counts[1];
    counts[1] = 0;
}

__global__ void wf_clear(uint32_t* counts) { counts[0] = 0; }

// ---------------------------------------------------------------------------
template <bool DIRECT, typename FilmT, int WIDTH, bool COUNT>
static int run_wavefront(Session& S, WavefrontState& W, void* target, uint64_t total) {
    hipStream_t st = S.stream();
    RenderArgs a = S.args(target);
    a.sample_base = S.samples_done;
    unsigned long long* counters = reinterpret_cast<unsigned long long*>(S.counters);
    const uint32_t C = W.capacity;
    const size_t lds = WIDTH == 8 ? (size_t)a.sc.stack_entries * NGI_WF_BLOCK * 4 : 0;
    uint64_t next = 0;          // next path of this render call
    // Path ids: per-pixel ids restart at 0 each call (sample_base carries the
    // sample offset); reference-uniform ids are global sample indices.
    const uint64_t id_base = S.p.sampling == NGI_SAMPLING_PER_PIXEL ? 0 : S.paths_done;
    uint32_t live = 0;          // paths in the current input queue
    int cur = 0;                // input queue index
    // Limits of one vertex or none: new paths end at the camera and their
    // placeholder rays are not counted as extension rays.
    const bool camera_only = a.max_nv >= 0 && a.max_nv <= 1;
    for (;;) {
        const uint32_t n_new = (uint32_t)std::min<uint64_t>(C - live, total - next);
        const uint32_t n = live + n_new;
        if (n == 0) break;
        TimedLaunch tl;
        int rc = S.timed_begin(K_SHADE, n_new, tl);
        if (rc) return rc;
        wf_clear<<<1, 1, 0, st>>>(W.counts);
        if (n_new) {
            wf_generate<DIRECT><<<(n_new + NGI_WF_BLOCK - 1) / NGI_WF_BLOCK, NGI_WF_BLOCK, 0, st>>>(a, id_base + next, n_new, live, W.ray[cur], W.st[cur],
                                                                                                W.shadow, W.counts);
        }
        if ((rc = S.timed_end(tl))) return rc;
        if ((rc = S.timed_begin(K_EXTEND, n, tl))) return rc;
        wf_extend<WIDTH, COUNT><<<(n + NGI_WF_BLOCK - 1) / NGI_WF_BLOCK, NGI_WF_BLOCK, lds, st>>>(a.sc, W.ray[cur], W.hit, n, counters);
        if ((rc = S.timed_end(tl))) return rc;
        if ((rc = S.timed_begin(K_SHADE, n, tl))) return rc;
        wf_shade<DIRECT, FilmT><<<(n + NGI_WF_BLOCK - 1) / NGI_WF_BLOCK, NGI_WF_BLOCK, 0, st>>>(a, W.ray[cur], W.st[cur], W.hit, n, W.ray[cur ^ 1],
                                                                                             W.st[cur ^ 1], W.shadow, W.counts);
        if ((rc = S.timed_end(tl))) return rc;
        if (DIRECT) {
            // At most one shadow ray per generated path and per shaded path.
            const uint32_t bound = n_new + n;
            if ((rc = S.timed_begin(K_SHADOW, bound, tl))) return rc;
            wf_shadow<WIDTH, FilmT, COUNT><<<(bound + NGI_WF_BLOCK - 1) / NGI_WF_BLOCK, NGI_WF_BLOCK, lds, st>>>(a.sc, W.shadow, W.counts, bound,
                                                                                                           target, counters);
            if ((rc = S.timed_end(tl))) return rc;
        }
        wf_finish<<<1, 1, 0, st>>>(W.counts, counters, camera_only ? live : n, n_new);
        NGI_HIP_CHECK(hipGetLastError());
        NGI_HIP_CHECK(hipMemcpyAsync(W.host_counts, W.counts, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        NGI_HIP_CHECK(hipStreamSynchronize(st));
        next += n_new;
        live = W.host_counts[0];
        if (live > C) return fail(NGI_ERR_QUEUE_OVERFLOW, "wavefront queue overflow (%u > %u)", live, C);
        cur ^= 1;
    }
    return NGI_OK;
}

template <bool DIRECT, typename FilmT>
static int run_wf2(Session& S, WavefrontState& W, void* target, uint64_t total, int width, bool count) {
    if (width == 8) return count ? run_wavefront<DIRECT, FilmT, 8, true>(S, W, target, total) : run_wavefront<DIRECT, FilmT, 8, false>(S, W, target, total);
    return count ? run_wavefront<DIRECT, FilmT, 2, true>(S, W, target, total) : run_wavefront<DIRECT, FilmT, 2, false>(S, W, target, total);
}

static int wf_alloc(WavefrontState& W, uint32_t cap) {
    W.capacity = cap;
    for (int k = 0; k < 2; ++k) {
        NGI_HIP_CHECK(hipMalloc(&W.ray[k], (size_t)cap * 32));
        NGI_HIP_CHECK(hipMalloc(&W.st[k], (size_t)cap * 32));
    }
    NGI_HIP_CHECK(hipMalloc(&W.hit, (size_t)cap * 16));
    NGI_HIP_CHECK(hipMalloc(&W.shadow, (size_t)cap * 2 * 48));
    NGI_HIP_CHECK(hipMalloc(&W.counts, 4 * sizeof(uint32_t)));
    NGI_HIP_CHECK(hipHostMalloc((void**)&W.host_counts, 4 * sizeof(uint32_t), hipHostMallocDefault));
    NGI_HIP_CHECK(hipMemset(W.counts, 0, 4 * sizeof(uint32_t)));
    return NGI_OK;
}

int Session::render_wavefront(void* target, uint64_t total) {
    if (p.precision == NGI_PRECISION_FP64_ALL)
        return fail(NGI_ERR_UNSUPPORTED, "the wavefront engine shades in FP32; FP64 shading runs on the megakernel");
    // Default capacity: the render's paths rounded up to a power of two,
    // between 2^16 and 2^25 (8 GB of queues); the queues grow when a later
    // render is larger.  Larger queues mean fewer iterations: 1,033 / 1,148 /
    // 1,216 / 1,245 M samples/s for 2^22 / 2^23 / 2^24 / 2^25 on the
    // 1.8M-triangle sphere field.
    // When the device cannot hold default queues, they halve (down to 2^16)
    // until they fit, and later renders stay at the size that fitted.
    uint32_t def = 1u << 16;
    while (def < wf_cap_max && def < total) def <<= 1;
    if (wf && p.paths_in_flight <= 0 && def > wf->capacity) {
        NGI_HIP_CHECK(hipStreamSynchronize(stream()));
        release_engine();
    }
    while (!wf) {
        wf = new WavefrontState();
        const uint32_t cap = std::max<uint32_t>(p.paths_in_flight > 0 ? (uint32_t)p.paths_in_flight : def, 256);
        const int st = wf_alloc(*wf, cap);
        if (!st) break;
        release_engine();          // frees what was allocated; wf = nullptr
        (void)hipGetLastError();   // the refused allocation must not fail the next launch check
        if (st != NGI_ERR_OOM || p.paths_in_flight > 0 || cap <= (1u << 16)) return st;
        wf_cap_max = def = cap / 2;
    }
    const bool direct = p.renderer == NGI_RENDERER_PTDIRECT;
    const int width = scene->dev->view.bvh_width;
    const bool count = p.collect_counters != 0;
    if (film64) return direct ? run_wf2<true, double>(*this, *wf, target, total, width, count) : run_wf2<false, double>(*this, *wf, target, total, width, count);
    return direct ? run_wf2<true, float>(*this, *wf, target, total, width, count) : run_wf2<false, float>(*this, *wf, target, total, width, count);
}

void Session::release_engine() {
    if (!wf) return;
    for (int k = 0; k < 2; ++k) {
        if (wf->ray[k]) (void)hipFree(wf->ray[k]);
        if (wf->st[k]) (void)hipFree(wf->st[k]);
    }
    if (wf->hit) (void)hipFree(wf->hit);
    if (wf->shadow) (void)hipFree(wf->shadow);
    if (wf->counts) (void)hipFree(wf->counts);
    if (wf->host_counts) (void)hipHostFree(wf->host_counts);
    delete wf;
    wf = nullptr;
}

}  // namespace ngi
