// Path-vertex building blocks shared by the megakernel, the wavefront
// stages, the adjoint and the batched renderer.
//
// One path sample follows Renderer::ProcessSample_PT (src/nanogi.cpp:446-607)
// and Renderer::ProcessSample_PTDirect (src/nanogi.cpp:609-802):
//   camera vertex (numVertices = 1) -> [NEE] -> sample direction -> trace ->
//   [emission on hit, pt only] -> Russian roulette -> next vertex,
// with random numbers drawn from the counter-based stream of ngi_common.h.
#pragma once

#include <float.h>
#include <stdlib.h>

#include <type_traits>

#include "ngi_shade.h"
#include "ngi_traverse.h"

namespace ngi {

#define NGI_EPS_F 1e-4f   // EpsF (include/nanogi/basic.hpp:85)
// Hard cap on path vertices so every lane terminates even with -m -1 and
// Russian roulette off (the reference would loop forever in a closed scene).
#define NGI_MAX_VERTICES 1024

enum { CNT_EXT_RAYS = 0, CNT_SHADOW_RAYS, CNT_EXT_NODES, CNT_EXT_TRIS, CNT_SH_NODES, CNT_SH_TRIS, CNT_SAMPLES, CNT_N };

struct RenderArgs {
    DevScene sc;
    void* film;                 // FilmT[W*H*3]
    int w, h;
    uint32_t npix;
    int sampling;               // NGI_SAMPLING_*
    uint64_t path_begin;        // first path id of the launch
    uint64_t path_count;
    uint32_t sample_base;       // per-pixel: sample index of path id 0
    int max_nv;
    int rr_depth;
    float rr_prob_f;
    double rr_prob_d;
    uint32_t seed, stream_id;
    unsigned long long* counters;   // CNT_N entries or null
    uint32_t path_chunk;            // persistent engines: path ids per counter atomic (0 = exactly the ids needed)
};

template <typename R> __device__ __forceinline__ R unif(uint32_t x);
template <> __device__ __forceinline__ float unif<float>(uint32_t x) { return u24f(x); }
template <> __device__ __forceinline__ double unif<double>(uint32_t x) { return u24d(x); }
template <typename R> __device__ __forceinline__ R rr_prob(const RenderArgs& a);
template <> __device__ __forceinline__ float rr_prob<float>(const RenderArgs& a) { return a.rr_prob_f; }
template <> __device__ __forceinline__ double rr_prob<double>(const RenderArgs& a) { return a.rr_prob_d; }

__device__ __forceinline__ U4 rng4(uint32_t c0, uint32_t c1, uint32_t vertex, uint32_t stream, uint32_t seed, uint32_t sid) {
    U4 c;
    c.x = c0, c.y = c1, c.z = vertex, c.w = stream;
    return philox4x32_10(c, seed, sid);
}

template <typename FilmT, typename R>
__device__ __forceinline__ void film_add(void* film, int pixel, V3<R> v) {
    FilmT* f = reinterpret_cast<FilmT*>(film) + 3 * (size_t)pixel;
    if (v.x != R(0)) unsafeAtomicAdd(f + 0, (FilmT)v.x);
    if (v.y != R(0)) unsafeAtomicAdd(f + 1, (FilmT)v.y);
    if (v.z != R(0)) unsafeAtomicAdd(f + 2, (FilmT)v.z);
}


template <int WIDTH, bool ANY, bool COUNT>
__device__ __forceinline__ bool trace(const DevScene& sc, const TraceRay& r, TraceHit& h, TraceCounters& c, uint32_t* stk, int stride) {
    if (sc.num_tris == 0) {   // no geometry (and no nodes): every query misses
        h.leaf = -1;
        return false;
    }
    if (WIDTH == 8) return trace_bvh8<ANY, COUNT>(sc, r, h, c, stk, stride);
    return trace_bvh2<ANY, COUNT>(sc, r, h, c);
}

// Deferred shadow rays of the persistent engines.  Only about half of the
// NEE samples of a vertex need a shadow ray, so tracing them where they are
// made runs the any-hit traversal with half of the wavefront idle.  Instead
// each wavefront appends its shadow rays (with their contribution and pixel)
// to its own ring of NGI_SHQ records in HBM and traces them 64 at a time.
// The film receives the same contributions; only the order of the FP64
// additions changes.  Record k of slot s of wavefront w is
// rec[(k * waves + w) * cap + s]: (o, tmax), (d, C.r), (C.g, C.b, pixel, -).
// A vertex step appends at most 64 rays to a queue holding at most 63.
#ifndef NGI_SHQ
#define NGI_SHQ 128
#endif
// Shadow rings of the resumable kernels on large scenes: NGI_SHQ_LONG records
// per wavefront, drained with refills once they cannot take another 64 rays
// (126 MB of rings at full occupancy).  Measured (M samples/s, sphere field /
// nave): rounds of 64 from a 128-ring 1,638 / 682; drains of 128-ring 1,621 /
// 663, 256: 1,721 / 704, 512: 1,737 / 717, 1024: 1,746 / 704.  Small scenes
// (chunked path feed) keep rounds of 64: the drain cost 4% on the Cornell box.
#ifndef NGI_SHQ_LONG
#define NGI_SHQ_LONG 512
#endif
struct ShadowQueue {
    float4* rec;
    uint32_t waves;
    uint32_t cap = NGI_SHQ;   // records per ring, a power of two >= NGI_SHQ
};
__device__ __forceinline__ float4* shq_at(const ShadowQueue& q, int k, uint32_t wave, uint32_t slot) {
    return q.rec + ((size_t)k * q.waves + wave) * q.cap + (slot & (q.cap - 1u));
}
// Appends this lane's shadow ray at `slot`.
__device__ __forceinline__ void shq_push(const ShadowQueue& q, uint32_t wave, uint32_t slot, const TraceRay& r, V3<float> C, int pixel) {
    *shq_at(q, 0, wave, slot) = make_float4(r.ox, r.oy, r.oz, r.tmax);
    *shq_at(q, 1, wave, slot) = make_float4(r.dx, r.dy, r.dz, C.x);
    *shq_at(q, 2, wave, slot) = make_float4(C.y, C.z, __int_as_float(pixel), 0.0f);
}
// FP64 shading queues its FP32 shadow rays with FP64 contributions: four
// float4s per record, (o, tmax), (d, pixel), (C.r, C.g), (C.b, -) at
// rec[(k * waves + w) * cap + s], k = 0..3.
template <typename CT> struct ShqWords { enum { value = 3 }; };
template <> struct ShqWords<double> { enum { value = 4 }; };
__device__ __forceinline__ float4 dbl2_bits(double x, double y) {
    return make_float4(__int_as_float(__double2loint(x)), __int_as_float(__double2hiint(x)), __int_as_float(__double2loint(y)),
                       __int_as_float(__double2hiint(y)));
}
__device__ __forceinline__ void shq_push(const ShadowQueue& q, uint32_t wave, uint32_t slot, const TraceRay& r, V3<double> C, int pixel) {
    *shq_at(q, 0, wave, slot) = make_float4(r.ox, r.oy, r.oz, r.tmax);
    *shq_at(q, 1, wave, slot) = make_float4(r.dx, r.dy, r.dz, __int_as_float(pixel));
    *shq_at(q, 2, wave, slot) = dbl2_bits(C.x, C.y);
    *shq_at(q, 3, wave, slot) = dbl2_bits(C.z, 0.0);
}
// The ray (tmin = EpsF), contribution and pixel of record `slot`.
template <typename CT>
__device__ __forceinline__ void shq_load(const ShadowQueue& q, uint32_t wave, uint32_t slot, TraceRay& sr, V3<CT>& C, int& pixel) {
    const float4 r0 = *shq_at(q, 0, wave, slot), r1 = *shq_at(q, 1, wave, slot), r2 = *shq_at(q, 2, wave, slot);
    sr.ox = r0.x, sr.oy = r0.y, sr.oz = r0.z, sr.tmax = r0.w;
    sr.dx = r1.x, sr.dy = r1.y, sr.dz = r1.z;
    sr.tmin = NGI_EPS_F;
    if constexpr (std::is_same<CT, double>::value) {
        const float4 r3 = *shq_at(q, 3, wave, slot);
        pixel = __float_as_int(r1.w);
        C = V3<double>(__hiloint2double(__float_as_int(r2.y), __float_as_int(r2.x)),
                       __hiloint2double(__float_as_int(r2.w), __float_as_int(r2.z)),
                       __hiloint2double(__float_as_int(r3.y), __float_as_int(r3.x)));
    } else {
        pixel = __float_as_int(r2.z);
        C = V3<float>(r1.w, r2.x, r2.y);
    }
}
// Lanes 0..n-1 trace the n oldest shadow rays from `head` and splat the
// unoccluded ones (Scene::Visible, rt.hpp:2251-2261; src/nanogi.cpp:700-706).
template <typename FilmT, int WIDTH, bool COUNT, typename CT = float>
__device__ __forceinline__ void shq_trace(const RenderArgs& a, const ShadowQueue& q, uint32_t wave, uint32_t head, uint32_t n,
                                          uint32_t* stk, int stride, uint32_t* cnt) {
    if (__lane_id() >= n) return;
    TraceRay sr;
    V3<CT> C;
    int pixel;
    shq_load<CT>(q, wave, head + __lane_id(), sr, C, pixel);
    TraceHit sh;
    TraceCounters tc{0, 0};
    const bool occluded = trace<WIDTH, true, COUNT>(a.sc, sr, sh, tc, stk, stride);
    cnt[CNT_SHADOW_RAYS]++;
    if (COUNT) cnt[CNT_SH_NODES] += tc.nodes, cnt[CNT_SH_TRIS] += tc.tris;
    if (!occluded) film_add<FilmT>(a.film, pixel, C);
}

// Traces the queued shadow rays in rounds of 64 (one ray per lane) while 64
// or more are waiting, or all of them when `all` (the wave is about to stop
// or to change scene).
template <typename FilmT, int WIDTH, bool COUNT, typename CT = float>
__device__ __forceinline__ void shq_flush(const RenderArgs& a, const ShadowQueue& q, uint32_t wave, uint32_t& head, uint32_t& count,
                                          bool all, uint32_t* stk, int stride, uint32_t* cnt) {
    while (count >= 64u || (all && count > 0u)) {
        const uint32_t n = count < 64u ? count : 64u;
        shq_trace<FilmT, WIDTH, COUNT, CT>(a, q, wave, head, n, stk, stride, cnt);
        head += n;
        count -= n;
    }
}

// Traces the queued shadow rays of a wavefront (wide BVH) with refills: each
// lane takes a queued ray, and lanes whose rays have finished take the next
// queued rays as soon as NGI_SHQ_REFILL of them are idle, so the wavefront
// waits for its slowest ray once per drain instead of after every 64.  Starts
// once the ring cannot take another 64 rays (or with any waiting, with `all`)
// and returns with the queue empty.  Visits, tests and contributions are
// those of shq_trace.
#ifndef NGI_SHQ_REFILL
#define NGI_SHQ_REFILL 16
#endif
template <typename FilmT, bool COUNT, typename CT = float>
__device__ __forceinline__ void shq_drain(const RenderArgs& a, const ShadowQueue& q, uint32_t wave, uint32_t& head, uint32_t& count,
                                          bool all, uint32_t* stk, int stride, uint32_t* cnt) {
    if (count + 64u < q.cap && !(all && count > 0)) return;
    const DevScene& sc = a.sc;
    const uint32_t lane = __lane_id();
    bool busy = false;
    RayPre r;
    float tmax = 0.0f;
    V3<CT> C;
    int pixel = 0;
    uint32_t cur = 0, mask = 0;
    int sp = 0;
    TraceCounters tc{0, 0};
    for (;;) {
        const unsigned long long idle = __ballot(!busy);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (count > 0 && (nidle >= NGI_SHQ_REFILL || nidle == 64u)) {
            const uint32_t take = nidle < count ? nidle : count;
            const uint32_t rank = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
            if (!busy && rank < take) {
                TraceRay sr;
                shq_load<CT>(q, wave, head + rank, sr, C, pixel);
                r = make_pre(sr);
                tmax = sr.tmax;
                cur = 0;
                sp = 0;
                mask = sc.num_tris ? visit8<COUNT>(sc, r, 0, NGI_EPS_F, tmax, tc) : 0u;
                busy = true;
                cnt[CNT_SHADOW_RAYS]++;
            }
            head += take;
            count -= take;
        }
        if (__ballot(busy) == 0) break;
        // One node visit or leaf per lane, after at most one pop (Scene::Visible,
        // rt.hpp:2251-2261; src/nanogi.cpp:700-706).
        if (busy && mask == 0) {
            if (sp == 0) {
                busy = false;   // nothing in the way
                film_add<FilmT>(a.film, pixel, C);
            } else {
                const uint32_t e = stk[(--sp) * stride];
                cur = e >> 8;
                mask = e & 0xffu;
            }
        }
        if (busy) {
            const uint32_t i = __builtin_ctz(mask);
            mask &= mask - 1u;
            const uint32_t slot = i ^ r.octant;
            const int32_t ref = reinterpret_cast<const int32_t*>(sc.nodes + 8 * (size_t)cur + 4)[slot];
            if (ref >= 0) {
                if (mask) stk[(sp++) * stride] = (cur << 8) | mask;
                cur = (uint32_t)ref;
                mask = visit8<COUNT>(sc, r, cur, NGI_EPS_F, tmax, tc);
            } else {
                TraceHit h;
                bool occluded = false;
                leaf_test<true, COUNT>(sc, r, ref, NGI_EPS_F, tmax, h, tc, occluded);
                if (occluded) busy = false;
            }
        }
    }
    if (COUNT) cnt[CNT_SH_NODES] += tc.nodes, cnt[CNT_SH_TRIS] += tc.tris;
}

// Surface at a hit: p = o + d t in the shading precision (rt.hpp:2197),
// geometric normal from the winding (2206), shading normal interpolated
// from vertex normals with weights (1 - u - v, u, v) computed in FP32 as
// Embree's u, v are (2209-2218), NaN -> geometric normal.
// The owning primitive's type bits come with the shading record, so a
// vertex needs the primitive's record only for what its BSDF or emission
// reads.
template <typename R>
__device__ __forceinline__ Surf<R> hit_surface(const DevScene& sc, V3<R> o, V3<R> d, const TraceHit& h, int& prim, int& type) {
    const float4* tv = sc.tris + 3 * (size_t)h.leaf;
    const float4* sh = sc.shade + 3 * (size_t)h.leaf;
    const float4 v0 = tv[0], v1 = tv[1], v2 = tv[2];
    const float4 n0 = sh[0], n1 = sh[1], n2 = sh[2];
    prim = __float_as_int(n0.w);
    type = __float_as_int(n1.w);
    Surf<R> s;
    s.p = o + d * (R)h.t;
    const V3<R> p1((R)v0.x, (R)v0.y, (R)v0.z), p2((R)v1.x, (R)v1.y, (R)v1.z), p3((R)v2.x, (R)v2.y, (R)v2.z);
    s.gn = vnorm(cross(p2 - p1, p3 - p1));
    const R w0 = (R)(1.0f - h.u - h.v);
    V3<R> sn = V3<R>((R)n0.x, (R)n0.y, (R)n0.z) * w0 + V3<R>((R)n1.x, (R)n1.y, (R)n1.z) * (R)h.u +
               V3<R>((R)n2.x, (R)n2.y, (R)n2.z) * (R)h.v;
    sn = vnorm(sn);
    if (Math<R>::isnan_(sn.x) || Math<R>::isnan_(sn.y) || Math<R>::isnan_(sn.z)) sn = s.gn;
    s.n = sn;
    s.degenerate = false;
    return s;
}
template <typename R>
__device__ __forceinline__ Surf<R> hit_surface(const DevScene& sc, V3<R> o, V3<R> d, const TraceHit& h, int& prim) {
    int type;
    return hit_surface<R>(sc, o, d, h, prim, type);
}

// Next-event estimation at one vertex (src/nanogi.cpp:656-710).  Returns the
// unoccluded contribution C (zero if nothing to connect) and fills the
// shadow ray and the film pixel it splats to.
template <typename R>
__device__ __forceinline__ V3<R> nee_contribution(const DevScene& sc, const Cam<R>& cam, const DevPrim<R>* prims, int type, int prim,
                                                  const Surf<R>& s, V3<R> wi, V3<R> T, U4 rn, int w, int h, int pixel, TraceRay& sr,
                                                  int& target) {
    target = pixel;
    if (sc.num_lights == 0) return V3<R>(R(0));
    const LightSample<R> ls = sample_light<R>(sc, unif<R>(rn.x), unif<R>(rn.y), unif<R>(rn.z));
    if (!ls.valid) return V3<R>(R(0));
    const V3<R> dv = ls.p - s.p;
    const R d2 = dot(dv, dv);
    const R len = Math<R>::sqrt_(d2);
    const V3<R> ppL = dv / len;
    V3<R> fsE;
    if (type == NGI_PRIM_E) {
        R rx = 0, ry = 0;
        const R imp = camera_importance(cam, ppL, rx, ry);
        if (imp == R(0)) return V3<R>(R(0));
        fsE = V3<R>(imp);
        target = pixel_index(rx, ry, w, h);
    } else {
        fsE = bsdf_eval_nee(prims[prim], type, s, wi, ppL);
    }
    if (!ls.degenerate && dot(ls.n, -ppL) <= R(0)) return V3<R>(R(0));   // area emitter back side
    R G = R(1);
    if (!s.degenerate) G *= Math<R>::abs_(dot(s.n, ppL));
    if (!ls.degenerate) G *= Math<R>::abs_(dot(ls.n, -ppL));
    G /= d2;
    const V3<R> C = T * fsE * ls.Le * (G * ls.inv_pdf);
    sr.ox = (float)s.p.x, sr.oy = (float)s.p.y, sr.oz = (float)s.p.z;
    sr.dx = (float)ppL.x, sr.dy = (float)ppL.y, sr.dz = (float)ppL.z;
    sr.tmin = NGI_EPS_F;
    sr.tmax = (float)len * (1.0f - NGI_EPS_F);   // Scene::Visible (rt.hpp:2251-2261)
    return C;
}

// Order in which per-pixel path ids visit the image: bands of NGI_TILE_ROWS
// rows, column by column within a band, so the 64 consecutive ids of one
// wavefront cover an 8 x 8 block (coherent camera rays) instead of a 64 x 1
// strip.  A bijection of [0, w h) for any image size; the random numbers of
// a path depend on its pixel and sample only, not on this order.
#ifndef NGI_TILE_ROWS
#define NGI_TILE_ROWS 8
#endif
__device__ __forceinline__ int pixel_of(uint32_t j, int w, int h) {
    const uint32_t r0 = j / (NGI_TILE_ROWS * (uint32_t)w) * NGI_TILE_ROWS;
    const uint32_t hb = (uint32_t)h - r0 < NGI_TILE_ROWS ? (uint32_t)h - r0 : NGI_TILE_ROWS;
    const uint32_t k = j - r0 * (uint32_t)w;
    return (int)((r0 + k % hb) * (uint32_t)w + k / hb);
}

// Per-path sample coordinates: RNG counter words and the pixel (per-pixel
// sampling) or -1 (pixel decided by the camera sample).
__device__ __forceinline__ void path_coords(const RenderArgs& a, uint64_t id, uint32_t& c0, uint32_t& c1, int& pixel) {
    if (a.sampling == NGI_SAMPLING_PER_PIXEL) {
        pixel = pixel_of((uint32_t)(id % a.npix), a.w, a.h);
        c0 = (uint32_t)pixel;
        c1 = a.sample_base + (uint32_t)(id / a.npix);
    } else {
        pixel = -1;
        c0 = (uint32_t)id;
        c1 = (uint32_t)(id >> 32) | NGI_UNIFORM_FLAG;
    }
}

// Camera vertex direction sample (rt.hpp:733-740); fs/pdf of the pinhole is
// exactly 1 (EvaluateDirection :955-978 equals EvaluateDirectionPDF
// :1189-1212), so the throughput stays 1.
template <typename R>
__device__ __forceinline__ V3<R> camera_sample(const RenderArgs& a, const Cam<R>& cam, R u0, R u1, int& pixel) {
    R rx, ry;
    if (a.sampling == NGI_SAMPLING_PER_PIXEL) {
        const int px = pixel % a.w, py = pixel / a.w;
        rx = ((R)px + u0) / (R)a.w;
        ry = ((R)py + u1) / (R)a.h;
    } else {
        rx = u0, ry = u1;
        pixel = pixel_index(rx, ry, a.w, a.h);
    }
    return camera_dir(cam, rx, ry);
}

// Path ids of a persistent wavefront, taken from the launch's counter.
// Every wavefront of the grid shares that counter.  With chunk = 0 a refill
// takes exactly the ids its idle lanes need, which keeps the lanes of a
// wavefront on paths from all over the current frontier; with chunk > 0 it
// takes at least `chunk` ids and hands out the rest on later refills, one
// atomic per chunk.  The state is wave-uniform; take() must be called by
// every lane of the wave.
struct PathFeed {
    unsigned long long nxt = 0, end = 0;   // reserved ids not handed out yet: [nxt, end)
    bool drained = false;                  // the counter has passed the path count

    __device__ __forceinline__ bool done() const { return drained && nxt >= end; }

    // Consecutive ids for the lanes set in `need` (rank order): this lane's
    // id, or ~0 when the ids ran out.  CHUNKED = false ignores `chunk`.
    template <bool CHUNKED>
    __device__ __forceinline__ unsigned long long take(unsigned long long* counter, unsigned long long count, unsigned long long need,
                                                       uint32_t chunk) {
        const uint32_t lane = __lane_id();
        const unsigned long long k = (unsigned long long)__popcll(need);
        if (!CHUNKED) {
            const int leader = __ffsll((long long)need) - 1;
            unsigned long long b = 0;
            if ((int)lane == leader) b = atomicAdd(counter, k);
            b = __shfl(b, leader);
            if (b + k >= count) drained = true;
            const unsigned long long id = b + (unsigned long long)__popcll(need & ((1ull << lane) - 1ull));
            return ((need >> lane) & 1ull) && id < count ? id : ~0ull;
        }
        const unsigned long long left = end - nxt;
        unsigned long long nb = 0, ne = 0;
        if (left < k && !drained) {
            const unsigned long long ask = k - left > chunk ? k - left : chunk;
            const int leader = __ffsll((long long)need) - 1;
            unsigned long long b = 0;
            if ((int)lane == leader) b = atomicAdd(counter, ask);
            b = __shfl(b, leader);
            if (b < count) nb = b, ne = b + ask < count ? b + ask : count;
            if (b + ask >= count) drained = true;
        }
        unsigned long long id = ~0ull;
        if ((need >> lane) & 1ull) {
            const unsigned long long r = (unsigned long long)__popcll(need & ((1ull << lane) - 1ull));
            if (r < left) id = nxt + r;
            else if (r - left < ne - nb) id = nb + (r - left);
        }
        if (k <= left) {
            nxt += k;
        } else if (ne > nb) {
            const unsigned long long used = k - left < ne - nb ? k - left : ne - nb;
            nxt = nb + used, end = ne;
        } else {
            nxt = end;
        }
        return id;
    }
};

// Per-lane path state between vertices of ProcessSample_PT/_PTDirect.
template <typename R>
struct PathState {
    V3<R> T;          // throughput
    Surf<R> s;        // current vertex
    V3<R> wi;         // direction towards the previous vertex
    int type, prim, nv, pixel;
    uint32_t c0, c1;  // RNG counter words of this path
};

// Camera vertex of path `id` (SampleEmitter(E) + SamplePosition, src/nanogi.cpp:448-480).
template <typename R>
__device__ __forceinline__ void path_start(const RenderArgs& a, uint64_t id, const Cam<R>& cam, PathState<R>& ps) {
    path_coords(a, id, ps.c0, ps.c1, ps.pixel);
    ps.T = V3<R>(R(1));
    ps.type = NGI_PRIM_E;
    ps.prim = a.sc.sensor_prim;
    ps.s.p = cam.eye;
    ps.s.degenerate = true;
    ps.s.gn = V3<R>(R(0));
    ps.s.n = V3<R>(R(0));
    ps.wi = V3<R>(R(0));
    ps.nv = 1;
}

// One iteration of the vertex loop (src/nanogi.cpp:483-606 / 645-801) in two
// halves around the extension ray.  vertex_begin: NEE (ptdirect) and the
// direction sample; returns false when the path ends here, else the
// direction wo and the roulette number urr of the vertex.  With QUEUE the NEE
// shadow ray is appended to the wavefront's shadow queue at slot qtail +
// (rank among the lanes appending) and *pushed is set, instead of being
// traced here.
template <typename R, typename FilmT, bool DIRECT, int WIDTH, bool COUNT, bool QUEUE>
__device__ __forceinline__ bool vertex_begin(const RenderArgs& a, const Cam<R>& cam, const DevPrim<R>* prims, PathState<R>& ps,
                                             uint32_t* stk, int stride, int max_nv, uint32_t* cnt, const ShadowQueue* q, uint32_t qwave,
                                             uint32_t qtail, bool* pushed, V3<R>& wo, R& urr) {
    const DevScene& sc = a.sc;
    if (ps.nv >= max_nv) return false;
    // FP32: the vertex's material, read before the light sample so that its
    // loads go out with the light's (the shadow-ray stores between NEE and
    // the direction sample would otherwise hold the sample's loads back).
    // FP64 keeps reading it where it is used (a copy spills).
    using MatRef = typename std::conditional<std::is_same<R, float>::value, const DevPrim<R>, const DevPrim<R>&>::type;
    MatRef mat = prims[ps.prim];
    if (DIRECT) {
        const U4 rn = rng4(ps.c0, ps.c1, (uint32_t)ps.nv, NGI_STREAM_NEE, a.seed, a.stream_id);
        TraceRay sr;
        int target;
        const V3<R> C = nee_contribution<R>(sc, cam, &mat, ps.type, 0, ps.s, ps.wi, ps.T, rn, a.w, a.h, ps.pixel, sr, target);
        if (QUEUE && !is_zero(C)) {
            const unsigned long long m = __ballot(1);
            shq_push(*q, qwave, qtail + (uint32_t)__popcll(m & ((1ull << __lane_id()) - 1ull)), sr, C, target);
            *pushed = true;
        } else if (!is_zero(C)) {
            TraceHit sh;
            TraceCounters tc{0, 0};
            const bool occluded = trace<WIDTH, true, COUNT>(sc, sr, sh, tc, stk, stride);
            cnt[CNT_SHADOW_RAYS]++;
            if (COUNT) cnt[CNT_SH_NODES] += tc.nodes, cnt[CNT_SH_TRIS] += tc.tris;
            if (!occluded) film_add<FilmT>(a.film, target, C);
        }
    }
    const U4 rx = rng4(ps.c0, ps.c1, (uint32_t)ps.nv, NGI_STREAM_EXT, a.seed, a.stream_id);
    const R u0 = unif<R>(rx.x), u1 = unif<R>(rx.y), uc = unif<R>(rx.z);
    urr = unif<R>(rx.w);
    if (ps.type == NGI_PRIM_E) {
        wo = camera_sample<R>(a, cam, u0, u1, ps.pixel);
    } else {
        V3<R> wgt;
        if (!bsdf_sample<R>(mat, ps.type, ps.s, ps.wi, u0, u1, uc, wo, wgt)) return false;
        ps.T = ps.T * wgt;
    }
    return true;
}

// Extension ray of a vertex: from the vertex along wo, hit interval (EpsF, inf).
template <typename R>
__device__ __forceinline__ TraceRay extension_ray(const PathState<R>& ps, V3<R> wo) {
    TraceRay er;
    er.ox = (float)ps.s.p.x, er.oy = (float)ps.s.p.y, er.oz = (float)ps.s.p.z;
    er.dx = (float)wo.x, er.dy = (float)wo.y, er.dz = (float)wo.z;
    er.tmin = NGI_EPS_F, er.tmax = FLT_MAX;
    return er;
}

// Second half: the hit of the extension ray (h.leaf = -1 for a miss),
// emission (pt), Russian roulette.  Returns false when the path ends.
template <typename R, typename FilmT, bool DIRECT>
__device__ __forceinline__ bool vertex_end(const RenderArgs& a, const DevPrim<R>* prims, PathState<R>& ps, V3<R> wo, R urr,
                                           const TraceHit& h) {
    const DevScene& sc = a.sc;
    if (h.leaf < 0) return false;
    int hp, htype;
    const Surf<R> ns = hit_surface<R>(sc, ps.s.p, wo, h, hp, htype);
    if (!DIRECT && (htype & NGI_PRIM_L)) {
        // Emission on hit (src/nanogi.cpp:568-575): area lights only.
        const DevPrim<R>& hm = prims[hp];
        const DevLight<R>* L = hm.light >= 0 ? &lights_of<R>(sc)[hm.light] : nullptr;
        if (L && L->type == NGI_LIGHT_AREA && dot(ns.n, -wo) > R(0))
            film_add<FilmT>(a.film, ps.pixel, ps.T * V3<R>(hm.Le[0], hm.Le[1], hm.Le[2]));
    }
    if (a.rr_depth >= 0 && ps.nv >= a.rr_depth) {   // src/nanogi.cpp:583-591
        const R q = rr_prob<R>(a);
        if (urr > q) return false;
        ps.T = ps.T / q;
    }
    ps.type = htype & ~(NGI_PRIM_L | NGI_PRIM_E);
    if (ps.type == 0) return false;   // [L]-only primitives absorb (rt.hpp:1146-1147)
    ps.s = ns;
    ps.prim = hp;
    ps.wi = -wo;
    ps.nv++;
    return true;
}

// The whole vertex with the extension ray traced in place.  Returns false
// when the path ends.
template <typename R, typename FilmT, bool DIRECT, int WIDTH, bool COUNT, bool QUEUE = false>
__device__ __forceinline__ bool path_vertex(const RenderArgs& a, const Cam<R>& cam, const DevPrim<R>* prims, PathState<R>& ps,
                                            uint32_t* stk, int stride, int max_nv, uint32_t* cnt, const ShadowQueue* q = nullptr,
                                            uint32_t qwave = 0, uint32_t qtail = 0, bool* pushed = nullptr) {
    V3<R> wo;
    R urr;
    if (!vertex_begin<R, FilmT, DIRECT, WIDTH, COUNT, QUEUE>(a, cam, prims, ps, stk, stride, max_nv, cnt, q, qwave, qtail, pushed, wo, urr))
        return false;
    TraceHit h;
    TraceCounters tc{0, 0};
    trace<WIDTH, false, COUNT>(a.sc, extension_ray(ps, wo), h, tc, stk, stride);
    cnt[CNT_EXT_RAYS]++;
    if (COUNT) cnt[CNT_EXT_NODES] += tc.nodes, cnt[CNT_EXT_TRIS] += tc.tris;
    return vertex_end<R, FilmT, DIRECT>(a, prims, ps, wo, urr, h);
}

// Lane state parked in LDS while a resumable traversal waits for the rest of
// its wavefront to shade (the shading code needs the registers): traversal
// state, direction and roulette number of the extension ray, one column per
// lane, S words apart.
enum { PK_CUR = 0, PK_MASK, PK_SP, PK_T, PK_U, PK_V, PK_LEAF, PK_WX, PK_WY, PK_WZ, PK_URR, PK_WORDS };
#define NGI_CU_LDS_BYTES (160u * 1024u)   // LDS per CU on gfx950
// Fewest blocks per CU (LDS-bound) a resumable kernel may run at on deep
// trees before the launchers fall back to the kernels that run every
// traversal to its end: 4, or the environment's NGI_RESUME_MIN_BLOCKS (1 to
// the kernel's register budget; read at every launch).  Measured on the
// 5.8M-triangle sphere field (wide BVH 10 levels deep, 33 KiB per block):
// ptdirect 1,229 -> 1,377 M samples/s, pt 1,970 -> 2,162 M, ltdirect
// 1,393 -> 1,723 M at 4 blocks against the fallback at 5.
#define NGI_RESUME_MIN_BLOCKS 4u
inline unsigned resume_min_blocks(unsigned budget) {
    unsigned m = NGI_RESUME_MIN_BLOCKS;
    if (const char* e = getenv("NGI_RESUME_MIN_BLOCKS")) {
        const long v = strtol(e, nullptr, 10);
        if (v >= 1 && v <= 64) m = (unsigned)v;
    }
    return budget < m ? budget : m;
}

template <int S>
__device__ __forceinline__ void park_trav(uint32_t* pk, const Trav8& t) {
    pk[PK_CUR * S] = t.cur;
    pk[PK_MASK * S] = t.mask;
    pk[PK_SP * S] = (uint32_t)t.sp;
    pk[PK_T * S] = __float_as_uint(t.h.t);
    pk[PK_U * S] = __float_as_uint(t.h.u);
    pk[PK_V * S] = __float_as_uint(t.h.v);
    pk[PK_LEAF * S] = (uint32_t)t.h.leaf;
}
template <int S>
__device__ __forceinline__ void park_dir(uint32_t* pk, V3<float> wo, float urr) {
    pk[PK_WX * S] = __float_as_uint(wo.x);
    pk[PK_WY * S] = __float_as_uint(wo.y);
    pk[PK_WZ * S] = __float_as_uint(wo.z);
    pk[PK_URR * S] = __float_as_uint(urr);
}
template <int S>
__device__ __forceinline__ TraceHit parked_hit(const uint32_t* pk) {
    TraceHit h;
    h.t = __uint_as_float(pk[PK_T * S]);
    h.u = __uint_as_float(pk[PK_U * S]);
    h.v = __uint_as_float(pk[PK_V * S]);
    h.leaf = (int32_t)pk[PK_LEAF * S];
    return h;
}
template <int S>
__device__ __forceinline__ Trav8 parked_trav(const uint32_t* pk) {
    Trav8 t;
    t.cur = pk[PK_CUR * S];
    t.mask = pk[PK_MASK * S];
    t.sp = (int)pk[PK_SP * S];
    t.h = parked_hit<S>(pk);
    return t;
}
template <int S>
__device__ __forceinline__ V3<float> parked_dir(const uint32_t* pk) {
    return V3<float>(__uint_as_float(pk[PK_WX * S]), __uint_as_float(pk[PK_WY * S]), __uint_as_float(pk[PK_WZ * S]));
}
template <int S>
__device__ __forceinline__ float parked_urr(const uint32_t* pk) { return __uint_as_float(pk[PK_URR * S]); }

// FP64 shading parks the direction and the roulette number as two words
// each: low words in rows PK_WX..PK_URR, high words in four more rows.
template <typename R> struct ParkWords { enum { value = PK_WORDS }; };
template <> struct ParkWords<double> { enum { value = PK_WORDS + 4 }; };
template <int S>
__device__ __forceinline__ void park_word2(uint32_t* pk, int lo_row, int hi_row, double x) {
    pk[lo_row * S] = (uint32_t)__double2loint(x);
    pk[hi_row * S] = (uint32_t)__double2hiint(x);
}
template <int S>
__device__ __forceinline__ double parked_word2(const uint32_t* pk, int lo_row, int hi_row) {
    return __hiloint2double((int)pk[hi_row * S], (int)pk[lo_row * S]);
}
template <int S>
__device__ __forceinline__ void park_dir(uint32_t* pk, V3<double> wo, double urr) {
    park_word2<S>(pk, PK_WX, PK_WORDS, wo.x);
    park_word2<S>(pk, PK_WY, PK_WORDS + 1, wo.y);
    park_word2<S>(pk, PK_WZ, PK_WORDS + 2, wo.z);
    park_word2<S>(pk, PK_URR, PK_WORDS + 3, urr);
}
template <int S, typename R>
__device__ __forceinline__ V3<R> parked_dir_r(const uint32_t* pk) {
    if constexpr (std::is_same<R, double>::value)
        return V3<R>(parked_word2<S>(pk, PK_WX, PK_WORDS), parked_word2<S>(pk, PK_WY, PK_WORDS + 1),
                     parked_word2<S>(pk, PK_WZ, PK_WORDS + 2));
    else
        return parked_dir<S>(pk);
}
template <int S, typename R>
__device__ __forceinline__ R parked_urr_r(const uint32_t* pk) {
    if constexpr (std::is_same<R, double>::value) return parked_word2<S>(pk, PK_URR, PK_WORDS + 3);
    else return parked_urr<S>(pk);
}

__device__ __forceinline__ int clamp_vertices(int m) { return (m < 0 || m > NGI_MAX_VERTICES) ? NGI_MAX_VERTICES : m; }

}  // namespace ngi
