// FP32 closest-hit and any-hit BVH traversal (device).
//
// Replaces rtcIntersect inside Scene::Intersect / Scene::Visible
// (include/nanogi/rt.hpp:2162-2261): two-sided triangles, hit interval
// (tmin, tmax) exclusive, barycentrics (u, v) = weights of vertices 1 and 2.
//
// Triangle test: watertight ray/triangle intersection (Woop, Benthin, Wald,
// "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013): shear to ray
// space, scaled edge functions with a double-precision recomputation when
// an edge function is exactly zero.  Box test: slabs with the far distance
// scaled up by 1 + 2*gamma(3) (Ize, "Robust BVH Ray Traversal", JCGT 2013)
// so traversal is conservative with respect to the triangle test.
#pragma once

#include "ngi_dev.h"

namespace ngi {

struct TraceRay {
    float ox, oy, oz;
    float dx, dy, dz;
    float tmin, tmax;
};

struct TraceHit {
    float t, u, v;
    int32_t leaf;   // leaf-order triangle index, -1 = miss
};

struct TraceCounters {
    uint32_t nodes;
    uint32_t tris;
};

// Per-ray constants for the watertight test and the slab test.
struct RayPre {
    float ox, oy, oz;
    float ix, iy, iz;        // 1 / d (guarded against zero components)
    float sx, sy, sz;        // shear constants
    int kx, ky, kz;
    float dx, dy, dz;
    uint32_t octant;         // bit k set if d[k] < 0
};

__device__ __forceinline__ float guard_dir(float d) {
    return fabsf(d) < 1e-20f ? copysignf(1e-20f, d) : d;
}

__device__ __forceinline__ float pick(float a, float b, float c, int k) { return k == 0 ? a : (k == 1 ? b : c); }

__device__ __forceinline__ RayPre make_pre(const TraceRay& r) {
    RayPre p;
    p.ox = r.ox, p.oy = r.oy, p.oz = r.oz;
    p.dx = r.dx, p.dy = r.dy, p.dz = r.dz;
    p.ix = 1.0f / guard_dir(r.dx);
    p.iy = 1.0f / guard_dir(r.dy);
    p.iz = 1.0f / guard_dir(r.dz);
    const float ax = fabsf(r.dx), ay = fabsf(r.dy), az = fabsf(r.dz);
    int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    // Component selection by value (no dynamically indexed private array).
    const float dz = pick(r.dx, r.dy, r.dz, kz);
    if (dz < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    p.kx = kx, p.ky = ky, p.kz = kz;
    p.sx = pick(r.dx, r.dy, r.dz, kx) / dz;
    p.sy = pick(r.dx, r.dy, r.dz, ky) / dz;
    p.sz = 1.0f / dz;
    p.octant = (r.dx < 0.0f ? 1u : 0u) | (r.dy < 0.0f ? 2u : 0u) | (r.dz < 0.0f ? 4u : 0u);
    return p;
}

// Watertight test; on a hit inside (tmin, tmax) writes t, u, v.
__device__ __forceinline__ bool tri_test(const RayPre& r, float4 a4, float4 b4, float4 c4, float tmin, float tmax, float& t,
                                         float& u, float& v) {
#pragma clang fp contract(off)
    const float Ax0 = a4.x - r.ox, Ay0 = a4.y - r.oy, Az0 = a4.z - r.oz;
    const float Bx0 = b4.x - r.ox, By0 = b4.y - r.oy, Bz0 = b4.z - r.oz;
    const float Cx0 = c4.x - r.ox, Cy0 = c4.y - r.oy, Cz0 = c4.z - r.oz;
    const float Akx = pick(Ax0, Ay0, Az0, r.kx), Aky = pick(Ax0, Ay0, Az0, r.ky), Akz = pick(Ax0, Ay0, Az0, r.kz);
    const float Bkx = pick(Bx0, By0, Bz0, r.kx), Bky = pick(Bx0, By0, Bz0, r.ky), Bkz = pick(Bx0, By0, Bz0, r.kz);
    const float Ckx = pick(Cx0, Cy0, Cz0, r.kx), Cky = pick(Cx0, Cy0, Cz0, r.ky), Ckz = pick(Cx0, Cy0, Cz0, r.kz);
    const float Ax = Akx - r.sx * Akz, Ay = Aky - r.sy * Akz;
    const float Bx = Bkx - r.sx * Bkz, By = Bky - r.sy * Bkz;
    const float Cx = Ckx - r.sx * Ckz, Cy = Cky - r.sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = r.sz * Akz, Bz = r.sz * Bkz, Cz = r.sz * Ckz;
    const float T = U * Az + V * Bz + W * Cz;
    // Sign-corrected range test: tmin < T/det < tmax.
    const float sT = det < 0.0f ? -T : T;
    const float sd = fabsf(det);
    if (sT <= tmin * sd || sT >= tmax * sd) return false;
    const float rd = 1.0f / det;
    t = T * rd;
    if (!(t > tmin && t < tmax)) return false;
    u = V * rd;
    v = W * rd;
    return true;
}

#ifndef NGI_PAIR_TRIS
#define NGI_PAIR_TRIS 1      // fetch leaf triangles two at a time
#endif

#define NGI_SLAB_SCALE 1.00000036f   // 1 + 2*gamma(3), gamma(n) = n*eps/(1-n*eps), eps = 2^-24

template <bool ANY, bool COUNT>
__device__ __forceinline__ void leaf_test(const DevScene& sc, const RayPre& r, int32_t ref, float tmin, float& tmax, TraceHit& h,
                                          TraceCounters& cnt, bool& done) {
    const uint32_t v = (uint32_t)(~ref);
    const uint32_t first = v >> 4, count = (v & 15u) + 1u;
#if !NGI_PAIR_TRIS
    for (uint32_t k = 0; k < count; ++k) {
        const uint32_t leaf = first + k;
        const float4 a = sc.tris[3 * leaf], b = sc.tris[3 * leaf + 1], c = sc.tris[3 * leaf + 2];
        if (COUNT) cnt.tris++;
        float t, u, w;
        if (tri_test(r, a, b, c, tmin, tmax, t, u, w)) {
            tmax = t;
            h.t = t, h.u = u, h.v = w, h.leaf = (int32_t)leaf;
            if (ANY) {
                done = true;
                return;
            }
        }
    }
    return;
#endif
    // Triangles are fetched two at a time so their loads overlap.
    for (uint32_t k = 0; k < count; k += 2) {
        const uint32_t l0 = first + k;
        const bool two = k + 1 < count;
        const uint32_t l1 = two ? l0 + 1 : l0;
        const float4 a0 = sc.tris[3 * l0], b0 = sc.tris[3 * l0 + 1], c0 = sc.tris[3 * l0 + 2];
        const float4 a1 = sc.tris[3 * l1], b1 = sc.tris[3 * l1 + 1], c1 = sc.tris[3 * l1 + 2];
        if (COUNT) cnt.tris += two ? 2u : 1u;
        float t, u, w;
        if (tri_test(r, a0, b0, c0, tmin, tmax, t, u, w)) {
            tmax = t;
            h.t = t, h.u = u, h.v = w, h.leaf = (int32_t)l0;
            if (ANY) {
                done = true;
                return;
            }
        }
        if (two && tri_test(r, a1, b1, c1, tmin, tmax, t, u, w)) {
            tmax = t;
            h.t = t, h.u = u, h.v = w, h.leaf = (int32_t)l1;
            if (ANY) {
                done = true;
                return;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Binary BVH, private stack (depth is bounded by the builder).
// ---------------------------------------------------------------------------
#define NGI_BVH2_STACK 64

template <bool ANY, bool COUNT>
__device__ bool trace_bvh2(const DevScene& sc, const TraceRay& ray, TraceHit& h, TraceCounters& cnt) {
    const RayPre r = make_pre(ray);
    const float tmin = ray.tmin;
    float tmax = ray.tmax;
    h.leaf = -1;
    int32_t stack[NGI_BVH2_STACK];
    int sp = 0;
    int32_t node = 0;
    bool done = false;
    for (;;) {
        if (node >= 0) {
            const float4* n = sc.nodes + 4 * (size_t)node;
            const float4 n0 = n[0], n1 = n[1], n2 = n[2], n3 = n[3];
            if (COUNT) cnt.nodes++;
            float x0 = (n0.x - r.ox) * r.ix, x1 = (n0.y - r.ox) * r.ix;
            float y0 = (n0.z - r.oy) * r.iy, y1 = (n0.w - r.oy) * r.iy;
            float z0 = (n2.x - r.oz) * r.iz, z1 = (n2.y - r.oz) * r.iz;
            const float l0 = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
            const float f0 = fminf(fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1)) * NGI_SLAB_SCALE, tmax);
            x0 = (n1.x - r.ox) * r.ix, x1 = (n1.y - r.ox) * r.ix;
            y0 = (n1.z - r.oy) * r.iy, y1 = (n1.w - r.oy) * r.iy;
            z0 = (n2.z - r.oz) * r.iz, z1 = (n2.w - r.oz) * r.iz;
            const float l1 = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
            const float f1 = fminf(fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1)) * NGI_SLAB_SCALE, tmax);
            const bool hit0 = l0 <= f0;
            const bool hit1 = l1 <= f1;
            const int32_t ch0 = __float_as_int(n3.x), ch1 = __float_as_int(n3.y);
            if (hit0 && hit1) {
                int32_t near = ch0, far = ch1;
                if (l1 < l0) {
                    near = ch1;
                    far = ch0;
                }
                stack[sp++] = far;
                node = near;
            } else if (hit0) {
                node = ch0;
            } else if (hit1) {
                node = ch1;
            } else {
                if (sp == 0) break;
                node = stack[--sp];
            }
        } else {
            leaf_test<ANY, COUNT>(sc, r, node, tmin, tmax, h, cnt, done);
            if (ANY && done) return true;
            if (sp == 0) break;
            node = stack[--sp];
        }
    }
    return h.leaf >= 0;
}

// ---------------------------------------------------------------------------
// 8-wide BVH.  Child slots are assigned at build time by the octant of the
// child's centroid relative to the parent's; a ray visits hit children in
// slot order XOR its direction octant (near-to-far along each axis).
// The stack holds one entry per tree level: (node << 8) | pending mask.
// Stack entries live in LDS, one column per lane (conflict-free).
// ---------------------------------------------------------------------------
#define NGI_BVH8_STACK 32

__device__ __forceinline__ float exp2_bits(uint32_t biased) { return __uint_as_float(biased << 23); }

// Box test of the eight children of `node`: a mask of the hit children in
// processing order (bit i = slot i ^ octant).
//
// The near and far planes of every child are chosen once per node from the
// sign of the inverse direction (for 1/d >= 0 the low bound is near): with
// t(q) = fma(q, a, b) and a >= 0, q_lo <= q_hi gives t(q_lo) <= t(q_hi)
// because rounding is monotone, so this equals min/max of the two planes
// bit for bit.  The hit bits are built in slot order with constant shifts,
// masked by the valid slots once, and permuted to processing order at the end.
template <bool COUNT>
__device__ __forceinline__ uint32_t visit8(const DevScene& sc, const RayPre& r, uint32_t node, float tmin, float tmax,
                                           TraceCounters& cnt) {
    const float4* n = sc.nodes + 8 * (size_t)node;
    const float4 h0 = n[0];
    const uint4 q0 = __builtin_bit_cast(uint4, n[1]);  // qlox[8], qloy[8]
    const uint4 q1 = __builtin_bit_cast(uint4, n[2]);  // qloz[8], qhix[8]
    const uint4 q2 = __builtin_bit_cast(uint4, n[3]);  // qhiy[8], qhiz[8]
    if (COUNT) cnt.nodes++;
    const uint32_t meta = __float_as_uint(h0.w);
    const float scx = exp2_bits(meta & 0xffu), scy = exp2_bits((meta >> 8) & 0xffu), scz = exp2_bits((meta >> 16) & 0xffu);
    // t(q) = (origin + q * scale - o) * invd = q * (scale * invd) + (origin - o) * invd
    const float ax = scx * r.ix, bx = (h0.x - r.ox) * r.ix;
    const float ay = scy * r.iy, by = (h0.y - r.oy) * r.iy;
    const float az = scz * r.iz, bz = (h0.z - r.oz) * r.iz;
    const bool fx = r.ix < 0.0f, fy = r.iy < 0.0f, fz = r.iz < 0.0f;
    // Near / far plane words for children 0-3 (.x) and 4-7 (.y).
    const uint2 nx = fx ? make_uint2(q1.z, q1.w) : make_uint2(q0.x, q0.y), ex = fx ? make_uint2(q0.x, q0.y) : make_uint2(q1.z, q1.w);
    const uint2 ny = fy ? make_uint2(q2.x, q2.y) : make_uint2(q0.z, q0.w), ey = fy ? make_uint2(q0.z, q0.w) : make_uint2(q2.x, q2.y);
    const uint2 nz = fz ? make_uint2(q2.z, q2.w) : make_uint2(q1.x, q1.y), ez = fz ? make_uint2(q1.x, q1.y) : make_uint2(q2.z, q2.w);
    uint32_t hits = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int sh = (k & 3) * 8;
        const uint32_t wnx = k < 4 ? nx.x : nx.y, wny = k < 4 ? ny.x : ny.y, wnz = k < 4 ? nz.x : nz.y;
        const uint32_t wex = k < 4 ? ex.x : ex.y, wey = k < 4 ? ey.x : ey.y, wez = k < 4 ? ez.x : ez.y;
        const float tnx = (float)((wnx >> sh) & 0xffu) * ax + bx, tfx = (float)((wex >> sh) & 0xffu) * ax + bx;
        const float tny = (float)((wny >> sh) & 0xffu) * ay + by, tfy = (float)((wey >> sh) & 0xffu) * ay + by;
        const float tnz = (float)((wnz >> sh) & 0xffu) * az + bz, tfz = (float)((wez >> sh) & 0xffu) * az + bz;
        const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tmin));
        const float tf = fminf(fminf(tfx, tfy), tfz) * NGI_SLAB_SCALE;
        if (tn <= fminf(tf, tmax)) hits |= 1u << k;
    }
    hits &= meta >> 24;   // valid slots
    // Slot order -> processing order: swap bit groups selected by the octant.
    const uint32_t o = r.octant;
    if (o & 1u) hits = ((hits & 0x55u) << 1) | ((hits >> 1) & 0x55u);
    if (o & 2u) hits = ((hits & 0x33u) << 2) | ((hits >> 2) & 0x33u);
    if (o & 4u) hits = ((hits & 0x0fu) << 4) | ((hits >> 4) & 0x0fu);
    return hits;
}

template <bool ANY, bool COUNT>
__device__ bool trace_bvh8(const DevScene& sc, const TraceRay& ray, TraceHit& h, TraceCounters& cnt, uint32_t* lds_stack,
                           int lane_stride) {
    const RayPre r = make_pre(ray);
    const float tmin = ray.tmin;
    float tmax = ray.tmax;
    h.leaf = -1;
    int sp = 0;
    uint32_t cur = 0;
    uint32_t mask = visit8<COUNT>(sc, r, 0, tmin, tmax, cnt);
    bool done = false;
    for (;;) {
        if (mask == 0) {
            if (sp == 0) break;
            const uint32_t e = lds_stack[(--sp) * lane_stride];
            cur = e >> 8;
            mask = e & 0xffu;
            continue;
        }
        const uint32_t i = __builtin_ctz(mask);
        mask &= mask - 1u;
        const uint32_t slot = i ^ r.octant;
        const int32_t ref = reinterpret_cast<const int32_t*>(sc.nodes + 8 * (size_t)cur + 4)[slot];
        if (ref >= 0) {
            if (mask) lds_stack[(sp++) * lane_stride] = (cur << 8) | mask;
            cur = (uint32_t)ref;
            mask = visit8<COUNT>(sc, r, cur, tmin, tmax, cnt);
        } else {
            leaf_test<ANY, COUNT>(sc, r, ref, tmin, tmax, h, cnt, done);
            if (ANY && done) return true;
        }
    }
    return h.leaf >= 0;
}

// Resumable closest-hit traversal of the wide BVH: the same visiting order
// and tests as trace_bvh8<false>, with the state kept by the caller so that a
// wavefront can stop traversing while some of its lanes are still inside the
// tree and resume them later (their stacks stay in LDS meanwhile).  The hit
// distance doubles as the interval's upper end: h.t starts at tmax.
struct Trav8 {
    uint32_t cur, mask;
    int sp;
    TraceHit h;
};

// A traversal about to visit the root (the root is visited by the first
// trav8_run, with the other lanes' next nodes).
#define NGI_TRAV8_FRESH 0x100u
__device__ __forceinline__ void trav8_begin(float tmax, Trav8& t) {
    t.h.t = tmax;
    t.h.u = t.h.v = 0.0f;
    t.h.leaf = -1;
    t.sp = 0;
    t.cur = 0;
    t.mask = NGI_TRAV8_FRESH;
}

// Advances the traversals of the lanes with `busy` set, one node visit or
// leaf per lane and iteration (a lane whose current node has no pending
// children first pops its stack: every stack entry holds at least one), until
// at most `keep` lanes of the wavefront are still busy (at least one
// iteration; keep = 0 finishes them all).  A lane clears `busy` when its
// traversal ends; t.h then holds its closest hit or t.h.leaf = -1.
template <bool COUNT>
__device__ __forceinline__ void trav8_run(const DevScene& sc, const RayPre& r, float tmin, Trav8& t, bool& busy, int keep,
                                          TraceCounters& cnt, uint32_t* lds_stack, int lane_stride) {
    if (busy && t.mask == NGI_TRAV8_FRESH) t.mask = visit8<COUNT>(sc, r, 0, tmin, t.h.t, cnt);
    do {
        if (busy && t.mask == 0) {
            if (t.sp == 0) {
                busy = false;
            } else {
                const uint32_t e = lds_stack[(--t.sp) * lane_stride];
                t.cur = e >> 8;
                t.mask = e & 0xffu;
            }
        }
        if (busy) {
            const uint32_t i = __builtin_ctz(t.mask);
            t.mask &= t.mask - 1u;
            const uint32_t slot = i ^ r.octant;
            const int32_t ref = reinterpret_cast<const int32_t*>(sc.nodes + 8 * (size_t)t.cur + 4)[slot];
            if (ref >= 0) {
                if (t.mask) lds_stack[(t.sp++) * lane_stride] = (t.cur << 8) | t.mask;
                t.cur = (uint32_t)ref;
                t.mask = visit8<COUNT>(sc, r, t.cur, tmin, t.h.t, cnt);
            } else {
                bool done = false;
                leaf_test<false, COUNT>(sc, r, ref, tmin, t.h.t, t.h, cnt, done);
            }
        }
    } while (__popcll(__ballot(busy)) > (unsigned)keep);
}

}  // namespace ngi
