// Light-path vertex loop of the light tracers: Renderer::ProcessSample_LT
// (src/nanogi.cpp:804-934) and ProcessSample_LTDirect (:936-1131).
//
// A path starts on a light chosen uniformly per light primitive
// (SampleEmitter, rt.hpp:2321-2336) at a position sampled by
// SamplePosition (rt.hpp:483-562), with throughput 1 / (pdfPL pdfL).  Each
// vertex, ltdirect first connects to the pinhole sensor; then a direction
// is sampled (the emitter's at the light vertex, the BSDF's with LE
// transport elsewhere), the extension ray is traced and Russian roulette
// is applied.  Random numbers: the light choice and position from stream 1
// at vertex 0, direction sampling and roulette from stream 0 at every
// vertex.
//
// lt's only film write is on a hit with the sensor, which adds
// We * EvaluatePosition(geom, false) (:897-903); for a pinhole that factor
// is zero (rt.hpp:590-593), so lt traces its paths and adds nothing.
#pragma once

#include "ngi_path.h"

namespace ngi {

template <typename R>
struct LightState {
    V3<R> T;          // throughput
    Surf<R> s;        // current vertex
    V3<R> wi;         // direction towards the previous vertex
    R pe;             // 1 / (the light's position pdf), see light_vertex
    int type, prim, nv, light;
    uint32_t c0, c1;  // RNG counter words of this path
};

// Light vertex of path `id`.  Returns false when the scene has no light.
template <typename R>
__device__ __forceinline__ bool light_start(const RenderArgs& a, uint64_t id, LightState<R>& ls) {
    const DevScene& sc = a.sc;
    int pixel;
    path_coords(a, id, ls.c0, ls.c1, pixel);
    if (sc.num_lights == 0) return false;
    const U4 r = rng4(ls.c0, ls.c1, 0u, NGI_STREAM_NEE, a.seed, a.stream_id);
    const R ul = unif<R>(r.x), ux = unif<R>(r.y), uy = unif<R>(r.z);
    ls.light = pick_light(sc, ul);
    const DevLight<R>& L = lights_of<R>(sc)[ls.light];
    const R n = (R)sc.num_lights;
    ls.prim = L.prim;
    ls.type = NGI_PRIM_L;
    ls.nv = 1;
    ls.wi = V3<R>(R(0));
    if (L.type == NGI_LIGHT_AREA) {
        V3<R> gn;
        area_light_point(sc, L, ux, uy, ls.s.p, gn);
        ls.s.gn = gn;
        ls.s.n = gn;
        ls.s.degenerate = false;
        ls.pe = R(1) / L.inv_area;
    } else if (L.type == NGI_LIGHT_POINT) {
        ls.s.p = V3<R>(L.pos[0], L.pos[1], L.pos[2]);
        ls.s.gn = V3<R>(R(0));
        ls.s.n = V3<R>(R(0));
        ls.s.degenerate = true;
        ls.pe = R(1);
    } else {
        // Point on the disk of radius r behind the scene (rt.hpp:549-561).
        const V3<R> d(L.dir[0], L.dir[1], L.dir[2]);
        ls.s.gn = d;
        ls.s.n = d;
        ls.s.degenerate = false;
        const Frame<R> fr = make_frame(d);
        R dx, dy;
        concentric_disk(ux, uy, dx, dy);
        const V3<R> c(L.pos[0], L.pos[1], L.pos[2]);
        ls.s.p = c - d * L.radius + (fr.s * (dx * L.radius) + fr.t * (dy * L.radius));
        ls.pe = R(1) / L.inv_area;
    }
    // EvaluatePosition(geomL, true) / pdfPL / pdfL.
    ls.T = V3<R>(n * ls.pe);
    return true;
}

// One iteration of the light-path loop in two halves around the extension
// ray.  light_begin: the sensor connection (ltdirect) and the direction
// sample; returns false when the path ends here, else the direction wo and
// the roulette number urr.  With QUEUE the sensor connection's shadow ray is
// appended to the wavefront's shadow queue (as in vertex_begin) instead of
// being traced here.
template <typename R, typename FilmT, bool DIRECT, int WIDTH, bool COUNT, bool QUEUE>
__device__ __forceinline__ bool light_begin(const RenderArgs& a, const Cam<R>& cam, const DevPrim<R>* prims, LightState<R>& ls,
                                            uint32_t* stk, int stride, int max_nv, uint32_t* cnt, const ShadowQueue* q, uint32_t qwave,
                                            uint32_t qtail, bool* pushed, V3<R>& wo, R& urr) {
    const DevScene& sc = a.sc;
    if (ls.nv >= max_nv) return false;
    // FP32: the vertex's material read up front, as in vertex_begin.
    using MatRef = typename std::conditional<std::is_same<R, float>::value, const DevPrim<R>, const DevPrim<R>&>::type;
    MatRef mat = prims[ls.prim];
    const DevLight<R>& L = lights_of<R>(sc)[ls.light];
    const V3<R> Le(L.Le[0], L.Le[1], L.Le[2]);
    if (DIRECT) {
        // Sensor connection (src/nanogi.cpp:1027-1079): the pinhole (pdfE =
        // 1, position = the eye).  The reference divides by
        // L->EvaluatePositionPDF(geomE, true) -- the light's position pdf,
        // not the sensor's (:1045) -- and multiplies by
        // L->EvaluatePosition(geomE, true) = 1; ls.pe carries the former.
        const V3<R> dv = cam.eye - ls.s.p;
        const R d2 = dot(dv, dv);
        const R len = Math<R>::sqrt_(d2);
        const V3<R> ppE = dv / len;
        V3<R> fsL;
        if (ls.type == NGI_PRIM_L) {
            // Emitter EvaluateDirection(.., LE, false) (rt.hpp:917-937).
            if (L.type == NGI_LIGHT_AREA) fsL = dot(ls.s.n, ppE) > R(0) ? Le : V3<R>(R(0));
            else if (L.type == NGI_LIGHT_POINT) fsL = Le;
            else fsL = V3<R>(R(0));
        } else {
            fsL = bsdf_eval_nee<R, true>(mat, ls.type, ls.s, ls.wi, ppE);
        }
        if (!is_zero(fsL)) {
            R rx = 0, ry = 0;
            const R We = camera_importance(cam, -ppE, rx, ry);
            if (We > R(0)) {
                R G = R(1) / d2;   // GeometryTerm (rt.hpp:2364-2374); the pinhole is degenerate
                if (!ls.s.degenerate) G *= Math<R>::abs_(dot(ls.s.n, ppE));
                const V3<R> C = ls.T * fsL * (G * We * ls.pe);
                if (!is_zero(C)) {
                    TraceRay sr;
                    sr.ox = (float)ls.s.p.x, sr.oy = (float)ls.s.p.y, sr.oz = (float)ls.s.p.z;
                    sr.dx = (float)ppE.x, sr.dy = (float)ppE.y, sr.dz = (float)ppE.z;
                    sr.tmin = NGI_EPS_F;
                    sr.tmax = (float)len * (1.0f - NGI_EPS_F);   // Scene::Visible (rt.hpp:2251-2261)
                    if (QUEUE) {
                        const unsigned long long m = __ballot(1);
                        shq_push(*q, qwave, qtail + (uint32_t)__popcll(m & ((1ull << __lane_id()) - 1ull)), sr, C,
                                 pixel_index(rx, ry, a.w, a.h));
                        *pushed = true;
                    } else {
                        TraceHit sh;
                        TraceCounters tc{0, 0};
                        const bool occluded = trace<WIDTH, true, COUNT>(sc, sr, sh, tc, stk, stride);
                        cnt[CNT_SHADOW_RAYS]++;
                        if (COUNT) cnt[CNT_SH_NODES] += tc.nodes, cnt[CNT_SH_TRIS] += tc.tris;
                        if (!occluded) film_add<FilmT>(a.film, pixel_index(rx, ry, a.w, a.h), C);
                    }
                }
            }
        }
    }
    const U4 rx = rng4(ls.c0, ls.c1, (uint32_t)ls.nv, NGI_STREAM_EXT, a.seed, a.stream_id);
    const R u0 = unif<R>(rx.x), u1 = unif<R>(rx.y), uc = unif<R>(rx.z);
    urr = unif<R>(rx.w);
    if (ls.type == NGI_PRIM_L) {
        // Emitter SampleDirection / EvaluateDirection(.., LE, true) /
        // EvaluateDirectionPDF (rt.hpp:695-716, 917-937, 1154-1171).
        if (L.type == NGI_LIGHT_AREA) {
            wo = to_world(make_frame(ls.s.n), cosine_hemisphere(u0, u1));
            if (dot(ls.s.n, wo) <= R(0)) return false;
            ls.T = ls.T * Le * R(NGI_PI);              // Le / (1/pi), projected solid angle
        } else if (L.type == NGI_LIGHT_POINT) {
            wo = uniform_sphere(u0, u1);
            ls.T = ls.T * Le * R(4.0 * NGI_PI);        // Le / (1/(4 pi))
        } else {
            wo = V3<R>(L.dir[0], L.dir[1], L.dir[2]);
            ls.T = ls.T * Le;                          // pdf 1
        }
    } else {
        V3<R> wgt;
        if (!bsdf_sample<R, true>(mat, ls.type, ls.s, ls.wi, u0, u1, uc, wo, wgt)) return false;
        ls.T = ls.T * wgt;
    }
    return true;
}

// light_end: the hit of the extension ray (h.leaf = -1 for a miss) and
// Russian roulette.  Returns false when the path ends.
template <typename R>
__device__ __forceinline__ bool light_end(const RenderArgs& a, const DevPrim<R>* prims, LightState<R>& ls, V3<R> wo, R urr, const TraceHit& h) {
    const DevScene& sc = a.sc;
    if (h.leaf < 0) return false;
    int hp, htype;
    const Surf<R> ns = hit_surface<R>(sc, ls.s.p, wo, h, hp, htype);
    if (a.rr_depth >= 0 && ls.nv >= a.rr_depth) {   // src/nanogi.cpp:910-921
        const R q = rr_prob<R>(a);
        if (urr > q) return false;
        ls.T = ls.T / q;
    }
    ls.type = htype & ~(NGI_PRIM_L | NGI_PRIM_E);
    if (ls.type == 0) return false;   // [L]-only surfaces: the next EvaluateDirection is zero
    ls.s = ns;
    ls.prim = hp;
    ls.wi = -wo;
    ls.nv++;
    return true;
}

// The whole light vertex with the extension ray traced in place.
template <typename R, typename FilmT, bool DIRECT, int WIDTH, bool COUNT, bool QUEUE = false>
__device__ __forceinline__ bool light_vertex(const RenderArgs& a, const Cam<R>& cam, const DevPrim<R>* prims, LightState<R>& ls,
                                             uint32_t* stk, int stride, int max_nv, uint32_t* cnt, const ShadowQueue* q = nullptr,
                                             uint32_t qwave = 0, uint32_t qtail = 0, bool* pushed = nullptr) {
    V3<R> wo;
    R urr;
    if (!light_begin<R, FilmT, DIRECT, WIDTH, COUNT, QUEUE>(a, cam, prims, ls, stk, stride, max_nv, cnt, q, qwave, qtail, pushed, wo, urr))
        return false;
    TraceRay er;
    er.ox = (float)ls.s.p.x, er.oy = (float)ls.s.p.y, er.oz = (float)ls.s.p.z;
    er.dx = (float)wo.x, er.dy = (float)wo.y, er.dz = (float)wo.z;
    er.tmin = NGI_EPS_F, er.tmax = FLT_MAX;
    TraceHit h;
    TraceCounters tc{0, 0};
    trace<WIDTH, false, COUNT>(a.sc, er, h, tc, stk, stride);
    cnt[CNT_EXT_RAYS]++;
    if (COUNT) cnt[CNT_EXT_NODES] += tc.nodes, cnt[CNT_EXT_TRIS] += tc.tris;
    return light_end<R>(a, prims, ls, wo, urr, h);
}

}  // namespace ngi
