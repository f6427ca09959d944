// Bidirectional path tracing (bdpt) on persistent lanes: one path sample per
// lane at a time, as Renderer::ProcessSample_BDPT (src/nanogi.cpp:1133-1189)
// with Path (include/nanogi/bdpt.hpp).
//
// A sample draws a light subpath and an eye subpath (SampleSubpath,
// bdpt.hpp:54-137) into per-lane vertex records in HBM, then evaluates every
// strategy (s, t) with n = s + t vertices: a shadow ray between the subpath
// ends (Connect, :139-191), the unweighted contribution
// alphaL(s) * c(s,t) * alphaE(t) (:210-333), the power-heuristic weight over
// all strategies that could produce the same path (:352-372) and the
// Russian-roulette selection probability (:193-208); the result goes to the
// pixel the camera end of the path projects to.  The strategies of all the
// samples of a wavefront are pooled and evaluated 64 at a time by whichever
// lane is free, reading the owning lane's records, so the connection rays and
// weights run with the whole wavefront busy although samples differ in their
// number of strategies.
//
// The weight is computed differently from the reference's
// EvaluatePowerHeuristicsMISWeightOpt, which recomputes every p_i(x) from
// scratch (EvaluatePDF, :491-528): here p_i / p_s is walked outwards from s,
// one vertex at a time, p_{i+1} / p_i = pL(x_i) / pE(x_i) with pL / pE the
// area-measure pdfs of generating vertex x_i from the light / eye side; the
// terms with c(i) = 0 are left out as in the reference (p_i = 0 there).  The
// ratios are accumulated in FP64.  Away from the connection every factor of
// the walk depends on one subpath only: after a sample's subpaths are drawn,
// each vertex record gets its generation pdf from the previous vertex, its
// pdf from the next one and whether the split before it has c != 0, so a
// strategy evaluates BSDFs only at the two vertices on either side of its
// connection (six pdf / c values) and reads the rest from the records.
//
// Random numbers: the light subpath's step k draws stream 1 at vertex k
// (step 0: light choice and position, as the light tracers), the eye
// subpath's step k stream 0 at vertex k (step 1: the camera ray, stratified
// in the pixel in per-pixel mode).
#include <algorithm>

#include "ngi_render.h"

namespace ngi {

#define NGI_BD_BLOCK 256
// Minimum waves per SIMD for the bdpt kernel's register allocation: 4
// (128 VGPRs, a little spilled) measured fastest against 3 (160 VGPRs) and
// 5: +9% on the Cornell box, +16% on the nave; with the stored MIS factors
// 4 / 3 / 5 waves give 201 / 189 / 174 M samples/s (Cornell) and 58.5 /
// 51.0 / 57.2 M (nave).
#ifndef NGI_BD_WAVES
#define NGI_BD_WAVES 4
#endif
// Strategies are pooled per path length when that takes at most this many
// quarters of the rounds of a single pool.  4 (no more rounds than one pool)
// measured best with the stored MIS factors: Cornell box 203 / 201 / 200 M
// samples/s and nave 61.5 / 58.5 / 58.4 M for 4 / 5 / 6 quarters.
#ifndef NGI_BD_POOL_QUARTERS
#define NGI_BD_POOL_QUARTERS 4u
#endif
// Vertices per subpath; the reference has no cap with -m -1 (with Russian
// roulette at 0.5 a longer subpath has probability 2^-62).
#define NGI_BDPT_MAX_VERTICES 64

// One subpath vertex.  type: NGI_PRIM_L (light end), NGI_PRIM_E (camera),
// else the surface's BSDF bits (0 on [L]-only surfaces, which absorb).
struct BVert {
    f3 p, gn, n;
    f3 alpha;   // throughput of the subpath up to and including this vertex
    int prim, type;
    bool deg;   // pinhole or point light
};

// Record of subpath vertex v_k (4 float4): (p, prim), (gn, type bits |
// 256 degenerate | 512 split flag), (n, fwd), (alpha, rev), with
//   fwd  = area pdf of generating v_k from v_{k-1} (incoming direction from
//          v_{k-2}); v_0: position pdf;
//   rev  = area pdf of generating v_k from v_{k+1} (incoming direction from
//          v_{k+2} where the subpath has it);
//   split flag = c != 0 for the split between the two vertices named below,
//          with their outer neighbours on the same subpath: light side, v_k
//          the split (v_{k-1}, v_k) (k = 0: v_0 emits towards v_1); eye side,
//          the split (v_{k+1}, v_k).
#define NGI_BD_SPLIT 512

struct BdptMem {
    float4* v;       // 4 float4 per (side, vertex, lane), [side][vertex][lane]
    uint32_t lanes;  // grid * block
    int maxv;        // vertices per subpath
    ShadowQueue q;   // wide BVH: per-wavefront rings of connection rays
};

__device__ __forceinline__ float4* brec(const BdptMem& m, int side, int k, uint32_t lane) {
    return m.v + ((size_t)(side * m.maxv + k) * m.lanes + lane) * 4;
}

__device__ __forceinline__ void bstore(const BdptMem& m, int side, int k, uint32_t lane, const BVert& v) {
    float4* q = brec(m, side, k, lane);
    q[0] = make_float4(v.p.x, v.p.y, v.p.z, __int_as_float(v.prim));
    q[1] = make_float4(v.gn.x, v.gn.y, v.gn.z, __int_as_float(v.type | (v.deg ? 256 : 0)));
    q[2] = make_float4(v.n.x, v.n.y, v.n.z, 0.0f);
    q[3] = make_float4(v.alpha.x, v.alpha.y, v.alpha.z, 0.0f);
}

// Geometry only (alpha left zero): the MIS walk never reads the throughput.
__device__ __forceinline__ BVert bload_geo(const BdptMem& m, int side, int k, uint32_t lane) {
    const float4* q = brec(m, side, k, lane);
    const float4 a = q[0], b = q[1], c = q[2];
    BVert v;
    v.p = f3(a.x, a.y, a.z);
    v.prim = __float_as_int(a.w);
    v.gn = f3(b.x, b.y, b.z);
    const int tb = __float_as_int(b.w);
    v.type = tb & 255;
    v.deg = (tb & 256) != 0;
    v.n = f3(c.x, c.y, c.z);
    v.alpha = f3(0.0f);
    return v;
}

// The strategy-independent MIS factors of vertex k (see the record layout).
struct BFac {
    float fwd, rev;
    bool split;
};

__device__ __forceinline__ BFac bload_fac(const BdptMem& m, int side, int k, uint32_t lane) {
    const float4* q = brec(m, side, k, lane);
    BFac f;
    f.split = (__float_as_int(q[1].w) & NGI_BD_SPLIT) != 0;
    f.fwd = q[2].w;
    f.rev = q[3].w;
    return f;
}

__device__ __forceinline__ void bstore_fac(const BdptMem& m, int side, int k, uint32_t lane, const BVert& v, float fwd, float rev,
                                           bool split) {
    float4* q = brec(m, side, k, lane);
    q[1] = make_float4(v.gn.x, v.gn.y, v.gn.z, __int_as_float(v.type | (v.deg ? 256 : 0) | (split ? NGI_BD_SPLIT : 0)));
    q[2] = make_float4(v.n.x, v.n.y, v.n.z, fwd);
    float4 a = q[3];
    a.w = rev;
    q[3] = a;
}

__device__ __forceinline__ f3 bload_pos(const BdptMem& m, int side, int k, uint32_t lane) {
    const float4 a = brec(m, side, k, lane)[0];
    return f3(a.x, a.y, a.z);
}

__device__ __forceinline__ BVert bload(const BdptMem& m, int side, int k, uint32_t lane) {
    const float4* q = brec(m, side, k, lane);
    const float4 a = q[0], b = q[1], c = q[2], d = q[3];
    BVert v;
    v.p = f3(a.x, a.y, a.z);
    v.prim = __float_as_int(a.w);
    v.gn = f3(b.x, b.y, b.z);
    const int tb = __float_as_int(b.w);
    v.type = tb & 255;
    v.deg = (tb & 256) != 0;
    v.n = f3(c.x, c.y, c.z);
    v.alpha = f3(d.x, d.y, d.z);
    return v;
}

struct BdptCtx {
    const DevScene* sc;
    const DevPrim<float>* prims;
    const DevLight<float>* lights;
    Cam<float> cam;
};

__device__ __forceinline__ Surf<float> surf_of(const BVert& v) {
    Surf<float> s;
    s.p = v.p, s.gn = v.gn, s.n = v.n, s.degenerate = v.deg;
    return s;
}

__device__ __forceinline__ const DevLight<float>& light_of(const BdptCtx& c, const BVert& v) { return c.lights[c.prims[v.prim].light]; }

// EvaluateDirection(geom, type, wi, wo, LE / EL, forceDegenerated = false)
// (rt.hpp:912-1148).  has_wi = false passes the reference's zero vector.
template <bool LE>
__device__ __forceinline__ f3 vert_eval(const BdptCtx& c, const BVert& v, bool has_wi, f3 wi, f3 wo) {
    if (v.type == NGI_PRIM_L) {
        const DevLight<float>& L = light_of(c, v);
        const f3 Le(L.Le[0], L.Le[1], L.Le[2]);
        if (L.type == NGI_LIGHT_AREA) return dot(v.n, wo) > 0.0f ? Le : f3(0.0f);
        if (L.type == NGI_LIGHT_POINT) return Le;
        return f3(0.0f);   // directional, not forced
    }
    if (v.type == NGI_PRIM_E) {
        float rx, ry;
        return f3(camera_importance(c.cam, wo, rx, ry));
    }
    return bsdf_eval_nee<float, LE>(c.prims[v.prim], v.type, surf_of(v), has_wi ? wi : f3(0.0f), wo);
}

// EvaluateDirectionPDF(geom, type, wi, wo, forceDegenerated = true) (rt.hpp:1150-1336).
__device__ __forceinline__ float vert_pdf(const BdptCtx& c, const BVert& v, bool has_wi, f3 wi, f3 wo) {
    if (v.type == NGI_PRIM_L) {
        const DevLight<float>& L = light_of(c, v);
        if (L.type == NGI_LIGHT_AREA) return dot(v.n, wo) > 0.0f ? (float)(1.0 / NGI_PI) : 0.0f;
        if (L.type == NGI_LIGHT_POINT) return (float)(0.25 / NGI_PI);
        return 1.0f;
    }
    if (v.type == NGI_PRIM_E) {
        float rx, ry;
        return camera_importance(c.cam, wo, rx, ry);
    }
    return bsdf_pdf<float>(c.prims[v.prim], v.type, surf_of(v), has_wi ? wi : f3(0.0f), wo);
}

// EvaluatePositionPDF(geom, true) * EvaluateEmitterPDF (rt.hpp:622-667, 2338-2352).
__device__ __forceinline__ float pos_pdf(const BdptCtx& c, const BVert& v) {
    if (v.type == NGI_PRIM_L) {
        const DevLight<float>& L = light_of(c, v);
        const float pa = L.type == NGI_LIGHT_POINT ? 1.0f : L.inv_area;
        return pa / (float)c.sc->num_lights;
    }
    return 1.0f;   // pinhole, the only sensor
}

// EvaluatePosition(geom, false) (rt.hpp:564-620): zero for the degenerate
// point light and pinhole, one otherwise.
__device__ __forceinline__ float pos_eval_unforced(const BdptCtx& c, const BVert& v) {
    if (v.type == NGI_PRIM_L) return light_of(c, v).type == NGI_LIGHT_POINT ? 0.0f : 1.0f;
    return 0.0f;
}

__device__ __forceinline__ float gterm(const BVert& a, const BVert& b) {
    f3 d = b.p - a.p;
    const float l2 = dot(d, d);
    d = d / sqrtf(l2);
    float g = 1.0f;
    if (!a.deg) g *= fabsf(dot(a.n, d));
    if (!b.deg) g *= fabsf(dot(b.n, -d));
    return g / l2;
}

__device__ __forceinline__ float rr_at(const RenderArgs& a, int step) {
    return (a.rr_depth >= 0 && step >= a.rr_depth) ? a.rr_prob_f : 1.0f;
}

// The connected path x_0 .. x_{n-1} of strategy s (light end first).
struct Conn {
    const BdptMem* m;
    uint32_t lane;
    int s, n;
    __device__ __forceinline__ BVert x(int k) const {
        BVert v = k < s ? bload(*m, 0, k, lane) : bload(*m, 1, n - 1 - k, lane);
        if (k == 0) v.type = NGI_PRIM_L;   // s = 0: the eye subpath's end on a light acts as the light (bdpt.hpp:150)
        return v;
    }
    __device__ __forceinline__ BVert xg(int k) const {   // x(k) without the throughput
        BVert v = k < s ? bload_geo(*m, 0, k, lane) : bload_geo(*m, 1, n - 1 - k, lane);
        if (k == 0) v.type = NGI_PRIM_L;
        return v;
    }
    __device__ __forceinline__ f3 xp(int k) const { return k < s ? bload_pos(*m, 0, k, lane) : bload_pos(*m, 1, n - 1 - k, lane); }
};

__device__ __forceinline__ f3 dirv(const BVert& from, const BVert& to) { return vnorm(to.p - from.p); }

// c_i(x): EvaluateCst (bdpt.hpp:210-244) for the split after x_{i-1}.
__device__ __forceinline__ f3 cst(const BdptCtx& c, const Conn& q, int i) {
    const int n = q.n;
    if (i == n) return f3(0.0f);   // t = 0: a pinhole is never hit (its EvaluatePosition(.., false) is zero)
    if (i == 0) {
        const BVert v0 = q.x(0), v1 = q.x(1);
        return vert_eval<false>(c, v0, false, f3(0.0f), dirv(v0, v1)) * pos_eval_unforced(c, v0);
    }
    const BVert vL = q.x(i - 1), vE = q.x(i);
    f3 wiL(0.0f), wiE(0.0f);
    if (i - 2 >= 0) wiL = dirv(vL, q.x(i - 2));
    if (i + 1 < n) wiE = dirv(vE, q.x(i + 1));
    const f3 fsL = vert_eval<true>(c, vL, i - 2 >= 0, wiL, dirv(vL, vE));
    if (is_zero(fsL)) return f3(0.0f);
    const f3 fsE = vert_eval<false>(c, vE, i + 1 < n, wiE, dirv(vE, vL));
    return fsL * fsE * gterm(vL, vE);
}

// Area-measure pdf of generating x_k from the light side (pL) or from the eye side (pE).
__device__ __forceinline__ double pdf_light_side(const BdptCtx& c, const Conn& q, int k) {
    if (k == 0) return (double)pos_pdf(c, q.x(0));
    const BVert a = q.x(k - 1), b = q.x(k);
    f3 wi(0.0f);
    if (k - 2 >= 0) wi = dirv(a, q.x(k - 2));
    return (double)vert_pdf(c, a, k - 2 >= 0, wi, dirv(a, b)) * (double)gterm(a, b);
}

__device__ __forceinline__ double pdf_eye_side(const BdptCtx& c, const Conn& q, int k) {
    const int n = q.n;
    if (k == n - 1) return (double)pos_pdf(c, q.x(n - 1));
    const BVert a = q.x(k + 1), b = q.x(k);
    f3 wi(0.0f);
    if (k + 2 <= n - 1) wi = dirv(a, q.x(k + 2));
    return (double)vert_pdf(c, a, k + 2 <= n - 1, wi, dirv(a, b)) * (double)gterm(a, b);
}

// c_i of cst() from the vertices around the split (vm2 = x_{i-2}.p when i >= 2,
// vp1 = x_{i+1}.p when i + 1 < n).
__device__ __forceinline__ f3 cst_at(const BdptCtx& c, int i, int n, f3 vm2, const BVert& vL, const BVert& vE, f3 vp1) {
    const f3 wiL = i - 2 >= 0 ? vnorm(vm2 - vL.p) : f3(0.0f);
    const f3 wiE = i + 1 < n ? vnorm(vp1 - vE.p) : f3(0.0f);
    const f3 fsL = vert_eval<true>(c, vL, i - 2 >= 0, wiL, dirv(vL, vE));
    if (is_zero(fsL)) return f3(0.0f);
    const f3 fsE = vert_eval<false>(c, vE, i + 1 < n, wiE, dirv(vE, vL));
    return fsL * fsE * gterm(vL, vE);
}

// p * G as stored in the records (one rounding of the FP64 product).
__device__ __forceinline__ float pdf_area(float pdf, float g) { return (float)((double)pdf * (double)g); }

// The split flag of cst_at() (c != 0 for the split vL | vE; vm2 = the
// outer neighbour of vL, vp1 = that of vE, each where it exists).
__device__ __forceinline__ bool split_nonzero(const BdptCtx& c, bool has_m2, f3 vm2, const BVert& vL, const BVert& vE, bool has_p1,
                                              f3 vp1) {
    const f3 wiL = has_m2 ? vnorm(vm2 - vL.p) : f3(0.0f);
    const f3 wiE = has_p1 ? vnorm(vp1 - vE.p) : f3(0.0f);
    const f3 fsL = vert_eval<true>(c, vL, has_m2, wiL, dirv(vL, vE));
    if (is_zero(fsL)) return false;
    const f3 fsE = vert_eval<false>(c, vE, has_p1, wiE, dirv(vE, vL));
    return !is_zero(fsL * fsE * gterm(vL, vE));
}

// Fills the MIS factors (fwd, rev, split flag) of the m vertices of one of
// this lane's subpaths.  A window of three vertices A = v_{k-1}, B = v_k,
// C = v_{k+1} and the positions v_{k-2}, v_{k+2} slides along the subpath.
__device__ __forceinline__ void subpath_factors(const BdptCtx& c, const BdptMem& mem, int side, int m, uint32_t gl) {
    if (m <= 0) return;
    BVert A, B = bload_geo(mem, side, 0, gl), C;
    f3 pA2(0.0f), pC2(0.0f);
    if (m > 1) C = bload_geo(mem, side, 1, gl);
    for (int k = 0; k < m; ++k) {
        if (k + 2 <= m - 1) pC2 = bload_pos(mem, side, k + 2, gl);
        // fwd: v_k generated from A = v_{k-1} (incoming from v_{k-2}).
        const float fwd = k == 0 ? pos_pdf(c, B)
                                 : pdf_area(vert_pdf(c, A, k >= 2, k >= 2 ? vnorm(pA2 - A.p) : f3(0.0f), dirv(A, B)), gterm(A, B));
        // rev: v_k generated from C = v_{k+1} (incoming from v_{k+2}).
        const float rev = k + 1 <= m - 1 ? pdf_area(vert_pdf(c, C, k + 2 <= m - 1, k + 2 <= m - 1 ? vnorm(pC2 - C.p) : f3(0.0f), dirv(C, B)),
                                                    gterm(C, B))
                                         : 0.0f;
        bool split = false;
        if (side == 0) {
            // Split (v_{k-1}, v_k) with v_{k+1} beyond v_k; k = 0: emission of v_0 towards v_1.
            if (k == 0) split = m > 1 && !is_zero(vert_eval<false>(c, B, false, f3(0.0f), dirv(B, C)) * pos_eval_unforced(c, B));
            else if (k + 1 <= m - 1) split = split_nonzero(c, k >= 2, pA2, A, B, true, C.p);
        } else if (k + 2 <= m - 1) {
            // Split (v_{k+1}, v_k): v_{k+1} on the light side of the path, v_{k+2} beyond it, v_{k-1} beyond v_k.
            split = split_nonzero(c, true, pC2, C, B, k >= 1, A.p);
        }
        bstore_fac(mem, side, k, gl, B, fwd, rev, split);
        pA2 = A.p;
        A = B;
        B = C;
        if (k + 2 <= m - 1) C = bload_geo(mem, side, k + 2, gl);
    }
}

// mis_weight() below from the records' factors: only the factors next to
// the connection (pE of x_{s-1} and x_{s-2}, pL of x_s and x_{s+1}, c_{s-1}
// and c_{s+1}) depend on both subpaths and are evaluated here, with the same
// arguments, hence the same values, as the definitions above.  Light
// vertices y_k = x_k, eye vertices z_j = x_{n-1-j}.
__device__ __forceinline__ double mis_weight_fac(const BdptCtx& c, const BdptMem& mem, uint32_t owner, int s, int n) {
    const int t = n - s;
    Conn q{&mem, owner, s, n};
    // The six connection-dependent factors, with at most three vertices
    // loaded at a time: pe1 = pE(x_{s-1}), pe2 = pE(x_{s-2}), pl0 = pL(x_s),
    // pl1 = pL(x_{s+1}), c_{s-1} and c_{s+1}.
    double pe1 = 0.0, pe2 = 0.0, pl0 = 0.0, pl1 = 0.0;
    bool cm1 = false, cp1 = false;
    {
        const BVert E0 = q.xg(s);                               // x_s = z_{t-1} (a light when s = 0)
        BVert L1;
        f3 pL2(0.0f);
        if (s >= 1) L1 = q.xg(s - 1);                           // x_{s-1} = y_{s-1}
        if (s >= 2) pL2 = q.xp(s - 2);
        if (s + 1 <= n - 1) {
            const BVert E1 = q.xg(s + 1);                       // x_{s+1} = z_{t-2}
            if (s >= 1) pe1 = (double)vert_pdf(c, E0, true, dirv(E0, E1), dirv(E0, L1)) * (double)gterm(E0, L1);
            if (s + 2 <= n - 1) pl1 = (double)vert_pdf(c, E0, s >= 1, s >= 1 ? dirv(E0, L1) : f3(0.0f), dirv(E0, E1)) * (double)gterm(E0, E1);
            const bool hp = s + 2 <= n - 1;
            const f3 pp = hp ? q.xp(s + 2) : f3(0.0f);
            cp1 = split_nonzero(c, s >= 1, s >= 1 ? L1.p : f3(0.0f), E0, E1, hp, pp);
        } else if (s >= 1) {
            pe1 = (double)vert_pdf(c, E0, false, f3(0.0f), dirv(E0, L1)) * (double)gterm(E0, L1);
        }
        if (s == 0) {
            pl0 = (double)pos_pdf(c, E0);
        } else {
            pl0 = (double)vert_pdf(c, L1, s >= 2, s >= 2 ? vnorm(pL2 - L1.p) : f3(0.0f), dirv(L1, E0)) * (double)gterm(L1, E0);
            if (s == 1) {
                cm1 = !is_zero(vert_eval<false>(c, L1, false, f3(0.0f), dirv(L1, E0)) * pos_eval_unforced(c, L1));
            } else {
                const BVert L2 = q.xg(s - 2);                   // x_{s-2} = y_{s-2}
                pe2 = (double)vert_pdf(c, L1, true, dirv(L1, E0), dirv(L1, L2)) * (double)gterm(L1, L2);
                const f3 pm3 = s >= 3 ? q.xp(s - 3) : f3(0.0f);
                cm1 = split_nonzero(c, s >= 3, pm3, L2, L1, true, E0.p);
            }
        }
    }
    double inv = 1.0, r = 1.0;
    // Towards the light, i = s - 1 .. 0: r *= pE(x_i) / pL(x_i).
    for (int i = s - 1; i >= 0; --i) {
        const BFac f = bload_fac(mem, 0, i, owner);
        const double pe = i == s - 1 ? pe1 : (i == s - 2 ? pe2 : (double)f.rev);
        if (pe == 0.0) break;
        const double pl = (double)f.fwd;   // pL(y_i); y_0: its position pdf
        r *= pe / pl;
        if (i == s - 1 ? cm1 : f.split) inv += r * r;
    }
    r = 1.0;
    // Towards the eye, i = s + 1 .. n - 1: r *= pL(x_{i-1}) / pE(x_{i-1}).
    for (int i = s + 1; i < n; ++i) {
        const BFac f = bload_fac(mem, 1, n - i, owner);         // x_{i-1} = z_{n-i}
        const double pl = i == s + 1 ? pl0 : (i == s + 2 ? pl1 : (double)f.rev);
        if (pl == 0.0) break;
        const double pe = (double)f.fwd;
        r *= pl / pe;
        if (i == s + 1 ? cp1 : bload_fac(mem, 1, n - 1 - i, owner).split) inv += r * r;
    }
    const double w = 1.0 / inv;
    return (w == w && inv < 1e300) ? w : 0.0;
}

// Power-heuristic weight of strategy s: 1 / sum_i (p_i / p_s)^2 over the
// strategies with c_i != 0.  Returns 0 when p_s itself vanishes.  (The
// definition; the kernel evaluates it through mis_weight_fac.)
__device__ __forceinline__ double mis_weight(const BdptCtx& c, const Conn& q) {
    const int s = q.s, n = q.n;
    double inv = 1.0, r = 1.0;
    for (int i = s - 1; i >= 0; --i) {   // p_i / p_{i+1} = pE(x_i) / pL(x_i)
        const double pe = pdf_eye_side(c, q, i);
        if (pe == 0.0) break;            // this and every smaller i have p_i = 0
        r *= pe / pdf_light_side(c, q, i);
        if (!is_zero(cst(c, q, i))) inv += r * r;
    }
    r = 1.0;
    for (int i = s + 1; i < n; ++i) {    // p_i / p_{i-1} = pL(x_{i-1}) / pE(x_{i-1}); c_n = 0 for a pinhole
        const double pl = pdf_light_side(c, q, i - 1);
        if (pl == 0.0) break;
        r *= pl / pdf_eye_side(c, q, i - 1);
        if (!is_zero(cst(c, q, i))) inv += r * r;
    }
    const double w = 1.0 / inv;
    return (w == w && inv < 1e300) ? w : 0.0;   // a strategy with p_s = 0 (never sampled) gets no weight
}

// Strategies (s, t = n - s >= 1) of a sample with nL light and nE eye
// vertices and at most nmax vertices per path, in the reference's order.
__device__ __forceinline__ int strategies_of_length(int n, int nL, int nE) {
    const int lo = n - nE > 0 ? n - nE : 0, hi = nL < n - 1 ? nL : n - 1;
    return hi >= lo ? hi - lo + 1 : 0;
}

__device__ __forceinline__ uint32_t strategy_count(int nL, int nE, int nmax) {
    uint32_t k = 0;
    for (int n = 2; n <= nmax; ++n) k += (uint32_t)strategies_of_length(n, nL, nE);
    return k;
}

// The k-th strategy: its path length n and light-vertex count s.
__device__ __forceinline__ void strategy_at(uint32_t k, int nL, int nE, int nmax, int& n, int& s) {
    for (n = 2; n <= nmax; ++n) {
        const uint32_t c = (uint32_t)strategies_of_length(n, nL, nE);
        if (k < c) {
            s = (n - nE > 0 ? n - nE : 0) + (int)k;
            return;
        }
        k -= c;
    }
    n = s = 0;   // not reached for k below strategy_count
}

// One strategy of the sample whose records live at `owner`: the shadow ray
// between the subpath ends (Connect, bdpt.hpp:139-191), the unweighted
// contribution alphaL * c(s,t) * alphaE, the power-heuristic weight and the
// selection probability; the result is splatted where the camera end of the
// path projects.  With QUEUE the connection ray goes to the wavefront's ring
// with its weighted contribution (at slot qtail + rank among the lanes
// appending; *pushed is set) and is traced when the ring is drained: a round
// of strategies then no longer waits for its slowest connection ray, and the
// weight of an occluded connection is computed where the round's other lanes
// compute theirs.
template <typename FilmT, int WIDTH, bool COUNT, bool QUEUE>
__device__ __forceinline__ void eval_strategy(const RenderArgs& a, const BdptCtx& c, const BdptMem& mem, uint32_t owner, int s, int n,
                                              uint32_t* my_stk, uint32_t* cnt, uint32_t qwave, uint32_t qtail, bool* pushed) {
    const int t = n - s;
    if (n < 2 || t < 1) return;
    Conn q{&mem, owner, s, n};
    TraceRay sr;
    if (s == 0) {
        const BVert e = bload(mem, 1, t - 1, owner);
        if (!(c.prims[e.prim].type & NGI_PRIM_L)) return;
    } else {
        const BVert l = bload(mem, 0, s - 1, owner), e = bload(mem, 1, t - 1, owner);
        const f3 dv = e.p - l.p;
        const float len = sqrtf(dot(dv, dv));
        const f3 d = dv / len;
        sr = TraceRay{l.p.x, l.p.y, l.p.z, d.x, d.y, d.z, NGI_EPS_F, len * (1.0f - NGI_EPS_F)};   // Scene::Visible
        if (!QUEUE) {
            TraceHit sh;
            TraceCounters tc{0, 0};
            const bool occluded = trace<WIDTH, true, COUNT>(*c.sc, sr, sh, tc, my_stk, NGI_BD_BLOCK);
            cnt[CNT_SHADOW_RAYS]++;
            if (COUNT) cnt[CNT_SH_NODES] += tc.nodes, cnt[CNT_SH_TRIS] += tc.tris;
            if (occluded) return;
        }
    }
    const f3 cs = cst(c, q, s);
    if (is_zero(cs)) return;
    const f3 aL = s > 0 ? bload(mem, 0, s - 1, owner).alpha : f3(1.0f);
    const f3 aE = bload(mem, 1, t - 1, owner).alpha;
    const f3 Cu = aL * cs * aE;
    if (is_zero(Cu)) return;
    const double w = mis_weight_fac(c, mem, owner, s, n);
    if (w == 0.0) return;
    float sel = 1.0f;   // SelectionProb (bdpt.hpp:193-208)
    for (int i = 1; i < s - 1; ++i) sel *= rr_at(a, i);
    for (int i = t - 2; i >= 1; --i) sel *= rr_at(a, i);
    const BVert cam_v = q.x(n - 1), prev = q.x(n - 2);
    float rx = 0.0f, ry = 0.0f;
    camera_importance(c.cam, dirv(cam_v, prev), rx, ry);
    const f3 C = Cu * (float)(w / (double)sel);
    if (QUEUE && s > 0) {
        if (is_zero(C)) return;
        const unsigned long long m = __ballot(1);
        shq_push(mem.q, qwave, qtail + (uint32_t)__popcll(m & ((1ull << __lane_id()) - 1ull)), sr, C, pixel_index(rx, ry, a.w, a.h));
        *pushed = true;
        return;
    }
    film_add<FilmT>(a.film, pixel_index(rx, ry, a.w, a.h), C);
}

// Subpath building blocks (bdpt.hpp:54-137): the first vertex of a side
// (0: light, LE; 1: eye, EL), a direction sample from the current vertex,
// and the vertex at a hit.
__device__ __forceinline__ uint32_t bd_stream(int side) { return side == 0 ? NGI_STREAM_NEE : NGI_STREAM_EXT; }

// The first vertex of `side`; false for the light side of a scene without lights.
__device__ __forceinline__ bool bd_first(const RenderArgs& a, const BdptCtx& c, int side, uint32_t c0, uint32_t c1, BVert& pv) {
    const DevScene& sc = *c.sc;
    if (side == 0) {
        if (sc.num_lights == 0) return false;
        const U4 r0 = rng4(c0, c1, 0u, bd_stream(0), a.seed, a.stream_id);
        const int li = pick_light(sc, u24f(r0.x));
        const DevLight<float>& L = c.lights[li];
        pv.prim = L.prim;
        pv.type = NGI_PRIM_L;
        if (L.type == NGI_LIGHT_AREA) {
            area_light_point(sc, L, u24f(r0.y), u24f(r0.z), pv.p, pv.gn);
            pv.n = pv.gn;
            pv.deg = false;
        } else if (L.type == NGI_LIGHT_POINT) {
            pv.p = f3(L.pos[0], L.pos[1], L.pos[2]);
            pv.gn = pv.n = f3(0.0f);
            pv.deg = true;
        } else {
            const f3 d(L.dir[0], L.dir[1], L.dir[2]);
            const Frame<float> fr = make_frame(d);
            float dx, dy;
            concentric_disk(u24f(r0.y), u24f(r0.z), dx, dy);
            pv.p = f3(L.pos[0], L.pos[1], L.pos[2]) - d * L.radius + (fr.s * (dx * L.radius) + fr.t * (dy * L.radius));
            pv.gn = pv.n = d;
            pv.deg = false;
        }
        pv.alpha = f3(1.0f / pos_pdf(c, pv));   // EvaluatePosition(.., true) = 1
    } else {
        pv.prim = sc.sensor_prim;
        pv.type = NGI_PRIM_E;
        pv.p = c.cam.eye;
        pv.gn = pv.n = f3(0.0f);
        pv.deg = true;
        pv.alpha = f3(1.0f);
    }
    return true;
}

// Direction sample at vertex `step - 1` of `side` (pv, with ppv_p the
// vertex before it); false when the subpath ends there.
__device__ __forceinline__ bool bd_direction(const RenderArgs& a, const BdptCtx& c, int side, int step, uint32_t c0, uint32_t c1, const BVert& pv,
                                             f3 ppv_p, int& pixel, f3& wo, f3& wgt, float& urr) {
    const U4 r = rng4(c0, c1, (uint32_t)step, bd_stream(side), a.seed, a.stream_id);
    const float u0 = u24f(r.x), u1 = u24f(r.y), uc = u24f(r.z);
    urr = u24f(r.w);
    const f3 wi = step >= 2 ? vnorm(ppv_p - pv.p) : f3(0.0f);
    if (pv.type == NGI_PRIM_L) {
        const DevLight<float>& L = light_of(c, pv);
        const f3 Le(L.Le[0], L.Le[1], L.Le[2]);
        if (L.type == NGI_LIGHT_AREA) {
            wo = to_world(make_frame(pv.n), cosine_hemisphere(u0, u1));
            if (dot(pv.n, wo) <= 0.0f) return false;
            wgt = Le * (float)NGI_PI;
        } else if (L.type == NGI_LIGHT_POINT) {
            wo = uniform_sphere(u0, u1);
            wgt = Le * (float)(4.0 * NGI_PI);
        } else {
            wo = f3(L.dir[0], L.dir[1], L.dir[2]);
            wgt = Le;
        }
        return true;
    }
    if (pv.type == NGI_PRIM_E) {
        wo = camera_sample<float>(a, c.cam, u0, u1, pixel);
        wgt = f3(1.0f);   // pinhole fs / pdf = 1
        return true;
    }
    return side == 0 ? bsdf_sample<float, true>(c.prims[pv.prim], pv.type, surf_of(pv), wi, u0, u1, uc, wo, wgt)
                     : bsdf_sample<float, false>(c.prims[pv.prim], pv.type, surf_of(pv), wi, u0, u1, uc, wo, wgt);
}

// The vertex where the ray from pv along wo hit (h.leaf >= 0); alpha is the
// subpath throughput up to it.
__device__ __forceinline__ BVert bd_vertex(const BdptCtx& c, const BVert& pv, f3 wo, const TraceHit& h, f3 alpha) {
    int hp, htype;
    const Surf<float> ns = hit_surface<float>(*c.sc, pv.p, wo, h, hp, htype);
    BVert v;
    v.p = ns.p, v.gn = ns.gn, v.n = ns.n, v.deg = false;
    v.prim = hp;
    v.type = htype & ~(NGI_PRIM_L | NGI_PRIM_E);
    v.alpha = alpha;
    return v;
}

// QUEUE (wide BVH, scenes of at least NGI_CHUNKED_FEED_TRIS triangles): the
// subpaths' traversals are resumable and the connection rays go through
// drained rings.  Measured (M samples/s, Cornell box / nave at 16 spp):
// traced in place 198 / 64.1, connection rays queued 191 / 79.8.  The keep
// of the subpath traversals matters little (nave at 8 spp: 4 / 8 / 12 / 16
// lanes: 82.1 / 82.1-82.4 / 82.6 / 82.4 M).
#ifndef NGI_BD_KEEP
#define NGI_BD_KEEP 8
#endif
template <typename FilmT, int WIDTH, bool COUNT, bool QUEUE>
__global__ void __launch_bounds__(NGI_BD_BLOCK, NGI_BD_WAVES) bdpt_kernel(RenderArgs a, unsigned long long* next, BdptMem mem) {
    extern __shared__ uint32_t stk[];   // a.sc.stack_entries x block (wide BVH only)
    const DevScene& sc = a.sc;
    uint32_t cnt[CNT_N] = {0, 0, 0, 0, 0, 0, 0};
    BdptCtx c;
    c.sc = &sc;
    c.prims = prims_of<float>(sc);
    c.lights = lights_of<float>(sc);
    c.cam = load_cam<float>(sc.cam);
    const uint32_t gl = blockIdx.x * NGI_BD_BLOCK + threadIdx.x;
    const uint32_t lane = __lane_id();
    uint32_t* my_stk = stk + threadIdx.x;
    const int maxv = mem.maxv;
    const uint32_t wave = blockIdx.x * (NGI_BD_BLOCK / 64) + (threadIdx.x >> 6);
    uint32_t qhead = 0, qcount = 0;   // wave-uniform: oldest slot, queued connection rays
    for (;;) {
        // One path id per active lane (one atomic per wavefront).
        const unsigned long long act = __ballot(true);
        const int leader = __ffsll((long long)act) - 1;
        unsigned long long base = 0;
        if ((int)lane == leader) base = atomicAdd(next, (unsigned long long)__popcll(act));
        base = __shfl(base, leader);
        const unsigned long long local = base + (unsigned long long)__popcll(act & ((1ull << lane) - 1ull));
        // Lanes without a path id left stay in the loop with no strategies, so
        // every lane of the wavefront takes part in the strategy rounds.
        const bool have = local < a.path_count;
        if (__ballot(have) == 0) {
            if (QUEUE) shq_drain<FilmT, COUNT>(a, mem.q, wave, qhead, qcount, true, my_stk, NGI_BD_BLOCK, cnt);
            break;
        }
        int nsub[2] = {0, 0};
        if (QUEUE) {
            // Subpaths with resumable traversals: every lane steps through its
            // light subpath, then its eye subpath; a wavefront stops
            // traversing once at most NGI_BD_KEEP lanes are inside the tree,
            // and the others store their vertices and start their next rays.
            uint32_t c0 = 0, c1 = 0;
            int pixel = 0;
            if (have) {
                path_coords(a, a.path_begin + local, c0, c1, pixel);
                cnt[CNT_SAMPLES]++;
            }
            uint32_t* pk = stk + (size_t)sc.stack_entries * NGI_BD_BLOCK + threadIdx.x;
            bool done = !have, busy = false, traced = false;
            int side = -1, step = 0, nv = 0;
            BVert pv;
            f3 ppv_p(0.0f);
            for (;;) {
                if (!done && !busy) {
                    bool side_over = side < 0;
                    if (traced) {   // the ray from vertex step - 1 has finished
                        traced = false;
                        const TraceHit h = parked_hit<NGI_BD_BLOCK>(pk);
                        if (h.leaf < 0) {
                            side_over = true;
                        } else {
                            const BVert v = bd_vertex(c, pv, parked_dir<NGI_BD_BLOCK>(pk), h, pv.alpha);
                            bstore(mem, side, step, gl, v);
                            nv = step + 1;
                            const float q = rr_at(a, step);
                            if (q < 1.0f && parked_urr<NGI_BD_BLOCK>(pk) > q) {
                                side_over = true;   // the vertex is kept (bdpt.hpp:118-123)
                            } else {
                                ppv_p = pv.p;
                                pv = v;
                                ++step;
                            }
                        }
                    }
                    for (;;) {   // close a finished side, open the next, sample a direction
                        if (side_over) {
                            if (side >= 0) nsub[side] = nv;
                            if (++side > 1) {
                                done = true;
                                break;
                            }
                            side_over = false;
                            nv = 0;
                            if (!bd_first(a, c, side, c0, c1, pv)) {
                                side_over = true;
                                continue;
                            }
                            bstore(mem, side, 0, gl, pv);
                            ppv_p = f3(0.0f);
                            nv = 1;
                            step = 1;
                        }
                        f3 wo, wgt;
                        float urr;
                        if (step >= maxv || !bd_direction(a, c, side, step, c0, c1, pv, ppv_p, pixel, wo, wgt, urr)) {
                            side_over = true;
                            continue;
                        }
                        pv.alpha = pv.alpha * wgt;   // pv's record is already stored
                        cnt[CNT_EXT_RAYS]++;
                        Trav8 tv;
                        trav8_begin(FLT_MAX, tv);
                        park_trav<NGI_BD_BLOCK>(pk, tv);
                        park_dir<NGI_BD_BLOCK>(pk, wo, urr);
                        busy = sc.num_tris != 0;
                        traced = true;
                        break;
                    }
                }
                const uint32_t working = (uint32_t)__popcll(__ballot(!done));
                if (working == 0) break;
                if (__ballot(busy)) {
                    Trav8 tv = parked_trav<NGI_BD_BLOCK>(pk);
                    const f3 d = parked_dir<NGI_BD_BLOCK>(pk);
                    const RayPre r = make_pre(TraceRay{pv.p.x, pv.p.y, pv.p.z, d.x, d.y, d.z, NGI_EPS_F, FLT_MAX});
                    // With few subpaths left, their traversals run to the end.
                    const int keep = __builtin_amdgcn_readfirstlane(working > 2u * NGI_BD_KEEP ? NGI_BD_KEEP : 0);
                    TraceCounters tc{0, 0};
                    trav8_run<COUNT>(sc, r, NGI_EPS_F, tv, busy, keep, tc, my_stk, NGI_BD_BLOCK);
                    if (COUNT) cnt[CNT_EXT_NODES] += tc.nodes, cnt[CNT_EXT_TRIS] += tc.tris;
                    park_trav<NGI_BD_BLOCK>(pk, tv);
                }
            }
            if (have) {
                subpath_factors(c, mem, 0, nsub[0], gl);
                subpath_factors(c, mem, 1, nsub[1], gl);
            }
        } else if (have) {
            // Small scenes: the subpaths traced vertex by vertex, written out
            // here rather than through bd_first / bd_direction / bd_vertex,
            // whose register allocation in this loop cost 6% on the Cornell box.
            uint32_t c0, c1;
            int pixel;
            path_coords(a, a.path_begin + local, c0, c1, pixel);
            cnt[CNT_SAMPLES]++;
            // ---- subpaths ----
            for (int side = 0; side < 2; ++side) {   // 0: light (LE), 1: eye (EL)
                const uint32_t stream = side == 0 ? NGI_STREAM_NEE : NGI_STREAM_EXT;
                BVert pv;
                f3 ppv_p(0.0f);
                if (side == 0) {
                    if (sc.num_lights == 0) continue;
                    const U4 r0 = rng4(c0, c1, 0u, stream, a.seed, a.stream_id);
                    const int li = pick_light(sc, u24f(r0.x));
                    const DevLight<float>& L = c.lights[li];
                    pv.prim = L.prim;
                    pv.type = NGI_PRIM_L;
                    if (L.type == NGI_LIGHT_AREA) {
                        area_light_point(sc, L, u24f(r0.y), u24f(r0.z), pv.p, pv.gn);
                        pv.n = pv.gn;
                        pv.deg = false;
                    } else if (L.type == NGI_LIGHT_POINT) {
                        pv.p = f3(L.pos[0], L.pos[1], L.pos[2]);
                        pv.gn = pv.n = f3(0.0f);
                        pv.deg = true;
                    } else {
                        const f3 d(L.dir[0], L.dir[1], L.dir[2]);
                        const Frame<float> fr = make_frame(d);
                        float dx, dy;
                        concentric_disk(u24f(r0.y), u24f(r0.z), dx, dy);
                        pv.p = f3(L.pos[0], L.pos[1], L.pos[2]) - d * L.radius + (fr.s * (dx * L.radius) + fr.t * (dy * L.radius));
                        pv.gn = pv.n = d;
                        pv.deg = false;
                    }
                    pv.alpha = f3(1.0f / pos_pdf(c, pv));   // EvaluatePosition(.., true) = 1
                } else {
                    pv.prim = sc.sensor_prim;
                    pv.type = NGI_PRIM_E;
                    pv.p = c.cam.eye;
                    pv.gn = pv.n = f3(0.0f);
                    pv.deg = true;
                    pv.alpha = f3(1.0f);
                }
                bstore(mem, side, 0, gl, pv);
                int nv = 1;
                for (int step = 1; step < maxv; ++step) {
                    const U4 r = rng4(c0, c1, (uint32_t)step, stream, a.seed, a.stream_id);
                    const float u0 = u24f(r.x), u1 = u24f(r.y), uc = u24f(r.z), urr = u24f(r.w);
                    const bool has_wi = step >= 2;
                    const f3 wi = has_wi ? vnorm(ppv_p - pv.p) : f3(0.0f);
                    f3 wo, wgt;
                    if (pv.type == NGI_PRIM_L) {
                        const DevLight<float>& L = light_of(c, pv);
                        const f3 Le(L.Le[0], L.Le[1], L.Le[2]);
                        if (L.type == NGI_LIGHT_AREA) {
                            wo = to_world(make_frame(pv.n), cosine_hemisphere(u0, u1));
                            if (dot(pv.n, wo) <= 0.0f) break;
                            wgt = Le * (float)NGI_PI;
                        } else if (L.type == NGI_LIGHT_POINT) {
                            wo = uniform_sphere(u0, u1);
                            wgt = Le * (float)(4.0 * NGI_PI);
                        } else {
                            wo = f3(L.dir[0], L.dir[1], L.dir[2]);
                            wgt = Le;
                        }
                    } else if (pv.type == NGI_PRIM_E) {
                        wo = camera_sample<float>(a, c.cam, u0, u1, pixel);
                        wgt = f3(1.0f);   // pinhole fs / pdf = 1
                    } else {
                        const bool ok = side == 0 ? bsdf_sample<float, true>(c.prims[pv.prim], pv.type, surf_of(pv), wi, u0, u1, uc, wo, wgt)
                                                  : bsdf_sample<float, false>(c.prims[pv.prim], pv.type, surf_of(pv), wi, u0, u1, uc, wo, wgt);
                        if (!ok) break;
                    }
                    TraceRay er{pv.p.x, pv.p.y, pv.p.z, wo.x, wo.y, wo.z, NGI_EPS_F, FLT_MAX};
                    TraceHit h;
                    TraceCounters tc{0, 0};
                    const bool hit = trace<WIDTH, false, COUNT>(sc, er, h, tc, my_stk, NGI_BD_BLOCK);
                    cnt[CNT_EXT_RAYS]++;
                    if (COUNT) cnt[CNT_EXT_NODES] += tc.nodes, cnt[CNT_EXT_TRIS] += tc.tris;
                    if (!hit) break;
                    int hp, htype;
                    const Surf<float> ns = hit_surface<float>(sc, pv.p, wo, h, hp, htype);
                    BVert v;
                    v.p = ns.p, v.gn = ns.gn, v.n = ns.n, v.deg = false;
                    v.prim = hp;
                    v.type = htype & ~(NGI_PRIM_L | NGI_PRIM_E);
                    v.alpha = pv.alpha * wgt;
                    bstore(mem, side, step, gl, v);
                    nv = step + 1;
                    const float q = rr_at(a, step);
                    if (q < 1.0f && urr > q) break;   // the vertex is kept (bdpt.hpp:118-123)
                    ppv_p = pv.p;
                    pv = v;
                }
                nsub[side] = nv;
            }
            subpath_factors(c, mem, 0, nsub[0], gl);
            subpath_factors(c, mem, 1, nsub[1], gl);
        }
        // Every lane reads the other lanes' vertex records below.
        __threadfence_block();
        // ---- strategies ----
        // The strategies (s, t) of all the wavefront's samples are pooled and
        // evaluated 64 at a time: lane j of a round takes pool entry base + j,
        // which belongs to lane `owner` (whose records it reads) as its k-th
        // strategy in the reference's order (n ascending, then s).  A round
        // keeps every lane on a connection regardless of how many strategies
        // each sample has.  Pooling per path length n gives rounds whose MIS
        // walks all take n steps (Cornell box 131 -> 158 M samples/s); with
        // Russian roulette the per-length pools are small, so a wave pools by
        // length only when that costs no more rounds than one pool.
        const int nmax = (a.max_nv >= 0 && a.max_nv < nsub[0] + nsub[1]) ? a.max_nv : nsub[0] + nsub[1];
        int wmax = have ? nmax : 0;   // longest path length of the wave
        for (int off = 32; off > 0; off >>= 1) wmax = max(wmax, __shfl_xor(wmax, off));
        auto wave_sum = [](uint32_t v) {
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            return v;
        };
        const uint32_t Kall = have ? strategy_count(nsub[0], nsub[1], nmax) : 0u;
        const uint32_t rounds_all = (wave_sum(Kall) + 63u) / 64u;
        uint32_t rounds_n = 0;
        for (int n = 2; n <= wmax; ++n)
            rounds_n += (wave_sum((have && n <= nmax) ? (uint32_t)strategies_of_length(n, nsub[0], nsub[1]) : 0u) + 63u) / 64u;
        const bool by_length = 4u * rounds_n <= NGI_BD_POOL_QUARTERS * rounds_all;
        for (int n = by_length ? 2 : 0; n <= (by_length ? wmax : 0); ++n) {   // n = 0: the single pool
            const uint32_t K = !have ? 0u
                               : (n == 0 ? Kall : (n <= nmax ? (uint32_t)strategies_of_length(n, nsub[0], nsub[1]) : 0u));
            uint32_t incl = K;   // inclusive prefix sum over the lanes
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t y = __shfl_up(incl, off);
                if ((int)lane >= off) incl += y;
            }
            const uint32_t total = __shfl(incl, 63);
            for (uint32_t base = 0; base < total; base += 64) {
                const uint32_t g = base + lane < total ? base + lane : total - 1;
                int owner = 0;   // the first lane whose prefix sum exceeds g
                for (int step = 32; step > 0; step >>= 1)
                    if (__shfl(incl, owner + step - 1) <= g) owner += step;
                const uint32_t k = g - (__shfl(incl, owner) - __shfl(K, owner));
                const int oL = __shfl(nsub[0], owner), oE = __shfl(nsub[1], owner), om = __shfl(nmax, owner);
                bool pushed = false;
                if (base + lane < total) {
                    int sn = n, ss = 0;
                    if (n == 0) strategy_at(k, oL, oE, om, sn, ss);
                    else ss = (n - oE > 0 ? n - oE : 0) + (int)k;
                    eval_strategy<FilmT, WIDTH, COUNT, QUEUE>(a, c, mem, gl - lane + (uint32_t)owner, ss, sn, my_stk, cnt, wave, qhead + qcount,
                                                              &pushed);
                }
                if (QUEUE) {
                    qcount += (uint32_t)__popcll(__ballot(pushed));   // at most q.cap - 65 + 64
                    shq_drain<FilmT, COUNT>(a, mem.q, wave, qhead, qcount, false, my_stk, NGI_BD_BLOCK, cnt);
                }
            }
        }
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

template <typename FilmT, int WIDTH, bool COUNT>
static int launch_bd(const RenderArgs& a, unsigned long long* next, LaneMem mem, Lanes ln, hipStream_t st) {
    const bool queue = WIDTH == 8 && a.sc.num_tris >= NGI_CHUNKED_FEED_TRIS;
    // The resumable subpaths park their lane state next to the stack.
    const size_t lds = WIDTH == 8 ? ((size_t)a.sc.stack_entries + (queue ? PK_WORDS : 0)) * NGI_BD_BLOCK * 4 : 0;
    auto kern = queue ? bdpt_kernel<FilmT, WIDTH, COUNT, WIDTH == 8> : bdpt_kernel<FilmT, WIDTH, COUNT, false>;
    unsigned grid = 0;
    int s = persistent_grid(reinterpret_cast<const void*>(kern), NGI_BD_BLOCK, lds, a.path_count, ln, &grid);
    if (s || !grid) return s;
    BdptMem m;
    m.lanes = grid * NGI_BD_BLOCK;
    m.maxv = (a.max_nv < 0 || a.max_nv > NGI_BDPT_MAX_VERTICES) ? NGI_BDPT_MAX_VERTICES : a.max_nv;
    // Every sample stores its subpaths' first vertices (the kernel has no
    // strategy for limits of one vertex or none, but still writes them).
    m.maxv = std::max(m.maxv, 1);
    const size_t bytes = (size_t)2 * m.maxv * m.lanes * 4 * sizeof(float4);
    // Wide BVH: connection-ray rings after the records.
    m.q.waves = grid * (NGI_BD_BLOCK / 64);
    m.q.cap = NGI_SHQ_LONG;
    const size_t qbytes = queue ? (size_t)3 * m.q.waves * m.q.cap * sizeof(float4) : 0;
    if ((s = mem.reserve(bytes + qbytes))) return s;
    m.v = reinterpret_cast<float4*>(*mem.ptr);
    if (!m.v) return fail(NGI_ERR_DEVICE, "bdpt: no subpath records");
    m.q.rec = queue ? m.v + bytes / sizeof(float4) : nullptr;
    kern<<<grid, NGI_BD_BLOCK, lds, st>>>(a, next, m);
    NGI_HIP_CHECK(hipGetLastError());
    return NGI_OK;
}

int launch_bdpt(const RenderArgs& a, unsigned long long* next, LaneMem mem, bool film64, int width, bool count, hipStream_t st, Lanes ln) {
    if (!film64) return width == 8 ? launch_bd<float, 8, false>(a, next, mem, ln, st) : launch_bd<float, 2, false>(a, next, mem, ln, st);
    if (count) return width == 8 ? launch_bd<double, 8, true>(a, next, mem, ln, st) : launch_bd<double, 2, true>(a, next, mem, ln, st);
    return width == 8 ? launch_bd<double, 8, false>(a, next, mem, ln, st) : launch_bd<double, 2, false>(a, next, mem, ln, st);
}

}  // namespace ngi
