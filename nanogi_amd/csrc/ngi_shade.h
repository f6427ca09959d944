// Device BSDF / emitter / sensor evaluation, templated on the shading
// precision (float or double).
//
// Restates the Primitive member functions of include/nanogi/rt.hpp for the
// EL transport direction of pt/ptdirect and the LE direction of lt/ltdirect:
//   OrthonormalBasis :55-59, CosineSampleHemisphere :87-109,
//   UniformSampleTriangle :129-133, PixelIndex :135-140,
//   SamplePosition :483-592, SampleDirection :692-910,
//   EvaluateDirection :912-1148 (shading-normal correction :994-1005),
//   EvaluateDirectionPDF :1150-1336, RasterPosition :1344-1399,
//   Beckmann D :1407-1414, shadow-masking :1423-1431,
//   conductor Fresnel :1433-1442, dielectric Fresnel :1450-1466,
//   GeometryTerm :2364-2374.
// Sampling returns the ratio fs/pdf directly; it is the reference's
// EvaluateDirection / EvaluateDirectionPDF quotient with common factors
// cancelled (1/pi for D, the Beckmann D for G, the branch probability for
// S), and is zero exactly when the reference's fs is zero.
#pragma once

#include "ngi_dev.h"
#include "ngi_guide.h"

namespace ngi {

template <typename R> struct Math;
template <> struct Math<float> {
    static __device__ __forceinline__ float sqrt_(float x) { return sqrtf(x); }
    static __device__ __forceinline__ float log_(float x) { return logf(x); }
    static __device__ __forceinline__ float exp_(float x) { return expf(x); }
    static __device__ __forceinline__ void sincos_(float x, float* s, float* c) { sincosf(x, s, c); }
    static __device__ __forceinline__ float abs_(float x) { return fabsf(x); }
    static __device__ __forceinline__ bool isnan_(float x) { return isnan(x); }
};
template <> struct Math<double> {
    static __device__ __forceinline__ double sqrt_(double x) { return sqrt(x); }
    static __device__ __forceinline__ double log_(double x) { return log(x); }
    static __device__ __forceinline__ double exp_(double x) { return exp(x); }
    static __device__ __forceinline__ void sincos_(double x, double* s, double* c) { sincos(x, s, c); }
    static __device__ __forceinline__ double abs_(double x) { return fabs(x); }
    static __device__ __forceinline__ bool isnan_(double x) { return isnan(x); }
};

#define NGI_PI 3.14159265358979323846

template <typename R>
struct Frame {
    V3<R> s, t, n;   // dpdu, dpdv, shading normal (ToWorld columns)
};

// OrthonormalBasis (rt.hpp:55-59) with a = n: dpdv first, then dpdu.
template <typename R>
__device__ __forceinline__ Frame<R> make_frame(V3<R> n) {
    typedef Math<R> M;
    Frame<R> f;
    f.n = n;
    V3<R> c = M::abs_(n.x) > M::abs_(n.y) ? V3<R>(n.z, R(0), -n.x) : V3<R>(R(0), n.z, -n.y);
    f.t = c / M::sqrt_(dot(c, c));
    V3<R> b = cross(f.t, n);
    f.s = b / M::sqrt_(dot(b, b));
    return f;
}

template <typename R>
__device__ __forceinline__ V3<R> to_local(const Frame<R>& f, V3<R> w) { return V3<R>(dot(f.s, w), dot(f.t, w), dot(f.n, w)); }
template <typename R>
__device__ __forceinline__ V3<R> to_world(const Frame<R>& f, V3<R> l) { return f.s * l.x + f.t * l.y + f.n * l.z; }

template <typename R>
__device__ __forceinline__ V3<R> vnorm(V3<R> v) { return v / Math<R>::sqrt_(dot(v, v)); }

// Shirley-Chiu concentric map of [0,1)^2 onto the unit disk (rt.hpp:87-103).
template <typename R>
__device__ __forceinline__ void concentric_disk(R u0, R u1, R& sx, R& sy) {
    const R vx = R(2) * u0 - R(1), vy = R(2) * u1 - R(1);
    sx = 0, sy = 0;
    if (vx == R(0) && vy == R(0)) return;
    R r, th;
    const R q = R(NGI_PI * 0.25);
    if (vx > -vy) {
        if (vx > vy) r = vx, th = q * (vy / vx);
        else r = vy, th = q * (R(2) - vx / vy);
    } else {
        if (vx < vy) r = -vx, th = q * (R(4) + vy / vx);
        else r = -vy, th = q * (R(6) - vx / vy);
    }
    R sn, cs;
    Math<R>::sincos_(th, &sn, &cs);
    sx = r * cs, sy = r * sn;
}

// Malley's projection of the concentric disk (rt.hpp:105-109).
template <typename R>
__device__ __forceinline__ V3<R> cosine_hemisphere(R u0, R u1) {
    R sx, sy;
    concentric_disk(u0, u1, sx, sy);
    const R z2 = R(1) - sx * sx - sy * sy;
    return V3<R>(sx, sy, Math<R>::sqrt_(z2 > R(0) ? z2 : R(0)));
}

// UniformSampleSphere (rt.hpp:116-122).
template <typename R>
__device__ __forceinline__ V3<R> uniform_sphere(R u0, R u1) {
    const R z = R(1) - R(2) * u0;
    const R r2 = R(1) - z * z;
    const R r = Math<R>::sqrt_(r2 > R(0) ? r2 : R(0));
    R sp, cp;
    Math<R>::sincos_(R(2.0 * NGI_PI) * u1, &sp, &cp);
    return V3<R>(r * cp, r * sp, z);
}

// A path vertex.  Only the shading normal is kept between vertices; the
// tangent frame (ComputeTangentSpace, rt.hpp:295-300) is rebuilt from it
// where a BSDF needs local coordinates, which keeps six fewer registers live
// across every ray traversal.
template <typename R>
struct Surf {
    V3<R> p, gn;
    V3<R> n;           // shading normal
    bool degenerate;   // pinhole / point light
};

// Shading-normal correction (rt.hpp:994-1005): zero when a direction's
// geometric and shading cosines disagree in sign; 1 for EL transport (pt,
// ptdirect), (wi.Ns)(wo.Ng) / ((wo.Ns)(wi.Ng)) for LE transport (lt, ltdirect).
template <typename R, bool LE = false>
__device__ __forceinline__ R sn_correction(const Surf<R>& s, V3<R> wi, V3<R> wo, R li_z, R lo_z) {
    const R wi_ng = dot(wi, s.gn), wo_ng = dot(wo, s.gn);
    if (wi_ng * li_z <= R(0) || wo_ng * lo_z <= R(0)) return R(0);
    return LE ? li_z * wo_ng / (lo_z * wi_ng) : R(1);
}

template <typename R>
__device__ __forceinline__ R local_tan(V3<R> v) {
    const R t = R(1) - v.z * v.z;
    return t <= R(0) ? R(0) : Math<R>::sqrt_(t) / v.z;
}

template <typename R>
__device__ __forceinline__ R beckmann_d(const DevPrim<R>& m, V3<R> H) {
    if (H.z <= R(0)) return R(0);
    const R ex = local_tan(H) / m.roughness;
    const R c2 = H.z * H.z;
    return Math<R>::exp_(-(ex * ex)) / (R(NGI_PI) * m.roughness * m.roughness * (c2 * c2));
}

// rt.hpp:1423-1431 (uses |wo.H| for both terms, as the reference does).
template <typename R>
__device__ __forceinline__ R shadow_masking(V3<R> li, V3<R> lo, V3<R> H) {
    const R woh = Math<R>::abs_(dot(lo, H));
    const R a = R(2) * H.z * lo.z / woh, b = R(2) * H.z * li.z / woh;
    R g = a < b ? a : b;
    return g < R(1) ? g : R(1);
}

template <typename R>
__device__ __forceinline__ V3<R> fresnel_conductor(const DevPrim<R>& m, R c) {
    V3<R> F;
    for (int k = 0; k < 3; ++k) {
        const R eta = m.eta[k], kk = m.k[k];
        const R e2k2 = eta * eta + kk * kk;
        const R tmp = e2k2 * (c * c);
        const R rpar = (tmp - eta * (R(2) * c) + R(1)) / (tmp + eta * (R(2) * c) + R(1));
        const R rperp = (e2k2 - eta * (R(2) * c) + c * c) / (e2k2 + eta * (R(2) * c) + c * c);
        F[k] = (rpar + rperp) * R(0.5);
    }
    return F;
}

template <typename R>
__device__ __forceinline__ R fresnel_dielectric(R cos_i, R etaI, R etaT) {
    const R eta = etaI / etaT;
    const R c2t = R(1) - eta * eta * (R(1) - cos_i * cos_i);
    if (c2t <= R(0)) return R(1);
    const R ci = Math<R>::abs_(cos_i), ct = Math<R>::sqrt_(c2t);
    const R rs = (etaI * ci - etaT * ct) / (etaI * ci + etaT * ct);
    const R rt = (etaI * ct - etaT * ci) / (etaI * ct + etaT * ci);
    return (rs * rs + rt * rt) * R(0.5);
}

template <typename R>
__device__ __forceinline__ V3<R> prim_R(const DevPrim<R>& m) { return V3<R>(m.R[0], m.R[1], m.R[2]); }

// EvaluateDirection(geom, type, wi, wo, EL or LE, forceDegenerated = false)
// for a surface vertex (next-event estimation, sensor connection).
// Specular components are zero.
template <typename R, bool LE = false>
__device__ __forceinline__ V3<R> bsdf_eval_nee(const DevPrim<R>& m, int type, const Surf<R>& s, V3<R> wi, V3<R> wo) {
    const Frame<R> fr = make_frame(s.n);
    const V3<R> li = to_local(fr, wi), lo = to_local(fr, wo);
    if (type & NGI_PRIM_D) {
        if (li.z <= R(0) || lo.z <= R(0)) return V3<R>(R(0));
        return prim_R(m) * (R(1.0 / NGI_PI) * sn_correction<R, LE>(s, wi, wo, li.z, lo.z));
    }
    if (type & NGI_PRIM_G) {
        if (li.z <= R(0) || lo.z <= R(0)) return V3<R>(R(0));
        const V3<R> H = vnorm(li + lo);
        const R D = beckmann_d(m, H);
        const R G = shadow_masking(li, lo, H);
        const V3<R> F = fresnel_conductor(m, dot(li, H));
        const R k = D * G / (R(4) * li.z) / lo.z * sn_correction<R, LE>(s, wi, wo, li.z, lo.z);
        return prim_R(m) * F * k;
    }
    return V3<R>(R(0));
}

// SampleDirection + EvaluateDirection/EvaluateDirectionPDF (forceDegenerated
// = true) at a surface vertex, EL or LE transport.  Returns false when the
// reference would terminate the path (wo left at zero, or fs == 0).
template <typename R, bool LE = false>
__device__ __forceinline__ bool bsdf_sample(const DevPrim<R>& m, int type, const Surf<R>& s, V3<R> wi, R u0, R u1, R ucomp,
                                            V3<R>& wo, V3<R>& weight) {
    typedef Math<R> M;
    const Frame<R> fr = make_frame(s.n);
    const V3<R> li = to_local(fr, wi);
    if (type & NGI_PRIM_D) {
        if (li.z <= R(0)) return false;
        const V3<R> lo = cosine_hemisphere(u0, u1);
        if (lo.z <= R(0)) return false;
        wo = to_world(fr, lo);
        const R corr = sn_correction<R, LE>(s, wi, wo, li.z, dot(s.n, wo));
        weight = prim_R(m) * corr;
        return !is_zero(weight);
    }
    if (type & NGI_PRIM_G) {
        if (li.z <= R(0)) return false;
        // Beckmann half-vector sample (rt.hpp:775-783).
        const R a = m.roughness;
        const R tan2 = -a * a * M::log_(R(1) - u0);
        const R ch = R(1) / M::sqrt_(R(1) + tan2);
        const R sh2 = R(1) - ch * ch;
        const R sh = M::sqrt_(sh2 > R(0) ? sh2 : R(0));
        R sp, cp;
        M::sincos_(R(2.0 * NGI_PI) * u1, &sp, &cp);
        const V3<R> H(sh * cp, sh * sp, ch);
        const R d = dot(-li, H);
        const V3<R> lo = -li - H * (R(2) * d);
        if (lo.z <= R(0)) return false;
        wo = to_world(fr, lo);
        const V3<R> lo2 = to_local(fr, wo);
        if (lo2.z <= R(0)) return false;
        const V3<R> H2 = vnorm(li + lo2);
        if (beckmann_d(m, H2) <= R(0)) return false;
        const R G = shadow_masking(li, lo2, H2);
        const V3<R> F = fresnel_conductor(m, dot(li, H2));
        const R corr = sn_correction<R, LE>(s, wi, wo, li.z, lo2.z);
        // fs/pdf = R D G F/(4 ci co) / (D hz/(4 (lo.H) co)) = R G F (lo.H)/(ci hz)
        weight = prim_R(m) * F * (G * dot(lo2, H2) / (li.z * H2.z) * corr);
        return !is_zero(weight);
    }
    if (type & NGI_PRIM_S) {
        V3<R> lo;
        R scale = R(1);   // eta^2 for refraction in EL transport only (rt.hpp:1096, 1131)
        if (m.spec_type == NGI_SPEC_REFLECTION) {
            if (li.z <= R(0)) return false;
            lo = V3<R>(-li.x, -li.y, li.z);
        } else {
            R etaI = m.s_eta1, etaT = m.s_eta2;
            if (li.z < R(0)) {
                const R t = etaI;
                etaI = etaT;
                etaT = t;
            }
            const R eta = etaI / etaT;
            const R c2t = R(1) - eta * eta * (R(1) - li.z * li.z);
            bool reflect;
            if (m.spec_type == NGI_SPEC_REFRACTION) {
                reflect = c2t <= R(0);
                if (!LE) scale = eta * eta;
            } else {
                const R Fr = fresnel_dielectric(li.z, etaI, etaT);
                reflect = ucomp <= Fr;
                if (!reflect && !LE) scale = eta * eta;
                // fs/pdf: reflection R Fr corr / Fr, refraction R (1-Fr) corr eta^2 / (1-Fr)
                if ((reflect && Fr <= R(0)) || (!reflect && Fr >= R(1))) return false;
            }
            if (reflect) {
                lo = V3<R>(-li.x, -li.y, li.z);
            } else {
                const R ct = M::sqrt_(c2t > R(0) ? c2t : R(0)) * (li.z > R(0) ? R(-1) : R(1));
                lo = V3<R>(-eta * li.x, -eta * li.y, ct);
            }
        }
        wo = to_world(fr, lo);
        const R lz = dot(s.n, wo);
        if (m.spec_type == NGI_SPEC_REFLECTION && (li.z <= R(0) || lz <= R(0))) return false;
        const R corr = sn_correction<R, LE>(s, wi, wo, li.z, lz);
        weight = prim_R(m) * (corr * scale);
        return !is_zero(weight);
    }
    return false;
}

// EvaluateDirectionPDF (forceDegenerated = true) of a surface vertex, in
// projected solid angle (rt.hpp:1213-1336): 1/pi for D, the Beckmann
// half-vector pdf for G, the branch probability for S.
template <typename R>
__device__ __forceinline__ R bsdf_pdf(const DevPrim<R>& m, int type, const Surf<R>& s, V3<R> wi, V3<R> wo) {
    const Frame<R> fr = make_frame(s.n);
    const V3<R> li = to_local(fr, wi), lo = to_local(fr, wo);
    if (type & NGI_PRIM_D) return (li.z <= R(0) || lo.z <= R(0)) ? R(0) : R(1.0 / NGI_PI);
    if (type & NGI_PRIM_G) {
        if (li.z <= R(0) || lo.z <= R(0)) return R(0);
        const V3<R> H = vnorm(li + lo);
        return beckmann_d(m, H) * H.z / (R(4) * dot(lo, H)) / lo.z;
    }
    if (type & NGI_PRIM_S) {
        if (m.spec_type == NGI_SPEC_REFLECTION) return (li.z <= R(0) || lo.z <= R(0)) ? R(0) : R(1);
        if (m.spec_type == NGI_SPEC_REFRACTION) return R(1);
        R etaI = m.s_eta1, etaT = m.s_eta2;
        if (li.z < R(0)) {
            const R t = etaI;
            etaI = etaT;
            etaT = t;
        }
        const R Fr = fresnel_dielectric(li.z, etaI, etaT);
        return li.z * lo.z >= R(0) ? Fr : R(1) - Fr;
    }
    return R(0);
}

// ---------------------------------------------------------------------------
// Pinhole sensor (rt.hpp:733-740, 955-978, 1189-1212, 1348-1378).
// ---------------------------------------------------------------------------
template <typename R>
struct Cam {
    V3<R> eye, vx, vy, vz;
    R tanf, aspect;
};

template <typename R>
__device__ __forceinline__ Cam<R> load_cam(const DevCamera& c) {
    Cam<R> k;
    k.eye = V3<R>((R)c.eye[0], (R)c.eye[1], (R)c.eye[2]);
    k.vx = V3<R>((R)c.vx[0], (R)c.vx[1], (R)c.vx[2]);
    k.vy = V3<R>((R)c.vy[0], (R)c.vy[1], (R)c.vy[2]);
    k.vz = V3<R>((R)c.vz[0], (R)c.vz[1], (R)c.vz[2]);
    k.tanf = (R)c.tan_fov;
    k.aspect = (R)c.aspect;
    return k;
}

template <typename R>
__device__ __forceinline__ V3<R> camera_dir(const Cam<R>& c, R rx, R ry) {
    const R px = R(2) * rx - R(1), py = R(2) * ry - R(1);
    const V3<R> we = vnorm(V3<R>(c.aspect * c.tanf * px, c.tanf * py, R(-1)));
    return c.vx * we.x + c.vy * we.y + c.vz * we.z;
}

// RasterPosition + importance We = 1/(A cos^3) (0 when off screen).
template <typename R>
__device__ __forceinline__ R camera_importance(const Cam<R>& c, V3<R> wo, R& rx, R& ry) {
    const R ex = dot(c.vx, wo), ey = dot(c.vy, wo), ez = dot(c.vz, wo);
    if (ez >= R(0)) return R(0);
    rx = (-ex / ez / c.tanf / c.aspect + R(1)) * R(0.5);
    ry = (-ey / ez / c.tanf + R(1)) * R(0.5);
    if (rx < R(0) || rx > R(1) || ry < R(0) || ry > R(1)) return R(0);
    const R ic = R(1) / (-ez);
    const R A = c.tanf * c.tanf * c.aspect * R(4);
    return ic * ic * ic / A;
}

template <typename R>
__device__ __forceinline__ int pixel_index(R rx, R ry, int w, int h) {
    int px = (int)(rx * (R)w), py = (int)(ry * (R)h);
    px = px < 0 ? 0 : (px > w - 1 ? w - 1 : px);
    py = py < 0 ? 0 : (py > h - 1 ? h - 1 : py);
    return py * w + px;
}

// ---------------------------------------------------------------------------
// Emitters: SampleEmitter (rt.hpp:2321-2336) and SamplePosition (483-562).
// ---------------------------------------------------------------------------
template <typename R>
struct LightSample {
    V3<R> p, n;
    R inv_pdf;      // 1 / (pdfL * pdfPL)
    V3<R> Le;
    bool degenerate;
    bool valid;
};

template <typename R> __device__ __forceinline__ const DevLight<R>* lights_of(const DevScene& sc);
template <> __device__ __forceinline__ const DevLight<float>* lights_of<float>(const DevScene& sc) { return sc.lights_f; }
template <> __device__ __forceinline__ const DevLight<double>* lights_of<double>(const DevScene& sc) { return sc.lights_d; }
template <typename R> __device__ __forceinline__ const R* cdf_of(const DevScene& sc);
template <> __device__ __forceinline__ const float* cdf_of<float>(const DevScene& sc) { return sc.cdf_f; }
template <> __device__ __forceinline__ const double* cdf_of<double>(const DevScene& sc) { return sc.cdf_d; }
template <typename R> __device__ __forceinline__ const uint32_t* guide_of(const DevScene& sc);
template <> __device__ __forceinline__ const uint32_t* guide_of<float>(const DevScene& sc) { return sc.guide_f; }
template <> __device__ __forceinline__ const uint32_t* guide_of<double>(const DevScene& sc) { return sc.guide_d; }
template <typename R> __device__ __forceinline__ const DevPrim<R>* prims_of(const DevScene& sc);
template <> __device__ __forceinline__ const DevPrim<float>* prims_of<float>(const DevScene& sc) { return sc.prims_f; }
template <> __device__ __forceinline__ const DevPrim<double>* prims_of<double>(const DevScene& sc) { return sc.prims_d; }

// A point on an area light: triangle by Distribution1D::SampleReuse
// (basic.hpp:469-475: upper_bound - 1, clamped) over the light's area CDF,
// searched through its guide table (the same index as the whole-CDF search),
// then UniformSampleTriangle (rt.hpp:129-133) with the reused coordinate;
// n = the triangle's winding normal (SampleTriangleMesh, rt.hpp:487-524).
template <typename R>
__device__ __forceinline__ void area_light_point(const DevScene& sc, const DevLight<R>& L, R u2x, R u2y, V3<R>& p, V3<R>& n) {
    const R* cdf = cdf_of<R>(sc) + L.cdf_begin;
    const uint32_t lo = guided_upper_bound(cdf, guide_of<R>(sc) + L.guide_begin, L.guide_bits, u2x);
    int i = (int)lo - 1;
    i = i < 0 ? 0 : (i > (int)L.tri_count - 1 ? (int)L.tri_count - 1 : i);
    const R c0 = cdf[i], c1 = cdf[i + 1];
    const R ux = (u2x - c0) / (c1 - c0);
    const R sq = Math<R>::sqrt_(ux > R(0) ? ux : R(0));
    const R bx = R(1) - sq, by = u2y * sq;
    const float4* t = sc.light_tris + 3 * (size_t)(L.tri_begin + i);
    const float4 a = t[0], b = t[1], c = t[2];
    const V3<R> p1((R)a.x, (R)a.y, (R)a.z), p2((R)b.x, (R)b.y, (R)b.z), p3((R)c.x, (R)c.y, (R)c.z);
    p = p1 * (R(1) - bx - by) + p2 * bx + p3 * by;
    n = vnorm(cross(p2 - p1, p3 - p1));
}

template <typename R>
__device__ __forceinline__ int pick_light(const DevScene& sc, R ul) {
    const int n = sc.num_lights;
    const int li = (int)(ul * (R)n);
    return li < 0 ? 0 : (li > n - 1 ? n - 1 : li);
}

template <typename R>
__device__ __forceinline__ LightSample<R> sample_light(const DevScene& sc, R ul, R u2x, R u2y) {
    LightSample<R> ls;
    ls.valid = false;
    const int n = sc.num_lights;
    const DevLight<R> L = lights_of<R>(sc)[pick_light(sc, ul)];
    ls.Le = V3<R>(L.Le[0], L.Le[1], L.Le[2]);
    if (L.type == NGI_LIGHT_AREA) {
        area_light_point(sc, L, u2x, u2y, ls.p, ls.n);
        ls.degenerate = false;
        ls.inv_pdf = (R)n / L.inv_area;
        ls.valid = true;
    } else if (L.type == NGI_LIGHT_POINT) {
        ls.p = V3<R>(L.pos[0], L.pos[1], L.pos[2]);
        ls.n = V3<R>(R(0));
        ls.degenerate = true;
        ls.inv_pdf = (R)n;
        ls.valid = true;
    }
    // Directional lights contribute zero to pt/ptdirect (rt.hpp:934-937).
    return ls;
}

}  // namespace ngi
