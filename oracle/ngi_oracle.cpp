// CPU FP64 oracle for nanogi's `pt` / `ptdirect` renderers.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library, and only as the
// checker; the product (nanogi_amd/libngi.so) never calls it.
//
// A line-by-line semantic restatement, in FP64, of:
//   src/nanogi.cpp:225-440   RenderProcess (sample loop, W*H/N scaling)
//   src/nanogi.cpp:446-607   ProcessSample_PT
//   src/nanogi.cpp:609-802   ProcessSample_PTDirect
//   include/nanogi/rt.hpp:51-144     sampling helpers
//   include/nanogi/rt.hpp:276-302    Ray, SurfaceGeometry
//   include/nanogi/rt.hpp:483-1470   Primitive sample/evaluate functions
//   include/nanogi/rt.hpp:2162-2261  Scene::Intersect, Scene::Visible
//   include/nanogi/rt.hpp:2321-2374  emitter sampling, GeometryTerm
//   include/nanogi/basic.hpp:440-497 Distribution1D
// with two documented substitutions: the RNG is the counter-based
// Philox-4x32-10 stream (24-bit uniforms) shared with the GPU renderer, and
// Embree's FP32 intersection is replaced by an FP64 Moller-Trumbore test
// over the oracle's own median-split BVH (or brute force).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../include/ngi.h"

namespace {

// ---------------------------------------------------------------------------
// FP64 vector helpers
// ---------------------------------------------------------------------------
struct V {
    double x = 0, y = 0, z = 0;
    V() {}
    V(double a, double b, double c) : x(a), y(b), z(c) {}
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
V operator+(V a, V b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
V operator-(V a, V b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
V operator-(V a) { return V(-a.x, -a.y, -a.z); }
V operator*(V a, double s) { return V(a.x * s, a.y * s, a.z * s); }
V operator*(V a, V b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
V operator/(V a, double s) { return V(a.x / s, a.y / s, a.z / s); }
double dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V cross(V a, V b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
double length(V a) { return std::sqrt(dot(a, a)); }
V normalize(V a) { return a / length(a); }
bool zero(V a) { return a.x == 0 && a.y == 0 && a.z == 0; }
V vec3(const double* p) { return V(p[0], p[1], p[2]); }

const double Pi = 3.14159265358979323846;
const double InvPi = 1.0 / Pi;
const float EpsF = 1e-4f;
const float InfF = 3.402823466e+38f;

// ---------------------------------------------------------------------------
// Philox-4x32-10 (Salmon et al., SC'11) and the shared stream layout:
// key (seed, stream_id), counter (path_lo, path_hi, vertex, stream).
// ---------------------------------------------------------------------------
struct Q {
    uint32_t v[4];
};
Q philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
    uint32_t c[4] = {c0, c1, c2, c3};
    for (int r = 0; r < 10; ++r) {
        const uint64_t a = (uint64_t)0xD2511F53u * c[0];
        const uint64_t b = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(b >> 32) ^ c[1] ^ k0;
        const uint32_t n1 = (uint32_t)b;
        const uint32_t n2 = (uint32_t)(a >> 32) ^ c[3] ^ k1;
        const uint32_t n3 = (uint32_t)a;
        c[0] = n0, c[1] = n1, c[2] = n2, c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return Q{{c[0], c[1], c[2], c[3]}};
}
double unif(uint32_t x) { return (double)(x >> 8) * (1.0 / 16777216.0); }

// ---------------------------------------------------------------------------
// SurfaceGeometry (rt.hpp:282-302) and OrthonormalBasis (rt.hpp:55-59)
// ---------------------------------------------------------------------------
struct Geom {
    bool degenerated = false;
    V p, sn, gn, dpdu, dpdv;
    void tangent_space() {
        // OrthonormalBasis(a = sn, b = dpdu, c = dpdv)
        const V& a = sn;
        dpdv = std::fabs(a.x) > std::fabs(a.y) ? normalize(V(a.z, 0, -a.x)) : normalize(V(0, a.z, -a.y));
        dpdu = normalize(cross(dpdv, a));
    }
    V to_local(V w) const { return V(dot(dpdu, w), dot(dpdv, w), dot(sn, w)); }
    V to_world(V l) const { return dpdu * l.x + dpdv * l.y + sn * l.z; }
};

double LocalCos(V v) { return v.z; }
double LocalTan(V v) {
    const double t = 1.0 - v.z * v.z;
    return t <= 0 ? 0 : std::sqrt(t) / v.z;
}
V LocalReflect(V wi) { return V(-wi.x, -wi.y, wi.z); }
V LocalRefract(V wi, double eta, double cosThetaT) { return V(-eta * wi.x, -eta * wi.y, cosThetaT); }

void ConcentricDisk(double ux, double uy, double& ox, double& oy) {
    const double vx = 2.0 * ux - 1.0, vy = 2.0 * uy - 1.0;
    if (vx == 0 && vy == 0) {
        ox = oy = 0;
        return;
    }
    double r, th;
    if (vx > -vy) {
        if (vx > vy) r = vx, th = (Pi * 0.25) * vy / vx;
        else r = vy, th = (Pi * 0.25) * (2.0 - vx / vy);
    } else {
        if (vx < vy) r = -vx, th = (Pi * 0.25) * (4.0 + vy / vx);
        else r = -vy, th = (Pi * 0.25) * (6.0 - vx / vy);
    }
    ox = r * std::cos(th);
    oy = r * std::sin(th);
}
V CosineSampleHemisphere(double ux, double uy) {
    double sx, sy;
    ConcentricDisk(ux, uy, sx, sy);
    return V(sx, sy, std::sqrt(std::max(0.0, 1.0 - sx * sx - sy * sy)));
}

int PixelIndex(double rx, double ry, int w, int h) {
    const int px = std::min(std::max((int)(rx * w), 0), w - 1);
    const int py = std::min(std::max((int)(ry * h), 0), h - 1);
    return py * w + px;
}

// ---------------------------------------------------------------------------
// Scene model
// ---------------------------------------------------------------------------
struct Prim {
    ngi_primitive p;
    std::vector<double> cdf;   // Distribution1D
    double inv_area = 0;
    // pinhole
    V eye, vx, vy, vz;
    double fov = 0, aspect = 1;
    // directional
    V dir_center;
    double dir_radius = 0, dir_inv_area = 0;
};

struct Node {
    double lo[3], hi[3];
    int left = -1, right = -1;   // internal if left >= 0
    uint32_t start = 0, count = 0;
};

struct Scene {
    const ngi_scene_desc* d = nullptr;
    std::vector<Prim> prims;
    std::vector<int> lights;
    int sensor = -1;
    std::vector<uint32_t> order;   // BVH leaf order of triangles
    std::vector<Node> nodes;
    bool use_bvh = true;

    V pos(uint32_t v) const { return vec3(d->positions + 3 * (size_t)v); }
    V nrm(uint32_t v) const { return vec3(d->normals + 3 * (size_t)v); }
    const uint32_t* face(uint32_t t) const { return d->faces + 3 * (size_t)t; }
};

void build_node(Scene& s, std::vector<double>& cen, uint32_t b, uint32_t e) {
    const int self = (int)s.nodes.size();
    s.nodes.push_back(Node());
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t t = s.order[i];
        for (int k = 0; k < 3; ++k) {
            const V p = s.pos(s.face(t)[k]);
            for (int a = 0; a < 3; ++a) lo[a] = std::min(lo[a], p[a]), hi[a] = std::max(hi[a], p[a]);
        }
        for (int a = 0; a < 3; ++a) clo[a] = std::min(clo[a], cen[3 * (size_t)t + a]), chi[a] = std::max(chi[a], cen[3 * (size_t)t + a]);
    }
    for (int a = 0; a < 3; ++a) s.nodes[self].lo[a] = lo[a], s.nodes[self].hi[a] = hi[a];
    if (e - b <= 4) {
        s.nodes[self].start = b;
        s.nodes[self].count = e - b;
        return;
    }
    int ax = 0;
    for (int a = 1; a < 3; ++a)
        if (chi[a] - clo[a] > chi[ax] - clo[ax]) ax = a;
    const uint32_t mid = b + (e - b) / 2;
    std::nth_element(s.order.begin() + b, s.order.begin() + mid, s.order.begin() + e, [&](uint32_t x, uint32_t y) {
        return cen[3 * (size_t)x + ax] < cen[3 * (size_t)y + ax] || (cen[3 * (size_t)x + ax] == cen[3 * (size_t)y + ax] && x < y);
    });
    const int l = (int)s.nodes.size();
    build_node(s, cen, b, mid);
    const int r = (int)s.nodes.size();
    build_node(s, cen, mid, e);
    s.nodes[self].left = l;
    s.nodes[self].right = r;
}

int load(Scene& s, const ngi_scene_desc* d, bool use_bvh) {
    s.d = d;
    s.use_bvh = use_bvh;
    if (!d->normals && d->num_triangles) return NGI_ERR_INVALID_ARG;
    s.prims.resize(d->num_primitives);
    double slo[3] = {INFINITY, INFINITY, INFINITY}, shi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t v = 0; v < d->num_vertices; ++v)
        for (int a = 0; a < 3; ++a) slo[a] = std::min(slo[a], d->positions[3 * v + a]), shi[a] = std::max(shi[a], d->positions[3 * v + a]);
    for (uint32_t i = 0; i < d->num_primitives; ++i) {
        Prim& P = s.prims[i];
        P.p = d->primitives[i];
        const int ty = P.p.type;
        if (ty & NGI_PRIM_E) s.sensor = (int)i;
        if (ty & NGI_PRIM_L) s.lights.push_back((int)i);
        if (((ty & NGI_PRIM_L) && P.p.light_type == NGI_LIGHT_AREA) || ((ty & NGI_PRIM_E) && P.p.sensor_type == NGI_SENSOR_AREA)) {
            // CreateTriangleAreaDist (rt.hpp:1747-1765)
            double sum = 0;
            P.cdf.assign(1, 0.0);
            for (uint32_t t = P.p.tri_begin; t < P.p.tri_begin + P.p.tri_count; ++t) {
                const V p1 = s.pos(s.face(t)[0]), p2 = s.pos(s.face(t)[1]), p3 = s.pos(s.face(t)[2]);
                const double a = length(cross(p2 - p1, p3 - p1)) * 0.5;
                P.cdf.push_back(P.cdf.back() + a);
                sum += a;
            }
            const double inv = 1.0 / P.cdf.back();
            for (double& c : P.cdf) c *= inv;
            P.inv_area = 1.0 / sum;
        }
        if ((ty & NGI_PRIM_E) && P.p.sensor_type == NGI_SENSOR_PINHOLE) {
            // rt.hpp:1882-1902
            P.eye = vec3(P.p.E_eye);
            P.fov = P.p.E_fov_deg * Pi / 180.0;
            P.vz = normalize(P.eye - vec3(P.p.E_center));
            P.vx = normalize(cross(vec3(P.p.E_up), P.vz));
            P.vy = cross(P.vz, P.vx);
            P.aspect = d->aspect;
        }
        if ((ty & NGI_PRIM_L) && P.p.light_type == NGI_LIGHT_DIRECTIONAL) {
            P.dir_center = V((shi[0] + slo[0]) * 0.5, (shi[1] + slo[1]) * 0.5, (shi[2] + slo[2]) * 0.5);
            P.dir_radius = length(P.dir_center - V(shi[0], shi[1], shi[2])) * 1.01;
            P.dir_inv_area = 1.0 / (2.0 * Pi * P.dir_radius * P.dir_radius);
        }
    }
    s.order.resize(d->num_triangles);
    for (uint32_t t = 0; t < d->num_triangles; ++t) s.order[t] = t;
    if (use_bvh && d->num_triangles) {
        std::vector<double> cen(3 * (size_t)d->num_triangles);
        for (uint32_t t = 0; t < d->num_triangles; ++t) {
            const V c = (s.pos(s.face(t)[0]) + s.pos(s.face(t)[1]) + s.pos(s.face(t)[2])) / 3.0;
            cen[3 * (size_t)t] = c.x, cen[3 * (size_t)t + 1] = c.y, cen[3 * (size_t)t + 2] = c.z;
        }
        s.nodes.reserve(d->num_triangles);
        build_node(s, cen, 0, d->num_triangles);
    }
    return NGI_OK;
}

// FP64 Moller-Trumbore, two-sided; hit iff tmin < t < tmax.
bool tri_hit(const Scene& s, uint32_t t, V o, V dir, double tmin, double tmax, double& th, double& u, double& v) {
    const V p0 = s.pos(s.face(t)[0]), p1 = s.pos(s.face(t)[1]), p2 = s.pos(s.face(t)[2]);
    const V e1 = p1 - p0, e2 = p2 - p0;
    const V pv = cross(dir, e2);
    const double det = dot(e1, pv);
    if (det == 0.0) return false;
    const double inv = 1.0 / det;
    const V tv = o - p0;
    const double uu = dot(tv, pv) * inv;
    if (uu < 0.0 || uu > 1.0) return false;
    const V qv = cross(tv, e1);
    const double vv = dot(dir, qv) * inv;
    if (vv < 0.0 || uu + vv > 1.0) return false;
    const double tt = dot(e2, qv) * inv;
    if (!(tt > tmin && tt < tmax)) return false;
    th = tt, u = uu, v = vv;
    return true;
}

bool box_hit(const Node& n, V o, V inv, double tmin, double tmax, double& tnear) {
    double t0 = tmin, t1 = tmax;
    for (int a = 0; a < 3; ++a) {
        double lo = (n.lo[a] - o[a]) * inv[a], hi = (n.hi[a] - o[a]) * inv[a];
        if (std::isnan(lo) || std::isnan(hi)) {  // zero direction with origin on the slab
            if (o[a] < n.lo[a] || o[a] > n.hi[a]) return false;
            continue;
        }
        if (lo > hi) std::swap(lo, hi);
        t0 = std::max(t0, lo);
        t1 = std::min(t1, hi * (1 + 1e-12));
        if (t0 > t1) return false;
    }
    tnear = t0;
    return true;
}

// Closest (or any) hit in (tmin, tmax): returns triangle or -1.
long intersect(const Scene& s, V o, V dir, double tmin, double tmax, bool any, double& th, double& u, double& v) {
    long best = -1;
    if (!s.use_bvh || s.nodes.empty()) {
        for (uint32_t t = 0; t < s.d->num_triangles; ++t) {
            double tt, uu, vv;
            if (tri_hit(s, t, o, dir, tmin, tmax, tt, uu, vv)) {
                tmax = tt, th = tt, u = uu, v = vv, best = t;
                if (any) return best;
            }
        }
        return best;
    }
    const V inv(1.0 / dir.x, 1.0 / dir.y, 1.0 / dir.z);
    int stack[256];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Node& n = s.nodes[stack[--sp]];
        double tn;
        if (!box_hit(n, o, inv, tmin, tmax, tn)) continue;
        if (n.left < 0) {
            for (uint32_t i = n.start; i < n.start + n.count; ++i) {
                const uint32_t t = s.order[i];
                double tt, uu, vv;
                if (tri_hit(s, t, o, dir, tmin, tmax, tt, uu, vv)) {   // tt < tmax: strictly closer
                    tmax = tt, th = tt, u = uu, v = vv, best = t;
                    if (any) return best;
                }
            }
        } else {
            double a = 0, b = 0;
            const bool ha = box_hit(s.nodes[n.left], o, inv, tmin, tmax, a);
            const bool hb = box_hit(s.nodes[n.right], o, inv, tmin, tmax, b);
            if (ha && hb) {
                if (a <= b) stack[sp++] = n.right, stack[sp++] = n.left;
                else stack[sp++] = n.left, stack[sp++] = n.right;
            } else if (ha) {
                stack[sp++] = n.left;
            } else if (hb) {
                stack[sp++] = n.right;
            }
        }
    }
    return best;
}

struct Isect {
    Geom geom;
    int prim = -1;
};

// Scene::Intersect (rt.hpp:2162-2244) with the FP64 intersector.
bool Intersect(const Scene& s, V o, V dir, Isect& is, double minT = EpsF, double maxT = InfF) {
    double t, u, v;
    const long tri = intersect(s, o, dir, minT, maxT, false, t, u, v);
    if (tri < 0) return false;
    is.prim = (int)s.d->tri_prim[tri];
    const uint32_t* f = s.face((uint32_t)tri);
    is.geom.p = o + dir * t;
    const V p1 = s.pos(f[0]), p2 = s.pos(f[1]), p3 = s.pos(f[2]);
    is.geom.gn = normalize(cross(p2 - p1, p3 - p1));
    const V n1 = s.nrm(f[0]), n2 = s.nrm(f[1]), n3 = s.nrm(f[2]);
    is.geom.sn = normalize(n1 * (1.0 - u - v) + n2 * u + n3 * v);
    if (std::isnan(is.geom.sn.x) || std::isnan(is.geom.sn.y) || std::isnan(is.geom.sn.z)) is.geom.sn = is.geom.gn;
    is.geom.degenerated = false;
    is.geom.tangent_space();
    return true;
}

// Scene::Visible (rt.hpp:2251-2261)
bool Visible(const Scene& s, V p1, V p2, long* rays) {
    const V d12 = p2 - p1;
    const double L = length(d12);
    const V dir = d12 / L;
    double t, u, v;
    if (rays) ++*rays;
    return intersect(s, p1, dir, EpsF, (double)((float)L * (1.0f - EpsF)), true, t, u, v) < 0;
}

// ---------------------------------------------------------------------------
// Primitive functions (rt.hpp:483-1466), L / E / D / G / S branches.
// ---------------------------------------------------------------------------
void SamplePosition(const Scene& s, const Prim& P, double ux, double uy, Geom& g) {
    const int ty = P.p.type;
    auto tri_mesh = [&]() {
        // Distribution1D::SampleReuse (basic.hpp:469-475)
        const auto it = std::upper_bound(P.cdf.begin(), P.cdf.end(), ux);
        int i = (int)(it - P.cdf.begin()) - 1;
        i = std::min(std::max(i, 0), (int)P.cdf.size() - 2);
        const double u2x = (ux - P.cdf[i]) / (P.cdf[i + 1] - P.cdf[i]);
        // UniformSampleTriangle (rt.hpp:129-133)
        const double sq = std::sqrt(std::max(0.0, u2x));
        const double bx = 1.0 - sq, by = uy * sq;
        const uint32_t* f = s.face(P.p.tri_begin + (uint32_t)i);
        const V p1 = s.pos(f[0]), p2 = s.pos(f[1]), p3 = s.pos(f[2]);
        g.p = p1 * (1.0 - bx - by) + p2 * bx + p3 * by;
        g.degenerated = false;
        g.gn = normalize(cross(p2 - p1, p3 - p1));
        g.sn = g.gn;
        g.tangent_space();
    };
    if (ty & NGI_PRIM_L) {
        if (P.p.light_type == NGI_LIGHT_AREA) return tri_mesh();
        if (P.p.light_type == NGI_LIGHT_POINT) {
            g.degenerated = true;
            g.p = vec3(P.p.light_position);
            return;
        }
        double dx, dy;
        ConcentricDisk(ux, uy, dx, dy);
        g.degenerated = false;
        g.gn = vec3(P.p.light_direction);
        g.sn = g.gn;
        g.tangent_space();
        g.p = P.dir_center - g.gn * P.dir_radius + (g.dpdu * (dx * P.dir_radius) + g.dpdv * (dy * P.dir_radius));
        return;
    }
    if (ty & NGI_PRIM_E) {
        g.degenerated = true;
        g.p = P.eye;
    }
}

double EvaluatePositionPDF(const Prim& P) {
    if (P.p.type & NGI_PRIM_L) {
        if (P.p.light_type == NGI_LIGHT_AREA) return P.inv_area;
        if (P.p.light_type == NGI_LIGHT_POINT) return 1;   // forceDegenerated
        return P.dir_inv_area;
    }
    return 1;
}

bool RasterPosition(const Prim& P, V wo, double& rx, double& ry) {
    const V we(dot(P.vx, wo), dot(P.vy, wo), dot(P.vz, wo));
    if (we.z >= 0) return false;
    const double tf = std::tan(P.fov * 0.5);
    rx = (-we.x / we.z / tf / P.aspect + 1.0) * 0.5;
    ry = (-we.y / we.z / tf + 1.0) * 0.5;
    return !(rx < 0 || rx > 1 || ry < 0 || ry > 1);
}

double BeckmannD(const Prim& P, V H) {
    if (LocalCos(H) <= 0) return 0.0;
    const double ex = LocalTan(H) / P.p.G_roughness;
    return std::exp(-(ex * ex)) / (Pi * P.p.G_roughness * P.p.G_roughness * std::pow(LocalCos(H), 4.0));
}
double ShadowMasking(V wi, V wo, V H) {
    const double woH = std::fabs(dot(wo, H));
    return std::min(1.0, std::min(2.0 * LocalCos(H) * LocalCos(wo) / woH, 2.0 * LocalCos(H) * LocalCos(wi) / woH));
}
V FrConductor(const Prim& P, double c) {
    double o[3];
    for (int k = 0; k < 3; ++k) {
        const double eta = P.p.G_eta[k], kk = P.p.G_k[k];
        const double tmp = (eta * eta + kk * kk) * (c * c);
        const double rpar = (tmp - eta * (2.0 * c) + 1.0) / (tmp + eta * (2.0 * c) + 1.0);
        const double tmpF = eta * eta + kk * kk;
        const double rperp = (tmpF - eta * (2.0 * c) + c * c) / (tmpF + eta * (2.0 * c) + c * c);
        o[k] = (rpar + rperp) * 0.5;
    }
    return V(o[0], o[1], o[2]);
}
double FresnelDielectric(V li, double etaI, double etaT) {
    const double c = LocalCos(li), eta = etaI / etaT;
    const double c2t = 1.0 - eta * eta * (1.0 - c * c);
    if (c2t <= 0) return 1;
    const double ci = std::fabs(c), ct = std::sqrt(c2t);
    const double rs = (etaI * ci - etaT * ct) / (etaI * ci + etaT * ct);
    const double rt = (etaI * ct - etaT * ci) / (etaI * ct + etaT * ci);
    return (rs * rs + rt * rt) * 0.5;
}

// SampleDirection (rt.hpp:692-910); wo untouched (zero) where the reference returns early.
void SampleDirection(const Prim& P, int type, const Geom& g, V wi, double u0, double u1, double uc, V& wo) {
    if (type & NGI_PRIM_L) {
        if (P.p.light_type == NGI_LIGHT_AREA) {
            wo = g.to_world(CosineSampleHemisphere(u0, u1));
        } else if (P.p.light_type == NGI_LIGHT_POINT) {
            // UniformSampleSphere (rt.hpp:116-122)
            const double z = 1.0 - 2.0 * u0, r = std::sqrt(std::max(0.0, 1.0 - z * z)), ph = 2.0 * Pi * u1;
            wo = V(r * std::cos(ph), r * std::sin(ph), z);
        } else {
            wo = vec3(P.p.light_direction);
        }
        return;
    }
    if (type & NGI_PRIM_E) {
        const double rx = 2.0 * u0 - 1.0, ry = 2.0 * u1 - 1.0;
        const double tf = std::tan(P.fov * 0.5);
        const V we = normalize(V(P.aspect * tf * rx, tf * ry, -1));
        wo = P.vx * we.x + P.vy * we.y + P.vz * we.z;
        return;
    }
    if (type & NGI_PRIM_D) {
        if (LocalCos(g.to_local(wi)) <= 0) return;
        wo = g.to_world(CosineSampleHemisphere(u0, u1));
        return;
    }
    if (type & NGI_PRIM_G) {
        const V li = g.to_local(wi);
        if (LocalCos(li) <= 0) return;
        const double a = P.p.G_roughness;
        const double tan2 = -a * a * std::log(1.0 - u0);
        const double ch = 1.0 / std::sqrt(1.0 + tan2);
        const double sh = std::sqrt(std::max(0.0, 1.0 - ch * ch));
        const double ph = 2.0 * Pi * u1;
        const V H(sh * std::cos(ph), sh * std::sin(ph), ch);
        const V lo = -li - H * (2.0 * dot(-li, H));
        if (LocalCos(lo) <= 0) return;
        wo = g.to_world(lo);
        return;
    }
    if (type & NGI_PRIM_S) {
        const V li = g.to_local(wi);
        if (P.p.spec_type == NGI_SPEC_REFLECTION) {
            if (LocalCos(li) <= 0) return;
            wo = g.to_world(LocalReflect(li));
            return;
        }
        double etaI = P.p.S_eta1, etaT = P.p.S_eta2;
        if (LocalCos(li) < 0) std::swap(etaI, etaT);
        const double c = LocalCos(li), eta = etaI / etaT;
        const double c2t = 1.0 - eta * eta * (1.0 - c * c);
        bool reflect;
        if (P.p.spec_type == NGI_SPEC_REFRACTION) reflect = c2t <= 0;
        else reflect = uc <= FresnelDielectric(li, etaI, etaT);
        if (reflect) wo = g.to_world(LocalReflect(li));
        else wo = g.to_world(LocalRefract(li, eta, std::sqrt(std::max(0.0, c2t)) * (c > 0 ? -1.0 : 1.0)));
    }
}

// rt.hpp:994-1005: zero when a direction's geometric and shading cosines
// disagree in sign; 1 for EL transport, (wi.Ns)(wo.Ng) / ((wo.Ns)(wi.Ng)) for LE.
double ShadingNormalCorrection(const Geom& g, V wi, V wo, bool le) {
    const double wiNg = dot(wi, g.gn), woNg = dot(wo, g.gn);
    const double wiNs = LocalCos(g.to_local(wi)), woNs = LocalCos(g.to_local(wo));
    if (wiNg * wiNs <= 0 || woNg * woNs <= 0) return 0.0;
    return le ? wiNs * woNg / (woNs * wiNg) : 1.0;
}

// EvaluateDirection (rt.hpp:912-1148); le selects TransportDirection::LE.
V EvaluateDirection(const Prim& P, int type, const Geom& g, V wi, V wo, bool forceDeg, bool le = false) {
    if (type & NGI_PRIM_L) {
        if (P.p.light_type == NGI_LIGHT_AREA) return LocalCos(g.to_local(wo)) <= 0 ? V() : vec3(P.p.Le);
        if (P.p.light_type == NGI_LIGHT_POINT) return vec3(P.p.Le);
        return forceDeg ? vec3(P.p.Le) : V();
    }
    if (type & NGI_PRIM_E) {
        double rx, ry;
        if (!RasterPosition(P, wo, rx, ry)) return V();
        const V we(dot(P.vx, wo), dot(P.vy, wo), dot(P.vz, wo));
        const double tf = std::tan(P.fov * 0.5);
        const double ic = 1.0 / -LocalCos(we);
        const double A = tf * tf * P.aspect * 4.0;
        const double W = ic * ic * ic / A;
        return V(W, W, W);
    }
    const double corr = ShadingNormalCorrection(g, wi, wo, le);
    const V li = g.to_local(wi), lo = g.to_local(wo);
    if (type & NGI_PRIM_D) {
        if (LocalCos(li) <= 0 || LocalCos(lo) <= 0) return V();
        return vec3(P.p.D_R) * (InvPi * corr);
    }
    if (type & NGI_PRIM_G) {
        if (LocalCos(li) <= 0 || LocalCos(lo) <= 0) return V();
        const V H = normalize(li + lo);
        const double D = BeckmannD(P, H), G = ShadowMasking(li, lo, H);
        const V F = FrConductor(P, dot(li, H));
        return vec3(P.p.G_R) * F * (D * G / (4.0 * LocalCos(li)) / LocalCos(lo) * corr);
    }
    if (type & NGI_PRIM_S) {
        if (!forceDeg) return V();
        if (P.p.spec_type == NGI_SPEC_REFLECTION) {
            if (LocalCos(li) <= 0 || LocalCos(lo) <= 0) return V();
            return vec3(P.p.S_R) * corr;
        }
        double etaI = P.p.S_eta1, etaT = P.p.S_eta2;
        if (LocalCos(li) < 0) std::swap(etaI, etaT);
        const double eta = le ? 1.0 : etaI / etaT;   // eta^2 radiance scaling for EL only (rt.hpp:1096, 1131)
        if (P.p.spec_type == NGI_SPEC_REFRACTION) return vec3(P.p.S_R) * (corr * eta * eta);
        const double Fr = FresnelDielectric(li, etaI, etaT);
        if (LocalCos(li) * LocalCos(lo) >= 0) return vec3(P.p.S_R) * (Fr * corr);
        return vec3(P.p.S_R) * ((1.0 - Fr) * corr * eta * eta);
    }
    return V();
}

// EvaluateDirectionPDF (rt.hpp:1150-1336), forceDegenerated = true.
double EvaluateDirectionPDF(const Prim& P, int type, const Geom& g, V wi, V wo) {
    if (type & NGI_PRIM_L) {
        if (P.p.light_type == NGI_LIGHT_AREA) return LocalCos(g.to_local(wo)) <= 0 ? 0.0 : InvPi;   // projected solid angle
        if (P.p.light_type == NGI_LIGHT_POINT) return InvPi * 0.25;
        return 1.0;
    }
    if (type & NGI_PRIM_E) {
        double rx, ry;
        if (!RasterPosition(P, wo, rx, ry)) return 0;
        const V we(dot(P.vx, wo), dot(P.vy, wo), dot(P.vz, wo));
        const double tf = std::tan(P.fov * 0.5);
        const double ic = 1.0 / -LocalCos(we);
        return ic * ic * ic / (tf * tf * P.aspect * 4.0);
    }
    const V li = g.to_local(wi), lo = g.to_local(wo);
    if (type & NGI_PRIM_D) {
        if (LocalCos(li) <= 0 || LocalCos(lo) <= 0) return 0;
        return InvPi;
    }
    if (type & NGI_PRIM_G) {
        if (LocalCos(li) <= 0 || LocalCos(lo) <= 0) return 0;
        const V H = normalize(li + lo);
        return BeckmannD(P, H) * LocalCos(H) / (4.0 * dot(lo, H)) / LocalCos(lo);
    }
    if (type & NGI_PRIM_S) {
        if (P.p.spec_type == NGI_SPEC_REFLECTION) return (LocalCos(li) <= 0 || LocalCos(lo) <= 0) ? 0 : 1;
        if (P.p.spec_type == NGI_SPEC_REFRACTION) return 1;
        double etaI = P.p.S_eta1, etaT = P.p.S_eta2;
        if (LocalCos(li) < 0) std::swap(etaI, etaT);
        const double Fr = FresnelDielectric(li, etaI, etaT);
        return LocalCos(li) * LocalCos(lo) >= 0 ? Fr : 1.0 - Fr;
    }
    return 0;
}

double GeometryTerm(const Geom& a, const Geom& b) {
    V d = b.p - a.p;
    const double L2 = dot(d, d), L = std::sqrt(L2);
    d = d / L;
    double t = 1.0;
    if (!a.degenerated) t *= std::fabs(dot(a.sn, d));
    if (!b.degenerated) t *= std::fabs(dot(b.sn, -d));
    return t / L2;
}

struct Counts {
    long ext = 0, shadow = 0;
};

// Adjoint bookkeeping of one path (see oracle_render_adjoint).  The path
// is replayed in a scene whose zero reflectance and radiance channels are
// set to 1; Rt / Let hold the true values (3 per primitive).
struct Adjoint {
    const double* dLdI = nullptr;   // W*H*3
    double scale = 1.0;             // W*H / N
    std::vector<double>* gR = nullptr;
    std::vector<double>* gLe = nullptr;
    const std::vector<double>* Rt = nullptr;
    const std::vector<double>* Let = nullptr;
    std::vector<int> rec;   // primitives whose reflectance multiplied the path, in order
    struct Contribution {
        size_t nrec;        // records before it
        int emitter;
        V gc;               // dL/dI * scale * replayed contribution
    };
    std::vector<Contribution> contrib;
};

// Product rule over one replayed path.  A contribution C~ with replayed
// factors R~_i (i < nrec) and Le~ is, in the true parameters,
//   C = C~ * prod_i (R_i / R~_i) * (Le / Le~),
// with R_i / R~_i = 1 for a non-zero channel and 0 for a zero one, so
//   dC/dR_j = C~ / R~_j * prod_{i != j} (R_i / R~_i) * (Le / Le~)
//   dC/dLe  = C~ / Le~  * prod_i (R_i / R~_i).
static void adj_finish_path(Adjoint& A, const Scene& s) {
    for (const Adjoint::Contribution& c : A.contrib) {
        for (int k = 0; k < 3; ++k) {
            int zeros = 0;
            for (size_t i = 0; i < c.nrec; ++i) zeros += (*A.Rt)[3 * (size_t)A.rec[i] + k] == 0.0;
            const double Le_t = (*A.Let)[3 * (size_t)c.emitter + k];
            const double Le_r = s.prims[c.emitter].p.Le[k];
            if (zeros == 0) (*A.gLe)[3 * (size_t)c.emitter + k] += c.gc[k] / Le_r;
            if (Le_t == 0.0) continue;
            for (size_t j = 0; j < c.nrec; ++j) {
                const int m = A.rec[j];
                const bool zj = (*A.Rt)[3 * (size_t)m + k] == 0.0;
                if (zeros - (zj ? 1 : 0) != 0) continue;
                const ngi_primitive& P = s.prims[m].p;
                const double R_r = ((P.type & NGI_PRIM_D) ? P.D_R : P.G_R)[k];
                (*A.gR)[3 * (size_t)m + k] += c.gc[k] / R_r;
            }
        }
    }
}

// One path sample: ProcessSample_PT (direct = false) or _PTDirect (true).
void process_sample(const Scene& s, const ngi_render_params& prm, uint64_t id, std::vector<double>& film, Counts& cnt,
                    Adjoint* adj = nullptr) {
    const bool direct = prm.renderer == NGI_RENDERER_PTDIRECT;
    const int W = prm.width, H = prm.height;
    const uint32_t npix = (uint32_t)W * (uint32_t)H;
    uint32_t c0, c1;
    int pixelIndex = -1;
    const bool per_pixel = prm.sampling == NGI_SAMPLING_PER_PIXEL;
    if (per_pixel) {
        pixelIndex = (int)(id % npix);
        c0 = (uint32_t)pixelIndex;
        c1 = (uint32_t)(id / npix);
    } else {
        c0 = (uint32_t)id;
        c1 = (uint32_t)(id >> 32) | 0x80000000u;
    }
    auto rng = [&](uint32_t vertex, uint32_t stream) { return philox(c0, c1, vertex, stream, prm.seed, prm.stream_id); };
    // Film write, or (adjoint) the gradient of dL/dI . I with respect to the
    // emitter radiance and the running sum S of weighted contributions.
    auto add = [&](int pix, V c, int emitter) {
        if (!adj) {
            film[3 * (size_t)pix] += c.x;
            film[3 * (size_t)pix + 1] += c.y;
            film[3 * (size_t)pix + 2] += c.z;
            return;
        }
        const double* g = adj->dLdI + 3 * (size_t)pix;
        const V gc(g[0] * c.x * adj->scale, g[1] * c.y * adj->scale, g[2] * c.z * adj->scale);
        adj->contrib.push_back(Adjoint::Contribution{adj->rec.size(), emitter, gc});
    };
    if (adj) {
        adj->rec.clear();
        adj->contrib.clear();
    }
    // Sensor (SampleEmitter(E): the last sensor, pdf 1) and its position.
    const Prim& E = s.prims[s.sensor];
    Geom geom;
    geom.degenerated = true;
    geom.p = E.eye;
    V throughput(1, 1, 1);   // EvaluatePosition / pdfPE / pdfE = 1
    const Prim* prim = &E;
    int type = NGI_PRIM_E;
    V wi;
    int numVertices = 1;
    // Same hard cap as the GPU renderer (1024 vertices) for -m -1 without RR.
    const int maxv = (prm.max_num_vertices < 0 || prm.max_num_vertices > 1024) ? 1024 : prm.max_num_vertices;
    for (;;) {
        if (numVertices >= maxv) break;
        if (adj && (type & (NGI_PRIM_D | NGI_PRIM_G))) adj->rec.push_back((int)(prim - &s.prims[0]));
        if (direct && !s.lights.empty()) {
            const Q q = rng((uint32_t)numVertices, 1);
            const int n = (int)s.lights.size();
            const int li = std::min(std::max((int)(unif(q.v[0]) * n), 0), n - 1);
            const Prim& L = s.prims[s.lights[li]];
            const double pdfL = 1.0 / n;
            Geom geomL;
            SamplePosition(s, L, unif(q.v[1]), unif(q.v[2]), geomL);
            const double pdfPL = EvaluatePositionPDF(L);
            const V ppL = normalize(geomL.p - geom.p);
            const V fsE = EvaluateDirection(*prim, type, geom, wi, ppL, false);
            // L->EvaluateDirection(geomL, L, 0, -ppL, LE, false)
            V fsL;
            if (L.p.light_type == NGI_LIGHT_AREA) fsL = LocalCos(geomL.to_local(-ppL)) <= 0 ? V() : vec3(L.p.Le);
            else if (L.p.light_type == NGI_LIGHT_POINT) fsL = vec3(L.p.Le);
            const double G = GeometryTerm(geom, geomL);
            const V Cu = throughput * fsE * fsL * (G / pdfL / pdfPL);   // EvaluatePosition(geomL, true) = 1
            if (!zero(Cu) && Visible(s, geom.p, geomL.p, &cnt.shadow)) {
                int index = pixelIndex;
                if (type == NGI_PRIM_E) {
                    double rx = 0, ry = 0;
                    RasterPosition(*prim, ppL, rx, ry);
                    index = PixelIndex(rx, ry, W, H);
                }
                add(index, Cu, s.lights[li]);
            }
        }
        const Q q = rng((uint32_t)numVertices, 0);
        const double u0 = unif(q.v[0]), u1 = unif(q.v[1]), uc = unif(q.v[2]), urr = unif(q.v[3]);
        V wo;
        if (type == NGI_PRIM_E && per_pixel) {
            // Per-pixel stratification: the raster sample lies in this pixel.
            const int px = pixelIndex % W, py = pixelIndex / W;
            SampleDirection(*prim, type, geom, wi, (px + u0) / W, (py + u1) / H, uc, wo);
        } else {
            SampleDirection(*prim, type, geom, wi, u0, u1, uc, wo);
        }
        const double pdfD = EvaluateDirectionPDF(*prim, type, geom, wi, wo);
        if (type == NGI_PRIM_E) {
            double rx, ry;
            if (!RasterPosition(*prim, wo, rx, ry)) break;
            if (!per_pixel) pixelIndex = PixelIndex(rx, ry, W, H);
        }
        const V fs = EvaluateDirection(*prim, type, geom, wi, wo, true);
        if (zero(fs)) break;
        throughput = throughput * fs / pdfD;
        Isect is;
        cnt.ext++;
        if (!Intersect(s, geom.p, wo, is)) break;
        const Prim& hp = s.prims[is.prim];
        if (!direct && (hp.p.type & NGI_PRIM_L)) {
            // EvaluateDirection(isect.geom, L, 0, -d, EL, false) * EvaluatePosition(isect.geom, false)
            if (hp.p.light_type == NGI_LIGHT_AREA && LocalCos(is.geom.to_local(-wo)) > 0) add(pixelIndex, throughput * vec3(hp.p.Le), is.prim);
        }
        if (prm.rr_depth >= 0 && numVertices >= prm.rr_depth) {
            if (urr > prm.rr_prob) break;
            throughput = throughput / prm.rr_prob;
        }
        geom = is.geom;
        prim = &hp;
        type = hp.p.type & ~(NGI_PRIM_L | NGI_PRIM_E);
        wi = -wo;
        numVertices++;
    }
    if (adj) adj_finish_path(*adj, s);
}

// One light path: ProcessSample_LT (direct = false, src/nanogi.cpp:804-934)
// or ProcessSample_LTDirect (true, :936-1131).  Random numbers: the light
// choice and its position from stream 1 at vertex 0, direction sampling and
// Russian roulette from stream 0 at every vertex, as the GPU renderer.
void process_light_sample(const Scene& s, const ngi_render_params& prm, uint64_t id, std::vector<double>& film, Counts& cnt) {
    const bool direct = prm.renderer == NGI_RENDERER_LTDIRECT;
    const int W = prm.width, H = prm.height;
    const uint32_t npix = (uint32_t)W * (uint32_t)H;
    uint32_t c0, c1;
    if (prm.sampling == NGI_SAMPLING_PER_PIXEL) {
        c0 = (uint32_t)(id % npix);
        c1 = (uint32_t)(id / npix);
    } else {
        c0 = (uint32_t)id;
        c1 = (uint32_t)(id >> 32) | 0x80000000u;
    }
    auto rng = [&](uint32_t vertex, uint32_t stream) { return philox(c0, c1, vertex, stream, prm.seed, prm.stream_id); };
    if (s.lights.empty()) return;
    // SampleEmitter(L) (uniform per light primitive) and SamplePosition.
    const Q ql = rng(0, 1);
    const int n = (int)s.lights.size();
    const int li = std::min(std::max((int)(unif(ql.v[0]) * n), 0), n - 1);
    const Prim& L = s.prims[s.lights[li]];
    const double pdfL = 1.0 / n;
    Geom geom;
    SamplePosition(s, L, unif(ql.v[1]), unif(ql.v[2]), geom);
    const double pdfPL = EvaluatePositionPDF(L);
    V throughput = V(1, 1, 1) * (1.0 / pdfPL / pdfL);   // EvaluatePosition(geomL, true) = 1
    const Prim* prim = &L;
    int type = NGI_PRIM_L;
    V wi;
    int numVertices = 1;
    const Prim& E = s.prims[s.sensor];
    const int maxv = (prm.max_num_vertices < 0 || prm.max_num_vertices > 1024) ? 1024 : prm.max_num_vertices;
    for (;;) {
        if (numVertices >= maxv) break;
        if (direct) {
            // Sensor connection (src/nanogi.cpp:1027-1079): the single pinhole,
            // pdfE = 1.  pdfPE is evaluated on the light primitive L, as the
            // reference does (:1045), and LeP = L->EvaluatePosition(.., true) = 1.
            Geom geomE;
            geomE.degenerated = true;
            geomE.p = E.eye;
            const double pdfPE = EvaluatePositionPDF(L);
            const V ppE = normalize(geomE.p - geom.p);
            const V fsL = EvaluateDirection(*prim, type, geom, wi, ppE, false, true);
            const V fsE = EvaluateDirection(E, NGI_PRIM_E, geomE, V(), -ppE, false);
            const double G = GeometryTerm(geom, geomE);
            const V C = throughput * fsL * fsE * (G / pdfPE);
            if (!zero(C) && Visible(s, geom.p, geomE.p, &cnt.shadow)) {
                double rx = 0, ry = 0;
                RasterPosition(E, -ppE, rx, ry);
                const int index = PixelIndex(rx, ry, W, H);
                film[3 * (size_t)index] += C.x;
                film[3 * (size_t)index + 1] += C.y;
                film[3 * (size_t)index + 2] += C.z;
            }
        }
        const Q q = rng((uint32_t)numVertices, 0);
        const double u0 = unif(q.v[0]), u1 = unif(q.v[1]), uc = unif(q.v[2]), urr = unif(q.v[3]);
        V wo;
        SampleDirection(*prim, type, geom, wi, u0, u1, uc, wo);
        const double pdfD = EvaluateDirectionPDF(*prim, type, geom, wi, wo);
        const V fs = EvaluateDirection(*prim, type, geom, wi, wo, true, true);
        if (zero(fs)) break;
        throughput = throughput * fs / pdfD;
        Isect is;
        cnt.ext++;
        if (!Intersect(s, geom.p, wo, is)) break;
        // A hit on the sensor adds throughput * We * EvaluatePosition(isect.geom,
        // false) (:897-903); a pinhole's EvaluatePosition without
        // forceDegenerated is zero (rt.hpp:590-593), so lt adds nothing.
        if (prm.rr_depth >= 0 && numVertices >= prm.rr_depth) {
            if (urr > prm.rr_prob) break;
            throughput = throughput / prm.rr_prob;
        }
        const Prim& hp = s.prims[is.prim];
        geom = is.geom;
        prim = &hp;
        type = hp.p.type & ~(NGI_PRIM_L | NGI_PRIM_E);
        wi = -wo;
        numVertices++;
        if (type == 0) break;   // [L]-only surfaces: the next EvaluateDirection is zero
    }
}

// ---------------------------------------------------------------------------
// Bidirectional path tracing: ProcessSample_BDPT (src/nanogi.cpp:1133-1189)
// with Path (include/nanogi/bdpt.hpp): SampleSubpath :54-137, Connect
// :139-191, EvaluateCst :210-244, EvaluateUnweightContribution :246-333,
// power-heuristic weight (the "Opt" form, every p_i recomputed) :352-372,
// EvaluatePDF :491-528, SelectionProb :193-208, RasterPosition :210-218.
// Random numbers: the eye subpath's step k draws stream 0 at vertex k, the
// light subpath's stream 1 at vertex k (its vertex 0: light choice and
// position, as the light tracers).
// ---------------------------------------------------------------------------
const int kBdptMaxVertices = 64;   // subpath cap; the reference's -m -1 has none

struct PV {
    int type = 0;
    Geom geom;
    const Prim* prim = nullptr;
};

// EvaluatePosition (rt.hpp:564-620).
double EvaluatePositionE(const Prim& P, bool forceDeg) {
    if (P.p.type & NGI_PRIM_L) return P.p.light_type == NGI_LIGHT_POINT ? (forceDeg ? 1.0 : 0.0) : 1.0;
    return P.p.sensor_type == NGI_SENSOR_PINHOLE ? (forceDeg ? 1.0 : 0.0) : 1.0;
}

// EvaluateEmitterPDF (rt.hpp:2338-2352): uniform over lights; one sensor.
double EmitterPDF(const Scene& s, const Prim& P) { return (P.p.type & NGI_PRIM_L) ? 1.0 / (double)s.lights.size() : 1.0; }

double rr_at(const ngi_render_params& prm, int step) { return (prm.rr_depth >= 0 && step >= prm.rr_depth) ? prm.rr_prob : 1.0; }

struct Bdpt {
    const Scene& s;
    const ngi_render_params& prm;
    std::vector<PV> x;   // connected path, light end first

    V dir(int from, int to) const { return normalize(x[to].geom.p - x[from].geom.p); }

    V Cst(int sl) const {
        const int n = (int)x.size(), t = n - sl;
        if (sl == 0 && t > 0) {
            const PV& v = x[0];
            return EvaluateDirection(*v.prim, v.type, v.geom, V(), dir(0, 1), false) * EvaluatePositionE(*v.prim, false);
        }
        if (sl > 0 && t == 0) {
            const PV& v = x[n - 1];
            return EvaluateDirection(*v.prim, v.type, v.geom, V(), dir(n - 1, n - 2), false, true) * EvaluatePositionE(*v.prim, false);
        }
        const PV& vL = x[sl - 1];
        const PV& vE = x[sl];
        const V fsL = EvaluateDirection(*vL.prim, vL.type, vL.geom, sl - 2 >= 0 ? dir(sl - 1, sl - 2) : V(), dir(sl - 1, sl), false, true);
        const V fsE = EvaluateDirection(*vE.prim, vE.type, vE.geom, sl + 1 < n ? dir(sl, sl + 1) : V(), dir(sl, sl - 1), false);
        return fsL * fsE * GeometryTerm(vL.geom, vE.geom);
    }

    V Unweighted(int sl) const {
        const int n = (int)x.size(), t = n - sl;
        auto local = [](V f, double p) { return zero(f) ? V() : f / p; };
        V aL(1, 1, 1), aE(1, 1, 1);
        if (sl > 0) {
            const PV& v0 = x[0];
            aL = local(V(1, 1, 1) * EvaluatePositionE(*v0.prim, true), EvaluatePositionPDF(*v0.prim) * EmitterPDF(s, *v0.prim));
            for (int i = 0; i < sl - 1; ++i) {
                const V wi = i >= 1 ? dir(i, i - 1) : V(), wo = dir(i, i + 1);
                aL = aL * local(EvaluateDirection(*x[i].prim, x[i].type, x[i].geom, wi, wo, true, true),
                                EvaluateDirectionPDF(*x[i].prim, x[i].type, x[i].geom, wi, wo));
            }
        }
        if (zero(aL)) return V();
        if (t > 0) {
            const PV& vn = x[n - 1];
            aE = local(V(1, 1, 1) * EvaluatePositionE(*vn.prim, true), EvaluatePositionPDF(*vn.prim) * EmitterPDF(s, *vn.prim));
            for (int i = n - 1; i > sl; --i) {
                const V wi = i < n - 1 ? dir(i, i + 1) : V(), wo = dir(i, i - 1);
                aE = aE * local(EvaluateDirection(*x[i].prim, x[i].type, x[i].geom, wi, wo, true),
                                EvaluateDirectionPDF(*x[i].prim, x[i].type, x[i].geom, wi, wo));
            }
        }
        if (zero(aE)) return V();
        const V c = Cst(sl);
        return zero(c) ? V() : aL * c * aE;
    }

    double PDF(int sl) const {
        if (zero(Cst(sl))) return 0;
        const int n = (int)x.size(), t = n - sl;
        double pdf = 1;
        if (sl > 0) {
            pdf *= EvaluatePositionPDF(*x[0].prim) * EmitterPDF(s, *x[0].prim);
            for (int i = 0; i < sl - 1; ++i) {
                pdf *= EvaluateDirectionPDF(*x[i].prim, x[i].type, x[i].geom, i >= 1 ? dir(i, i - 1) : V(), dir(i, i + 1));
                pdf *= GeometryTerm(x[i].geom, x[i + 1].geom);
            }
        }
        if (t > 0) {
            pdf *= EvaluatePositionPDF(*x[n - 1].prim) * EmitterPDF(s, *x[n - 1].prim);
            for (int i = n - 1; i >= sl + 1; --i) {
                pdf *= EvaluateDirectionPDF(*x[i].prim, x[i].type, x[i].geom, i + 1 < n ? dir(i, i + 1) : V(), dir(i, i - 1));
                pdf *= GeometryTerm(x[i].geom, x[i - 1].geom);
            }
        }
        return pdf;
    }

    double MIS(int sl) const {
        const int n = (int)x.size();
        const double ps = PDF(sl);
        double inv = 0;
        for (int i = 0; i <= n; ++i) {
            const double pi = PDF(i);
            if (pi > 0) inv += (pi / ps) * (pi / ps);
        }
        return 1.0 / inv;
    }

    double Selection(int sl) const {
        const int t = (int)x.size() - sl;
        double p = 1;
        for (int i = 1; i < sl - 1; ++i) p *= rr_at(prm, i);
        for (int i = t - 2; i >= 1; --i) p *= rr_at(prm, i);
        return p;
    }
};

void sample_subpath(const Scene& s, const ngi_render_params& prm, uint32_t c0, uint32_t c1, bool le, int pixel, std::vector<PV>& out,
                    Counts& cnt) {
    out.clear();
    const int W = prm.width, H = prm.height;
    const int maxv = (prm.max_num_vertices < 0 || prm.max_num_vertices > kBdptMaxVertices) ? kBdptMaxVertices : prm.max_num_vertices;
    const uint32_t stream = le ? 1u : 0u;
    for (int step = 0; step < maxv; ++step) {
        const Q q = philox(c0, c1, (uint32_t)step, stream, prm.seed, prm.stream_id);
        PV v;
        if (step == 0) {
            if (le) {
                const int n = (int)s.lights.size();
                if (n == 0) return;
                const int li = std::min(std::max((int)(unif(q.v[0]) * n), 0), n - 1);
                v.prim = &s.prims[s.lights[li]];
                v.type = NGI_PRIM_L;
                SamplePosition(s, *v.prim, unif(q.v[1]), unif(q.v[2]), v.geom);
            } else {
                v.prim = &s.prims[s.sensor];
                v.type = NGI_PRIM_E;
                v.geom.degenerated = true;
                v.geom.p = v.prim->eye;
            }
            out.push_back(v);
            continue;
        }
        const PV& pv = out.back();
        const V wi = out.size() > 1 ? normalize(out[out.size() - 2].geom.p - pv.geom.p) : V();
        const double u0 = unif(q.v[0]), u1 = unif(q.v[1]), uc = unif(q.v[2]), urr = unif(q.v[3]);
        V wo;
        if (pv.type == NGI_PRIM_E && pixel >= 0) {
            // Per-pixel mode: the raster sample lies in this path's pixel.
            SampleDirection(*pv.prim, pv.type, pv.geom, wi, (pixel % W + u0) / W, (pixel / W + u1) / H, uc, wo);
        } else {
            SampleDirection(*pv.prim, pv.type, pv.geom, wi, u0, u1, uc, wo);
        }
        if (zero(EvaluateDirection(*pv.prim, pv.type, pv.geom, wi, wo, true, le))) break;
        Isect is;
        cnt.ext++;
        if (!Intersect(s, pv.geom.p, wo, is)) break;
        v.geom = is.geom;
        v.prim = &s.prims[is.prim];
        v.type = v.prim->p.type & ~(NGI_PRIM_L | NGI_PRIM_E);
        out.push_back(v);
        const double q_rr = rr_at(prm, step);
        if (q_rr < 1.0 && urr > q_rr) break;   // the vertex is kept (bdpt.hpp:118-123)
    }
}

void process_bdpt_sample(const Scene& s, const ngi_render_params& prm, uint64_t id, std::vector<double>& film, Counts& cnt) {
    const int W = prm.width, H = prm.height;
    const uint32_t npix = (uint32_t)W * (uint32_t)H;
    uint32_t c0, c1;
    int pixel = -1;
    if (prm.sampling == NGI_SAMPLING_PER_PIXEL) {
        pixel = (int)(id % npix);
        c0 = (uint32_t)pixel;
        c1 = (uint32_t)(id / npix);
    } else {
        c0 = (uint32_t)id;
        c1 = (uint32_t)(id >> 32) | 0x80000000u;
    }
    std::vector<PV> L, E;
    sample_subpath(s, prm, c0, c1, true, -1, L, cnt);
    sample_subpath(s, prm, c0, c1, false, pixel, E, cnt);
    const int nL = (int)L.size(), nE = (int)E.size();
    const int maxv = (prm.max_num_vertices < 0 || prm.max_num_vertices > 2 * kBdptMaxVertices) ? 2 * kBdptMaxVertices : prm.max_num_vertices;
    Bdpt b{s, prm, {}};
    for (int n = 2; n <= nE + nL; ++n) {
        if (n > maxv) continue;
        for (int sl = std::max(0, n - nE); sl <= std::min(nL, n); ++sl) {
            const int t = n - sl;
            b.x.clear();
            if (sl == 0) {
                if (!(E[t - 1].prim->p.type & NGI_PRIM_L)) continue;
                for (int i = t - 1; i >= 0; --i) b.x.push_back(E[i]);
                b.x.front().type = NGI_PRIM_L;
            } else if (t == 0) {
                if (!(L[sl - 1].prim->p.type & NGI_PRIM_E)) continue;
                for (int i = 0; i < sl; ++i) b.x.push_back(L[i]);
                b.x.back().type = NGI_PRIM_E;
            } else {
                if (!Visible(s, L[sl - 1].geom.p, E[t - 1].geom.p, &cnt.shadow)) continue;
                for (int i = 0; i < sl; ++i) b.x.push_back(L[i]);
                for (int i = t - 1; i >= 0; --i) b.x.push_back(E[i]);
            }
            const V Cu = b.Unweighted(sl);
            if (zero(Cu)) continue;
            const V C = Cu * (b.MIS(sl) / b.Selection(sl));
            double rx = 0, ry = 0;
            RasterPosition(*b.x[n - 1].prim, b.dir(n - 1, n - 2), rx, ry);
            const int index = PixelIndex(rx, ry, W, H);
            film[3 * (size_t)index] += C.x;
            film[3 * (size_t)index + 1] += C.y;
            film[3 * (size_t)index + 2] += C.z;
        }
    }
}

void sample_path(const Scene& s, const ngi_render_params& prm, uint64_t id, std::vector<double>& film, Counts& cnt) {
    if (prm.renderer == NGI_RENDERER_LT || prm.renderer == NGI_RENDERER_LTDIRECT) process_light_sample(s, prm, id, film, cnt);
    else if (prm.renderer == NGI_RENDERER_BDPT) process_bdpt_sample(s, prm, id, film, cnt);
    else process_sample(s, prm, id, film, cnt);
}

}  // namespace

extern "C" {

// Hits of rays against the scene (FP64), brute force or via the oracle's BVH.
int oracle_trace(const ngi_scene_desc* d, const ngi_ray* rays, size_t n, ngi_hit* hits, int any_hit, int use_bvh) {
    Scene s;
    int st = load(s, d, use_bvh != 0);
    if (st) return st;
    for (size_t i = 0; i < n; ++i) {
        const ngi_ray& r = rays[i];
        double t = 0, u = 0, v = 0;
        const long tri = intersect(s, V(r.o[0], r.o[1], r.o[2]), V(r.d[0], r.d[1], r.d[2]), r.tmin, r.tmax, any_hit != 0, t, u, v);
        hits[i].tri = tri < 0 ? NGI_MISS : (uint32_t)tri;
        hits[i].t = (float)t;
        hits[i].u = (float)u;
        hits[i].v = (float)v;
    }
    return NGI_OK;
}

// Render path ids [path_begin, path_begin + path_count) (per-pixel: id ->
// pixel id % (W*H), sample id / (W*H); uniform: global sample index) into
// film (W*H*3, unnormalised sums).  counts[0..1] += extension / shadow rays.
int oracle_render_paths(const ngi_scene_desc* d, const ngi_render_params* prm, uint64_t path_begin, uint64_t path_count, double* film,
                        int num_threads, long long* counts, int use_bvh) {
    if (!d || !prm || !film) return NGI_ERR_INVALID_ARG;
    if (prm->renderer < NGI_RENDERER_PT || prm->renderer > NGI_RENDERER_BDPT) return NGI_ERR_UNSUPPORTED;
    Scene s;
    int st = load(s, d, use_bvh != 0);
    if (st) return st;
    if (s.sensor < 0 || s.prims[s.sensor].p.sensor_type != NGI_SENSOR_PINHOLE) return NGI_ERR_UNSUPPORTED;
    const size_t np = (size_t)prm->width * prm->height * 3;
    int T = num_threads > 0 ? num_threads : (int)std::thread::hardware_concurrency();
    T = std::max(1, std::min<int>(T, (int)std::max<uint64_t>(1, path_count / 64)));
    std::vector<std::vector<double>> films(T, std::vector<double>(np, 0.0));
    std::vector<Counts> cnts(T);
    std::atomic<uint64_t> next(0);
    const uint64_t grain = 256;
    auto work = [&](int k) {
        for (;;) {
            const uint64_t b = next.fetch_add(grain);
            if (b >= path_count) break;
            const uint64_t e = std::min(path_count, b + grain);
            for (uint64_t i = b; i < e; ++i) sample_path(s, *prm, path_begin + i, films[k], cnts[k]);
        }
    };
    std::vector<std::thread> th;
    for (int k = 1; k < T; ++k) th.emplace_back(work, k);
    work(0);
    for (auto& t : th) t.join();
    for (int k = 0; k < T; ++k) {
        for (size_t i = 0; i < np; ++i) film[i] += films[k][i];
        if (counts) counts[0] += cnts[k].ext, counts[1] += cnts[k].shadow;
    }
    return NGI_OK;
}

// Full render, normalised by W*H/N as RenderProcess does (src/nanogi.cpp:429-437).
int oracle_render(const ngi_scene_desc* d, const ngi_render_params* prm, double* film, int num_threads) {
    const uint64_t npix = (uint64_t)prm->width * prm->height;
    const uint64_t total = prm->sampling == NGI_SAMPLING_PER_PIXEL ? npix * (uint64_t)prm->spp : (uint64_t)prm->num_samples;
    std::fill(film, film + 3 * npix, 0.0);
    const int st = oracle_render_paths(d, prm, 0, total, film, num_threads, nullptr, 1);
    if (st) return st;
    const double scale = (double)npix / (double)total;
    for (uint64_t i = 0; i < 3 * npix; ++i) film[i] *= scale;
    return NGI_OK;
}

// Philox stream as used by the renderers (for known-answer tests).
void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    const Q q = philox(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
    for (int k = 0; k < 4; ++k) out[k] = q.v[k];
}

}  // extern "C"

extern "C" {

// Test hook: SampleDirection / EvaluateDirectionPDF / EvaluateDirection of
// one primitive at a surface whose geometric and shading normals are +z
// (so local and world frames coincide).  out = {wo.x, wo.y, wo.z, pdf,
// fs.r, fs.g, fs.b}; force_degenerated selects the sampling-time branch and
// le the transport direction of EvaluateDirection (TransportDirection::LE).
void oracle_bsdf(const ngi_primitive* prim, int type, const double wi[3], const double u[3], const double wo_in[3], int use_wo_in,
                 int force_degenerated, int le, double out[7]) {
    Prim P;
    P.p = *prim;
    Geom g;
    g.degenerated = false;
    g.p = V();
    g.gn = g.sn = V(0, 0, 1);
    g.tangent_space();
    const V w(wi[0], wi[1], wi[2]);
    V wo;
    if (use_wo_in) wo = V(wo_in[0], wo_in[1], wo_in[2]);
    else SampleDirection(P, type, g, w, u[0], u[1], u[2], wo);
    const double pdf = EvaluateDirectionPDF(P, type, g, w, wo);
    const V fs = EvaluateDirection(P, type, g, w, wo, force_degenerated != 0, le != 0);
    out[0] = wo.x, out[1] = wo.y, out[2] = wo.z, out[3] = pdf, out[4] = fs.x, out[5] = fs.y, out[6] = fs.z;
}

}  // extern "C"

extern "C" {

// Loaded oracle scene (BVH built once; the desc arrays must outlive it).
void* oracle_scene_new(const ngi_scene_desc* d) {
    Scene* s = new Scene();
    if (load(*s, d, true)) {
        delete s;
        return nullptr;
    }
    return s;
}

void oracle_scene_free(void* h) { delete reinterpret_cast<Scene*>(h); }

// Pinhole sensor of a loaded scene: raster (u0, u1) -> SampleDirection
// (rt.hpp:733-740) -> out = {wo.xyz, RasterPosition rx, ry (rt.hpp:1348-1378),
// We (EvaluateDirection, rt.hpp:955-978), pdf (EvaluateDirectionPDF,
// rt.hpp:1189-1212)}.
int oracle_scene_camera(void* h, double u0, double u1, double out[7]) {
    const Scene& s = *reinterpret_cast<Scene*>(h);
    if (s.sensor < 0 || s.prims[s.sensor].p.sensor_type != NGI_SENSOR_PINHOLE) return NGI_ERR_UNSUPPORTED;
    const Prim& P = s.prims[s.sensor];
    Geom g;
    g.degenerated = true;
    g.p = P.eye;
    V wo;
    SampleDirection(P, NGI_PRIM_E, g, V(), u0, u1, 0.0, wo);
    double rx = -1.0, ry = -1.0;
    RasterPosition(P, wo, rx, ry);
    const V W = EvaluateDirection(P, NGI_PRIM_E, g, V(), wo, false);
    out[0] = wo.x, out[1] = wo.y, out[2] = wo.z, out[3] = rx, out[4] = ry, out[5] = W.x;
    out[6] = EvaluateDirectionPDF(P, NGI_PRIM_E, g, V(), wo);
    return NGI_OK;
}

// Light k (in scene order) of a loaded scene: SamplePosition (rt.hpp:483-562)
// at (ux, uy) -> out = {p.xyz, gn.xyz, EvaluatePositionPDF (rt.hpp:645-662),
// emitter selection pdf (rt.hpp:2338-2352)}.
int oracle_scene_light(void* h, int k, double ux, double uy, double out[8]) {
    const Scene& s = *reinterpret_cast<Scene*>(h);
    if (k < 0 || k >= (int)s.lights.size()) return NGI_ERR_INVALID_ARG;
    const Prim& P = s.prims[s.lights[k]];
    Geom g;
    SamplePosition(s, P, ux, uy, g);
    out[0] = g.p.x, out[1] = g.p.y, out[2] = g.p.z, out[3] = g.gn.x, out[4] = g.gn.y, out[5] = g.gn.z;
    out[6] = EvaluatePositionPDF(P);
    out[7] = EmitterPDF(s, P);
    return NGI_OK;
}

// As oracle_render_paths, on a loaded scene: only the sample loop is timed
// by callers (the region the reference times, src/nanogi.cpp:201-216).
int oracle_scene_render_paths(void* h, const ngi_render_params* prm, uint64_t path_begin, uint64_t path_count, double* film,
                              int num_threads, long long* counts) {
    const Scene& s = *reinterpret_cast<Scene*>(h);
    if (s.sensor < 0 || s.prims[s.sensor].p.sensor_type != NGI_SENSOR_PINHOLE) return NGI_ERR_UNSUPPORTED;
    const size_t np = (size_t)prm->width * prm->height * 3;
    int T = num_threads > 0 ? num_threads : (int)std::thread::hardware_concurrency();
    T = std::max(1, std::min<int>(T, (int)std::max<uint64_t>(1, path_count / 64)));
    std::vector<std::vector<double>> films(T, std::vector<double>(np, 0.0));
    std::vector<Counts> cnts(T);
    std::atomic<uint64_t> next(0);
    auto work = [&](int k) {
        for (;;) {
            const uint64_t b = next.fetch_add(256);
            if (b >= path_count) break;
            const uint64_t e = std::min(path_count, b + 256);
            for (uint64_t i = b; i < e; ++i) sample_path(s, *prm, path_begin + i, films[k], cnts[k]);
        }
    };
    std::vector<std::thread> th;
    for (int k = 1; k < T; ++k) th.emplace_back(work, k);
    work(0);
    for (auto& t : th) t.join();
    for (int k = 0; k < T; ++k) {
        for (size_t i = 0; i < np; ++i) film[i] += films[k][i];
        if (counts) counts[0] += cnts[k].ext, counts[1] += cnts[k].shadow;
    }
    return NGI_OK;
}

}  // extern "C"

extern "C" {

// FP64 reference gradient of L = sum_p dLdI_p . I_p with respect to each
// primitive's reflectance (D.R or G.R) and radiance (Le), by replaying the
// same paths as oracle_render (see ngi_adjoint.hip for the estimator).
int oracle_render_adjoint(const ngi_scene_desc* d, const ngi_render_params* prm, const double* dLdI, double* gR, double* gLe, int num_threads) {
    if (!d || !prm || !dLdI || !gR || !gLe) return NGI_ERR_INVALID_ARG;
    if (prm->renderer != NGI_RENDERER_PT && prm->renderer != NGI_RENDERER_PTDIRECT) return NGI_ERR_UNSUPPORTED;
    Scene s;
    int st = load(s, d, true);
    if (st) return st;
    if (s.sensor < 0) return NGI_ERR_UNSUPPORTED;
    const uint64_t npix = (uint64_t)prm->width * prm->height;
    const uint64_t total = prm->sampling == NGI_SAMPLING_PER_PIXEL ? npix * (uint64_t)prm->spp : (uint64_t)prm->num_samples;
    const size_t np = d->num_primitives;
    // Replay scene: zero reflectance / radiance channels set to 1 (path
    // sampling does not depend on them); the true values kept aside.
    std::vector<double> Rt(3 * np, 0.0), Let(3 * np, 0.0);
    for (size_t i = 0; i < np; ++i) {
        ngi_primitive& P = s.prims[i].p;
        double* R = (P.type & NGI_PRIM_D) ? P.D_R : ((P.type & NGI_PRIM_G) ? P.G_R : nullptr);
        for (int k = 0; k < 3; ++k) {
            if (R) {
                Rt[3 * i + k] = R[k];
                if (R[k] == 0.0) R[k] = 1.0;
            }
            if (P.type & NGI_PRIM_L) {
                Let[3 * i + k] = P.Le[k];
                if (P.Le[k] == 0.0) P.Le[k] = 1.0;
            }
        }
    }
    int T = num_threads > 0 ? num_threads : (int)std::thread::hardware_concurrency();
    T = std::max(1, T);
    std::vector<std::vector<double>> R(T, std::vector<double>(3 * np, 0.0)), L(T, std::vector<double>(3 * np, 0.0));
    std::atomic<uint64_t> next(0);
    auto work = [&](int k) {
        Adjoint adj;
        adj.dLdI = dLdI;
        adj.scale = (double)npix / (double)total;
        adj.gR = &R[k];
        adj.gLe = &L[k];
        adj.Rt = &Rt;
        adj.Let = &Let;
        std::vector<double> dummy;
        Counts cnt;
        for (;;) {
            const uint64_t b = next.fetch_add(256);
            if (b >= total) break;
            for (uint64_t i = b; i < std::min(total, b + 256); ++i) process_sample(s, *prm, i, dummy, cnt, &adj);
        }
    };
    std::vector<std::thread> th;
    for (int k = 1; k < T; ++k) th.emplace_back(work, k);
    work(0);
    for (auto& t : th) t.join();
    for (size_t i = 0; i < 3 * np; ++i) {
        gR[i] = 0;
        gLe[i] = 0;
        for (int k = 0; k < T; ++k) gR[i] += R[k][i], gLe[i] += L[k][i];
    }
    return NGI_OK;
}

}  // extern "C"
