// Seeded procedural scenes for the benchmark workloads.  The reference has
// no procedural scenes; these produce the same data model Scene::Load
// builds (include/nanogi/rt.hpp:1519-2154): one primitive per material or
// emitter, Lambertian (D), Beckmann-conductor glossy (G), area lights (L)
// and one pinhole camera (E).
//
// Every position and normal is rounded to the nearest float, so the FP32
// device path and the FP64 oracle see bit-identical geometry.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

#include "ngi_internal.h"

namespace ngi {
namespace {

inline double qf(double v) { return (double)(float)v; }

struct Rng64 {  // splitmix64 stream for scene parameters
    uint64_t s;
    explicit Rng64(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x1234567ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double range(double a, double b) { return a + (b - a) * uniform(); }
};

// 3-D value noise: hashed lattice values in [-1,1], quintic-fade trilinear
// interpolation; fbm() sums octaves with lacunarity 2 and gain 1/2,
// normalised to [-1,1].
struct ValueNoise {
    uint64_t seed;
    double lattice(int64_t x, int64_t y, int64_t z) const {
        uint64_t h = seed ^ ((uint64_t)x * 0x8CB92BA72F3D8DD7ull) ^ ((uint64_t)y * 0xC2B2AE3D27D4EB4Full) ^ ((uint64_t)z * 0x165667B19E3779F9ull);
        h ^= h >> 33;
        h *= 0xFF51AFD7ED558CCDull;
        h ^= h >> 33;
        h *= 0xC4CEB9FE1A85EC53ull;
        h ^= h >> 33;
        return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
    static double fade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
    double noise(double x, double y, double z) const {
        const double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
        const int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
        const double tx = fade(x - fx), ty = fade(y - fy), tz = fade(z - fz);
        double c[2][2][2];
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d);
        auto lerp = [](double a, double b, double t) { return a + (b - a) * t; };
        double x00 = lerp(c[0][0][0], c[1][0][0], tx), x10 = lerp(c[0][1][0], c[1][1][0], tx);
        double x01 = lerp(c[0][0][1], c[1][0][1], tx), x11 = lerp(c[0][1][1], c[1][1][1], tx);
        return lerp(lerp(x00, x10, ty), lerp(x01, x11, ty), tz);
    }
    double fbm(double x, double y, double z, int octaves) const {
        double sum = 0, amp = 1, norm = 0, f = 1;
        for (int o = 0; o < octaves; ++o) {
            sum += amp * noise(x * f, y * f, z * f);
            norm += amp;
            amp *= 0.5;
            f *= 2.0;
        }
        return sum / norm;
    }
};

struct Builder {
    HostScene& s;
    int cur = -1;
    explicit Builder(HostScene& sc) : s(sc) {}

    static ngi_primitive blank() {
        ngi_primitive p;
        memset(&p, 0, sizeof(p));
        return p;
    }
    int begin(ngi_primitive p) {
        p.tri_begin = (uint32_t)(s.faces.size() / 3);
        p.tri_count = 0;
        s.prims.push_back(p);
        cur = (int)s.prims.size() - 1;
        return cur;
    }
    void end() {
        ngi_primitive& p = s.prims[cur];
        p.tri_count = (uint32_t)(s.faces.size() / 3) - p.tri_begin;
        cur = -1;
    }
    uint32_t vertex(d3 p, d3 n) {
        const double l = length(n);
        if (l > 0) n = n / l;
        s.positions.insert(s.positions.end(), {qf(p.x), qf(p.y), qf(p.z)});
        s.normals.insert(s.normals.end(), {qf(n.x), qf(n.y), qf(n.z)});
        return (uint32_t)(s.positions.size() / 3 - 1);
    }
    d3 pos(uint32_t i) const { return d3(s.positions[3 * i], s.positions[3 * i + 1], s.positions[3 * i + 2]); }
    void tri(uint32_t a, uint32_t b, uint32_t c) {
        s.faces.insert(s.faces.end(), {a, b, c});
        s.tri_prim.push_back((uint32_t)cur);
    }
    // Triangle whose winding normal points along `hint`.
    void tri_oriented(uint32_t a, uint32_t b, uint32_t c, d3 hint) {
        if (dot(cross(pos(b) - pos(a), pos(c) - pos(a)), hint) < 0) std::swap(b, c);
        tri(a, b, c);
    }
    // Flat rectangle: centre c, unit normal n, unit in-plane axis a, half
    // extents ha (along a) and hb (along n x a); winding normal = n.
    void rect(d3 c, d3 n, d3 a, double ha, double hb) {
        n = normalize(n);
        a = normalize(a);
        d3 b = cross(n, a);
        uint32_t v0 = vertex(c - a * ha - b * hb, n), v1 = vertex(c + a * ha - b * hb, n);
        uint32_t v2 = vertex(c + a * ha + b * hb, n), v3 = vertex(c - a * ha + b * hb, n);
        tri_oriented(v0, v1, v2, n);
        tri_oriented(v0, v2, v3, n);
    }
    // Grid of nx x nz quads on a flat rectangle (same orientation rule).
    void grid(d3 c, d3 n, d3 a, double ha, double hb, int na, int nb) {
        n = normalize(n);
        a = normalize(a);
        d3 b = cross(n, a);
        std::vector<uint32_t> idx((size_t)(na + 1) * (nb + 1));
        for (int j = 0; j <= nb; ++j)
            for (int i = 0; i <= na; ++i)
                idx[(size_t)j * (na + 1) + i] = vertex(c + a * (ha * (2.0 * i / na - 1)) + b * (hb * (2.0 * j / nb - 1)), n);
        for (int j = 0; j < nb; ++j)
            for (int i = 0; i < na; ++i) {
                uint32_t v00 = idx[(size_t)j * (na + 1) + i], v10 = idx[(size_t)j * (na + 1) + i + 1];
                uint32_t v01 = idx[(size_t)(j + 1) * (na + 1) + i], v11 = idx[(size_t)(j + 1) * (na + 1) + i + 1];
                tri_oriented(v00, v10, v11, n);
                tri_oriented(v00, v11, v01, n);
            }
    }
    // Box standing on y = base_y, centred at (cx, cz), rotated about +y by
    // `deg` degrees; outward normals, 12 triangles.
    void box(double cx, double base_y, double cz, double hx, double hy, double hz, double deg) {
        const double t = deg * M_PI / 180.0;
        d3 ex(std::cos(t), 0, -std::sin(t)), ey(0, 1, 0), ez(std::sin(t), 0, std::cos(t));
        d3 c(cx, base_y + hy, cz);
        rect(c + ex * hx, ex, ez, hz, hy);
        rect(c - ex * hx, -ex, ez, hz, hy);
        rect(c + ey * hy, ey, ex, hx, hz);
        rect(c - ey * hy, -ey, ex, hx, hz);
        rect(c + ez * hz, ez, ex, hx, hy);
        rect(c - ez * hz, -ez, ex, hx, hy);
    }
    // Open cylinder side (smooth radial normals) plus flat top cap.
    void column(d3 base, double r, double h, int seg, int rings) {
        std::vector<uint32_t> idx((size_t)(seg + 1) * (rings + 1));
        for (int j = 0; j <= rings; ++j)
            for (int i = 0; i <= seg; ++i) {
                const double ph = 2 * M_PI * i / seg;
                d3 n(std::cos(ph), 0, std::sin(ph));
                idx[(size_t)j * (seg + 1) + i] = vertex(base + n * r + d3(0, h * j / rings, 0), n);
            }
        for (int j = 0; j < rings; ++j)
            for (int i = 0; i < seg; ++i) {
                uint32_t a = idx[(size_t)j * (seg + 1) + i], b = idx[(size_t)j * (seg + 1) + i + 1];
                uint32_t c = idx[(size_t)(j + 1) * (seg + 1) + i + 1], d = idx[(size_t)(j + 1) * (seg + 1) + i];
                const double ph = 2 * M_PI * (i + 0.5) / seg;
                d3 hint(std::cos(ph), 0, std::sin(ph));
                tri_oriented(a, b, c, hint);
                tri_oriented(a, c, d, hint);
            }
        d3 up(0, 1, 0);
        uint32_t ctr = vertex(base + d3(0, h, 0), up);
        std::vector<uint32_t> rim(seg);
        for (int i = 0; i < seg; ++i) {
            const double ph = 2 * M_PI * i / seg;
            rim[i] = vertex(base + d3(r * std::cos(ph), h, r * std::sin(ph)), up);
        }
        for (int i = 0; i < seg; ++i) tri_oriented(ctr, rim[i], rim[(i + 1) % seg], up);
    }
    // Half torus: centre line c + R(cos(phi) e1 + sin(phi) e2), phi in [0,pi],
    // tube radius rho; smooth normals.
    void half_torus(d3 c, d3 e1, d3 e2, double R, double rho, int seg, int tube) {
        d3 bn = normalize(cross(e1, e2));
        std::vector<uint32_t> idx((size_t)(seg + 1) * tube);
        std::vector<d3> nrm((size_t)(seg + 1) * tube);
        for (int i = 0; i <= seg; ++i) {
            const double ph = M_PI * i / seg;
            d3 rad = e1 * std::cos(ph) + e2 * std::sin(ph);
            d3 ctr = c + rad * R;
            for (int j = 0; j < tube; ++j) {
                const double ps = 2 * M_PI * j / tube;
                d3 n = rad * std::cos(ps) + bn * std::sin(ps);
                idx[(size_t)i * tube + j] = vertex(ctr + n * rho, n);
                nrm[(size_t)i * tube + j] = n;
            }
        }
        for (int i = 0; i < seg; ++i)
            for (int j = 0; j < tube; ++j) {
                const size_t a = (size_t)i * tube + j, b = (size_t)(i + 1) * tube + j;
                const size_t c2 = (size_t)(i + 1) * tube + (j + 1) % tube, d = (size_t)i * tube + (j + 1) % tube;
                d3 hint = nrm[a] + nrm[b] + nrm[c2] + nrm[d];
                tri_oriented(idx[a], idx[b], idx[c2], hint);
                tri_oriented(idx[a], idx[c2], idx[d], hint);
            }
    }
};

// Icosphere by recursive 4-way subdivision, emitted depth-first so that each
// aligned run of 4^k triangles is one level-(levels-k) ancestor triangle.
// radius_fn maps a unit direction to a radius; normals are area-weighted
// averages of the displaced faces (smooth shading).
template <typename RadiusFn>
void icosphere(Builder& B, d3 center, int levels, RadiusFn radius_fn, std::vector<uint32_t>* patch_starts = nullptr) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<d3> dirs = {d3(-1, t, 0), d3(1, t, 0), d3(-1, -t, 0), d3(1, -t, 0), d3(0, -1, t), d3(0, 1, t),
                            d3(0, -1, -t), d3(0, 1, -t), d3(t, 0, -1), d3(t, 0, 1), d3(-t, 0, -1), d3(-t, 0, 1)};
    for (auto& d : dirs) d = normalize(d);
    const int base[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                             {11, 10, 2}, {10, 7, 6}, {7, 1, 8},   {3, 9, 4},  {3, 4, 2},   {3, 2, 6}, {3, 6, 8},
                             {3, 8, 9},  {4, 9, 5},  {2, 4, 11},  {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
    std::unordered_map<uint64_t, uint32_t> mid;
    mid.reserve((size_t)30 << (2 * levels));
    auto midpoint = [&](uint32_t a, uint32_t b) -> uint32_t {
        const uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
        auto it = mid.find(key);
        if (it != mid.end()) return it->second;
        dirs.push_back(normalize(dirs[a] + dirs[b]));
        const uint32_t id = (uint32_t)dirs.size() - 1;
        mid.emplace(key, id);
        return id;
    };
    std::vector<uint32_t> tris;
    tris.reserve((size_t)60 << (2 * levels));
    struct Frame { uint32_t a, b, c; int level; };
    std::vector<Frame> stack;
    for (int f = 19; f >= 0; --f) stack.push_back({(uint32_t)base[f][0], (uint32_t)base[f][1], (uint32_t)base[f][2], 0});
    while (!stack.empty()) {
        Frame fr = stack.back();
        stack.pop_back();
        if (fr.level == levels) {
            tris.insert(tris.end(), {fr.a, fr.b, fr.c});
            continue;
        }
        uint32_t ab = midpoint(fr.a, fr.b), bc = midpoint(fr.b, fr.c), ca = midpoint(fr.c, fr.a);
        const int l = fr.level + 1;
        // Pushed in reverse so children are emitted in order a, b, c, centre.
        stack.push_back({ab, bc, ca, l});
        stack.push_back({ca, bc, fr.c, l});
        stack.push_back({ab, fr.b, bc, l});
        stack.push_back({fr.a, ab, ca, l});
    }
    // Displaced positions (float-rounded) and smooth normals.
    std::vector<d3> P(dirs.size()), N(dirs.size(), d3(0, 0, 0));
    for (size_t i = 0; i < dirs.size(); ++i) {
        d3 p = center + dirs[i] * radius_fn(dirs[i]);
        P[i] = d3(qf(p.x), qf(p.y), qf(p.z));
    }
    for (size_t k = 0; k < tris.size(); k += 3) {
        d3 n = cross(P[tris[k + 1]] - P[tris[k]], P[tris[k + 2]] - P[tris[k]]);
        for (int j = 0; j < 3; ++j) N[tris[k + j]] = N[tris[k + j]] + n;
    }
    const uint32_t voff = (uint32_t)(B.s.positions.size() / 3);
    for (size_t i = 0; i < dirs.size(); ++i) B.vertex(P[i], length(N[i]) > 0 ? N[i] : dirs[i]);
    (void)patch_starts;
    for (size_t k = 0; k < tris.size(); k += 3) B.tri(voff + tris[k], voff + tris[k + 1], voff + tris[k + 2]);
}

ngi_primitive diffuse(double r, double g, double b) {
    ngi_primitive p = Builder::blank();
    p.type = NGI_PRIM_D;
    p.D_R[0] = r, p.D_R[1] = g, p.D_R[2] = b;
    return p;
}

ngi_primitive area_light(double r, double g, double b) {
    ngi_primitive p = Builder::blank();
    p.type = NGI_PRIM_L;
    p.light_type = NGI_LIGHT_AREA;
    p.Le[0] = r, p.Le[1] = g, p.Le[2] = b;
    return p;
}

void add_camera(Builder& B, d3 eye, d3 center, double fov) {
    ngi_primitive p = Builder::blank();
    p.type = NGI_PRIM_E;
    p.sensor_type = NGI_SENSOR_PINHOLE;
    p.E_We[0] = p.E_We[1] = p.E_We[2] = 1.0;
    p.E_eye[0] = eye.x, p.E_eye[1] = eye.y, p.E_eye[2] = eye.z;
    p.E_center[0] = center.x, p.E_center[1] = center.y, p.E_center[2] = center.z;
    p.E_up[1] = 1.0;
    p.E_fov_deg = fov;
    B.begin(p);
    B.end();
}

// Unit Cornell-style room [0,1]^3 open towards +z: 5 walls (10 triangles).
void cornell_room(Builder& B, const double left[3], const double right[3], const double grey[3]) {
    const d3 X(1, 0, 0), Y(0, 1, 0), Z(0, 0, 1);
    B.begin(diffuse(grey[0], grey[1], grey[2]));   // floor
    B.rect(d3(0.5, 0, 0.5), Y, X, 0.5, 0.5);
    B.end();
    B.begin(diffuse(grey[0], grey[1], grey[2]));   // ceiling
    B.rect(d3(0.5, 1, 0.5), -Y, X, 0.5, 0.5);
    B.end();
    B.begin(diffuse(grey[0], grey[1], grey[2]));   // back
    B.rect(d3(0.5, 0.5, 0), Z, X, 0.5, 0.5);
    B.end();
    B.begin(diffuse(left[0], left[1], left[2]));   // left (x = 0)
    B.rect(d3(0, 0.5, 0.5), X, Z, 0.5, 0.5);
    B.end();
    B.begin(diffuse(right[0], right[1], right[2]));  // right (x = 1)
    B.rect(d3(1, 0.5, 0.5), -X, Z, 0.5, 0.5);
    B.end();
}

// Cornell-style box: 10 room + 2x12 box + 2 light triangles = 36.
int gen_cornell(HostScene& s) {
    Builder B(s);
    const double left[3] = {0.63, 0.065, 0.05}, right[3] = {0.14, 0.45, 0.09}, grey[3] = {0.73, 0.73, 0.73};
    cornell_room(B, left, right, grey);
    B.begin(diffuse(0.73, 0.73, 0.73));
    B.box(0.33, 0.0, 0.37, 0.15, 0.30, 0.15, 15.0);    // tall box
    B.end();
    B.begin(diffuse(0.73, 0.73, 0.73));
    B.box(0.67, 0.0, 0.65, 0.15, 0.15, 0.15, -18.0);   // short box
    B.end();
    B.begin(area_light(17, 12, 4));                     // 0.25 x 0.25 quad, facing down
    B.rect(d3(0.5, 0.999, 0.5), d3(0, -1, 0), d3(1, 0, 0), 0.125, 0.125);
    B.end();
    add_camera(B, d3(0.5, 0.5, 2.2), d3(0.5, 0.5, 0.0), 45.0);
    return NGI_OK;
}

// Nave interior, 40 x 12 x 16 m: 200x80 floor grid, 18 columns of
// 32-segment cylinders, 16 longitudinal and 9 transverse half-torus arches,
// plinths, walls and ceiling; six seeded albedos in [0.2, 0.8]; four 2x2 m
// quad lights of radiance 20.
int gen_nave(HostScene& s, uint64_t seed) {
    Builder B(s);
    Rng64 rng(seed);
    double alb[6][3];
    for (auto& a : alb)
        for (double& c : a) c = rng.range(0.2, 0.8);
    const d3 X(1, 0, 0), Y(0, 1, 0), Z(0, 0, 1);
    B.begin(diffuse(alb[0][0], alb[0][1], alb[0][2]));   // floor tiles
    B.grid(d3(0, 0, 0), Y, X, 20.0, 8.0, 200, 80);
    B.end();
    B.begin(diffuse(alb[1][0], alb[1][1], alb[1][2]));   // ceiling
    B.rect(d3(0, 12, 0), -Y, X, 20.0, 8.0);
    B.end();
    B.begin(diffuse(alb[2][0], alb[2][1], alb[2][2]));   // walls
    B.rect(d3(0, 6, -8), Z, X, 20.0, 6.0);
    B.rect(d3(0, 6, 8), -Z, X, 20.0, 6.0);
    B.rect(d3(-20, 6, 0), X, Z, 8.0, 6.0);
    B.rect(d3(20, 6, 0), -X, Z, 8.0, 6.0);
    B.end();
    B.begin(diffuse(alb[3][0], alb[3][1], alb[3][2]));   // columns
    for (int side = -1; side <= 1; side += 2)
        for (int i = 0; i < 9; ++i) B.column(d3(-16.0 + 4.0 * i, 0.5, 5.0 * side), 0.4, 6.5, 32, 12);
    B.end();
    B.begin(diffuse(alb[4][0], alb[4][1], alb[4][2]));   // arches
    for (int side = -1; side <= 1; side += 2)
        for (int i = 0; i < 8; ++i) B.half_torus(d3(-14.0 + 4.0 * i, 7.0, 5.0 * side), X, Y, 2.0, 0.3, 32, 16);
    for (int i = 0; i < 9; ++i) B.half_torus(d3(-16.0 + 4.0 * i, 7.0, 0.0), Z, Y, 5.0, 0.3, 48, 16);
    B.end();
    B.begin(diffuse(alb[5][0], alb[5][1], alb[5][2]));   // plinths
    for (int side = -1; side <= 1; side += 2)
        for (int i = 0; i < 9; ++i) B.box(-16.0 + 4.0 * i, 0.0, 5.0 * side, 0.5, 0.25, 0.5, 0.0);
    B.end();
    for (int i = 0; i < 4; ++i) {
        B.begin(area_light(20, 20, 20));
        B.rect(d3(-15.0 + 10.0 * i, 10.0, 0.0), -Y, X, 1.0, 1.0);
        B.end();
    }
    add_camera(B, d3(-18.5, 2.0, 0.0), d3(10.0, 5.0, 0.0), 60.0);
    return NGI_OK;
}

// Displaced icosphere (8 subdivisions, 1,310,720 triangles; 4-octave value
// noise of amplitude 0.1 r) on a 512x512-vertex heightfield (522,242
// triangles).  Each aligned run of 64 sphere triangles is one glossy
// primitive with Phong exponent n ~ logU[10,1000] mapped to Beckmann
// alpha = sqrt(2/(n+2)) (Walter et al. 2007).  Lights: one icosphere area
// light and 16 small quads.
int gen_sphere_field(HostScene& s, uint64_t seed, int levels) {
    Builder B(s);
    Rng64 rng(seed);
    ValueNoise noise{seed * 7919 + 17};
    const double R = 1.0;
    const d3 C(0, R, 0);
    // Sphere geometry first into a temporary primitive, then split into patches.
    ngi_primitive g = Builder::blank();
    g.type = NGI_PRIM_G;
    const int first = B.begin(g);
    icosphere(B, C, levels, [&](d3 d) { return R * (1.0 + 0.1 * noise.fbm(2.0 * d.x + 11.3, 2.0 * d.y + 5.7, 2.0 * d.z + 3.1, 4)); });
    B.end();
    const uint32_t sphere_begin = s.prims[first].tri_begin, sphere_count = s.prims[first].tri_count;
    s.prims.pop_back();
    const uint32_t patch = 64;
    const double eta[3] = {1.657, 0.880, 0.521}, kk[3] = {9.224, 6.270, 4.837};   // aluminium
    for (uint32_t t0 = 0; t0 < sphere_count; t0 += patch) {
        ngi_primitive p = Builder::blank();
        p.type = NGI_PRIM_G;
        const double n = std::exp(rng.range(std::log(10.0), std::log(1000.0)));
        p.G_roughness = std::sqrt(2.0 / (n + 2.0));
        for (int k = 0; k < 3; ++k) {
            p.G_R[k] = rng.range(0.6, 1.0);
            p.G_eta[k] = eta[k];
            p.G_k[k] = kk[k];
        }
        p.tri_begin = sphere_begin + t0;
        p.tri_count = std::min(patch, sphere_count - t0);
        const uint32_t id = (uint32_t)s.prims.size();
        for (uint32_t t = p.tri_begin; t < p.tri_begin + p.tri_count; ++t) s.tri_prim[t] = id;
        s.prims.push_back(p);
    }
    // Heightfield plane.
    B.begin(diffuse(0.5, 0.5, 0.5));
    {
        // 512^2 vertices at full detail; reduced sphere detail (test scenes)
        // uses a 64^2 grid over the same extent.
        const int nv = levels >= 8 ? 512 : 64;
        const double ext = 6.0;
        auto h = [&](double x, double z) { return 0.05 * noise.fbm(0.8 * x + 101.0, 0.0, 0.8 * z + 37.0, 2); };
        std::vector<d3> P((size_t)nv * nv);
        for (int j = 0; j < nv; ++j)
            for (int i = 0; i < nv; ++i) {
                const double x = -ext + 2 * ext * i / (nv - 1), z = -ext + 2 * ext * j / (nv - 1);
                P[(size_t)j * nv + i] = d3(qf(x), qf(h(x, z)), qf(z));
            }
        std::vector<d3> N(P.size(), d3(0, 0, 0));
        auto quad_tris = [&](int i, int j, uint32_t out[6]) {
            uint32_t a = j * nv + i, b = j * nv + i + 1, c = (j + 1) * nv + i + 1, d = (j + 1) * nv + i;
            // (a, d, c) and (a, c, b) wind towards +y for x right, z forward.
            out[0] = a, out[1] = d, out[2] = c, out[3] = a, out[4] = c, out[5] = b;
        };
        for (int j = 0; j + 1 < nv; ++j)
            for (int i = 0; i + 1 < nv; ++i) {
                uint32_t t[6];
                quad_tris(i, j, t);
                for (int k = 0; k < 6; k += 3) {
                    d3 n = cross(P[t[k + 1]] - P[t[k]], P[t[k + 2]] - P[t[k]]);
                    for (int m = 0; m < 3; ++m) N[t[k + m]] = N[t[k + m]] + n;
                }
            }
        const uint32_t voff = (uint32_t)(s.positions.size() / 3);
        for (size_t v = 0; v < P.size(); ++v) B.vertex(P[v], N[v]);
        for (int j = 0; j + 1 < nv; ++j)
            for (int i = 0; i + 1 < nv; ++i) {
                uint32_t t[6];
                quad_tris(i, j, t);
                B.tri(voff + t[0], voff + t[1], voff + t[2]);
                B.tri(voff + t[3], voff + t[4], voff + t[5]);
            }
    }
    B.end();
    // Spherical area light.
    B.begin(area_light(12.0, 11.0, 9.0));
    icosphere(B, d3(-1.8, 2.6, 1.2), 2, [](d3) { return 0.35; });
    B.end();
    // Ring of 16 small quads facing the sphere.
    for (int i = 0; i < 16; ++i) {
        const double a = 2 * M_PI * (i + 0.5) / 16;
        d3 c(3.2 * std::cos(a), 2.2, 3.2 * std::sin(a));
        d3 n = normalize(C - c);
        d3 ax = normalize(cross(d3(0, 1, 0), n));
        const double hue = (double)i / 16.0;
        B.begin(area_light(6.0 + 4.0 * std::cos(2 * M_PI * hue), 6.0 + 4.0 * std::cos(2 * M_PI * (hue - 1.0 / 3)),
                           6.0 + 4.0 * std::cos(2 * M_PI * (hue - 2.0 / 3))));
        B.rect(c, n, ax, 0.125, 0.125);
        B.end();
    }
    add_camera(B, d3(0.0, 2.2, 5.5), d3(0.0, 0.9, 0.0), 40.0);
    return NGI_OK;
}

// Cornell-style room plus seeded boxes and icospheres until about `target`
// triangles (10 room + 2 light + content).
int gen_batch_room(HostScene& s, uint64_t seed, int target) {
    Builder B(s);
    Rng64 rng(seed);
    double left[3], right[3], grey[3];
    for (int k = 0; k < 3; ++k) left[k] = rng.range(0.2, 0.8), right[k] = rng.range(0.2, 0.8), grey[k] = rng.range(0.5, 0.8);
    cornell_room(B, left, right, grey);
    B.begin(area_light(rng.range(10, 20), rng.range(10, 20), rng.range(10, 20)));
    B.rect(d3(0.5, 0.999, 0.5), d3(0, -1, 0), d3(1, 0, 0), 0.15, 0.15);
    B.end();
    int count = 12;
    while (count + 12 <= target) {
        const int remaining = target - count;
        B.begin(diffuse(rng.range(0.2, 0.9), rng.range(0.2, 0.9), rng.range(0.2, 0.9)));
        const double cx = rng.range(0.15, 0.85), cz = rng.range(0.15, 0.85);
        if (remaining >= 80 && rng.uniform() < 0.5) {
            int lv = remaining >= 1280 ? 3 : (remaining >= 320 ? 2 : 1);
            const double r = rng.range(0.04, 0.12);
            icosphere(B, d3(cx, rng.range(r, 0.8), cz), lv, [r](d3) { return r; });
            count += 20 << (2 * lv);
        } else {
            B.box(cx, 0.0, cz, rng.range(0.02, 0.1), rng.range(0.02, 0.2), rng.range(0.02, 0.1), rng.range(-45, 45));
            count += 12;
        }
        B.end();
    }
    add_camera(B, d3(0.5, 0.5, 2.2), d3(0.5, 0.5, 0.0), 45.0);
    return NGI_OK;
}

}  // namespace

int generate_scene(int kind, uint64_t seed, double aspect, int detail, HostScene& s) {
    s = HostScene();
    s.aspect = aspect;
    switch (kind) {
        case NGI_GEN_CORNELL: return gen_cornell(s);
        case NGI_GEN_NAVE: return gen_nave(s, seed);
        case NGI_GEN_SPHERE_FIELD: return gen_sphere_field(s, seed, detail > 0 ? std::min(detail, 9) : 8);
        case NGI_GEN_BATCH_ROOM: return gen_batch_room(s, seed, detail > 0 ? detail : 1000);
        default: return fail(NGI_ERR_INVALID_ARG, "unknown generator %d", kind);
    }
}

}  // namespace ngi
