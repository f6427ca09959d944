// Scene creation, validation and YAML ingest/export.
//
// ngi_scene_load_yaml restates Scene::Load (include/nanogi/rt.hpp:1519-2154):
// version check (1535-1540), primitive type bits (1587-1616), mesh ingest
// (1620-1738), per-type parameters (1806-2040) and the directional-light
// post-configuration (2059-2079).  The Embree build (2087-2143) is replaced
// by ngi_scene_build.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <sys/stat.h>

#include "ngi_internal.h"

using namespace ngi;

namespace ngi {

// Flat normals for scenes given without normals: every triangle gets its
// own three vertices carrying the winding normal (Assimp GenNormals).
void ensure_normals(HostScene& s) {
    if (!s.normals.empty()) return;
    const size_t nt = s.faces.size() / 3;
    std::vector<double> P, N, T;
    std::vector<uint32_t> F;
    P.reserve(nt * 9);
    N.reserve(nt * 9);
    const bool tex = !s.texcoords.empty();
    for (size_t t = 0; t < nt; ++t) {
        uint32_t i[3] = {s.faces[3 * t], s.faces[3 * t + 1], s.faces[3 * t + 2]};
        d3 p[3];
        for (int k = 0; k < 3; ++k) p[k] = d3(s.positions[3 * i[k]], s.positions[3 * i[k] + 1], s.positions[3 * i[k] + 2]);
        d3 n = cross(p[1] - p[0], p[2] - p[0]);
        double l = length(n);
        n = l > 0 ? n / l : d3(0, 0, 1);
        for (int k = 0; k < 3; ++k) {
            F.push_back((uint32_t)(P.size() / 3));
            P.insert(P.end(), {p[k].x, p[k].y, p[k].z});
            N.insert(N.end(), {n.x, n.y, n.z});
            if (tex) T.insert(T.end(), {s.texcoords[2 * i[k]], s.texcoords[2 * i[k] + 1]});
        }
    }
    s.positions.swap(P);
    s.normals.swap(N);
    s.texcoords.swap(T);
    s.faces.swap(F);
}

int prepare_scene(HostScene& s) {
    const size_t np = s.prims.size();
    const size_t nt = s.faces.size() / 3;
    const size_t nv = s.positions.size() / 3;
    if (s.faces.size() % 3 || s.positions.size() % 3) return fail(NGI_ERR_INVALID_ARG, "array sizes are not multiples of 3");
    if (s.tri_prim.size() != nt) return fail(NGI_ERR_INVALID_ARG, "tri_prim size %zu != triangles %zu", s.tri_prim.size(), nt);
    if (!s.normals.empty() && s.normals.size() != s.positions.size()) return fail(NGI_ERR_INVALID_ARG, "normals size mismatch");
    if (!s.texcoords.empty() && s.texcoords.size() != 2 * nv) return fail(NGI_ERR_INVALID_ARG, "texcoords size mismatch");
    for (size_t t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k)
            if (s.faces[3 * t + k] >= nv) return fail(NGI_ERR_INVALID_ARG, "face %zu references vertex %u >= %zu", t, s.faces[3 * t + k], nv);
    s.sensor_prim = -1;
    s.light_prims.clear();
    s.area_dist.assign(np, AreaDist());
    for (size_t i = 0; i < np; ++i) {
        const ngi_primitive& p = s.prims[i];
        const int ty = p.type;
        // rt.hpp:1600-1604: empty type or both L and E is invalid.
        if ((ty & ~31) || ty == 0 || ((ty & NGI_PRIM_L) && (ty & NGI_PRIM_E)))
            return fail(NGI_ERR_INVALID_TYPE, "primitive %zu: invalid primitive type %d", i, ty);
        if ((uint64_t)p.tri_begin + p.tri_count > nt) return fail(NGI_ERR_INVALID_ARG, "primitive %zu: triangle range out of bounds", i);
        for (uint32_t t = p.tri_begin; t < p.tri_begin + p.tri_count; ++t)
            if (s.tri_prim[t] != i) return fail(NGI_ERR_INVALID_ARG, "triangle %u: tri_prim %u disagrees with primitive %zu range", t, s.tri_prim[t], i);
        if (ty & NGI_PRIM_E) s.sensor_prim = (int)i;   // the last sensor wins (rt.hpp:1606-1610)
        if (ty & NGI_PRIM_L) s.light_prims.push_back((int)i);
        const bool area_l = (ty & NGI_PRIM_L) && p.light_type == NGI_LIGHT_AREA;
        const bool area_e = (ty & NGI_PRIM_E) && p.sensor_type == NGI_SENSOR_AREA;
        if ((ty & NGI_PRIM_L) && (p.light_type < 0 || p.light_type > 2)) return fail(NGI_ERR_INVALID_SCENE, "primitive %zu: bad light type", i);
        if ((ty & NGI_PRIM_E) && (p.sensor_type < 0 || p.sensor_type > 1)) return fail(NGI_ERR_INVALID_SCENE, "primitive %zu: bad sensor type", i);
        if ((ty & NGI_PRIM_S) && (p.spec_type < 0 || p.spec_type > 2)) return fail(NGI_ERR_INVALID_SCENE, "primitive %zu: bad specular type", i);
        if (area_l && p.tri_count == 0) return fail(NGI_ERR_INVALID_SCENE, "Area light must be associated with mesh (primitive %zu)", i);
        if (area_e && (p.tri_count == 0 || s.texcoords.empty()))
            return fail(NGI_ERR_INVALID_SCENE, "Raw sensor must be associated with mesh with UV coordinates (primitive %zu)", i);
        if (area_l || area_e) {
            // CreateTriangleAreaDist (rt.hpp:1747-1765)
            AreaDist& d = s.area_dist[i];
            d.cdf.assign(1, 0.0);
            double sum = 0;
            for (uint32_t t = p.tri_begin; t < p.tri_begin + p.tri_count; ++t) {
                const uint32_t* f = &s.faces[3 * t];
                d3 p1(s.positions[3 * f[0]], s.positions[3 * f[0] + 1], s.positions[3 * f[0] + 2]);
                d3 p2(s.positions[3 * f[1]], s.positions[3 * f[1] + 1], s.positions[3 * f[1] + 2]);
                d3 p3(s.positions[3 * f[2]], s.positions[3 * f[2] + 1], s.positions[3 * f[2] + 2]);
                const double a = length(cross(p2 - p1, p3 - p1)) * 0.5;
                d.cdf.push_back(d.cdf.back() + a);
                sum += a;
            }
            const double inv = 1.0 / d.cdf.back();
            for (double& c : d.cdf) c *= inv;
            d.inv_area = 1.0 / sum;
        }
    }
    // Scene bound over every vertex of every loaded mesh (rt.hpp:1697).
    for (int k = 0; k < 3; ++k) {
        s.scene_lo[k] = nv ? HUGE_VAL : 0;
        s.scene_hi[k] = nv ? -HUGE_VAL : 0;
    }
    for (size_t v = 0; v < nv; ++v)
        for (int k = 0; k < 3; ++k) {
            s.scene_lo[k] = std::fmin(s.scene_lo[k], s.positions[3 * v + k]);
            s.scene_hi[k] = std::fmax(s.scene_hi[k], s.positions[3 * v + k]);
        }
    s.dir_center.assign(3 * np, 0.0);
    s.dir_radius.assign(np, 0.0);
    s.dir_inv_area.assign(np, 0.0);
    for (size_t i = 0; i < np; ++i) {
        const ngi_primitive& p = s.prims[i];
        if ((p.type & NGI_PRIM_L) && p.light_type == NGI_LIGHT_DIRECTIONAL) {
            // rt.hpp:2067-2074 (note the reference's 2*pi*r^2 "area")
            d3 c, mx(s.scene_hi[0], s.scene_hi[1], s.scene_hi[2]);
            for (int k = 0; k < 3; ++k) s.dir_center[3 * i + k] = (s.scene_hi[k] + s.scene_lo[k]) * 0.5;
            c = d3(s.dir_center[3 * i], s.dir_center[3 * i + 1], s.dir_center[3 * i + 2]);
            const double r = length(c - mx) * 1.01;
            s.dir_radius[i] = r;
            s.dir_inv_area[i] = 1.0 / (2.0 * M_PI * r * r);
        }
    }
    return NGI_OK;
}

}  // namespace ngi

// ---------------------------------------------------------------------------
// YAML scene loading
// ---------------------------------------------------------------------------
namespace {

struct LoadError {
    int status;
    std::string msg;
};

const YNode& need(const YNode* n, const char* what) {
    if (!n || n->kind == YNode::Null) throw LoadError{NGI_ERR_PARSE, std::string("missing key '") + what + "'"};
    return *n;
}

double as_double(const YNode& n, const char* what) {
    if (n.kind != YNode::Scalar) throw LoadError{NGI_ERR_PARSE, std::string("'") + what + "' is not a scalar"};
    char* e = nullptr;
    double v = strtod(n.scalar.c_str(), &e);
    if (n.scalar.empty() || *e) throw LoadError{NGI_ERR_PARSE, std::string("'") + what + "' is not a number: " + n.scalar};
    return v;
}

int as_int(const YNode& n, const char* what) {
    if (n.kind != YNode::Scalar) throw LoadError{NGI_ERR_PARSE, std::string("'") + what + "' is not a scalar"};
    char* e = nullptr;
    long v = strtol(n.scalar.c_str(), &e, 10);
    if (n.scalar.empty() || *e) throw LoadError{NGI_ERR_PARSE, std::string("'") + what + "' is not an integer: " + n.scalar};
    return (int)v;
}

bool as_bool(const YNode& n, const char* what) {
    if (n.kind == YNode::Scalar) {
        const std::string& s = n.scalar;
        if (s == "true" || s == "True" || s == "TRUE" || s == "yes" || s == "on") return true;
        if (s == "false" || s == "False" || s == "FALSE" || s == "no" || s == "off") return false;
    }
    throw LoadError{NGI_ERR_PARSE, std::string("'") + what + "' is not a boolean"};
}

void vec3(const YNode* n, const char* what, double out[3]) {
    const YNode& v = need(n, what);
    if (v.kind != YNode::Seq || v.seq.size() < 3) throw LoadError{NGI_ERR_PARSE, std::string("'") + what + "' is not a 3-vector"};
    for (int k = 0; k < 3; ++k) out[k] = as_double(v.seq[k], what);
}

std::string dirname_of(const std::string& p) {
    size_t s = p.find_last_of('/');
    return s == std::string::npos ? std::string() : p.substr(0, s + 1);
}

int load_yaml_scene(const char* path, double aspect, HostScene& s) {
    std::ifstream f(path);
    if (!f) return fail(NGI_ERR_IO, "cannot open scene '%s'", path);
    std::stringstream ss;
    ss << f.rdbuf();
    YNode root;
    int st = yaml_parse(ss.str(), root);
    if (st) return st;
    const std::string base = dirname_of(path);
    try {
        const int version = as_int(need(root.get("version"), "version"), "version");
        if (version < 3 || version > 5)
            return fail(NGI_ERR_VERSION, "Invalid config version [Min 3, Max 5, Actual %d]", version);
        const YNode& sceneNode = need(root.get("scene"), "scene");
        const YNode& prims = need(sceneNode.get("primitives"), "primitives");
        if (prims.kind != YNode::Seq) throw LoadError{NGI_ERR_PARSE, "'primitives' is not a sequence"};
        for (size_t i = 0; i < prims.seq.size(); ++i) {
            const YNode& pn = prims.seq[i];
            ngi_primitive p;
            memset(&p, 0, sizeof(p));
            const YNode& tn = need(pn.get("type"), "type");
            if (tn.kind != YNode::Seq) throw LoadError{NGI_ERR_PARSE, "'type' is not a sequence"};
            for (const YNode& e : tn.seq) {
                if (e.kind != YNode::Scalar) continue;
                if (e.scalar == "D") p.type |= NGI_PRIM_D;
                else if (e.scalar == "G") p.type |= NGI_PRIM_G;
                else if (e.scalar == "S") p.type |= NGI_PRIM_S;
                else if (e.scalar == "L") p.type |= NGI_PRIM_L;
                else if (e.scalar == "E") p.type |= NGI_PRIM_E;
            }
            if (p.type == 0 || ((p.type & NGI_PRIM_L) && (p.type & NGI_PRIM_E)))
                return fail(NGI_ERR_INVALID_TYPE, "Invalid primitive type (primitive %zu)", i);
            const uint32_t prim_index = (uint32_t)s.prims.size();
            // Mesh (rt.hpp:1620-1738).
            const YNode* mn = pn.get("mesh");
            p.tri_begin = (uint32_t)(s.faces.size() / 3);
            if (mn && mn->kind != YNode::Null) {
                const std::string local = need(mn->get("path"), "path").scalar;
                bool smooth = false;
                if (const YNode* pp = mn->get("postprocess")) {
                    if (const YNode* g = pp->get("generate_smooth_normals")) smooth = as_bool(*g, "generate_smooth_normals");
                    if (const YNode* g = pp->get("generate_normals")) (void)as_bool(*g, "generate_normals");
                }
                ObjMesh m;
                st = obj_load(base + local, smooth, m);
                if (st) return st;
                const uint32_t voff = (uint32_t)(s.positions.size() / 3);
                s.positions.insert(s.positions.end(), m.positions.begin(), m.positions.end());
                s.normals.insert(s.normals.end(), m.normals.begin(), m.normals.end());
                if (!m.texcoords.empty() || !s.texcoords.empty()) {
                    s.texcoords.resize(2 * voff, 0.0);
                    if (m.texcoords.empty()) s.texcoords.resize(2 * (voff + m.positions.size() / 3), 0.0);
                    else s.texcoords.insert(s.texcoords.end(), m.texcoords.begin(), m.texcoords.end());
                }
                for (uint32_t idx : m.faces) s.faces.push_back(idx + voff);
                p.tri_count = (uint32_t)(m.faces.size() / 3);
                s.tri_prim.insert(s.tri_prim.end(), p.tri_count, prim_index);
            }
            const YNode& params = need(pn.get("params"), "params");
            if (p.type & NGI_PRIM_L) {
                const YNode& L = need(params.get("L"), "L");
                const std::string ty = need(L.get("type"), "L.type").scalar;
                if (ty == "area") {
                    p.light_type = NGI_LIGHT_AREA;
                    vec3(need(L.get("area"), "L.area").get("Le"), "L.area.Le", p.Le);
                    if (p.tri_count == 0) return fail(NGI_ERR_INVALID_SCENE, "Area light must be associated with mesh");
                } else if (ty == "point") {
                    p.light_type = NGI_LIGHT_POINT;
                    const YNode& n = need(L.get("point"), "L.point");
                    vec3(n.get("Le"), "L.point.Le", p.Le);
                    vec3(n.get("position"), "L.point.position", p.light_position);
                } else if (ty == "directional") {
                    p.light_type = NGI_LIGHT_DIRECTIONAL;
                    const YNode& n = need(L.get("directional"), "L.directional");
                    vec3(n.get("Le"), "L.directional.Le", p.Le);
                    vec3(n.get("direction"), "L.directional.direction", p.light_direction);
                } else {
                    return fail(NGI_ERR_INVALID_SCENE, "unknown light type '%s'", ty.c_str());
                }
            }
            if (p.type & NGI_PRIM_E) {
                const YNode& E = need(params.get("E"), "E");
                const std::string ty = need(E.get("type"), "E.type").scalar;
                if (ty == "pinhole") {
                    p.sensor_type = NGI_SENSOR_PINHOLE;
                    const YNode& ph = need(E.get("pinhole"), "E.pinhole");
                    const YNode& view = need(ph.get("view"), "E.pinhole.view");
                    vec3(view.get("eye"), "eye", p.E_eye);
                    vec3(view.get("center"), "center", p.E_center);
                    vec3(view.get("up"), "up", p.E_up);
                    vec3(ph.get("We"), "We", p.E_We);
                    p.E_fov_deg = as_double(need(need(ph.get("perspective"), "perspective").get("fov"), "fov"), "fov");
                } else if (ty == "area") {
                    p.sensor_type = NGI_SENSOR_AREA;
                    vec3(need(E.get("area"), "E.area").get("We"), "E.area.We", p.E_We);
                    if (p.tri_count == 0 || s.texcoords.empty())
                        return fail(NGI_ERR_INVALID_SCENE, "Raw sensor must be associated with mesh with UV coordinates");
                } else {
                    return fail(NGI_ERR_INVALID_SCENE, "unknown sensor type '%s'", ty.c_str());
                }
            }
            if (p.type & NGI_PRIM_D) {
                const YNode& D = need(params.get("D"), "D");
                if (D.get("R")) vec3(D.get("R"), "D.R", p.D_R);
                else if (D.get("TexR")) return fail(NGI_ERR_UNSUPPORTED, "textures (D.TexR) are not supported");
                else return fail(NGI_ERR_INVALID_SCENE, "D requires R (primitive %zu)", i);
            }
            if (p.type & NGI_PRIM_G) {
                const YNode& G = need(params.get("G"), "G");
                vec3(G.get("Eta"), "G.Eta", p.G_eta);
                vec3(G.get("K"), "G.K", p.G_k);
                p.G_roughness = as_double(need(G.get("Roughness"), "G.Roughness"), "G.Roughness");
                if (G.get("R")) vec3(G.get("R"), "G.R", p.G_R);
                else if (G.get("TexR")) return fail(NGI_ERR_UNSUPPORTED, "textures (G.TexR) are not supported");
                else return fail(NGI_ERR_INVALID_SCENE, "G requires R (primitive %zu)", i);
            }
            if (p.type & NGI_PRIM_S) {
                const YNode& S = need(params.get("S"), "S");
                const std::string ty = need(S.get("type"), "S.type").scalar;
                const char* key = nullptr;
                if (ty == "reflection") p.spec_type = NGI_SPEC_REFLECTION, key = "reflection";
                else if (ty == "refraction") p.spec_type = NGI_SPEC_REFRACTION, key = "refraction";
                else if (ty == "fresnel") p.spec_type = NGI_SPEC_FRESNEL, key = "fresnel";
                else return fail(NGI_ERR_INVALID_SCENE, "unknown specular type '%s'", ty.c_str());
                const YNode& n = need(S.get(key), key);
                vec3(n.get("R"), "S.R", p.S_R);
                if (p.spec_type != NGI_SPEC_REFLECTION) {
                    p.S_eta1 = as_double(need(n.get("eta1"), "eta1"), "eta1");
                    p.S_eta2 = as_double(need(n.get("eta2"), "eta2"), "eta2");
                }
            }
            s.prims.push_back(p);
        }
    } catch (const LoadError& e) {
        return fail(e.status, "YAML exception: %s", e.msg.c_str());
    }
    s.aspect = aspect;
    return NGI_OK;
}

void write_vec(std::ostream& o, const double* v) {
    char b[160];
    snprintf(b, sizeof(b), "[%.17g, %.17g, %.17g]", v[0], v[1], v[2]);
    o << b;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int ngi_scene_load_yaml(const char* path, double aspect, ngi_scene** out) {
    clear_error();
    if (!path || !out) return fail(NGI_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (!(aspect > 0)) return fail(NGI_ERR_INVALID_ARG, "aspect must be positive");
    ngi_scene* s = new ngi_scene();
    int st = load_yaml_scene(path, aspect, *s);
    if (!st) {
        ensure_normals(*s);
        st = prepare_scene(*s);
    }
    if (st) {
        delete s;
        return st;
    }
    *out = s;
    return NGI_OK;
}

int ngi_scene_create(const ngi_scene_desc* d, ngi_scene** out) {
    clear_error();
    if (!d || !out) return fail(NGI_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (!(d->aspect > 0)) return fail(NGI_ERR_INVALID_ARG, "aspect must be positive");
    if ((d->num_vertices && !d->positions) || (d->num_triangles && (!d->faces || !d->tri_prim)) || (d->num_primitives && !d->primitives))
        return fail(NGI_ERR_INVALID_ARG, "null array with nonzero count");
    ngi_scene* s = new ngi_scene();
    s->positions.assign(d->positions, d->positions + 3 * (size_t)d->num_vertices);
    if (d->normals) s->normals.assign(d->normals, d->normals + 3 * (size_t)d->num_vertices);
    if (d->texcoords) s->texcoords.assign(d->texcoords, d->texcoords + 2 * (size_t)d->num_vertices);
    s->faces.assign(d->faces, d->faces + 3 * (size_t)d->num_triangles);
    s->tri_prim.assign(d->tri_prim, d->tri_prim + d->num_triangles);
    s->prims.assign(d->primitives, d->primitives + d->num_primitives);
    s->aspect = d->aspect;
    int st = prepare_scene(*s);   // validate indices before de-indexing
    if (!st && s->normals.empty()) {
        ensure_normals(*s);
        st = prepare_scene(*s);
    }
    if (st) {
        delete s;
        return st;
    }
    *out = s;
    return NGI_OK;
}

int ngi_scene_generate(int kind, uint64_t seed, double aspect, int detail, ngi_scene** out) {
    clear_error();
    if (!out) return fail(NGI_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (!(aspect > 0)) return fail(NGI_ERR_INVALID_ARG, "aspect must be positive");
    ngi_scene* s = new ngi_scene();
    int st = generate_scene(kind, seed, aspect, detail, *s);
    if (!st) st = prepare_scene(*s);
    if (st) {
        delete s;
        return st;
    }
    *out = s;
    return NGI_OK;
}

void ngi_scene_destroy(ngi_scene* s) {
    if (!s) return;
    device_release(*s);
    delete s;
}

int ngi_scene_get_desc(const ngi_scene* s, ngi_scene_desc* d) {
    clear_error();
    if (!s || !d) return fail(NGI_ERR_INVALID_ARG, "null argument");
    d->num_vertices = (uint32_t)(s->positions.size() / 3);
    d->positions = s->positions.data();
    d->normals = s->normals.empty() ? nullptr : s->normals.data();
    d->texcoords = s->texcoords.empty() ? nullptr : s->texcoords.data();
    d->num_triangles = (uint32_t)(s->faces.size() / 3);
    d->faces = s->faces.data();
    d->tri_prim = s->tri_prim.data();
    d->num_primitives = (uint32_t)s->prims.size();
    d->primitives = s->prims.data();
    d->aspect = s->aspect;
    return NGI_OK;
}

int ngi_scene_set_params(ngi_scene* s, const double* refl, const double* Le) {
    clear_error();
    if (!s) return fail(NGI_ERR_INVALID_ARG, "null scene");
    for (size_t i = 0; i < s->prims.size(); ++i) {
        ngi_primitive& p = s->prims[i];
        for (int k = 0; k < 3; ++k) {
            if (refl && (p.type & NGI_PRIM_D)) p.D_R[k] = refl[3 * i + k];
            else if (refl && (p.type & NGI_PRIM_G)) p.G_R[k] = refl[3 * i + k];
            if (Le && (p.type & NGI_PRIM_L)) p.Le[k] = Le[3 * i + k];
        }
    }
    s->params_version++;
    return NGI_OK;
}

int ngi_scene_save_yaml(const ngi_scene* s, const char* dir) {
    clear_error();
    if (!s || !dir) return fail(NGI_ERR_INVALID_ARG, "null argument");
    mkdir(dir, 0755);
    const std::string base = std::string(dir) + "/";
    std::ofstream y(base + "scene.yml");
    if (!y) return fail(NGI_ERR_IO, "cannot write '%sscene.yml'", base.c_str());
    y << "version: 5\nscene:\n  primitives:\n";
    for (size_t i = 0; i < s->prims.size(); ++i) {
        const ngi_primitive& p = s->prims[i];
        y << "    - type: [";
        const char* names[5] = {"D", "G", "S", "L", "E"};
        bool first = true;
        for (int b = 4; b >= 0; --b)   // emitter letters first, as the sample scenes do
            if (p.type & (1 << b)) {
                y << (first ? "" : ", ") << names[b];
                first = false;
            }
        y << "]\n";
        if (p.tri_count) {
            char fn[64];
            snprintf(fn, sizeof(fn), "prim%05zu.obj", i);
            std::ofstream o(base + fn);
            if (!o) return fail(NGI_ERR_IO, "cannot write mesh '%s'", fn);
            o.precision(17);
            // Vertices are written per face corner so each OBJ is self-contained.
            for (uint32_t t = p.tri_begin; t < p.tri_begin + p.tri_count; ++t)
                for (int k = 0; k < 3; ++k) {
                    uint32_t v = s->faces[3 * t + k];
                    o << "v " << s->positions[3 * v] << " " << s->positions[3 * v + 1] << " " << s->positions[3 * v + 2] << "\n";
                }
            for (uint32_t t = p.tri_begin; t < p.tri_begin + p.tri_count; ++t)
                for (int k = 0; k < 3; ++k) {
                    uint32_t v = s->faces[3 * t + k];
                    o << "vn " << s->normals[3 * v] << " " << s->normals[3 * v + 1] << " " << s->normals[3 * v + 2] << "\n";
                }
            for (uint32_t t = 0; t < p.tri_count; ++t)
                o << "f " << 3 * t + 1 << "//" << 3 * t + 1 << " " << 3 * t + 2 << "//" << 3 * t + 2 << " " << 3 * t + 3 << "//" << 3 * t + 3 << "\n";
            y << "      mesh:\n        path: '" << fn << "'\n";
        }
        y << "      params:\n";
        if (p.type & NGI_PRIM_L) {
            y << "        L:\n";
            if (p.light_type == NGI_LIGHT_AREA) {
                y << "          type: area\n          area:\n            Le: ";
                write_vec(y, p.Le);
                y << "\n";
            } else if (p.light_type == NGI_LIGHT_POINT) {
                y << "          type: point\n          point:\n            Le: ";
                write_vec(y, p.Le);
                y << "\n            position: ";
                write_vec(y, p.light_position);
                y << "\n";
            } else {
                y << "          type: directional\n          directional:\n            Le: ";
                write_vec(y, p.Le);
                y << "\n            direction: ";
                write_vec(y, p.light_direction);
                y << "\n";
            }
        }
        if (p.type & NGI_PRIM_E) {
            y << "        E:\n";
            if (p.sensor_type == NGI_SENSOR_PINHOLE) {
                y << "          type: pinhole\n          pinhole:\n            We: ";
                write_vec(y, p.E_We);
                y << "\n            view:\n              eye: ";
                write_vec(y, p.E_eye);
                y << "\n              center: ";
                write_vec(y, p.E_center);
                y << "\n              up: ";
                write_vec(y, p.E_up);
                char b[64];
                snprintf(b, sizeof(b), "%.17g", p.E_fov_deg);
                y << "\n            perspective:\n              fov: " << b << "\n";
            } else {
                y << "          type: area\n          area:\n            We: ";
                write_vec(y, p.E_We);
                y << "\n";
            }
        }
        if (p.type & NGI_PRIM_D) {
            y << "        D:\n          R: ";
            write_vec(y, p.D_R);
            y << "\n";
        }
        if (p.type & NGI_PRIM_G) {
            y << "        G:\n          R: ";
            write_vec(y, p.G_R);
            y << "\n          Eta: ";
            write_vec(y, p.G_eta);
            y << "\n          K: ";
            write_vec(y, p.G_k);
            char b[64];
            snprintf(b, sizeof(b), "%.17g", p.G_roughness);
            y << "\n          Roughness: " << b << "\n";
        }
        if (p.type & NGI_PRIM_S) {
            static const char* st[3] = {"reflection", "refraction", "fresnel"};
            y << "        S:\n          type: " << st[p.spec_type] << "\n          " << st[p.spec_type] << ":\n            R: ";
            write_vec(y, p.S_R);
            y << "\n";
            if (p.spec_type != NGI_SPEC_REFLECTION) {
                char b[128];
                snprintf(b, sizeof(b), "            eta1: %.17g\n            eta2: %.17g\n", p.S_eta1, p.S_eta2);
                y << b;
            }
        }
    }
    return y.good() ? NGI_OK : fail(NGI_ERR_IO, "write failed");
}

int ngi_renderer_from_string(const char* name) {
    // RendererType_String (src/nanogi.cpp:61-69)
    static const char* names[] = {"pt", "ptdirect", "lt", "ltdirect", "bdpt", "ptmnee"};
    if (!name) return -1;
    for (int i = 0; i < 6; ++i)
        if (!strcmp(name, names[i])) return i;
    return -1;
}

void ngi_render_params_default(ngi_render_params* p) {
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->renderer = NGI_RENDERER_PT;
    p->width = 1280;                      // src/nanogi.cpp:2006
    p->height = 720;                      // src/nanogi.cpp:2007
    p->sampling = NGI_SAMPLING_REFERENCE_UNIFORM;
    p->num_samples = 10000000LL;          // src/nanogi.cpp:2004
    p->spp = 1;
    p->max_num_vertices = -1;             // src/nanogi.cpp:2005
    p->rr_depth = 1;                      // RR after every intersection (src/nanogi.cpp:583-591)
    p->rr_prob = 0.5;
    p->seed = 1;
    p->stream_id = 0;
    p->precision = NGI_PRECISION_FP32_FP64ACC;
    p->engine = NGI_ENGINE_PERSISTENT;     // the fastest engine on every measured workload
}

}  // extern "C"
