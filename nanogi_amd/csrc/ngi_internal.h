// Host-side scene model shared by the loaders, the BVH builder and the
// device layer.  Mirrors nanogi's Scene (include/nanogi/rt.hpp:1488-1500):
// meshes are concatenated into one FP64 vertex/face array, primitives keep
// their per-primitive material/emitter record.
#pragma once

#include <string>
#include <vector>

#include "../../include/ngi.h"
#include "ngi_common.h"

namespace ngi {

// Per-area-emitter triangle distribution (rt.hpp:1747-1765): cdf[0] = 0,
// cdf[i+1] = normalised running sum of triangle areas.
struct AreaDist {
    std::vector<double> cdf;
    double inv_area = 0;
};

struct Bvh {
    int width = 0;
    std::vector<uint8_t> nodes;          // flattened node records
    uint32_t num_nodes = 0;
    uint32_t node_bytes = 0;
    uint32_t num_leaves = 0;
    uint32_t max_depth = 0;
    std::vector<uint32_t> leaf_to_tri;   // leaf-order slot -> original triangle
    double sah_cost = 0;
    double build_ms = 0;
};

struct DeviceScene;  // device-resident copy, defined by the HIP layer

struct HostScene {
    std::vector<double> positions;   // 3 per vertex
    std::vector<double> normals;     // 3 per vertex
    std::vector<double> texcoords;   // 2 per vertex or empty
    std::vector<uint32_t> faces;     // 3 per triangle
    std::vector<uint32_t> tri_prim;  // owner primitive per triangle
    std::vector<ngi_primitive> prims;
    double aspect = 1.0;

    // Derived by prepare_scene().
    int sensor_prim = -1;            // last E primitive, as Scene::Load (rt.hpp:1606-1610)
    std::vector<int> light_prims;    // LightPrimitiveIndices (rt.hpp:1612-1615)
    std::vector<AreaDist> area_dist; // indexed like prims (empty for non-area)
    double scene_lo[3] = {0, 0, 0}, scene_hi[3] = {0, 0, 0};
    // Directional-light disk (rt.hpp:2059-2079), per primitive.
    std::vector<double> dir_center, dir_radius, dir_inv_area;

    Bvh bvh;
    bool built = false;
    uint64_t params_version = 0;     // bumped by ngi_scene_set_params

    DeviceScene* dev = nullptr;
    // Streams other than the scene's own whose kernels read its device
    // arrays (one per live batch holding the scene).  Re-uploads and
    // releases wait for them first.
    std::vector<void*> user_streams;
};

// Error reporting (thread-local message returned by ngi_last_error).
int fail(int status, const char* fmt, ...);
void clear_error();

// Validate type bits / ranges and compute the derived emitter data.
int prepare_scene(HostScene& s);
// Generate per-vertex normals where none were given (flat, de-indexed).
void ensure_normals(HostScene& s);

// YAML subset reader used by the scene loader.
struct YNode {
    enum Kind { Null, Scalar, Seq, Map } kind = Null;
    std::string scalar;
    std::vector<YNode> seq;
    std::vector<std::pair<std::string, YNode>> map;
    const YNode* get(const char* key) const;
    size_t size() const { return kind == Seq ? seq.size() : (kind == Map ? map.size() : 0); }
};
// Returns NGI_OK or NGI_ERR_PARSE (message via fail()).
int yaml_parse(const std::string& text, YNode& out);

// Wavefront-OBJ reader: first object, first material run (the first Assimp
// mesh, rt.hpp:1679).  Positions/normals/texcoords per unique (v,vt,vn).
struct ObjMesh {
    std::vector<double> positions, normals, texcoords;
    std::vector<uint32_t> faces;
};
int obj_load(const std::string& path, bool smooth_normals, ObjMesh& out);

// BVH construction (ngi_bvh.cpp).
int bvh_build(const HostScene& s, const ngi_bvh_opts& opts, Bvh& out);

// Procedural scenes (ngi_generate.cpp).
int generate_scene(int kind, uint64_t seed, double aspect, int detail, HostScene& out);

// Device layer hooks (ngi_device.hip).
void device_release(HostScene& s);

}  // namespace ngi

struct ngi_scene : ngi::HostScene {};
