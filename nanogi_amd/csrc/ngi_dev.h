// Device-resident scene layout and HIP plumbing shared by the kernels.
//
// HBM layout (all arrays 16-byte aligned, leaf order = BVH leaf order):
//   nodes      : Bvh2Node[] (64 B) or Bvh8Node[] (128 B)
//   tris       : 3 x float4 per leaf triangle: (v0, orig id), (v1, 0), (v2, 0)
//   shade      : 3 x float4 per leaf triangle: (n0, prim id), (n1, prim type bits), (n2, 0)
//   prims_f/_d : DevPrim<float>/<double> per primitive
//   lights_f/_d: DevLight<T> per light primitive (Scene::LightPrimitiveIndices)
//   light_tris : 3 x float4 per area-light triangle, per light contiguous
//   cdf_f/_d   : per area light, cdf[0..n] (rt.hpp:1747-1765)
//   guide_f/_d : per area light, 2^guide_bits + 1 bucket starts of cdf_f/_d
#pragma once

#include <hip/hip_runtime.h>

#include "ngi_bvh.h"
#include "ngi_common.h"
#include "ngi_internal.h"

#define NGI_HIP_CHECK(call)                                                                        \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return ::ngi::fail(e_ == hipErrorOutOfMemory ? NGI_ERR_OOM : NGI_ERR_DEVICE, "%s: %s", \
                               #call, hipGetErrorString(e_));                                      \
    } while (0)

namespace ngi {

template <typename T>
struct DevPrim {
    int type;         // NGI_PRIM_* bits
    int light;        // index into the light table or -1
    int spec_type;
    int pad;
    T R[3];           // D.R, G.R or S.R
    T roughness;      // G.Roughness
    T eta[3];         // G.Eta
    T k[3];           // G.K
    T s_eta1, s_eta2; // S eta1/eta2
    T Le[3];          // L.*.Le
};

template <typename T>
struct DevLight {
    int prim;
    int type;         // NGI_LIGHT_*
    uint32_t tri_begin;  // first row in light_tris
    uint32_t tri_count;
    uint32_t cdf_begin;  // first entry in cdf (tri_count + 1 entries)
    uint32_t guide_begin;  // first entry in guide (2^guide_bits + 1 entries)
    uint32_t guide_bits;
    uint32_t pad;
    T inv_area;       // area light: 1 / area; directional: the reference's 1 / (2 pi r^2)
    T Le[3];
    T pos[3];         // point light position; directional: centre of the scene bound
    T dir[3];         // directional: direction as given (not normalised, as rt.hpp:1861)
    T radius;         // directional: disk radius (rt.hpp:2071)
};

struct DevCamera {
    double eye[3], vx[3], vy[3], vz[3];
    double tan_fov, aspect;
};

struct DevScene {
    const float4* nodes = nullptr;
    const float4* tris = nullptr;
    const float4* shade = nullptr;
    const DevPrim<float>* prims_f = nullptr;
    const DevPrim<double>* prims_d = nullptr;
    const DevLight<float>* lights_f = nullptr;
    const DevLight<double>* lights_d = nullptr;
    const float4* light_tris = nullptr;
    const float* cdf_f = nullptr;
    const double* cdf_d = nullptr;
    const uint32_t* guide_f = nullptr;
    const uint32_t* guide_d = nullptr;
    int num_lights = 0;
    int any_glossy = 0;           // a primitive with a glossy or specular component
    int bvh_width = 2;
    uint32_t stack_entries = 0;   // LDS traversal-stack entries per lane (wide BVH depth + 1)
    uint32_t num_tris = 0;
    int sensor_prim = -1;
    DevCamera cam;
};

// Owning host-side record of the device allocations of one scene.
struct DeviceScene {
    int device = 0;
    hipStream_t stream = nullptr;
    DevScene view;
    std::vector<void*> allocs;
    uint64_t params_version = ~0ull;
    void* prim_f_mem = nullptr;
    void* prim_d_mem = nullptr;
    void* light_f_mem = nullptr;
    void* light_d_mem = nullptr;
    // Scratch used by ngi_trace for host<->device staging.
    void* stage = nullptr;
    size_t stage_bytes = 0;
};

// Upload (or refresh) the scene on the current device; builds the default
// BVH if the caller did not.
int device_prepare(HostScene& s);
int ensure_stage(DeviceScene& d, size_t bytes);

}  // namespace ngi
