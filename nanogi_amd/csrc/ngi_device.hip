// Device scene upload and the ray-query entry points (ngi_trace*).
//
// The upload replaces the Embree buffers of the reference (FP32 de-indexed
// triangle soup, include/nanogi/rt.hpp:2110-2135) with leaf-ordered FP32
// triangle records plus a parallel shading record per triangle (vertex
// normals and owning primitive, used to rebuild the FP64 surface of
// Scene::Intersect, rt.hpp:2197-2233, in the shade stage).
#include <cmath>
#include <cstring>

#include "ngi_dev.h"
#include "ngi_guide.h"
#include "ngi_traverse.h"

namespace ngi {

static int device_ready() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(NGI_ERR_DEVICE, "no HIP device available");
    NGI_HIP_CHECK(hipSetDevice(0));
    return NGI_OK;
}

template <typename T>
static int upload(DeviceScene& d, const T* host, size_t count, const T** out) {
    *out = nullptr;
    if (count == 0) return NGI_OK;
    void* p = nullptr;
    NGI_HIP_CHECK(hipMalloc(&p, count * sizeof(T)));
    d.allocs.push_back(p);
    NGI_HIP_CHECK(hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice));
    *out = reinterpret_cast<const T*>(p);
    return NGI_OK;
}

template <typename T>
static void fill_prim(const ngi_primitive& p, int light, DevPrim<T>& o) {
    memset(&o, 0, sizeof(o));
    o.type = p.type;
    o.light = light;
    o.spec_type = p.spec_type;
    for (int k = 0; k < 3; ++k) {
        o.R[k] = (T)((p.type & NGI_PRIM_D) ? p.D_R[k] : (p.type & NGI_PRIM_G) ? p.G_R[k] : p.S_R[k]);
        o.eta[k] = (T)p.G_eta[k];
        o.k[k] = (T)p.G_k[k];
        o.Le[k] = (T)p.Le[k];
    }
    o.roughness = (T)p.G_roughness;
    o.s_eta1 = (T)p.S_eta1;
    o.s_eta2 = (T)p.S_eta2;
}

// Guide bits of an area light's triangle CDF (FP64; its FP32 rounding is
// sorted and NaN-free whenever the FP64 CDF is).
static uint32_t light_guide_bits(const std::vector<double>& cdf) {
    const uint32_t m = (uint32_t)cdf.size();
    return m >= 2 && guide_sorted(cdf.data(), m) ? guide_bits(m - 1) : 0u;
}

// Wait for every kernel that may read the scene's device arrays: the
// scene's stream and the streams of the batches holding it.
static int wait_users(HostScene& s) {
    if (s.dev && s.dev->stream) NGI_HIP_CHECK(hipStreamSynchronize(s.dev->stream));
    for (void* u : s.user_streams) NGI_HIP_CHECK(hipStreamSynchronize(reinterpret_cast<hipStream_t>(u)));
    return NGI_OK;
}

// (Re)build the parameter-dependent tables: primitives and lights.
static int upload_params(HostScene& s, DeviceScene& d) {
    const size_t np = s.prims.size();
    std::vector<int> light_of(np, -1);
    for (size_t i = 0; i < s.light_prims.size(); ++i) light_of[s.light_prims[i]] = (int)i;
    std::vector<DevPrim<float>> pf(np);
    std::vector<DevPrim<double>> pd(np);
    for (size_t i = 0; i < np; ++i) {
        fill_prim(s.prims[i], light_of[i], pf[i]);
        fill_prim(s.prims[i], light_of[i], pd[i]);
    }
    std::vector<DevLight<float>> lf(s.light_prims.size());
    std::vector<DevLight<double>> ld(s.light_prims.size());
    uint32_t tri_row = 0, cdf_row = 0, guide_row = 0;
    for (size_t i = 0; i < s.light_prims.size(); ++i) {
        const int pi = s.light_prims[i];
        const ngi_primitive& p = s.prims[pi];
        DevLight<double> L;
        memset(&L, 0, sizeof(L));
        L.prim = pi;
        L.type = p.light_type;
        for (int k = 0; k < 3; ++k) L.Le[k] = p.Le[k], L.pos[k] = p.light_position[k];
        if (p.light_type == NGI_LIGHT_DIRECTIONAL) {
            for (int k = 0; k < 3; ++k) L.pos[k] = s.dir_center[3 * (size_t)pi + k], L.dir[k] = p.light_direction[k];
            L.radius = s.dir_radius[pi];
            L.inv_area = s.dir_inv_area[pi];
        }
        if (p.light_type == NGI_LIGHT_AREA) {
            L.tri_begin = tri_row;
            L.tri_count = p.tri_count;
            L.cdf_begin = cdf_row;
            L.guide_begin = guide_row;
            L.guide_bits = light_guide_bits(s.area_dist[pi].cdf);
            L.inv_area = s.area_dist[pi].inv_area;
            tri_row += p.tri_count;
            cdf_row += p.tri_count + 1;
            guide_row += (1u << L.guide_bits) + 1u;
        }
        ld[i] = L;
        DevLight<float>& F = lf[i];
        F.prim = L.prim, F.type = L.type, F.tri_begin = L.tri_begin, F.tri_count = L.tri_count, F.cdf_begin = L.cdf_begin;
        F.guide_begin = L.guide_begin, F.guide_bits = L.guide_bits, F.pad = 0;
        F.inv_area = (float)L.inv_area;
        for (int k = 0; k < 3; ++k) F.Le[k] = (float)L.Le[k], F.pos[k] = (float)L.pos[k], F.dir[k] = (float)L.dir[k];
        F.radius = (float)L.radius;
    }
    if (!d.prim_f_mem) {
        NGI_HIP_CHECK(hipMalloc(&d.prim_f_mem, std::max<size_t>(np, 1) * sizeof(DevPrim<float>)));
        NGI_HIP_CHECK(hipMalloc(&d.prim_d_mem, std::max<size_t>(np, 1) * sizeof(DevPrim<double>)));
        NGI_HIP_CHECK(hipMalloc(&d.light_f_mem, std::max<size_t>(lf.size(), 1) * sizeof(DevLight<float>)));
        NGI_HIP_CHECK(hipMalloc(&d.light_d_mem, std::max<size_t>(ld.size(), 1) * sizeof(DevLight<double>)));
        d.allocs.insert(d.allocs.end(), {d.prim_f_mem, d.prim_d_mem, d.light_f_mem, d.light_d_mem});
    }
    int st = wait_users(s);   // a running frame must not see a half-updated table
    if (st) return st;
    if (np) {
        NGI_HIP_CHECK(hipMemcpy(d.prim_f_mem, pf.data(), np * sizeof(pf[0]), hipMemcpyHostToDevice));
        NGI_HIP_CHECK(hipMemcpy(d.prim_d_mem, pd.data(), np * sizeof(pd[0]), hipMemcpyHostToDevice));
    }
    if (!lf.empty()) {
        NGI_HIP_CHECK(hipMemcpy(d.light_f_mem, lf.data(), lf.size() * sizeof(lf[0]), hipMemcpyHostToDevice));
        NGI_HIP_CHECK(hipMemcpy(d.light_d_mem, ld.data(), ld.size() * sizeof(ld[0]), hipMemcpyHostToDevice));
    }
    d.view.prims_f = reinterpret_cast<const DevPrim<float>*>(d.prim_f_mem);
    d.view.prims_d = reinterpret_cast<const DevPrim<double>*>(d.prim_d_mem);
    d.view.lights_f = reinterpret_cast<const DevLight<float>*>(d.light_f_mem);
    d.view.lights_d = reinterpret_cast<const DevLight<double>*>(d.light_d_mem);
    d.view.num_lights = (int)s.light_prims.size();
    d.view.any_glossy = 0;
    for (const ngi_primitive& p : s.prims)
        if (p.type & (NGI_PRIM_G | NGI_PRIM_S)) d.view.any_glossy = 1;
    d.params_version = s.params_version;
    return NGI_OK;
}

int device_prepare(HostScene& s) {
    int st = device_ready();
    if (st) return st;
    if (!s.built) {
        ngi_bvh_opts o;
        memset(&o, 0, sizeof(o));
        st = bvh_build(s, o, s.bvh);
        if (st) return st;
        s.built = true;
    }
    if (s.dev) {
        if (s.dev->params_version != s.params_version) return upload_params(s, *s.dev);
        return NGI_OK;
    }
    DeviceScene* d = new DeviceScene();
    s.dev = d;
    if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
        d->stream = nullptr;
        return fail(NGI_ERR_DEVICE, "hipStreamCreate failed");
    }
    const size_t nt = s.faces.size() / 3;
    DevScene& v = d->view;
    v.bvh_width = s.bvh.width;
    v.stack_entries = s.bvh.width == 8 ? s.bvh.max_depth + 1 : 0;
    v.num_tris = (uint32_t)nt;
    v.sensor_prim = s.sensor_prim;
    // Geometry in leaf order.
    std::vector<float4> tris(3 * nt), shade(3 * nt);
    for (size_t j = 0; j < nt; ++j) {
        const uint32_t t = s.bvh.leaf_to_tri[j];
        for (int k = 0; k < 3; ++k) {
            const uint32_t vi = s.faces[3 * (size_t)t + k];
            float w = 0.0f, wn = 0.0f;
            if (k == 0) {
                uint32_t bits = t, pbits = s.tri_prim[t];
                memcpy(&w, &bits, 4);
                memcpy(&wn, &pbits, 4);
            } else if (k == 1) {
                const uint32_t tbits = (uint32_t)s.prims[s.tri_prim[t]].type;   // fixed by the scene (set_params keeps it)
                memcpy(&wn, &tbits, 4);
            }
            tris[3 * j + k] = make_float4((float)s.positions[3 * (size_t)vi], (float)s.positions[3 * (size_t)vi + 1],
                                          (float)s.positions[3 * (size_t)vi + 2], w);
            shade[3 * j + k] = make_float4((float)s.normals[3 * (size_t)vi], (float)s.normals[3 * (size_t)vi + 1],
                                           (float)s.normals[3 * (size_t)vi + 2], wn);
        }
    }
    if ((st = upload(*d, tris.data(), tris.size(), &v.tris))) return st;
    if ((st = upload(*d, shade.data(), shade.size(), &v.shade))) return st;
    const float4* nodes = nullptr;
    if ((st = upload(*d, reinterpret_cast<const float4*>(s.bvh.nodes.data()), s.bvh.nodes.size() / 16, &nodes))) return st;
    v.nodes = nodes;
    // Area-light triangles and CDFs, per light primitive.
    std::vector<float4> ltris;
    std::vector<float> cf;
    std::vector<double> cd;
    std::vector<uint32_t> gf, gd;
    for (int pi : s.light_prims) {
        const ngi_primitive& p = s.prims[pi];
        if (p.light_type != NGI_LIGHT_AREA) continue;
        for (uint32_t t = p.tri_begin; t < p.tri_begin + p.tri_count; ++t)
            for (int k = 0; k < 3; ++k) {
                const uint32_t vi = s.faces[3 * (size_t)t + k];
                ltris.push_back(make_float4((float)s.positions[3 * (size_t)vi], (float)s.positions[3 * (size_t)vi + 1],
                                            (float)s.positions[3 * (size_t)vi + 2], 0.0f));
            }
        const std::vector<double>& c = s.area_dist[pi].cdf;
        const size_t f0 = cf.size();
        for (double x : c) {
            cd.push_back(x);
            cf.push_back((float)x);
        }
        // The guides follow upload_params' order and light_guide_bits.
        const uint32_t bits = light_guide_bits(c), g0 = (uint32_t)gf.size();
        gf.resize(g0 + (1u << bits) + 1u);
        gd.resize(gf.size());
        guide_build(cf.data() + f0, (uint32_t)c.size(), bits, gf.data() + g0);
        guide_build(c.data(), (uint32_t)c.size(), bits, gd.data() + g0);
    }
    if ((st = upload(*d, ltris.data(), ltris.size(), &v.light_tris))) return st;
    if ((st = upload(*d, cf.data(), cf.size(), &v.cdf_f))) return st;
    if ((st = upload(*d, cd.data(), cd.size(), &v.cdf_d))) return st;
    if ((st = upload(*d, gf.data(), gf.size(), &v.guide_f))) return st;
    if ((st = upload(*d, gd.data(), gd.size(), &v.guide_d))) return st;
    // Pinhole camera (rt.hpp:1882-1902).
    memset(&v.cam, 0, sizeof(v.cam));
    if (s.sensor_prim >= 0 && s.prims[s.sensor_prim].sensor_type == NGI_SENSOR_PINHOLE) {
        const ngi_primitive& p = s.prims[s.sensor_prim];
        d3 eye(p.E_eye[0], p.E_eye[1], p.E_eye[2]), ctr(p.E_center[0], p.E_center[1], p.E_center[2]), up(p.E_up[0], p.E_up[1], p.E_up[2]);
        d3 vz = normalize(eye - ctr);
        d3 vx = normalize(cross(up, vz));
        d3 vy = cross(vz, vx);
        for (int k = 0; k < 3; ++k) v.cam.eye[k] = eye[k], v.cam.vx[k] = vx[k], v.cam.vy[k] = vy[k], v.cam.vz[k] = vz[k];
        v.cam.tan_fov = std::tan(p.E_fov_deg * M_PI / 180.0 * 0.5);
        v.cam.aspect = s.aspect;
    }
    return upload_params(s, *d);
}

void device_release(HostScene& s) {
    if (!s.dev) return;
    (void)wait_users(s);
    DeviceScene* d = s.dev;
    if (d->stream) {
        (void)hipStreamSynchronize(d->stream);
        (void)hipStreamDestroy(d->stream);
    }
    for (void* p : d->allocs) (void)hipFree(p);
    if (d->stage) (void)hipFree(d->stage);
    delete d;
    s.dev = nullptr;
}

int ensure_stage(DeviceScene& d, size_t bytes) {
    if (d.stage_bytes >= bytes) return NGI_OK;
    if (d.stage) {
        NGI_HIP_CHECK(hipStreamSynchronize(d.stream));
        NGI_HIP_CHECK(hipFree(d.stage));
        d.stage = nullptr;
        d.stage_bytes = 0;
    }
    NGI_HIP_CHECK(hipMalloc(&d.stage, bytes));
    d.stage_bytes = bytes;
    return NGI_OK;
}

// ---------------------------------------------------------------------------
// Ray-query kernel: one lane per ray, rays and hits in the public layout.
// ---------------------------------------------------------------------------
#define NGI_TRACE_BLOCK 256

template <int WIDTH, bool ANY>
__global__ void __launch_bounds__(NGI_TRACE_BLOCK) trace_kernel(DevScene sc, const ngi_ray* __restrict__ rays, size_t n,
                                                                ngi_hit* __restrict__ hits) {
    extern __shared__ uint32_t stk[];   // sc.stack_entries x block (wide BVH only)
    const size_t i = (size_t)blockIdx.x * NGI_TRACE_BLOCK + threadIdx.x;
    if (i >= n) return;
    const ngi_ray r = rays[i];
    TraceRay tr{r.o[0], r.o[1], r.o[2], r.d[0], r.d[1], r.d[2], r.tmin, r.tmax};
    TraceHit h;
    TraceCounters c{0, 0};
    bool hit;
    if (WIDTH == 8) hit = trace_bvh8<ANY, false>(sc, tr, h, c, stk + threadIdx.x, NGI_TRACE_BLOCK);
    else hit = trace_bvh2<ANY, false>(sc, tr, h, c);
    ngi_hit o;
    if (!hit) {
        o.t = r.tmax, o.u = 0.0f, o.v = 0.0f, o.tri = NGI_MISS;
    } else {
        o.t = h.t, o.u = h.u, o.v = h.v;
        o.tri = __float_as_uint(sc.tris[3 * (size_t)h.leaf].w);
    }
    hits[i] = o;
}

int launch_trace(HostScene& s, const ngi_ray* d_rays, size_t n, ngi_hit* d_hits, int any_hit) {
    DeviceScene& d = *s.dev;
    if (n == 0) return NGI_OK;
    if (d.view.num_tris == 0) {
        // No geometry: every query misses (fill on the device).
        std::vector<ngi_hit> miss(n);
        for (size_t i = 0; i < n; ++i) miss[i] = ngi_hit{0.0f, 0.0f, 0.0f, NGI_MISS};
        NGI_HIP_CHECK(hipMemcpyAsync(d_hits, miss.data(), n * sizeof(ngi_hit), hipMemcpyHostToDevice, d.stream));
        NGI_HIP_CHECK(hipStreamSynchronize(d.stream));
        return NGI_OK;
    }
    const unsigned grid = (unsigned)((n + NGI_TRACE_BLOCK - 1) / NGI_TRACE_BLOCK);
    if (d.view.bvh_width == 8) {
        const size_t lds = (size_t)d.view.stack_entries * NGI_TRACE_BLOCK * 4;
        if (any_hit) trace_kernel<8, true><<<grid, NGI_TRACE_BLOCK, lds, d.stream>>>(d.view, d_rays, n, d_hits);
        else trace_kernel<8, false><<<grid, NGI_TRACE_BLOCK, lds, d.stream>>>(d.view, d_rays, n, d_hits);
    } else {
        if (any_hit) trace_kernel<2, true><<<grid, NGI_TRACE_BLOCK, 0, d.stream>>>(d.view, d_rays, n, d_hits);
        else trace_kernel<2, false><<<grid, NGI_TRACE_BLOCK, 0, d.stream>>>(d.view, d_rays, n, d_hits);
    }
    NGI_HIP_CHECK(hipGetLastError());
    return NGI_OK;
}

}  // namespace ngi

using namespace ngi;

extern "C" {

int ngi_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int ngi_trace(ngi_scene* s, const ngi_ray* rays, size_t n, ngi_hit* hits, int any_hit) {
    clear_error();
    if (!s || (n && (!rays || !hits))) return fail(NGI_ERR_INVALID_ARG, "null argument");
    int st = device_prepare(*s);
    if (st) return st;
    if (n == 0) return NGI_OK;
    DeviceScene& d = *s->dev;
    const size_t rb = n * sizeof(ngi_ray), hb = n * sizeof(ngi_hit);
    if ((st = ensure_stage(d, rb + hb))) return st;
    ngi_ray* dr = reinterpret_cast<ngi_ray*>(d.stage);
    ngi_hit* dh = reinterpret_cast<ngi_hit*>(reinterpret_cast<char*>(d.stage) + rb);
    NGI_HIP_CHECK(hipMemcpyAsync(dr, rays, rb, hipMemcpyHostToDevice, d.stream));
    if ((st = launch_trace(*s, dr, n, dh, any_hit))) return st;
    NGI_HIP_CHECK(hipMemcpyAsync(hits, dh, hb, hipMemcpyDeviceToHost, d.stream));
    NGI_HIP_CHECK(hipStreamSynchronize(d.stream));
    return NGI_OK;
}

int ngi_trace_device(ngi_scene* s, const ngi_ray* d_rays, size_t n, ngi_hit* d_hits, int any_hit) {
    clear_error();
    if (!s || (n && (!d_rays || !d_hits))) return fail(NGI_ERR_INVALID_ARG, "null argument");
    int st = device_prepare(*s);
    if (st) return st;
    return launch_trace(*s, d_rays, n, d_hits, any_hit);
}

}  // extern "C"
