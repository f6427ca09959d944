// Adjoint of pt / ptdirect by path replay (no reference counterpart).
//
// Gradient of  L = sum_p dL/dI_p . I_p  with respect to every primitive's
// reflectance (D.R, G.R) and every emitter's radiance (L.*.Le).
//
// Path sampling in ProcessSample_PT/_PTDirect (src/nanogi.cpp:446-802) does
// not depend on these parameters: directions come from cosine / Beckmann
// sampling, lights are picked uniformly and positions by area, and Russian
// roulette uses a constant probability.  Replaying a path from its
// counter-RNG seed therefore reproduces the same vertices, and every
// contribution C_k of the path is a product that is linear in the
// reflectance of each surface vertex before it (the BSDF weight fs/pdf, and
// the NEE factor fsE at the vertex itself) and in the radiance of its
// emitter.  With S = sum_k g_k C_k over the whole path (g_k = dL/dI at the
// contribution's pixel times the film scale W*H/N) and P_j the value of that
// sum when vertex j's reflectance first multiplies the path,
//     dL/dR_m += sum_{j : prim_j = m} (S - P_j) / R_m      (per channel)
//     dL/dLe_l += sum_{k : emitter_k = l} g_k C_k / Le_l.
// A parameter channel that is zero takes no part in a division: the replay
// runs with every zero reflectance or radiance channel replaced by 1 (path
// sampling does not depend on them, so the vertices are unchanged), and each
// path counts per channel the zero-reflectance vertices it has passed, z.
// The true value of a replayed contribution C~_k is C~_k if z_k = 0 and its
// emitter's Le is non-zero, else 0.  By the product rule:
//   - a record j with non-zero R gets (S_A - P_A,j) / R, S_A summing the
//     contributions with z = 0 and non-zero Le;
//   - the first zero-R record of a channel gets S_B, the sum of the
//     contributions with z = 1 and non-zero Le (its own factor is the one
//     that vanished; the replayed factor 1 leaves the rest);
//   - later zero-R records get 0 (another vanished factor remains);
//   - an emitter channel gets g C~ / Le~ from every contribution with z = 0
//     (Le~ = 1 where Le = 0).
//
// Layout: the persistent engine's regenerating lanes; per-lane vertex
// records (prim, P_j) in HBM as [record][lane] float4; per-block FP64 LDS
// accumulators for the first NGI_ADJ_LDS_PRIMS primitives, flushed with one
// FP64 global atomic per value and block; the rest go straight to global
// FP64 atomics.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "ngi_render.h"

namespace ngi {

#define NGI_ADJ_BLOCK 256
// Ended paths are replaced at once: waiting for 16 of them, as the forward
// persistent kernels do, made the 64-spp replay on the nave 35.83 ms against
// 34.58 ms.
// Minimum waves per SIMD for the adjoint kernel's register allocation: 4
// (128 VGPRs) measured fastest on the nave (38.2 ms per 64-spp replay against
// 44.8 ms unconstrained at 140 VGPRs / 3 waves, 38.6 ms at 5 waves).
#ifndef NGI_ADJ_WAVES
#define NGI_ADJ_WAVES 4
#endif
#define NGI_ADJ_LDS_PRIMS 256

// Zero-parameter mask per primitive: bits 0-2 reflectance channel is 0,
// bits 3-5 radiance channel is 0 (the replay tables hold 1 there).
#define NGI_ADJ_RZERO 1u
#define NGI_ADJ_LEZERO 8u
// Record word w: primitive in the low 29 bits, bits 29-31 = this record is
// the first zero-reflectance vertex of channel 0/1/2 on its path.
#define NGI_ADJ_PRIM_MASK 0x1fffffffu

struct AdjArgs {
    const float* dLdI;     // W*H*3
    float scale;           // W*H / N
    const uint8_t* zmask;  // per primitive, NGI_ADJ_RZERO / NGI_ADJ_LEZERO bits
    float4* rec;           // [max_rec][lanes]
    uint32_t lanes;
    int max_rec;
    double* gR;            // nprims*3
    double* gLe;           // nprims*3
    int nprims;
};

__device__ __forceinline__ void grad_add(double* lds, double* glob, int nlds, int prim, int c, float v) {
    if (v == 0.0f) return;
    if (prim < nlds) atomicAdd(lds + 3 * prim + c, (double)v);
    else unsafeAtomicAdd(glob + 3 * prim + c, (double)v);
}

// Adjoint sums of one path: SA / SB = weighted contributions with no / one
// zero-reflectance vertex before them, zc = per-channel count of such
// vertices (2 bits each, saturating at 2), nrec = vertex records written.
struct AdjPath {
    f3 SA, SB;
    uint32_t zc;
    int nrec;
};

// A contribution C at pixel `pix` from emitter primitive `emitter_prim`.
template <bool ZEROS>
__device__ __forceinline__ void adj_contribute(const AdjArgs& g, const DevPrim<float>* prims, double* sL, int nlds, AdjPath& q, int pix,
                                               f3 C, int emitter_prim) {
    const float* gp = g.dLdI + 3 * (size_t)pix;
    const float gC[3] = {gp[0] * C.x * g.scale, gp[1] * C.y * g.scale, gp[2] * C.z * g.scale};
    const DevPrim<float>& E = prims[emitter_prim];
    const uint32_t zm = ZEROS ? g.zmask[emitter_prim] : 0u;
    float sa[3] = {q.SA.x, q.SA.y, q.SA.z}, sb[3] = {q.SB.x, q.SB.y, q.SB.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const uint32_t z = (q.zc >> (2 * c)) & 3u;
        const bool le_zero = (zm & (NGI_ADJ_LEZERO << c)) != 0;
        if (z == 0) {
            grad_add(sL, g.gLe, nlds, emitter_prim, c, gC[c] / E.Le[c]);   // Le~ = 1 where Le = 0
            if (!le_zero) sa[c] += gC[c];
        } else if (z == 1 && !le_zero) {
            sb[c] += gC[c];
        }
    }
    q.SA = f3(sa[0], sa[1], sa[2]);
    q.SB = f3(sb[0], sb[1], sb[2]);
}

// The record of a surface vertex whose reflectance enters the path here.
template <bool ZEROS>
__device__ __forceinline__ void adj_record(const AdjArgs& g, uint32_t gid, AdjPath& q, int prim) {
    const uint32_t zm = ZEROS ? g.zmask[prim] : 0u;
    uint32_t w = (uint32_t)prim;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (!(zm & (NGI_ADJ_RZERO << c))) continue;
        const uint32_t z = (q.zc >> (2 * c)) & 3u;
        if (z == 0) w |= 1u << (29 + c);
        if (z < 2) q.zc += 1u << (2 * c);
    }
    g.rec[(size_t)q.nrec * g.lanes + gid] = make_float4(q.SA.x, q.SA.y, q.SA.z, __uint_as_float(w));
    q.nrec++;
}

// End of a path: the reflectance gradients of its records.
template <bool ZEROS>
__device__ __forceinline__ void adj_finish(const AdjArgs& g, const DevPrim<float>* prims, double* sR, int nlds, uint32_t gid, const AdjPath& q) {
    const float sa[3] = {q.SA.x, q.SA.y, q.SA.z}, sb[3] = {q.SB.x, q.SB.y, q.SB.z};
    for (int r = 0; r < q.nrec; ++r) {
        const float4 e = g.rec[(size_t)r * g.lanes + gid];
        const uint32_t w = __float_as_uint(e.w);
        const int m = (int)(w & NGI_ADJ_PRIM_MASK);
        const DevPrim<float>& P = prims[m];
        const uint32_t zm = ZEROS ? g.zmask[m] : 0u;
        const float pa[3] = {e.x, e.y, e.z};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (!(zm & (NGI_ADJ_RZERO << c))) grad_add(sR, g.gR, nlds, m, c, (sa[c] - pa[c]) / P.R[c]);
            else if (w & (1u << (29 + c))) grad_add(sR, g.gR, nlds, m, c, sb[c]);
        }
    }
}

// The light picked by a NEE sample (the draw of sample_light).
__device__ __forceinline__ int nee_light(const DevScene& sc, const DevLight<float>* lights, U4 rn) {
    int li = (int)(u24f(rn.x) * (float)sc.num_lights);
    li = li < 0 ? 0 : (li > sc.num_lights - 1 ? sc.num_lights - 1 : li);
    return lights[li].prim;
}

// ZEROS: some parameter channel is zero (the zero bookkeeping is compiled in
// only then; without it every mask is 0 and z stays 0).
template <bool DIRECT, int WIDTH, bool ZEROS>
__global__ void __launch_bounds__(NGI_ADJ_BLOCK, NGI_ADJ_WAVES) adjoint_kernel(RenderArgs a, AdjArgs g, unsigned long long* next) {
    extern __shared__ double adj_smem[];
    const int nlds = min(g.nprims, NGI_ADJ_LDS_PRIMS);
    double* sR = adj_smem;
    double* sL = adj_smem + 3 * nlds;
    uint32_t* stk = reinterpret_cast<uint32_t*>(adj_smem + 6 * nlds);
    for (int i = threadIdx.x; i < 6 * nlds; i += NGI_ADJ_BLOCK) adj_smem[i] = 0.0;
    __syncthreads();

    const DevScene& sc = a.sc;
    const Cam<float> cam = load_cam<float>(sc.cam);
    const DevPrim<float>* prims = prims_of<float>(sc);
    const DevLight<float>* lights = lights_of<float>(sc);
    const int max_nv = (a.max_nv < 0 || a.max_nv > NGI_MAX_VERTICES) ? NGI_MAX_VERTICES : a.max_nv;
    const uint32_t lane = __lane_id();
    const uint32_t gid = blockIdx.x * NGI_ADJ_BLOCK + threadIdx.x;
    uint32_t* my_stk = stk + threadIdx.x;

    // Path state (as ProcessSample_PT/_PTDirect) plus the adjoint sums.
    f3 T(0.0f), wi(0.0f);
    AdjPath q{f3(0.0f), f3(0.0f), 0u, 0};
    Surf<float> s;
    int type = 0, prim = 0, nv = 0, pixel = 0;
    uint32_t c0 = 0, c1 = 0;
    bool active = false, exhausted = false;

    for (;;) {
        if (!exhausted) {
            const unsigned long long need = __ballot(!active);
            if (need) {
                const int leader = __ffsll((long long)need) - 1;
                unsigned long long base = 0;
                if ((int)lane == leader) base = atomicAdd(next, (unsigned long long)__popcll(need));
                base = __shfl(base, leader);
                if (base + __popcll(need) >= a.path_count) exhausted = true;
                if (!active) {
                    const unsigned long long local = base + (unsigned long long)__popcll(need & ((1ull << lane) - 1ull));
                    if (local < a.path_count) {
                        path_coords(a, a.path_begin + local, c0, c1, pixel);
                        T = f3(1.0f);
                        type = NGI_PRIM_E;
                        prim = sc.sensor_prim;
                        s.p = cam.eye;
                        s.degenerate = true;
                        wi = f3(0.0f);
                        nv = 1;
                        q = AdjPath{f3(0.0f), f3(0.0f), 0u, 0};
                        active = true;
                    }
                }
            }
        }
        if (__ballot(active) == 0) break;
        if (!active) continue;
        // ---- one vertex (src/nanogi.cpp:483-606 / 645-801) ----
        bool alive = nv < max_nv;
        // Vertex reflectance enters here (NEE fsE, then the BSDF weight).  (A
        // register copy of the material taken before the record stores, as the
        // forward kernels take one, spills 6 VGPRs here: 35.08 against 34.38 ms
        // per 64-spp replay on the nave.)
        if (alive && (type & (NGI_PRIM_D | NGI_PRIM_G)) && q.nrec < g.max_rec) adj_record<ZEROS>(g, gid, q, prim);
        if (alive && DIRECT && sc.num_lights > 0) {
            const U4 rn = rng4(c0, c1, (uint32_t)nv, NGI_STREAM_NEE, a.seed, a.stream_id);
            TraceRay sr;
            int target;
            const f3 C = nee_contribution<float>(sc, cam, prims, type, prim, s, wi, T, rn, a.w, a.h, pixel, sr, target);
            if (!is_zero(C)) {
                TraceHit sh;
                TraceCounters tc{0, 0};
                if (!trace<WIDTH, true, false>(sc, sr, sh, tc, my_stk, NGI_ADJ_BLOCK))
                    adj_contribute<ZEROS>(g, prims, sL, nlds, q, target, C, nee_light(sc, lights, rn));
            }
        }
        f3 wo(0.0f);
        if (alive) {
            const U4 rx = rng4(c0, c1, (uint32_t)nv, NGI_STREAM_EXT, a.seed, a.stream_id);
            if (type == NGI_PRIM_E) {
                wo = camera_sample<float>(a, cam, u24f(rx.x), u24f(rx.y), pixel);
            } else {
                f3 wgt;
                alive = bsdf_sample<float>(prims[prim], type, s, wi, u24f(rx.x), u24f(rx.y), u24f(rx.z), wo, wgt);
                if (alive) T = T * wgt;
            }
            if (alive) {
                TraceRay er{s.p.x, s.p.y, s.p.z, wo.x, wo.y, wo.z, NGI_EPS_F, FLT_MAX};
                TraceHit h;
                TraceCounters tc{0, 0};
                alive = trace<WIDTH, false, false>(sc, er, h, tc, my_stk, NGI_ADJ_BLOCK);
                if (alive) {
                    int hp, htype;
                    const Surf<float> ns = hit_surface<float>(sc, s.p, wo, h, hp, htype);
                    if (!DIRECT && (htype & NGI_PRIM_L)) {
                        const DevPrim<float>& hm = prims[hp];
                        if (hm.light >= 0 && lights[hm.light].type == NGI_LIGHT_AREA && dot(ns.n, -wo) > 0.0f)
                            adj_contribute<ZEROS>(g, prims, sL, nlds, q, pixel, T * f3(hm.Le[0], hm.Le[1], hm.Le[2]), hp);
                    }
                    if (a.rr_depth >= 0 && nv >= a.rr_depth) {
                        if (u24f(rx.w) > a.rr_prob_f) alive = false;
                        else T = T / a.rr_prob_f;
                    }
                    type = htype & ~(NGI_PRIM_L | NGI_PRIM_E);
                    if (type == 0) alive = false;
                    s = ns;
                    prim = hp;
                    wi = -wo;
                    nv++;
                }
            }
        }
        if (!alive) {
            adj_finish<ZEROS>(g, prims, sR, nlds, gid, q);
            active = false;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * nlds; i += NGI_ADJ_BLOCK) {
        if (sR[i] != 0.0) unsafeAtomicAdd(g.gR + i, sR[i]);
        if (sL[i] != 0.0) unsafeAtomicAdd(g.gLe + i, sL[i]);
    }
}

// The adjoint kernel with resumable extension traversals (wide BVH): a
// wavefront stops traversing once at most NGI_ADJ_KEEP of its lanes are still
// inside the tree; the others complete their vertices and start their next
// rays while those wait, their traversal state, direction and roulette number
// parked in LDS.  NEE shadow rays are traced where they are made (their
// contributions enter the adjoint sums before the next vertex record), on a
// second LDS stack.  Every path takes the same vertices as in adjoint_kernel.
#ifndef NGI_ADJ_KEEP
#define NGI_ADJ_KEEP 8
#endif
template <bool DIRECT, bool ZEROS>
__global__ void __launch_bounds__(NGI_ADJ_BLOCK, NGI_ADJ_WAVES) adjoint_resume_kernel(RenderArgs a, AdjArgs g, unsigned long long* next) {
    // LDS: the FP64 accumulators, then extension stacks, shadow stacks
    // (a.sc.stack_entries words per lane each) and the parked lane state.
    extern __shared__ double adj_smem[];
    const int nlds = min(g.nprims, NGI_ADJ_LDS_PRIMS);
    double* sR = adj_smem;
    double* sL = adj_smem + 3 * nlds;
    uint32_t* words = reinterpret_cast<uint32_t*>(adj_smem + 6 * nlds);
    for (int i = threadIdx.x; i < 6 * nlds; i += NGI_ADJ_BLOCK) adj_smem[i] = 0.0;
    __syncthreads();

    const DevScene& sc = a.sc;
    const Cam<float> cam = load_cam<float>(sc.cam);
    const DevPrim<float>* prims = prims_of<float>(sc);
    const DevLight<float>* lights = lights_of<float>(sc);
    const int max_nv = (a.max_nv < 0 || a.max_nv > NGI_MAX_VERTICES) ? NGI_MAX_VERTICES : a.max_nv;
    const uint32_t lane = __lane_id();
    const uint32_t gid = blockIdx.x * NGI_ADJ_BLOCK + threadIdx.x;
    uint32_t* ext_stk = words + threadIdx.x;
    uint32_t* sh_stk = words + (size_t)sc.stack_entries * NGI_ADJ_BLOCK + threadIdx.x;
    uint32_t* pk = words + (size_t)2 * sc.stack_entries * NGI_ADJ_BLOCK + threadIdx.x;
    const bool geometry = sc.num_tris != 0;

    f3 T(0.0f), wi(0.0f);
    AdjPath q{f3(0.0f), f3(0.0f), 0u, 0};
    Surf<float> s;
    int type = 0, prim = 0, nv = 0, pixel = 0;
    uint32_t c0 = 0, c1 = 0;
    bool active = false;      // the lane holds a path
    bool busy = false;        // ... whose extension ray is being traced
    bool exhausted = false;   // wave-uniform: no path ids left
    for (;;) {
        // Vertices whose extension rays have finished (hit, emission for pt,
        // Russian roulette).
        if (active && !busy) {
            const TraceHit h = parked_hit<NGI_ADJ_BLOCK>(pk);
            bool alive = h.leaf >= 0;
            if (alive) {
                const f3 wo = parked_dir<NGI_ADJ_BLOCK>(pk);
                int hp, htype;
                const Surf<float> ns = hit_surface<float>(sc, s.p, wo, h, hp, htype);
                if (!DIRECT && (htype & NGI_PRIM_L)) {
                    const DevPrim<float>& hm = prims[hp];
                    if (hm.light >= 0 && lights[hm.light].type == NGI_LIGHT_AREA && dot(ns.n, -wo) > 0.0f)
                        adj_contribute<ZEROS>(g, prims, sL, nlds, q, pixel, T * f3(hm.Le[0], hm.Le[1], hm.Le[2]), hp);
                }
                if (a.rr_depth >= 0 && nv >= a.rr_depth) {
                    if (parked_urr<NGI_ADJ_BLOCK>(pk) > a.rr_prob_f) alive = false;
                    else T = T / a.rr_prob_f;
                }
                type = htype & ~(NGI_PRIM_L | NGI_PRIM_E);
                if (type == 0) alive = false;
                s = ns;
                prim = hp;
                wi = -wo;
                nv++;
            }
            if (!alive) {
                adj_finish<ZEROS>(g, prims, sR, nlds, gid, q);
                active = false;
            }
        }
        if (!exhausted) {
            const unsigned long long need = __ballot(!active);
            if (need) {
                const int leader = __ffsll((long long)need) - 1;
                unsigned long long base = 0;
                if ((int)lane == leader) base = atomicAdd(next, (unsigned long long)__popcll(need));
                base = __shfl(base, leader);
                if (base + __popcll(need) >= a.path_count) exhausted = true;
                if (!active) {
                    const unsigned long long local = base + (unsigned long long)__popcll(need & ((1ull << lane) - 1ull));
                    if (local < a.path_count) {
                        path_coords(a, a.path_begin + local, c0, c1, pixel);
                        T = f3(1.0f);
                        type = NGI_PRIM_E;
                        prim = sc.sensor_prim;
                        s.p = cam.eye;
                        s.degenerate = true;
                        wi = f3(0.0f);
                        nv = 1;
                        q = AdjPath{f3(0.0f), f3(0.0f), 0u, 0};
                        active = true;
                    }
                }
            }
        }
        // Next vertex of every path not inside the tree: record, NEE,
        // direction sample, extension ray.
        if (active && !busy) {
            bool alive = nv < max_nv;
            if (alive && (type & (NGI_PRIM_D | NGI_PRIM_G)) && q.nrec < g.max_rec) adj_record<ZEROS>(g, gid, q, prim);
            if (alive && DIRECT && sc.num_lights > 0) {
                const U4 rn = rng4(c0, c1, (uint32_t)nv, NGI_STREAM_NEE, a.seed, a.stream_id);
                TraceRay sr;
                int target;
                const f3 C = nee_contribution<float>(sc, cam, prims, type, prim, s, wi, T, rn, a.w, a.h, pixel, sr, target);
                if (!is_zero(C)) {
                    TraceHit sh;
                    TraceCounters tc{0, 0};
                    if (!trace<8, true, false>(sc, sr, sh, tc, sh_stk, NGI_ADJ_BLOCK))
                        adj_contribute<ZEROS>(g, prims, sL, nlds, q, target, C, nee_light(sc, lights, rn));
                }
            }
            if (alive) {
                const U4 rx = rng4(c0, c1, (uint32_t)nv, NGI_STREAM_EXT, a.seed, a.stream_id);
                f3 wo(0.0f);
                if (type == NGI_PRIM_E) {
                    wo = camera_sample<float>(a, cam, u24f(rx.x), u24f(rx.y), pixel);
                } else {
                    f3 wgt;
                    alive = bsdf_sample<float>(prims[prim], type, s, wi, u24f(rx.x), u24f(rx.y), u24f(rx.z), wo, wgt);
                    if (alive) T = T * wgt;
                }
                if (alive) {
                    Trav8 tv;
                    trav8_begin(FLT_MAX, tv);
                    park_trav<NGI_ADJ_BLOCK>(pk, tv);
                    park_dir<NGI_ADJ_BLOCK>(pk, wo, u24f(rx.w));
                    busy = geometry;
                }
            }
            if (!alive) {
                adj_finish<ZEROS>(g, prims, sR, nlds, gid, q);
                active = false;
            }
        }
        if (__ballot(active) == 0) {
            if (exhausted) break;
            continue;
        }
        if (__ballot(busy)) {
            // Lanes without a ray load and store back what they parked last.
            Trav8 tv = parked_trav<NGI_ADJ_BLOCK>(pk);
            const f3 d = parked_dir<NGI_ADJ_BLOCK>(pk);
            const RayPre r = make_pre(TraceRay{s.p.x, s.p.y, s.p.z, d.x, d.y, d.z, NGI_EPS_F, FLT_MAX});
            const int keep = __builtin_amdgcn_readfirstlane(exhausted ? 0 : NGI_ADJ_KEEP);
            TraceCounters tc{0, 0};
            trav8_run<false>(sc, r, NGI_EPS_F, tv, busy, keep, tc, ext_stk, NGI_ADJ_BLOCK);
            park_trav<NGI_ADJ_BLOCK>(pk, tv);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * nlds; i += NGI_ADJ_BLOCK) {
        if (sR[i] != 0.0) unsafeAtomicAdd(g.gR + i, sR[i]);
        if (sL[i] != 0.0) unsafeAtomicAdd(g.gLe + i, sL[i]);
    }
}

template <bool DIRECT, int WIDTH, bool ZEROS, bool RESUME>
static const void* adjoint_fn() {
    if (RESUME) return reinterpret_cast<const void*>(adjoint_resume_kernel<DIRECT, ZEROS>);
    return reinterpret_cast<const void*>(adjoint_kernel<DIRECT, WIDTH, ZEROS>);
}

template <bool DIRECT, int WIDTH, bool ZEROS, bool RESUME>
static int launch_adjoint(const RenderArgs& a, const AdjArgs& g, unsigned long long* next, unsigned grid, size_t lds, hipStream_t st) {
    if (RESUME) adjoint_resume_kernel<DIRECT, ZEROS><<<grid, NGI_ADJ_BLOCK, lds, st>>>(a, g, next);
    else adjoint_kernel<DIRECT, WIDTH, ZEROS><<<grid, NGI_ADJ_BLOCK, lds, st>>>(a, g, next);
    NGI_HIP_CHECK(hipGetLastError());
    return NGI_OK;
}

template <bool DIRECT, int WIDTH, bool ZEROS, bool RESUME>
static int adjoint_grid(size_t lds, unsigned& grid) {
    int dev = 0, cus = 0, per_cu = 0;
    NGI_HIP_CHECK(hipGetDevice(&dev));
    NGI_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    NGI_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, adjoint_fn<DIRECT, WIDTH, ZEROS, RESUME>(), NGI_ADJ_BLOCK, lds));
    grid = (unsigned)std::max(1, cus * std::max(1, per_cu));
    return NGI_OK;
}

// One of the eight kernel variants, chosen at run time.
template <bool RESUME>
static int adjoint_dispatch(bool direct, int width, bool zeros, bool grid_only, const RenderArgs& a, const AdjArgs& g, unsigned long long* next,
                            unsigned& grid, size_t lds, hipStream_t st) {
#define NGI_ADJ_CASE(D, W, Z) \
    return grid_only ? adjoint_grid<D, W, Z, RESUME>(lds, grid) : launch_adjoint<D, W, Z, RESUME>(a, g, next, grid, lds, st)
    if (direct) {
        if (width == 8) { if (zeros) NGI_ADJ_CASE(true, 8, true); else NGI_ADJ_CASE(true, 8, false); }
        if (zeros) NGI_ADJ_CASE(true, 2, true);
        NGI_ADJ_CASE(true, 2, false);
    }
    if (width == 8) { if (zeros) NGI_ADJ_CASE(false, 8, true); else NGI_ADJ_CASE(false, 8, false); }
    if (zeros) NGI_ADJ_CASE(false, 2, true);
    NGI_ADJ_CASE(false, 2, false);
#undef NGI_ADJ_CASE
}

int render_adjoint(ngi_scene* scene, const ngi_render_params& p, const double* dLdI, double* gR_out, double* gLe_out, ngi_stats* stats) {
    if (p.renderer != NGI_RENDERER_PT && p.renderer != NGI_RENDERER_PTDIRECT)
        return fail(NGI_ERR_UNSUPPORTED, "the adjoint differentiates pt and ptdirect");
    Session S;
    ngi_render_params q = p;
    q.engine = NGI_ENGINE_PERSISTENT;
    int st = S.init(scene, q);
    if (st) {
        S.destroy();
        return st;
    }
    if (p.precision == NGI_PRECISION_FP64_ALL) {
        S.destroy();
        return fail(NGI_ERR_UNSUPPORTED, "the adjoint shades in FP32 (FP64 is used for the gradient reductions)");
    }

    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t str = S.stream();
    const uint32_t npix = S.npix;
    const uint64_t total = p.sampling == NGI_SAMPLING_PER_PIXEL ? (uint64_t)npix * (uint64_t)std::max(p.spp, 0) : (uint64_t)p.num_samples;
    const int nprims = (int)scene->prims.size();
    const bool direct = p.renderer == NGI_RENDERER_PTDIRECT;
    const int width = scene->dev->view.bvh_width;
    const int nlds = std::min(nprims, NGI_ADJ_LDS_PRIMS);
    // Resumable traversals (wide BVH) while their LDS leaves the kernel its
    // occupancy of NGI_ADJ_WAVES blocks per CU.
    const size_t resume_lds = (size_t)6 * nlds * 8 + ((size_t)2 * scene->dev->view.stack_entries + PK_WORDS) * NGI_ADJ_BLOCK * 4;
    const bool resume = width == 8 && resume_lds * NGI_ADJ_WAVES <= NGI_CU_LDS_BYTES;
    const size_t lds = resume ? resume_lds : (size_t)6 * nlds * 8 + (width == 8 ? (size_t)scene->dev->view.stack_entries * NGI_ADJ_BLOCK * 4 : 0);
    unsigned grid = 1;
    std::vector<void*> mem;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(str);
        for (void* m : mem) (void)hipFree(m);
        S.destroy();
    };
    const int max_nv = (p.max_num_vertices < 0 || p.max_num_vertices > NGI_MAX_VERTICES) ? NGI_MAX_VERTICES : p.max_num_vertices;
    AdjArgs g;
    g.max_rec = max_nv;
    g.nprims = nprims;
    g.scale = total ? (float)((double)npix / (double)total) : 0.0f;
    std::vector<float> dl(3 * (size_t)npix);
    for (size_t i = 0; i < dl.size(); ++i) dl[i] = (float)dLdI[i];
    if ((uint32_t)nprims > NGI_ADJ_PRIM_MASK) {
        cleanup();
        return fail(NGI_ERR_INVALID_ARG, "adjoint: at most 2^29 - 1 primitives");
    }
    // Replay tables: the scene's FP32 primitive and light records with every
    // zero reflectance (D/G) or radiance (L) channel set to 1, and the mask of
    // those channels.
    const DevScene& view = scene->dev->view;
    std::vector<DevPrim<float>> rp(std::max(nprims, 1));
    std::vector<DevLight<float>> rl(std::max(view.num_lights, 1));
    std::vector<uint8_t> zm(std::max(nprims, 1), 0);
    if (hipMemcpy(rp.data(), view.prims_f, (size_t)nprims * sizeof(DevPrim<float>), hipMemcpyDeviceToHost) != hipSuccess ||
        (view.num_lights &&
         hipMemcpy(rl.data(), view.lights_f, (size_t)view.num_lights * sizeof(DevLight<float>), hipMemcpyDeviceToHost) != hipSuccess)) {
        cleanup();
        return fail(NGI_ERR_DEVICE, "adjoint: cannot read the parameter tables");
    }
    for (int i = 0; i < nprims; ++i) {
        const int ty = scene->prims[i].type;
        for (int c = 0; c < 3; ++c) {
            if ((ty & (NGI_PRIM_D | NGI_PRIM_G)) && rp[i].R[c] == 0.0f) zm[i] |= (uint8_t)(NGI_ADJ_RZERO << c), rp[i].R[c] = 1.0f;
            if ((ty & NGI_PRIM_L) && rp[i].Le[c] == 0.0f) zm[i] |= (uint8_t)(NGI_ADJ_LEZERO << c), rp[i].Le[c] = 1.0f;
        }
    }
    for (int l = 0; l < view.num_lights; ++l)
        for (int c = 0; c < 3; ++c)
            if (zm[rl[l].prim] & (NGI_ADJ_LEZERO << c)) rl[l].Le[c] = 1.0f;
    const bool zeros = std::any_of(zm.begin(), zm.end(), [](uint8_t v) { return v != 0; });
    RenderArgs a = S.args(nullptr);
    st = resume ? adjoint_dispatch<true>(direct, width, zeros, true, a, g, nullptr, grid, lds, str)
                : adjoint_dispatch<false>(direct, width, zeros, true, a, g, nullptr, grid, lds, str);
    if (st) {
        cleanup();
        return st;
    }
    g.lanes = grid * NGI_ADJ_BLOCK;
    void *d_dl = nullptr, *d_rec = nullptr, *d_gR = nullptr, *d_gL = nullptr, *d_next = nullptr;
    void *d_rp = nullptr, *d_rl = nullptr, *d_zm = nullptr;
#define NGI_ADJ_ALLOC(ptr, bytes)                                   \
    do {                                                            \
        hipError_t e_ = hipMalloc(&ptr, bytes);                     \
        if (e_ != hipSuccess) {                                     \
            cleanup();                                              \
            return fail(NGI_ERR_OOM, "adjoint: hipMalloc failed");  \
        }                                                           \
        mem.push_back(ptr);                                         \
    } while (0)
    NGI_ADJ_ALLOC(d_dl, dl.size() * 4);
    NGI_ADJ_ALLOC(d_rec, (size_t)g.lanes * g.max_rec * 16);
    NGI_ADJ_ALLOC(d_gR, (size_t)nprims * 3 * 8 + 8);
    NGI_ADJ_ALLOC(d_gL, (size_t)nprims * 3 * 8 + 8);
    NGI_ADJ_ALLOC(d_next, 8);
    NGI_ADJ_ALLOC(d_rp, rp.size() * sizeof(DevPrim<float>));
    NGI_ADJ_ALLOC(d_rl, rl.size() * sizeof(DevLight<float>));
    NGI_ADJ_ALLOC(d_zm, zm.size());
#undef NGI_ADJ_ALLOC
    if (hipMemcpyAsync(d_dl, dl.data(), dl.size() * 4, hipMemcpyHostToDevice, str) != hipSuccess ||
        hipMemcpyAsync(d_rp, rp.data(), rp.size() * sizeof(DevPrim<float>), hipMemcpyHostToDevice, str) != hipSuccess ||
        hipMemcpyAsync(d_rl, rl.data(), rl.size() * sizeof(DevLight<float>), hipMemcpyHostToDevice, str) != hipSuccess ||
        hipMemcpyAsync(d_zm, zm.data(), zm.size(), hipMemcpyHostToDevice, str) != hipSuccess) {
        cleanup();
        return fail(NGI_ERR_DEVICE, "adjoint: upload failed");
    }
    st = S.zero(d_gR, (size_t)nprims * 3, true);
    if (!st) st = S.zero(d_gL, (size_t)nprims * 3, true);
    if (!st) st = S.zero(d_next, 1, true);
    g.dLdI = reinterpret_cast<const float*>(d_dl);
    g.rec = reinterpret_cast<float4*>(d_rec);
    g.gR = reinterpret_cast<double*>(d_gR);
    g.gLe = reinterpret_cast<double*>(d_gL);
    g.zmask = reinterpret_cast<const uint8_t*>(d_zm);
    a.sc.prims_f = reinterpret_cast<const DevPrim<float>*>(d_rp);
    a.sc.lights_f = reinterpret_cast<const DevLight<float>*>(d_rl);
    a.path_begin = 0;
    a.path_count = total;
    a.sample_base = 0;
    if (!st && total) {
        unsigned long long* nx = reinterpret_cast<unsigned long long*>(d_next);
        st = resume ? adjoint_dispatch<true>(direct, width, zeros, false, a, g, nx, grid, lds, str)
                    : adjoint_dispatch<false>(direct, width, zeros, false, a, g, nx, grid, lds, str);
    }
    if (!st && hipMemcpyAsync(gR_out, d_gR, (size_t)nprims * 3 * 8, hipMemcpyDeviceToHost, str) != hipSuccess) st = fail(NGI_ERR_DEVICE, "adjoint: read-back failed");
    if (!st && hipMemcpyAsync(gLe_out, d_gL, (size_t)nprims * 3 * 8, hipMemcpyDeviceToHost, str) != hipSuccess) st = fail(NGI_ERR_DEVICE, "adjoint: read-back failed");
    if (!st && hipStreamSynchronize(str) != hipSuccess) st = fail(NGI_ERR_DEVICE, "adjoint: kernel failed");
    if (!st) {
        // Only parameters that exist: reflectance of D/G primitives, radiance of emitters.
        for (int i = 0; i < nprims; ++i) {
            const int ty = scene->prims[i].type;
            if (!(ty & (NGI_PRIM_D | NGI_PRIM_G))) gR_out[3 * i] = gR_out[3 * i + 1] = gR_out[3 * i + 2] = 0.0;
            if (!(ty & NGI_PRIM_L)) gLe_out[3 * i] = gLe_out[3 * i + 1] = gLe_out[3 * i + 2] = 0.0;
        }
        if (stats) {
            memset(stats, 0, sizeof(*stats));
            stats->samples = (int64_t)total;
            stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
    }
    cleanup();
    return st;
}

}  // namespace ngi

using namespace ngi;

extern "C" int ngi_render_adjoint(ngi_scene* scene, const ngi_render_params* p, const double* dL_dimage, double* grad_reflectance,
                                  double* grad_Le, ngi_stats* stats) {
    clear_error();
    if (!scene || !p || !dL_dimage || !grad_reflectance || !grad_Le) return fail(NGI_ERR_INVALID_ARG, "null argument");
    return render_adjoint(scene, *p, dL_dimage, grad_reflectance, grad_Le, stats);
}
