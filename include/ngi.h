/*
 * ngi.h - C ABI of the MI355X-native `pt` / `ptdirect` renderer (and the
 * `lt` / `ltdirect` light tracers and the `bdpt` bidirectional path tracer).
 *
 * This is the drop-in boundary for nanogi's path- and light-tracing renderers.
 * Every entry point names the reference interface it replaces
 * (paths relative to the nanogi source tree).  Plain pointers and sizes
 * only; no C++ or torch types cross this boundary.  All output buffers are
 * caller-allocated.  Every function returns an ngi_status (0 = OK) unless
 * stated otherwise; on failure ngi_last_error() holds a message.
 *
 * Threading: a scene handle owns one HIP stream and is not reentrant
 * (callers serialise calls on one handle).  Distinct handles may be used
 * from distinct threads.
 */
#ifndef NGI_H
#define NGI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NGI_ABI_VERSION 1

/* ------------------------------------------------------------------ */
/* Status codes.  The reference returns `bool` from Scene::Load and    */
/* logs (include/nanogi/rt.hpp:1519-2154, src/nanogi.cpp:2123-2151);   */
/* each false-return site maps to one code here.                       */
/* ------------------------------------------------------------------ */
typedef enum ngi_status {
    NGI_OK = 0,
    NGI_ERR_IO = 1,              /* file missing / unreadable (YAML::LoadFile, Assimp ReadFile rt.hpp:1525,1643-1647) */
    NGI_ERR_PARSE = 2,           /* YAML / OBJ syntax error (YAML::Exception rt.hpp:2147-2151) */
    NGI_ERR_VERSION = 3,         /* version outside [3,5] (rt.hpp:1482-1486, 1535-1540) */
    NGI_ERR_INVALID_TYPE = 4,    /* empty type or [L,E] (rt.hpp:1600-1604) */
    NGI_ERR_INVALID_SCENE = 5,   /* area light w/o mesh (rt.hpp:1826-1830), sensor w/o UV (1921-1925), no mesh in file (1649-1652), missing D/G reflectance (1943-1946) */
    NGI_ERR_UNSUPPORTED = 6,     /* ptmnee, textures, area sensor, unsupported precision/engine combinations */
    NGI_ERR_INVALID_ARG = 7,
    NGI_ERR_DEVICE = 8,          /* no GPU / HIP runtime error */
    NGI_ERR_OOM = 9,
    NGI_ERR_QUEUE_OVERFLOW = 10
} ngi_status;

/* Thread-local message describing the last failure ("" if none). */
const char* ngi_last_error(void);

/* Number of visible HIP devices (0 when none); never fails. */
int ngi_device_count(void);

/* ------------------------------------------------------------------ */
/* Scene data model: include/nanogi/rt.hpp:148-155 (Mesh),             */
/* :338-377 (type bits and sub-type enums), :379-477 (Primitive).     */
/* ------------------------------------------------------------------ */
#define NGI_PRIM_D 1   /* diffuse          PrimitiveType::D */
#define NGI_PRIM_G 2   /* glossy           PrimitiveType::G */
#define NGI_PRIM_S 4   /* specular         PrimitiveType::S */
#define NGI_PRIM_L 8   /* light            PrimitiveType::L */
#define NGI_PRIM_E 16  /* sensor (camera)  PrimitiveType::E */

enum { NGI_LIGHT_AREA = 0, NGI_LIGHT_POINT = 1, NGI_LIGHT_DIRECTIONAL = 2 };       /* LType rt.hpp:353-358 */
enum { NGI_SENSOR_AREA = 0, NGI_SENSOR_PINHOLE = 1 };                                /* EType rt.hpp:360-364 */
enum { NGI_SPEC_REFLECTION = 0, NGI_SPEC_REFRACTION = 1, NGI_SPEC_FRESNEL = 2 };     /* SType rt.hpp:366-371 */

/* One primitive = one mesh + one material/emitter record
 * (Primitive::Params, rt.hpp:389-477).  Triangles of the primitive are the
 * contiguous range [tri_begin, tri_begin + tri_count) of the scene's
 * triangle array; tri_count == 0 means "no mesh" (pinhole, point light). */
typedef struct ngi_primitive {
    int32_t type;          /* OR of NGI_PRIM_* */
    int32_t light_type;    /* NGI_LIGHT_*  (valid if type & L) */
    int32_t sensor_type;   /* NGI_SENSOR_* (valid if type & E) */
    int32_t spec_type;     /* NGI_SPEC_*   (valid if type & S) */
    uint32_t tri_begin;
    uint32_t tri_count;
    double Le[3];          /* L.Area.Le / L.Point.Le / L.Directional.Le */
    double light_position[3];   /* L.Point.Position */
    double light_direction[3];  /* L.Directional.Direction */
    double D_R[3];         /* D.R */
    double G_R[3];         /* G.R */
    double G_eta[3];       /* G.Eta */
    double G_k[3];         /* G.K */
    double G_roughness;    /* G.Roughness (Beckmann alpha, rt.hpp:1407-1414) */
    double S_R[3];         /* S.{Reflection,Refraction,Fresnel}.R */
    double S_eta1, S_eta2; /* S.{Refraction,Fresnel}.Eta1/Eta2 */
    double E_We[3];        /* E.Pinhole.We / E.Area.We (parsed, unused by pinhole importance) */
    double E_eye[3], E_center[3], E_up[3];
    double E_fov_deg;      /* vertical field of view in degrees (rt.hpp:1897) */
} ngi_primitive;

/* Flat scene description (the FP64 arrays of rt.hpp Mesh, concatenated).
 * faces index into positions/normals; tri_prim[i] is the primitive owning
 * triangle i and must agree with the primitives' [tri_begin, tri_count)
 * ranges.  normals may be NULL (flat normals are then generated). */
typedef struct ngi_scene_desc {
    uint32_t num_vertices;
    const double* positions;     /* 3 * num_vertices */
    const double* normals;       /* 3 * num_vertices, or NULL */
    const double* texcoords;     /* 2 * num_vertices, or NULL */
    uint32_t num_triangles;
    const uint32_t* faces;       /* 3 * num_triangles */
    const uint32_t* tri_prim;    /* num_triangles */
    uint32_t num_primitives;
    const ngi_primitive* primitives;
    double aspect;               /* W/H passed to Scene::Load (src/nanogi.cpp:2069) */
} ngi_scene_desc;

typedef struct ngi_scene ngi_scene;

/* Scene::Load(path, aspect) (rt.hpp:1519).  YAML version 3..5 scene per
 * schema.yml with OBJ meshes; builds nothing on the GPU yet. */
int ngi_scene_load_yaml(const char* path, double aspect, ngi_scene** out);

/* Create a scene from flat arrays (copied). */
int ngi_scene_create(const ngi_scene_desc* desc, ngi_scene** out);

/* Procedural scenes used by the benchmarks (no reference counterpart). */
enum {
    NGI_GEN_CORNELL = 1,      /* 36-triangle Cornell-style box */
    NGI_GEN_NAVE = 2,         /* ~75k-triangle interior, 6 albedos, 4 quad lights */
    NGI_GEN_SPHERE_FIELD = 3, /* displaced 8x-subdivided icosphere on a 512^2 heightfield */
    NGI_GEN_BATCH_ROOM = 4    /* Cornell-style room with seeded boxes/icospheres; detail = target triangles */
};
int ngi_scene_generate(int kind, uint64_t seed, double aspect, int detail, ngi_scene** out);

/* Write a scene as YAML version 5 + one OBJ per meshed primitive into
 * directory dir (the inverse of ngi_scene_load_yaml). */
int ngi_scene_save_yaml(const ngi_scene* scene, const char* dir);

void ngi_scene_destroy(ngi_scene* scene);

/* Pointers into the scene's own host arrays (valid while the scene lives). */
int ngi_scene_get_desc(const ngi_scene* scene, ngi_scene_desc* out);

/* Replace per-primitive reflectance (D.R, or G.R for glossy) and emitter
 * radiance (L.*.Le): arrays of num_primitives*3 doubles, either may be NULL.
 * Entries of primitives without the matching type bit are ignored.  The
 * device tables are rewritten at the next render, after every enqueued
 * render or batch frame that reads them has finished, so a frame enqueued
 * before this call renders with the old values. */
int ngi_scene_set_params(ngi_scene* scene, const double* reflectance, const double* Le);

/* ------------------------------------------------------------------ */
/* Acceleration structure: replaces the Embree build (rt.hpp:2087-2143)*/
/* ------------------------------------------------------------------ */
enum { NGI_BVH_2 = 2, NGI_BVH_8 = 8 };

typedef struct ngi_bvh_opts {
    int width;             /* NGI_BVH_2 or NGI_BVH_8 (0 = default 8) */
    int max_leaf_size;     /* triangles per leaf, at most 16; 0 = default (4) */
    int sah_bins;          /* 0 = default (32) */
    int num_threads;       /* 0 = hardware concurrency */
    double traversal_cost; /* SAH cost of a node visit relative to one triangle test (0 = default: 1.2 for
                            * the binary BVH's splits and leaves, 1.0 for the wide BVH's collapse) */
} ngi_bvh_opts;

typedef struct ngi_bvh_info {
    int width;
    uint32_t num_nodes;
    uint32_t num_leaves;
    uint32_t num_triangles;
    uint32_t max_depth;
    uint32_t node_bytes;     /* bytes per node record */
    double sah_cost;
    double build_ms;
} ngi_bvh_info;

/* Build the host BVH (deterministic for a given scene and options).
 * Rebuilding releases the scene's device copy after waiting for every
 * kernel that reads it (its own stream and the streams of batches holding
 * the scene).  Sessions of the scene stay valid: their films keep their
 * contents and the scene is uploaded again at the session's next call. */
int ngi_scene_build(ngi_scene* scene, const ngi_bvh_opts* opts);
int ngi_scene_bvh_info(const ngi_scene* scene, ngi_bvh_info* out);
/* Copy out the flattened node records (info.num_nodes * info.node_bytes
 * bytes) and the leaf-order triangle permutation (num_triangles uint32). */
int ngi_scene_bvh_export(const ngi_scene* scene, void* nodes, uint32_t* leaf_to_tri);

/* ------------------------------------------------------------------ */
/* Ray queries: Scene::Intersect (rt.hpp:2162-2249) and Scene::Visible */
/* (rt.hpp:2251-2261).  FP32, two-sided, watertight.                   */
/* ------------------------------------------------------------------ */
typedef struct ngi_ray {
    float o[3];
    float tmin;            /* reference default 1e-4 (EpsF, basic.hpp:85) */
    float d[3];
    float tmax;            /* reference default FLT_MAX (InfF) */
} ngi_ray;

typedef struct ngi_hit {
    float t;
    float u, v;            /* barycentric weights of vertex 1 and 2 (Embree convention, rt.hpp:2212) */
    uint32_t tri;          /* global triangle index, 0xFFFFFFFF = miss (any-hit: 0 = occluded marker) */
} ngi_hit;

#define NGI_MISS 0xFFFFFFFFu

/* Host buffers in and out (copies to and from the device). any_hit != 0
 * performs the occlusion query of Visible: hit.tri != NGI_MISS iff some
 * triangle is hit in (tmin, tmax); t/u/v are then unspecified. */
int ngi_trace(ngi_scene* scene, const ngi_ray* rays, size_t n, ngi_hit* hits, int any_hit);
/* Device-resident buffers, enqueued on the scene's stream (no sync). */
int ngi_trace_device(ngi_scene* scene, const ngi_ray* d_rays, size_t n, ngi_hit* d_hits, int any_hit);

/* ------------------------------------------------------------------ */
/* Rendering: RendererType + Renderer::Render (src/nanogi.cpp:51-71,   */
/* 182-221), RenderProcess (225-440), ProcessSample_PT (446-607),      */
/* ProcessSample_PTDirect (609-802), ProcessSample_LT (804-934),       */
/* ProcessSample_LTDirect (936-1131), ProcessSample_BDPT (1133-1189).  */
/* ------------------------------------------------------------------ */
enum { NGI_RENDERER_PT = 0, NGI_RENDERER_PTDIRECT = 1, NGI_RENDERER_LT = 2,
       NGI_RENDERER_LTDIRECT = 3, NGI_RENDERER_BDPT = 4, NGI_RENDERER_PTMNEE = 5 };
/* RendererType_String lookup (src/nanogi.cpp:61-69); -1 if unknown. */
int ngi_renderer_from_string(const char* name);

enum { NGI_SAMPLING_PER_PIXEL = 0,            /* spp stratified by pixel */
       NGI_SAMPLING_REFERENCE_UNIFORM = 1 };  /* raster uniform over the image (rt.hpp:733-740) */
enum { NGI_PRECISION_FP32_FP64ACC = 0,        /* FP32 traversal+shading, FP64 film */
       NGI_PRECISION_FP32_ALL = 1,            /* FP32 film too */
       NGI_PRECISION_FP64_ALL = 2 };          /* FP64 shading (traversal stays FP32, as Embree) */
enum { NGI_ENGINE_MEGAKERNEL = 0,   /* one path per lane */
       NGI_ENGINE_WAVEFRONT = 1,    /* staged queues: generate / extend / shade / shadow */
       NGI_ENGINE_PERSISTENT = 2 }; /* persistent lanes with path regeneration */

typedef struct ngi_render_params {
    int renderer;            /* NGI_RENDERER_PT, _PTDIRECT, _LT, _LTDIRECT or _BDPT */
    int width, height;       /* -w/-h (src/nanogi.cpp:2006-2007) */
    int sampling;            /* NGI_SAMPLING_* */
    int64_t num_samples;     /* -n: total paths, REFERENCE_UNIFORM mode */
    int spp;                 /* samples per pixel, PER_PIXEL mode */
    int max_num_vertices;    /* -m: -1 = unlimited (src/nanogi.cpp:2005) */
    int rr_depth;            /* Russian roulette applies once numVertices >= rr_depth; -1 = off; 1 = reference (src/nanogi.cpp:583-591) */
    double rr_prob;          /* continuation probability; reference 0.5 */
    uint32_t seed;           /* counter-RNG key */
    uint32_t stream_id;      /* second key word (scene id in batched renders) */
    int precision;           /* NGI_PRECISION_* */
    int engine;              /* NGI_ENGINE_* */
    int paths_in_flight;     /* concurrent paths: wavefront queue capacity, or lanes of the
                              * persistent kernels (0 = auto: full device fill; wavefront
                              * queues sized to the render, 2^16..2^25 paths, growing) */
    int block_size;          /* threads per block (0 = auto) */
    int collect_counters;    /* count node visits / triangle tests (slower) */
    int time_kernels;        /* record per-kernel HIP-event times */
} ngi_render_params;

/* Reference defaults: pt, 1280x720, 10,000,000 samples uniform over the
 * image, -m -1, RR (1, 0.5) (src/nanogi.cpp:2000-2019, 583-591); FP32
 * shading with an FP64 film on the persistent engine. */
void ngi_render_params_default(ngi_render_params* p);

typedef struct ngi_stats {
    int64_t samples;           /* paths started ("# of samples", src/nanogi.cpp:421) */
    int64_t extension_rays;
    int64_t shadow_rays;
    double ms_total;           /* wall time of the render region (src/nanogi.cpp:201-216) */
    double ms_trace;           /* sum of extension+shadow traversal kernel times (time_kernels) */
    int64_t trace_launches;
    double ms_extend;          /* extension traversal kernel time */
    int64_t extend_launches;
    int64_t extend_rays_timed; /* rays processed by the timed extension launches */
    double ms_shadow;
    int64_t shadow_launches;
    double ms_shade;
    double node_visits;        /* totals when collect_counters */
    double tri_tests;
    double shadow_node_visits;
    double shadow_tri_tests;
} ngi_stats;

/* One-shot render into caller-allocated film_out (W*H*3 doubles, RGB,
 * row-major, row 0 = bottom, already scaled by W*H/N as in
 * src/nanogi.cpp:429-437).  stats may be NULL. */
int ngi_render(ngi_scene* scene, const ngi_render_params* p, double* film_out, ngi_stats* stats);

/* Render without allocating device memory.  No counterpart in the
 * reference: for callers that cannot call hipMalloc (code under stream or
 * graph capture, tight serving loops, harnesses that forbid allocation).
 * Renders *p like ngi_render into the caller's device buffer d_film
 * (W*H*3 doubles, or floats with NGI_PRECISION_FP32_ALL), already scaled by
 * W*H/N, enqueued on the scene's stream.  Counters, the path counter and
 * the shadow-ray queues live in one static device array of 30 MiB per GPU,
 * held until the process ends; the lanes in flight are capped at 327,680
 * (5,120 wavefront queues), and a larger paths_in_flight is rejected with
 * NGI_ERR_INVALID_ARG.  pt, ptdirect, lt and ltdirect on the persistent
 * engine or the megakernel; the wavefront engine and bdpt are rejected with
 * NGI_ERR_UNSUPPORTED.  The scene must already be resident with its current
 * parameters (after ngi_render / ngi_trace / a session render since the
 * last build or set_params), else NGI_ERR_INVALID_ARG.  stats NULL: no
 * synchronisation; non-NULL: waits for the render and fills the counts.
 * Every call on one GPU uses the same array: calls that overlap on one GPU
 * (other scenes' streams, other threads) race, so serialise them. */
int ngi_render_device_noalloc(ngi_scene* scene, const ngi_render_params* p, void* d_film, ngi_stats* stats);

/* Incremental rendering (the reference's time-budget chunks,
 * src/nanogi.cpp:243-418). */
typedef struct ngi_session ngi_session;
int ngi_session_create(ngi_scene* scene, const ngi_render_params* p, ngi_session** out);
/* Enqueue `spp` more samples per pixel (PER_PIXEL) or `spp` more full
 * passes of num_samples paths (REFERENCE_UNIFORM); asynchronous. */
int ngi_session_render(ngi_session* s, int spp);
/* Same, rendering into a fresh batch film that is then folded into the
 * per-pixel first and second moments used for the variance estimate. */
int ngi_session_render_batch(ngi_session* s, int spp);
int ngi_session_sync(ngi_session* s);
/* Normalised film (W*H*3 doubles); var_out (optional) = per-pixel variance
 * of the batch means (sample variance over batches / number of batches). */
int ngi_session_get_film(ngi_session* s, double* film_out, double* var_out);
int ngi_session_get_stats(ngi_session* s, ngi_stats* out);
int ngi_session_reset(ngi_session* s);
void ngi_session_destroy(ngi_session* s);

/* ------------------------------------------------------------------ */
/* Launch tuning for one scene (no counterpart in the reference, whose */
/* TBB loop has no engine to choose; it is the single-scene side of    */
/* the launch autotuner).  Renders tune_paths paths (0 = 2^24) of the   */
/* workload *p describes with each candidate -- persistent lanes at    */
/* the full, 3/4 and 1/2 device fill; for pt/ptdirect also wavefront   */
/* queues of 2^22 to 2^25 paths (FP32 shading) and the megakernel --   */
/* as one warm-up and two timed renders (wall time, the                */
/* better one counts), then writes the fastest into p->engine and      */
/* p->paths_in_flight.  Every candidate draws the same random numbers, */
/* so the choice changes speed, not the estimate.  trials (optional,   */
/* max_trials entries) receives the timings in candidate order and     */
/* *num_trials (optional) the number of candidates.                    */
/* ------------------------------------------------------------------ */
typedef struct ngi_tune_trial {
    int engine;              /* NGI_ENGINE_* */
    int paths_in_flight;     /* 0 = full device fill */
    double ms;               /* wall time of one render of the sample */
    double samples_per_s;
} ngi_tune_trial;
int ngi_autotune(ngi_scene* scene, ngi_render_params* p, int64_t tune_paths,
                 ngi_tune_trial* trials, int max_trials, int* num_trials);

/* ------------------------------------------------------------------ */
/* Film output: SaveImage (include/nanogi/basic.hpp:506-672).  Writes  */
/* .hdr (Radiance RGBE), .exr (float B,G,R), .png (gamma 1/2.2) or     */
/* .pfm by extension, upright; film as returned by ngi_render.        */
/* Parent directories are created; other extensions fail.             */
/* ------------------------------------------------------------------ */
int ngi_image_save(const char* path, const double* film, int width, int height);

/* ------------------------------------------------------------------ */
/* Adjoint (no reference counterpart): gradients of sum(dL_dimage *    */
/* image) with respect to per-primitive reflectance (D.R) and emitter  */
/* radiance (L.Area.Le), by path replay from the counter-RNG seeds.   */
/* ------------------------------------------------------------------ */
int ngi_render_adjoint(ngi_scene* scene, const ngi_render_params* p, const double* dL_dimage,
                       double* grad_reflectance, double* grad_Le, ngi_stats* stats);

/* ------------------------------------------------------------------ */
/* Batched rendering of many independent scenes in one persistent      */
/* launch (no reference counterpart; the reference runs one scene per  */
/* process).  films[i] receives scene i's W_i*H_i*3 normalised film.   */
/* Every scene renders pt, or every scene ptdirect, in FP32_FP64ACC    */
/* with BVHs of one width; other engines and precisions are rejected.  */
/* Scene i's paths are the ones ngi_render draws for params[i] (same   */
/* seed and stream id give the same image).  scene_ms (optional,       */
/* num_scenes doubles) receives each scene's completion time in ms     */
/* after the launch started, from the device wall clock.               */
/* ------------------------------------------------------------------ */
int ngi_render_batch(ngi_scene* const* scenes, const ngi_render_params* params, int num_scenes,
                     double* const* films, double* scene_ms, ngi_stats* stats);

/* Reusable batch (no reference counterpart): device films and per-scene   */
/* arguments stay allocated, and each ngi_batch_render replays one         */
/* captured hipGraph (film/counter fill kernel + persistent launch).       */
/* Scenes must outlive the batch; a scene rebuilt with a deeper BVH or     */
/* another BVH width than at creation makes ngi_batch_render fail.         */
typedef struct ngi_batch_opts {
    int block_size;      /* 64, 128 or 256 threads; 0 = autotune */
    int blocks_per_cu;   /* resident blocks per CU (clamped to occupancy); 0 = autotune */
    int use_graph;       /* 1 = replay a captured hipGraph, 0 = plain launches */
    int chunk_paths;     /* paths a wavefront takes from one scene at a time (<= 2^24); 0 = about
                          * 16 chunks per resident wave, a power of two in [256, 8192] */
    int64_t tune_paths;  /* paths per autotuner trial, drawn proportionally from every scene */
} ngi_batch_opts;
typedef struct ngi_batch ngi_batch;
void ngi_batch_opts_default(ngi_batch_opts* o);   /* autotune both, graph on, 2^24 paths per trial */
int ngi_batch_create(ngi_scene* const* scenes, const ngi_render_params* params, int num_scenes,
                     const ngi_batch_opts* opts, ngi_batch** out);
/* Enqueue one frame (asynchronous).  Scene i uses seed params[i].seed +  */
/* frame, so frame 0 reproduces ngi_render for params[i].                  */
int ngi_batch_render(ngi_batch* b, uint32_t frame);
int ngi_batch_sync(ngi_batch* b);
/* Normalised film of scene `scene` from the last frame (W*H*3 doubles). */
int ngi_batch_get_film(ngi_batch* b, int scene, double* film_out);
/* Per-scene completion times (ms after the launch started, device wall  */
/* clock) and totals of the last frame; either output may be NULL.       */
int ngi_batch_get_times(ngi_batch* b, double* scene_ms, ngi_stats* stats);
/* Launch configuration in use (after autotuning) and the tuning time.    */
int ngi_batch_get_opts(ngi_batch* b, ngi_batch_opts* chosen, double* tune_ms);
void ngi_batch_destroy(ngi_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* NGI_H */
