"""Python mirror of the renderer-selection and scene interface of nanogi.

This module is a thin ctypes binding over ``libngi.so`` (the C ABI declared
in ``include/ngi.h``).  It mirrors the reference's operator surface for the
two path tracers: ``Scene.load_yaml`` is ``Scene::Load``
(include/nanogi/rt.hpp:1519), ``render`` is ``Renderer::Render`` with the
``-r pt|ptdirect`` selection (src/nanogi.cpp:51-71, 182-221) and
``Scene.trace`` is ``Scene::Intersect`` / ``Scene::Visible``
(rt.hpp:2162-2261).  Every call runs the native library; when it is missing
the import fails loudly -- there is no Python or CPU fallback.
"""
import copy
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("NGI_LIB_PATH") or os.path.join(_HERE, "libngi.so")

c_int, c_uint32, c_int64, c_double, c_float, c_size_t = (
    ctypes.c_int, ctypes.c_uint32, ctypes.c_int64, ctypes.c_double, ctypes.c_float, ctypes.c_size_t)

# Status codes (ngi_status)
OK, ERR_IO, ERR_PARSE, ERR_VERSION, ERR_INVALID_TYPE, ERR_INVALID_SCENE, ERR_UNSUPPORTED, \
    ERR_INVALID_ARG, ERR_DEVICE, ERR_OOM, ERR_QUEUE_OVERFLOW = range(11)

PRIM_D, PRIM_G, PRIM_S, PRIM_L, PRIM_E = 1, 2, 4, 8, 16
LIGHT_AREA, LIGHT_POINT, LIGHT_DIRECTIONAL = 0, 1, 2
SENSOR_AREA, SENSOR_PINHOLE = 0, 1
SPEC_REFLECTION, SPEC_REFRACTION, SPEC_FRESNEL = 0, 1, 2
GEN_CORNELL, GEN_NAVE, GEN_SPHERE_FIELD, GEN_BATCH_ROOM = 1, 2, 3, 4
BVH_2, BVH_8 = 2, 8
RENDERER_PT, RENDERER_PTDIRECT, RENDERER_LT, RENDERER_LTDIRECT, RENDERER_BDPT, RENDERER_PTMNEE = range(6)
SAMPLING_PER_PIXEL, SAMPLING_REFERENCE_UNIFORM = 0, 1
PRECISION_FP32_FP64ACC, PRECISION_FP32_ALL, PRECISION_FP64_ALL = 0, 1, 2
ENGINE_MEGAKERNEL, ENGINE_WAVEFRONT, ENGINE_PERSISTENT = 0, 1, 2
MISS = 0xFFFFFFFF

_d3 = c_double * 3


class Primitive(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("light_type", ctypes.c_int32), ("sensor_type", ctypes.c_int32),
                ("spec_type", ctypes.c_int32), ("tri_begin", c_uint32), ("tri_count", c_uint32),
                ("Le", _d3), ("light_position", _d3), ("light_direction", _d3), ("D_R", _d3),
                ("G_R", _d3), ("G_eta", _d3), ("G_k", _d3), ("G_roughness", c_double),
                ("S_R", _d3), ("S_eta1", c_double), ("S_eta2", c_double), ("E_We", _d3),
                ("E_eye", _d3), ("E_center", _d3), ("E_up", _d3), ("E_fov_deg", c_double)]


PRIM_DTYPE = np.dtype(Primitive)


class SceneDesc(ctypes.Structure):
    _fields_ = [("num_vertices", c_uint32), ("positions", ctypes.POINTER(c_double)),
                ("normals", ctypes.POINTER(c_double)), ("texcoords", ctypes.POINTER(c_double)),
                ("num_triangles", c_uint32), ("faces", ctypes.POINTER(c_uint32)),
                ("tri_prim", ctypes.POINTER(c_uint32)), ("num_primitives", c_uint32),
                ("primitives", ctypes.POINTER(Primitive)), ("aspect", c_double)]


class BvhOpts(ctypes.Structure):
    _fields_ = [("width", c_int), ("max_leaf_size", c_int), ("sah_bins", c_int), ("num_threads", c_int),
                ("traversal_cost", c_double)]


class BvhInfo(ctypes.Structure):
    _fields_ = [("width", c_int), ("num_nodes", c_uint32), ("num_leaves", c_uint32), ("num_triangles", c_uint32),
                ("max_depth", c_uint32), ("node_bytes", c_uint32), ("sah_cost", c_double), ("build_ms", c_double)]


class Ray(ctypes.Structure):
    _fields_ = [("o", c_float * 3), ("tmin", c_float), ("d", c_float * 3), ("tmax", c_float)]


class Hit(ctypes.Structure):
    _fields_ = [("t", c_float), ("u", c_float), ("v", c_float), ("tri", c_uint32)]


RAY_DTYPE = np.dtype(Ray)
HIT_DTYPE = np.dtype(Hit)


class RenderParams(ctypes.Structure):
    _fields_ = [("renderer", c_int), ("width", c_int), ("height", c_int), ("sampling", c_int),
                ("num_samples", c_int64), ("spp", c_int), ("max_num_vertices", c_int), ("rr_depth", c_int),
                ("rr_prob", c_double), ("seed", c_uint32), ("stream_id", c_uint32), ("precision", c_int),
                ("engine", c_int), ("paths_in_flight", c_int), ("block_size", c_int),
                ("collect_counters", c_int), ("time_kernels", c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("samples", c_int64), ("extension_rays", c_int64), ("shadow_rays", c_int64),
                ("ms_total", c_double), ("ms_trace", c_double), ("trace_launches", c_int64),
                ("ms_extend", c_double), ("extend_launches", c_int64), ("extend_rays_timed", c_int64),
                ("ms_shadow", c_double), ("shadow_launches", c_int64), ("ms_shade", c_double),
                ("node_visits", c_double), ("tri_tests", c_double), ("shadow_node_visits", c_double),
                ("shadow_tri_tests", c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class TuneTrial(ctypes.Structure):
    _fields_ = [("engine", c_int), ("paths_in_flight", c_int), ("ms", c_double), ("samples_per_s", c_double)]


class BatchOpts(ctypes.Structure):
    _fields_ = [("block_size", c_int), ("blocks_per_cu", c_int), ("use_graph", c_int), ("chunk_paths", c_int),
                ("tune_paths", c_int64)]


# Every entry point declared in include/ngi.h (checked by the CPU tests).
EXPORTS = [
    "ngi_last_error", "ngi_device_count", "ngi_scene_load_yaml", "ngi_scene_create", "ngi_scene_generate",
    "ngi_scene_save_yaml", "ngi_scene_destroy", "ngi_scene_get_desc", "ngi_scene_set_params", "ngi_scene_build",
    "ngi_scene_bvh_info", "ngi_scene_bvh_export", "ngi_trace", "ngi_trace_device", "ngi_renderer_from_string",
    "ngi_render_params_default", "ngi_render", "ngi_render_device_noalloc", "ngi_session_create", "ngi_session_render",
    "ngi_session_render_batch", "ngi_session_sync", "ngi_session_get_film", "ngi_session_get_stats",
    "ngi_session_reset", "ngi_session_destroy", "ngi_render_adjoint", "ngi_render_batch", "ngi_image_save",
    "ngi_batch_opts_default", "ngi_batch_create", "ngi_batch_render", "ngi_batch_sync", "ngi_batch_get_film",
    "ngi_batch_get_times", "ngi_batch_get_opts", "ngi_batch_destroy", "ngi_autotune",
]


class NgiError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"ngi status {status}: {message}")
        self.status = status
        self.message = message


_lib = None


def lib():
    """Load libngi.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m nanogi_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    P = ctypes.POINTER
    sigs = {
        "ngi_last_error": ([], ctypes.c_char_p),
        "ngi_device_count": ([], c_int),
        "ngi_scene_load_yaml": ([ctypes.c_char_p, c_double, P(vp)], c_int),
        "ngi_scene_create": ([P(SceneDesc), P(vp)], c_int),
        "ngi_scene_generate": ([c_int, ctypes.c_uint64, c_double, c_int, P(vp)], c_int),
        "ngi_scene_save_yaml": ([vp, ctypes.c_char_p], c_int),
        "ngi_scene_destroy": ([vp], None),
        "ngi_scene_get_desc": ([vp, P(SceneDesc)], c_int),
        "ngi_scene_set_params": ([vp, vp, vp], c_int),
        "ngi_scene_build": ([vp, P(BvhOpts)], c_int),
        "ngi_scene_bvh_info": ([vp, P(BvhInfo)], c_int),
        "ngi_scene_bvh_export": ([vp, vp, vp], c_int),
        "ngi_trace": ([vp, vp, c_size_t, vp, c_int], c_int),
        "ngi_trace_device": ([vp, vp, c_size_t, vp, c_int], c_int),
        "ngi_renderer_from_string": ([ctypes.c_char_p], c_int),
        "ngi_render_params_default": ([P(RenderParams)], None),
        "ngi_render": ([vp, P(RenderParams), vp, P(Stats)], c_int),
        "ngi_render_device_noalloc": ([vp, P(RenderParams), vp, P(Stats)], c_int),
        "ngi_session_create": ([vp, P(RenderParams), P(vp)], c_int),
        "ngi_session_render": ([vp, c_int], c_int),
        "ngi_session_render_batch": ([vp, c_int], c_int),
        "ngi_session_sync": ([vp], c_int),
        "ngi_session_get_film": ([vp, vp, vp], c_int),
        "ngi_session_get_stats": ([vp, P(Stats)], c_int),
        "ngi_session_reset": ([vp], c_int),
        "ngi_session_destroy": ([vp], None),
        "ngi_render_adjoint": ([vp, P(RenderParams), vp, vp, vp, P(Stats)], c_int),
        "ngi_render_batch": ([vp, P(RenderParams), c_int, vp, vp, P(Stats)], c_int),
        "ngi_image_save": ([ctypes.c_char_p, vp, c_int, c_int], c_int),
        "ngi_batch_opts_default": ([P(BatchOpts)], None),
        "ngi_batch_create": ([vp, P(RenderParams), c_int, P(BatchOpts), P(vp)], c_int),
        "ngi_batch_render": ([vp, c_uint32], c_int),
        "ngi_batch_sync": ([vp], c_int),
        "ngi_batch_get_film": ([vp, c_int, vp], c_int),
        "ngi_batch_get_times": ([vp, vp, P(Stats)], c_int),
        "ngi_batch_get_opts": ([vp, P(BatchOpts), P(c_double)], c_int),
        "ngi_batch_destroy": ([vp], None),
        "ngi_autotune": ([vp, P(RenderParams), ctypes.c_int64, P(TuneTrial), c_int, P(c_int)], c_int),
    }
    for name, (args, res) in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _check(status):
    if status != OK:
        raise NgiError(status, lib().ngi_last_error().decode("utf-8", "replace"))


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def device_count():
    return lib().ngi_device_count()


def renderer_from_string(name):
    return lib().ngi_renderer_from_string(name.encode())


def default_params(**kw):
    p = RenderParams()
    lib().ngi_render_params_default(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(f"unknown render parameter {k!r}")
        if k == "renderer" and isinstance(v, str):
            v = renderer_from_string(v)
            if v < 0:
                raise NgiError(ERR_UNSUPPORTED, f"unknown renderer {kw['renderer']!r}")
        setattr(p, k, v)
    return p


def make_primitive(**fields):
    """A zero-initialised primitive record with the given fields."""
    p = np.zeros(1, dtype=PRIM_DTYPE)[0]
    for k, v in fields.items():
        p[k] = v
    return p


class Scene:
    """Owning handle of an ngi_scene (host arrays, BVH, device copy)."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle)

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().ngi_scene_destroy(self._h)
            self._h = ctypes.c_void_p()

    @property
    def handle(self):
        return self._h

    # --- construction ---------------------------------------------------
    @classmethod
    def load_yaml(cls, path, aspect=1280 / 720):
        h = ctypes.c_void_p()
        _check(lib().ngi_scene_load_yaml(str(path).encode(), float(aspect), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def generate(cls, kind, seed=1, aspect=1.0, detail=0):
        h = ctypes.c_void_p()
        _check(lib().ngi_scene_generate(int(kind), int(seed), float(aspect), int(detail), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def create(cls, positions, faces, tri_prim, primitives, normals=None, texcoords=None, aspect=1.0):
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        fc = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
        tp = np.ascontiguousarray(tri_prim, dtype=np.uint32).reshape(-1)
        prims = np.ascontiguousarray(primitives, dtype=PRIM_DTYPE).reshape(-1)
        nrm = None if normals is None else np.ascontiguousarray(normals, dtype=np.float64).reshape(-1, 3)
        tex = None if texcoords is None else np.ascontiguousarray(texcoords, dtype=np.float64).reshape(-1, 2)
        d = SceneDesc()
        d.num_vertices = pos.shape[0]
        d.positions = pos.ctypes.data_as(ctypes.POINTER(c_double))
        d.normals = nrm.ctypes.data_as(ctypes.POINTER(c_double)) if nrm is not None else None
        d.texcoords = tex.ctypes.data_as(ctypes.POINTER(c_double)) if tex is not None else None
        d.num_triangles = fc.shape[0]
        d.faces = fc.ctypes.data_as(ctypes.POINTER(c_uint32))
        d.tri_prim = tp.ctypes.data_as(ctypes.POINTER(c_uint32))
        d.num_primitives = prims.shape[0]
        d.primitives = prims.ctypes.data_as(ctypes.POINTER(Primitive))
        d.aspect = float(aspect)
        h = ctypes.c_void_p()
        _check(lib().ngi_scene_create(ctypes.byref(d), ctypes.byref(h)))
        return cls(h.value)

    # --- inspection -------------------------------------------------------
    def desc(self):
        """Copies of the scene arrays (the FP64 data the oracle consumes)."""
        d = SceneDesc()
        _check(lib().ngi_scene_get_desc(self._h, ctypes.byref(d)))
        nv, nt, npr = d.num_vertices, d.num_triangles, d.num_primitives

        def arr(ptr, n, dtype):
            if n == 0 or not ptr:
                return np.zeros(n, dtype=dtype)
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)

        out = {
            "positions": arr(d.positions, 3 * nv, np.float64).reshape(-1, 3),
            "normals": arr(d.normals, 3 * nv, np.float64).reshape(-1, 3) if d.normals else None,
            "texcoords": arr(d.texcoords, 2 * nv, np.float64).reshape(-1, 2) if d.texcoords else None,
            "faces": arr(d.faces, 3 * nt, np.uint32).reshape(-1, 3),
            "tri_prim": arr(d.tri_prim, nt, np.uint32),
            "aspect": d.aspect,
        }
        prims = np.zeros(npr, dtype=PRIM_DTYPE)
        if npr:
            ctypes.memmove(prims.ctypes.data, d.primitives, npr * PRIM_DTYPE.itemsize)
        out["primitives"] = prims
        return out

    def raw_desc(self):
        """The SceneDesc pointing at the scene's own arrays (valid while alive)."""
        d = SceneDesc()
        _check(lib().ngi_scene_get_desc(self._h, ctypes.byref(d)))
        return d

    @property
    def num_triangles(self):
        return self.raw_desc().num_triangles

    def set_params(self, reflectance=None, Le=None):
        r = None if reflectance is None else np.ascontiguousarray(reflectance, dtype=np.float64)
        e = None if Le is None else np.ascontiguousarray(Le, dtype=np.float64)
        _check(lib().ngi_scene_set_params(self._h, _ptr(r), _ptr(e)))

    def save_yaml(self, directory):
        _check(lib().ngi_scene_save_yaml(self._h, str(directory).encode()))

    # --- acceleration structure --------------------------------------------
    def build(self, width=BVH_8, max_leaf_size=0, sah_bins=0, num_threads=0, traversal_cost=0.0):
        o = BvhOpts(width, max_leaf_size, sah_bins, num_threads, traversal_cost)
        _check(lib().ngi_scene_build(self._h, ctypes.byref(o)))
        return self.bvh_info()

    def bvh_info(self):
        i = BvhInfo()
        _check(lib().ngi_scene_bvh_info(self._h, ctypes.byref(i)))
        return {name: getattr(i, name) for name, _ in BvhInfo._fields_}

    def bvh_export(self):
        info = self.bvh_info()
        nodes = np.zeros(info["num_nodes"] * info["node_bytes"], dtype=np.uint8)
        perm = np.zeros(info["num_triangles"], dtype=np.uint32)
        _check(lib().ngi_scene_bvh_export(self._h, _ptr(nodes), _ptr(perm)))
        return nodes, perm

    # --- ray queries -------------------------------------------------------
    def trace(self, rays, any_hit=False):
        r = np.ascontiguousarray(rays, dtype=RAY_DTYPE).reshape(-1)
        hits = np.zeros(r.shape[0], dtype=HIT_DTYPE)
        _check(lib().ngi_trace(self._h, _ptr(r), r.shape[0], _ptr(hits), 1 if any_hit else 0))
        return hits


def make_rays(o, d, tmin=1e-4, tmax=np.finfo(np.float32).max):
    o = np.asarray(o, dtype=np.float32).reshape(-1, 3)
    d = np.asarray(d, dtype=np.float32).reshape(-1, 3)
    r = np.zeros(o.shape[0], dtype=RAY_DTYPE)
    r["o"] = o
    r["d"] = d
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


class Session:
    """Incremental render (ngi_session_*)."""

    def __init__(self, scene, params):
        self.scene = scene
        self.params = params
        h = ctypes.c_void_p()
        _check(lib().ngi_session_create(scene.handle, ctypes.byref(params), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().ngi_session_destroy(self._h)
            self._h = ctypes.c_void_p()

    def render(self, spp=1):
        _check(lib().ngi_session_render(self._h, int(spp)))

    def render_batch(self, spp=1):
        _check(lib().ngi_session_render_batch(self._h, int(spp)))

    def sync(self):
        _check(lib().ngi_session_sync(self._h))

    def film(self, variance=False):
        n = self.params.width * self.params.height * 3
        f = np.zeros(n, dtype=np.float64)
        v = np.zeros(n, dtype=np.float64) if variance else None
        _check(lib().ngi_session_get_film(self._h, _ptr(f), _ptr(v)))
        shape = (self.params.height, self.params.width, 3)
        return (f.reshape(shape), v.reshape(shape)) if variance else f.reshape(shape)

    def stats(self):
        s = Stats()
        _check(lib().ngi_session_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def reset(self):
        _check(lib().ngi_session_reset(self._h))


def render(scene, params=None, **kw):
    """Render and return (film[H, W, 3] float64, stats dict); row 0 = bottom."""
    p = params if params is not None else default_params(**kw)
    # The library validates the size; a bad one must not fail here first.
    film = np.zeros(max(p.width * p.height * 3, 1), dtype=np.float64)
    st = Stats()
    _check(lib().ngi_render(scene.handle, ctypes.byref(p), _ptr(film), ctypes.byref(st)))
    return film[:p.width * p.height * 3].reshape(p.height, p.width, 3), st.as_dict()


def autotune(scene, params, tune_paths=0):
    """Time every engine / concurrency candidate on a sample of `tune_paths`
    paths (0 = 2^24) of the workload `params` describes (ngi_autotune).

    Returns (tuned, trials): a copy of `params` with the fastest engine and
    paths_in_flight, and one dict per candidate (engine, paths_in_flight, ms,
    samples_per_s) in candidate order."""
    q = RenderParams()
    ctypes.pointer(q)[0] = params
    trials = (TuneTrial * 16)()
    n = c_int(0)
    _check(lib().ngi_autotune(scene.handle, ctypes.byref(q), int(tune_paths), trials, 16, ctypes.byref(n)))
    return q, [{name: getattr(trials[i], name) for name, _ in TuneTrial._fields_} for i in range(min(n.value, 16))]


def render_batch(scenes, params):
    """Render many scenes in one persistent launch (ngi_render_batch).

    `params` is one RenderParams per scene.  Returns (films, scene_ms, stats):
    films[i] is scene i's film[H_i, W_i, 3] (as render() returns it) and
    scene_ms[i] its completion time in ms after the launch started."""
    n = len(scenes)
    if n != len(params):
        raise ValueError("one RenderParams per scene")
    films = [np.zeros(max(p.width * p.height * 3, 1), dtype=np.float64) for p in params]
    ptrs = (ctypes.c_void_p * max(n, 1))(*[f.ctypes.data for f in films])
    handles = (ctypes.c_void_p * max(n, 1))(*[s.handle for s in scenes])
    parr = (RenderParams * max(n, 1))(*params)
    ms = np.zeros(max(n, 1))
    st = Stats()
    _check(lib().ngi_render_batch(handles, parr, n, ptrs, _ptr(ms), ctypes.byref(st)))
    out = [f[:p.width * p.height * 3].reshape(p.height, p.width, 3) for f, p in zip(films, params)]
    return out, ms[:n], st.as_dict()


def batch_opts(**kw):
    """BatchOpts with the library defaults (autotuned launch, graph replay), then `kw`."""
    o = BatchOpts()
    lib().ngi_batch_opts_default(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Batch:
    """Reusable many-scene render (ngi_batch_*): one graph replay per frame.

    Holds references to its scenes so they outlive the device batch."""

    def __init__(self, scenes, params, opts=None):
        if len(scenes) != len(params) or not scenes:
            raise ValueError("one RenderParams per scene, at least one scene")
        self.scenes = list(scenes)
        self.params = [copy.copy(p) for p in params]
        n = len(scenes)
        handles = (ctypes.c_void_p * n)(*[s.handle for s in scenes])
        parr = (RenderParams * n)(*self.params)
        h = ctypes.c_void_p()
        _check(lib().ngi_batch_create(handles, parr, n, ctypes.byref(opts) if opts is not None else None, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        self.close()

    def close(self):
        h = getattr(self, "_h", None)
        if h and h.value and _lib is not None:
            _lib.ngi_batch_destroy(h)
            self._h = ctypes.c_void_p()

    def render(self, frame=0):
        _check(lib().ngi_batch_render(self._h, int(frame)))

    def sync(self):
        _check(lib().ngi_batch_sync(self._h))

    def film(self, i):
        p = self.params[i]
        f = np.zeros(p.width * p.height * 3)
        _check(lib().ngi_batch_get_film(self._h, int(i), _ptr(f)))
        return f.reshape(p.height, p.width, 3)

    def times(self):
        """(per-scene completion ms of the last frame, stats dict)."""
        ms = np.zeros(len(self.scenes))
        st = Stats()
        _check(lib().ngi_batch_get_times(self._h, _ptr(ms), ctypes.byref(st)))
        return ms, st.as_dict()

    def config(self):
        o = BatchOpts()
        t = c_double()
        _check(lib().ngi_batch_get_opts(self._h, ctypes.byref(o), ctypes.byref(t)))
        return {"block_size": o.block_size, "blocks_per_cu": o.blocks_per_cu, "use_graph": o.use_graph,
                "chunk_paths": o.chunk_paths, "tune_paths": o.tune_paths, "tune_ms": t.value}


def save_image(path, film):
    """SaveImage (include/nanogi/basic.hpp:506): .hdr/.exr/.png/.pfm by extension."""
    f = np.ascontiguousarray(film, dtype=np.float64)
    h, w = f.shape[0], f.shape[1]
    _check(lib().ngi_image_save(str(path).encode(), _ptr(f), int(w), int(h)))


def render_adjoint(scene, params, dL_dimage):
    """Path-replay gradients of sum(dL_dimage * image) (ngi_render_adjoint).

    Returns (grad_reflectance, grad_Le, stats) with one row of 3 per primitive."""
    g = np.ascontiguousarray(dL_dimage, dtype=np.float64).reshape(params.height, params.width, 3)
    n = scene.raw_desc().num_primitives
    gR = np.zeros((n, 3))
    gL = np.zeros((n, 3))
    st = Stats()
    _check(lib().ngi_render_adjoint(scene.handle, ctypes.byref(params), _ptr(g), _ptr(gR), _ptr(gL), ctypes.byref(st)))
    return gR, gL, st.as_dict()
