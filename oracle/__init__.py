"""ctypes binding of the CPU FP64 oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this package, and only as the checker.  See ngi_oracle.cpp for the
reference functions it restates.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `python oracle/build.py`")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.oracle_trace.argtypes = [vp, vp, ctypes.c_size_t, vp, ctypes.c_int, ctypes.c_int]
        L.oracle_trace.restype = ctypes.c_int
        L.oracle_render_paths.argtypes = [vp, vp, ctypes.c_uint64, ctypes.c_uint64, vp, ctypes.c_int, vp, ctypes.c_int]
        L.oracle_render_paths.restype = ctypes.c_int
        L.oracle_render.argtypes = [vp, vp, vp, ctypes.c_int]
        L.oracle_render.restype = ctypes.c_int
        L.oracle_philox.argtypes = [vp, vp, vp]
        L.oracle_philox.restype = None
        _lib = L
    return _lib


def _check(st):
    if st != 0:
        raise RuntimeError(f"oracle status {st}")


def trace(scene, rays, any_hit=False, use_bvh=True):
    """FP64 hits for rays (RAY_DTYPE) against a nanogi_amd.Scene's arrays."""
    import nanogi_amd as ngi
    d = scene.raw_desc()
    r = np.ascontiguousarray(rays, dtype=ngi.RAY_DTYPE)
    hits = np.zeros(r.shape[0], dtype=ngi.HIT_DTYPE)
    _check(lib().oracle_trace(ctypes.addressof(d), r.ctypes.data, r.shape[0], hits.ctypes.data, int(any_hit), int(use_bvh)))
    return hits


def render(scene, params, num_threads=0):
    """Normalised FP64 film[H, W, 3] of the full render described by params."""
    d = scene.raw_desc()
    film = np.zeros(params.width * params.height * 3, dtype=np.float64)
    _check(lib().oracle_render(ctypes.addressof(d), ctypes.addressof(params), film.ctypes.data, int(num_threads)))
    return film.reshape(params.height, params.width, 3)


def render_paths(scene, params, path_begin, path_count, num_threads=0):
    """Unnormalised film sums of a range of path ids and the ray counts."""
    d = scene.raw_desc()
    film = np.zeros(params.width * params.height * 3, dtype=np.float64)
    counts = np.zeros(2, dtype=np.int64)
    _check(lib().oracle_render_paths(ctypes.addressof(d), ctypes.addressof(params), int(path_begin), int(path_count),
                                     film.ctypes.data, int(num_threads), counts.ctypes.data, 1))
    return film.reshape(params.height, params.width, 3), counts


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().oracle_philox(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def bsdf(prim, type_bits, wi, u=(0.0, 0.0, 0.0), wo=None, force_degenerated=True, le=False):
    """(wo, pdf, fs) of the oracle's SampleDirection/EvaluateDirection[PDF]
    at a surface with normal +z; wo given -> evaluate only; le selects the
    light-to-eye transport direction of EvaluateDirection."""
    import nanogi_amd as ngi
    L = lib()
    if not hasattr(L.oracle_bsdf, "_set"):
        L.oracle_bsdf.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.oracle_bsdf.restype = None
        L.oracle_bsdf._set = True
    p = np.ascontiguousarray(np.asarray(prim, dtype=ngi.PRIM_DTYPE).reshape(1))
    wi_a = np.asarray(wi, dtype=np.float64)
    u_a = np.asarray(u, dtype=np.float64)
    wo_a = np.asarray(wo if wo is not None else (0, 0, 0), dtype=np.float64)
    out = np.zeros(7, dtype=np.float64)
    L.oracle_bsdf(p.ctypes.data, int(type_bits), wi_a.ctypes.data, u_a.ctypes.data, wo_a.ctypes.data,
                  int(wo is not None), int(force_degenerated), int(le), out.ctypes.data)
    return out[:3], out[3], out[4:]


class LoadedScene:
    """Oracle scene with its BVH built once (keeps the source scene alive)."""

    def __init__(self, scene):
        L = lib()
        L.oracle_scene_new.argtypes = [ctypes.c_void_p]
        L.oracle_scene_new.restype = ctypes.c_void_p
        L.oracle_scene_free.argtypes = [ctypes.c_void_p]
        L.oracle_scene_render_paths.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                                ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.oracle_scene_render_paths.restype = ctypes.c_int
        self.scene = scene
        self.desc = scene.raw_desc()
        self.h = L.oracle_scene_new(ctypes.addressof(self.desc))
        if not self.h:
            raise RuntimeError("oracle could not load the scene")

    def render_paths(self, params, path_begin, path_count, num_threads=0):
        film = np.zeros(params.width * params.height * 3, dtype=np.float64)
        counts = np.zeros(2, dtype=np.int64)
        _check(lib().oracle_scene_render_paths(self.h, ctypes.addressof(params), int(path_begin), int(path_count),
                                               film.ctypes.data, int(num_threads), counts.ctypes.data))
        return film.reshape(params.height, params.width, 3), counts

    def camera(self, u0, u1):
        """(wo, (rx, ry), We, pdf) of the pinhole for raster position (u0, u1)."""
        L = lib()
        L.oracle_scene_camera.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
        L.oracle_scene_camera.restype = ctypes.c_int
        out = np.zeros(7, dtype=np.float64)
        _check(L.oracle_scene_camera(self.h, float(u0), float(u1), out.ctypes.data))
        return out[0:3], out[3:5], out[5], out[6]

    def light(self, k, ux, uy):
        """(p, gn, position pdf, selection pdf) of light k sampled at (ux, uy)."""
        L = lib()
        L.oracle_scene_light.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
        L.oracle_scene_light.restype = ctypes.c_int
        out = np.zeros(8, dtype=np.float64)
        _check(L.oracle_scene_light(self.h, int(k), float(ux), float(uy), out.ctypes.data))
        return out[0:3], out[3:6], out[6], out[7]

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_scene_free(self.h)
            self.h = None


def render_adjoint(scene, params, dLdI, num_threads=0):
    """FP64 reference gradients (reflectance, Le), each (num_primitives, 3)."""
    L = lib()
    L.oracle_render_adjoint.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_int]
    L.oracle_render_adjoint.restype = ctypes.c_int
    d = scene.raw_desc()
    g = np.ascontiguousarray(dLdI, dtype=np.float64)
    n = d.num_primitives
    gR = np.zeros((n, 3))
    gL = np.zeros((n, 3))
    _check(L.oracle_render_adjoint(ctypes.addressof(d), ctypes.addressof(params), g.ctypes.data, gR.ctypes.data,
                                   gL.ctypes.data, int(num_threads)))
    return gR, gL
