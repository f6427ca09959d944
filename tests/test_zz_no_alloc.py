"""ngi_render_device_noalloc (renders without device allocations, through a
file-scope scratch array) against ngi_render on the same parameters.

Same kernels, same path ids and random numbers: the films agree up to the
order of the FP64 film additions (relative 1e-9 of the image maximum; FP32
films 1e-5).  Two calls in a row, the second smaller, so counters, the path
counter and the shadow queues left by the first call must not leak into the
second.
"""
import ctypes
import os

import numpy as np
import pytest

import nanogi_amd as ngi

pytestmark = pytest.mark.gpu


def _hip_runtime():
    """The HIP runtime libngi.so is bound to.  A process can map two copies
    (the ROCm one and torch's; whichever was loaded first under the soname
    libamdhip64.so.7 serves libngi.so), and a buffer from the other copy's
    hipMalloc would not belong to libngi.so's runtime.  RTLD_NOLOAD returns
    the copy already loaded under that soname and never loads a new one."""
    ngi.lib()
    return ctypes.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD)


class _Device:
    """hipMalloc'd film buffers for the test (the caller's memory)."""

    def __init__(self):
        self.hip = _hip_runtime()
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
        return p

    def free(self, p):
        assert self.hip.hipFree(p) == 0

    def to_host(self, p, n, dtype):
        out = np.empty(n, dtype=dtype)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, 2) == 0   # device to host
        return out


def _noalloc(scene, p, dev, stats=True):
    L = ngi.lib()
    f = L.ngi_render_device_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.POINTER(ngi.RenderParams), ctypes.c_void_p, ctypes.POINTER(ngi.Stats)]
    f.restype = ctypes.c_int
    film64 = p.precision != ngi.PRECISION_FP32_ALL
    n = p.width * p.height * 3
    buf = dev.alloc(n * (8 if film64 else 4))
    try:
        st = ngi.Stats()
        status = f(scene.handle, ctypes.byref(p), buf, ctypes.byref(st) if stats else None)
        if status != ngi.OK:
            raise ngi.NgiError(status, L.ngi_last_error().decode())
        film = dev.to_host(buf, n, np.float64 if film64 else np.float32).astype(np.float64)
    finally:
        dev.free(buf)
    return film.reshape(p.height, p.width, 3), st.as_dict()


def _params(renderer, w, h, spp, engine=ngi.ENGINE_PERSISTENT, precision=ngi.PRECISION_FP32_FP64ACC, seed=3):
    return ngi.default_params(renderer=renderer, width=w, height=h, sampling=ngi.SAMPLING_PER_PIXEL, spp=spp,
                              max_num_vertices=9, rr_depth=3, seed=seed, engine=engine, precision=precision)


def _agree(a, b, rtol):
    scale = max(np.abs(b).max(), 1e-30)
    np.testing.assert_allclose(a, b, rtol=rtol, atol=rtol * scale)


CASES = [("pt", ngi.ENGINE_PERSISTENT, ngi.PRECISION_FP32_FP64ACC, 1e-9),
         ("ptdirect", ngi.ENGINE_PERSISTENT, ngi.PRECISION_FP32_FP64ACC, 1e-9),
         ("lt", ngi.ENGINE_PERSISTENT, ngi.PRECISION_FP32_FP64ACC, 1e-9),
         ("ltdirect", ngi.ENGINE_PERSISTENT, ngi.PRECISION_FP32_FP64ACC, 1e-9),
         ("ltdirect", ngi.ENGINE_PERSISTENT, ngi.PRECISION_FP64_ALL, 1e-9),
         ("ptdirect", ngi.ENGINE_PERSISTENT, ngi.PRECISION_FP64_ALL, 1e-9),
         ("ptdirect", ngi.ENGINE_PERSISTENT, ngi.PRECISION_FP32_ALL, 1e-5),
         ("pt", ngi.ENGINE_MEGAKERNEL, ngi.PRECISION_FP32_FP64ACC, 1e-9),
         ("ptdirect", ngi.ENGINE_MEGAKERNEL, ngi.PRECISION_FP32_FP64ACC, 1e-9)]


# Cornell box and the 13k-triangle sphere field: chunked path feed, shadow
# rays in rounds of 64; the 29k-triangle sphere field: exact refills,
# resumable traversals and shadow drains from the 128-record rings the
# scratch array holds.
@pytest.mark.parametrize("kind,detail", [(ngi.GEN_CORNELL, 0), (ngi.GEN_SPHERE_FIELD, 4), (ngi.GEN_SPHERE_FIELD, 5)])
@pytest.mark.parametrize("renderer,engine,precision,rtol", CASES)
def test_noalloc_matches_render_twice(kind, detail, renderer, engine, precision, rtol):
    dev = _Device()
    s = ngi.Scene.generate(kind, 1, 4 / 3, detail)
    big = _params(renderer, 64, 48, 8, engine, precision)
    small = _params(renderer, 32, 24, 3, engine, precision, seed=5)
    ref_big, st_big = ngi.render(s, big)        # also makes the scene resident
    ref_small, st_small = ngi.render(s, small)
    for p, ref, st_ref in ((big, ref_big, st_big), (small, ref_small, st_small)):
        print(f"noalloc case: {renderer} engine {engine} precision {precision} {p.width}x{p.height} spp {p.spp}", flush=True)
        film, st = _noalloc(s, p, dev)
        if renderer == "lt":
            # Light paths cannot hit a pinhole: lt adds nothing (rt.hpp:590-593).
            assert not ref.any() and not film.any()
        else:
            assert np.abs(ref).max() > 0
            _agree(film, ref, rtol)
        assert st["samples"] == st_ref["samples"] == p.width * p.height * p.spp
        assert abs(st["extension_rays"] - st_ref["extension_rays"]) <= 1e-3 * st_ref["extension_rays"]


def test_noalloc_reference_sampling_and_no_stats():
    dev = _Device()
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    p = ngi.default_params(renderer="ptdirect", width=40, height=40, num_samples=70001, max_num_vertices=6, seed=8,
                           engine=ngi.ENGINE_PERSISTENT)
    ref, _ = ngi.render(s, p)
    film, _ = _noalloc(s, p, dev, stats=False)
    _agree(film, ref, 1e-9)


def test_noalloc_rejections():
    dev = _Device()
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    p = _params("pt", 16, 16, 1)
    with pytest.raises(ngi.NgiError) as e:        # never uploaded: the entry point may not upload it
        _noalloc(s, p, dev)
    assert e.value.status == ngi.ERR_INVALID_ARG
    ngi.render(s, p)
    for bad, status in ((_params("pt", 16, 16, 1, ngi.ENGINE_WAVEFRONT), ngi.ERR_UNSUPPORTED),
                        (_params("bdpt", 16, 16, 1), ngi.ERR_UNSUPPORTED)):
        with pytest.raises(ngi.NgiError) as e:
            _noalloc(s, bad, dev)
        assert e.value.status == status
    q = _params("ptdirect", 16, 16, 1)
    q.paths_in_flight = 5120 * 64 + 1
    with pytest.raises(ngi.NgiError) as e:
        _noalloc(s, q, dev)
    assert e.value.status == ngi.ERR_INVALID_ARG
    # Parameters changed since the upload: rejected until an allocating call refreshes them.
    s.set_params(None, s.desc()["primitives"]["Le"] * 2.0)
    with pytest.raises(ngi.NgiError) as e:
        _noalloc(s, p, dev)
    assert e.value.status == ngi.ERR_INVALID_ARG
