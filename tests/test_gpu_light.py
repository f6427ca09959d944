"""GPU light tracers (lt, ltdirect) against the CPU FP64 oracle.

Both draw the light choice, the light position and every direction from the
same counter-based streams, so each light path is the same path on the GPU
and in the oracle up to floating-point rounding; the bar is the one of the
path tracers (test_gpu_render.py): at least 97% of pixels within a relative
2e-4 to 5e-4 and image means within 2e-3 (5e-3 with glossy surfaces).
"""
import numpy as np
import pytest

import nanogi_amd as ngi
import oracle
from tests import scenes

pytestmark = pytest.mark.gpu


def _p(renderer, w, h, n, m, seed=1, rr_depth=-1, precision=ngi.PRECISION_FP32_FP64ACC):
    p = ngi.default_params(renderer=renderer, width=w, height=h, sampling=ngi.SAMPLING_REFERENCE_UNIFORM,
                           max_num_vertices=m, rr_depth=rr_depth, seed=seed, precision=precision)
    p.num_samples = n
    return p


def _crn_close(gpu, ref, pix_rtol=2e-4, frac=0.97, mean_rtol=2e-3):
    err = np.abs(gpu - ref).max(axis=2)
    scale = np.abs(ref).max(axis=2) + 1e-3 * np.abs(ref).max()
    ok = err <= pix_rtol * scale
    assert ok.mean() >= frac, f"only {ok.mean():.4f} of pixels agree (worst {np.max(err / scale):.3g})"
    np.testing.assert_allclose(gpu.reshape(-1, 3).mean(0), ref.reshape(-1, 3).mean(0), rtol=mean_rtol)


@pytest.mark.parametrize("width", [ngi.BVH_2, ngi.BVH_8])
@pytest.mark.parametrize("precision", [ngi.PRECISION_FP32_FP64ACC, ngi.PRECISION_FP32_ALL, ngi.PRECISION_FP64_ALL])
def test_ltdirect_cornell_common_random_numbers(width, precision):
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    s.build(width)
    p = _p("ltdirect", 32, 32, 200000, 9, precision=precision)
    g, st = ngi.render(s, p)
    o = oracle.render(s, p)
    assert np.abs(o).max() > 0 and st["samples"] == 200000 and st["shadow_rays"] > 0
    # FP32 film sums (FP32_ALL) round at ~1e-7 per addition over ~200 splats per pixel.
    _crn_close(g, o, 2e-4 if precision != ngi.PRECISION_FP32_ALL else 5e-4)


def test_ltdirect_shading_normals_glossy_and_roulette():
    s = scenes.tilted_normals_box()
    p = _p("ltdirect", 24, 24, 150000, 7, seed=4, rr_depth=2)
    _crn_close(ngi.render(s, p)[0], oracle.render(s, p), 5e-4, mean_rtol=5e-3)


@pytest.mark.parametrize("kind", ["point", "directional"])
def test_ltdirect_point_and_directional_lights(kind):
    s = scenes.plane_with_light(kind)
    p = _p("ltdirect", 16, 16, 100000, 3, seed=2)
    _crn_close(ngi.render(s, p)[0], oracle.render(s, p))


def test_ltdirect_per_pixel_sampling_and_batches():
    """Per-pixel mode numbers light paths by (pixel, sample) counter words;
    sessions continue the sample index across batches as the oracle's ids do."""
    s = ngi.Scene.generate(ngi.GEN_BATCH_ROOM, 5, 1.0, 1500)
    p = ngi.default_params(renderer="ltdirect", width=16, height=16, sampling=ngi.SAMPLING_PER_PIXEL, spp=64,
                           max_num_vertices=6, rr_depth=3, seed=9)
    se = ngi.Session(s, p)
    for _ in range(4):
        se.render(16)
    g = se.film()
    se.close()
    _crn_close(g, oracle.render(s, p), 5e-4)


def test_lt_traces_and_adds_nothing_with_a_pinhole():
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    g, st = ngi.render(s, _p("lt", 16, 16, 50000, 6))
    assert np.all(g == 0) and st["extension_rays"] > 50000 and st["shadow_rays"] == 0


def test_ltdirect_counters_and_engine_choice():
    """The light tracers run on the persistent engine whatever engine is
    asked for; traversal counters are collected as for the path tracers."""
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    out = []
    for eng in (ngi.ENGINE_MEGAKERNEL, ngi.ENGINE_WAVEFRONT, ngi.ENGINE_PERSISTENT):
        p = _p("ltdirect", 16, 16, 30000, 5, seed=3)
        p.engine = eng
        p.collect_counters = 1
        out.append(ngi.render(s, p))
    for g, st in out[1:]:
        np.testing.assert_allclose(g, out[0][0], rtol=1e-9, atol=1e-12 * np.abs(out[0][0]).max())
        assert st["node_visits"] > st["extension_rays"] > 0


@pytest.mark.parametrize("spec", [ngi.SPEC_REFRACTION, ngi.SPEC_FRESNEL])
def test_ltdirect_through_glass(spec):
    """Light paths enter the glass room through its top and reach the floor
    (light-to-eye transport: no eta^2 scaling at the interface); every sensor
    connection from inside the glass reaches the camera inside it."""
    s = scenes.glass_room(spec)
    p = _p("ltdirect", 24, 24, 300000, 8, seed=3, rr_depth=3)
    g, _ = ngi.render(s, p)
    o = oracle.render(s, p)
    assert (o.sum(axis=2) > 0).mean() > 0.2
    _crn_close(g, o, 5e-4, frac=0.95, mean_rtol=5e-3)


@pytest.mark.parametrize("precision", [ngi.PRECISION_FP32_FP64ACC, ngi.PRECISION_FP64_ALL])
@pytest.mark.parametrize("kind,detail,pix_rtol", [(ngi.GEN_NAVE, 0, 5e-3), (ngi.GEN_SPHERE_FIELD, 3, 5e-4)])
def test_ltdirect_nave_and_glossy_sphere_field(kind, detail, pix_rtol, precision):
    """Light paths of up to 10 vertices over thousands of triangles, with
    FP32 and FP64 shading (both scenes are large enough for the sensor
    connections to go through the shadow rings).  Each light path splats
    into up to nine pixels, so the few paths whose FP32 rays take other
    geometry than the FP64 ones (57 of 1.18 M extension rays on the nave)
    touch many pixels by about 1/500 each: on the nave the per-pixel bar is
    5e-3; the image means agree to 2e-5 there."""
    s = ngi.Scene.generate(kind, 1, 1.0, detail)
    p = _p("ltdirect", 24, 24, 300000, 10, seed=5, rr_depth=3, precision=precision)
    g, st = ngi.render(s, p)
    o = oracle.render(s, p)
    assert np.abs(o).max() > 0 and st["shadow_rays"] > 0
    _crn_close(g, o, pix_rtol, frac=0.95, mean_rtol=2e-3)
