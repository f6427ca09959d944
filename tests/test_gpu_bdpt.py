"""GPU bidirectional path tracer (bdpt) against the CPU FP64 oracle.

Both draw every subpath from the same counter-based streams, so each sample
is the same pair of subpaths on the GPU and in the oracle up to rounding.
The GPU computes the power-heuristic weights by walking pdf ratios outwards
from each strategy; the oracle recomputes every strategy's pdf from scratch
as the reference does -- the two agree to the bar of the other renderers:
97% of pixels within a relative 2e-4 to 5e-4 and image means within 2e-3
(5e-3 with glossy or specular surfaces).
"""
import os

import numpy as np
import pytest

import nanogi_amd as ngi
import oracle
from tests import scenes

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _p(w, h, spp, m, rr_depth=1, seed=1, precision=ngi.PRECISION_FP32_FP64ACC, sampling=ngi.SAMPLING_PER_PIXEL):
    return ngi.default_params(renderer="bdpt", width=w, height=h, sampling=sampling, spp=spp, max_num_vertices=m,
                              rr_depth=rr_depth, seed=seed, precision=precision)


def _crn_close(gpu, ref, pix_rtol=2e-4, frac=0.97, mean_rtol=2e-3):
    err = np.abs(gpu - ref).max(axis=2)
    scale = np.abs(ref).max(axis=2) + 1e-3 * np.abs(ref).max()
    ok = err <= pix_rtol * scale
    assert ok.mean() >= frac, f"only {ok.mean():.4f} of pixels agree (worst {np.max(err / scale):.3g})"
    np.testing.assert_allclose(gpu.reshape(-1, 3).mean(0), ref.reshape(-1, 3).mean(0), rtol=mean_rtol)


@pytest.mark.parametrize("width", [ngi.BVH_2, ngi.BVH_8])
@pytest.mark.parametrize("precision", [ngi.PRECISION_FP32_FP64ACC, ngi.PRECISION_FP32_ALL])
def test_bdpt_cornell_common_random_numbers(width, precision):
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    s.build(width)
    p = _p(32, 32, 16, 9, precision=precision)
    g, st = ngi.render(s, p)
    o = oracle.render(s, p)
    assert np.abs(o).max() > 0 and st["shadow_rays"] > 0
    _crn_close(g, o, 2e-4 if precision != ngi.PRECISION_FP32_ALL else 5e-4)


def test_bdpt_shading_normals_and_glossy():
    s = scenes.tilted_normals_box()
    p = _p(24, 24, 16, 6, seed=4)
    _crn_close(ngi.render(s, p)[0], oracle.render(s, p), 5e-4, mean_rtol=5e-3)


def test_bdpt_specular_scene_and_uniform_sampling():
    """The reference's sample scene with Fresnel spheres; raster-uniform sampling."""
    s = ngi.Scene.load_yaml(os.path.join(GOLDEN, "runc_data", "scene.yml"), 1.0)
    p = _p(24, 24, 1, 7, seed=2, sampling=ngi.SAMPLING_REFERENCE_UNIFORM)
    p.num_samples = 24 * 24 * 16
    _crn_close(ngi.render(s, p)[0], oracle.render(s, p), 5e-4, frac=0.95, mean_rtol=5e-3)


@pytest.mark.parametrize("kind", ["point", "directional"])
def test_bdpt_point_and_directional_lights(kind):
    s = scenes.plane_with_light(kind, fov=60.0, half=1.0)
    p = _p(16, 16, 32, 4, seed=6)
    _crn_close(ngi.render(s, p)[0], oracle.render(s, p))


def test_bdpt_unbounded_length_with_roulette_and_engine_choice():
    """-m -1 (subpaths capped at 64 vertices) with roulette from the first
    bounce; bdpt runs on the persistent engine whatever engine is asked for."""
    s = ngi.Scene.generate(ngi.GEN_BATCH_ROOM, 5, 1.0, 1500)
    out = []
    for eng in (ngi.ENGINE_PERSISTENT, ngi.ENGINE_WAVEFRONT):
        p = _p(16, 16, 8, -1, seed=8)
        p.engine = eng
        out.append(ngi.render(s, p)[0])
    np.testing.assert_allclose(out[1], out[0], rtol=1e-9, atol=1e-12 * np.abs(out[0]).max())
    _crn_close(out[0], oracle.render(s, _p(16, 16, 8, -1, seed=8)), 5e-4)


@pytest.mark.parametrize("spec", [ngi.SPEC_REFRACTION, ngi.SPEC_FRESNEL])
def test_bdpt_through_glass(spec):
    """Camera and floor inside a glass box, the light outside: every strategy
    that connects across the glass is blocked, so the image comes from the
    strategies whose subpaths cross the refracting interface (their c = 0
    neighbours drop out of the MIS sum)."""
    s = scenes.glass_room(spec)
    p = _p(24, 24, 32, 8, seed=5)
    g, _ = ngi.render(s, p)
    o = oracle.render(s, p)
    assert (o.sum(axis=2) > 0).mean() > 0.2
    _crn_close(g, o, 5e-4, frac=0.95, mean_rtol=5e-3)


@pytest.mark.parametrize("kind,detail", [(ngi.GEN_NAVE, 0), (ngi.GEN_SPHERE_FIELD, 3)])
def test_bdpt_nave_and_glossy_sphere_field(kind, detail):
    """Long paths (up to 12 vertices, roulette from 3) over thousands of
    triangles, diffuse (nave, 4 lights) and Beckmann-glossy patches (sphere
    field, 17 lights).  A subpath whose FP32 ray hits other geometry than
    the FP64 one changes every strategy of its sample, so on the nave's long
    paths 95% of the pixels is the bar (96.5% agree here)."""
    s = ngi.Scene.generate(kind, 1, 1.0, detail)
    p = _p(24, 24, 8, 12, rr_depth=3, seed=7)
    g, st = ngi.render(s, p)
    o = oracle.render(s, p)
    assert np.abs(o).max() > 0 and st["shadow_rays"] > 0
    _crn_close(g, o, 5e-4, frac=0.95, mean_rtol=5e-3)
