"""Launch tuning for one scene (ngi_autotune) and the concurrency knob it sets.

paths_in_flight caps the lanes of the persistent kernels; lanes take the
next path from a shared counter, so every cap renders the same paths with
the same random numbers and only the order of the FP64 film additions can
change.  The tuner's choice must be one of its candidates, the fastest of
them, and must render the image the default engine renders.
"""
import numpy as np
import pytest

import nanogi_amd as ngi
from tests import scenes

pytestmark = pytest.mark.gpu


def _p(renderer, w=32, h=32, spp=8, m=8, **kw):
    return ngi.default_params(renderer=renderer, width=w, height=h, sampling=ngi.SAMPLING_PER_PIXEL, spp=spp,
                              max_num_vertices=m, seed=5, engine=ngi.ENGINE_PERSISTENT, **kw)


def _same(a, b):
    np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12 * np.abs(a).max())


@pytest.mark.parametrize("renderer", ["pt", "ptdirect", "lt", "ltdirect", "bdpt"])
@pytest.mark.parametrize("lanes", [256, 1000, 20000])
def test_lane_caps_render_the_same_paths(renderer, lanes):
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    full, st0 = ngi.render(s, _p(renderer))
    p = _p(renderer)
    p.paths_in_flight = lanes
    capped, st1 = ngi.render(s, p)
    # lt never reaches a pinhole camera (its image is zero, as the reference's).
    assert (np.abs(full).max() > 0) == (renderer != "lt")
    _same(capped, full)
    assert st1["samples"] == st0["samples"] == 32 * 32 * 8


@pytest.mark.parametrize("renderer", ["ptdirect", "bdpt"])
def test_autotune_picks_its_fastest_candidate(renderer):
    s = ngi.Scene.generate(ngi.GEN_BATCH_ROOM, 3, 1.0, 2000)
    p = _p(renderer, 64, 64, 16, 10, rr_depth=3)
    tuned, trials = ngi.autotune(s, p, tune_paths=1 << 18)
    engines = {t["engine"] for t in trials}
    if renderer == "bdpt":
        assert len(trials) == 3 and engines == {ngi.ENGINE_PERSISTENT}
    else:
        assert len(trials) == 8 and engines == {ngi.ENGINE_PERSISTENT, ngi.ENGINE_WAVEFRONT, ngi.ENGINE_MEGAKERNEL}
    assert all(t["ms"] > 0 and t["samples_per_s"] > 0 for t in trials)
    best = min(trials, key=lambda t: t["ms"])
    assert (tuned.engine, tuned.paths_in_flight) == (best["engine"], best["paths_in_flight"])
    # Everything but the launch shape is left as it was.
    for name, _ in ngi.RenderParams._fields_:
        if name not in ("engine", "paths_in_flight"):
            assert getattr(tuned, name) == getattr(p, name), name
    # The tuned render draws the default render's paths: equal up to the
    # FP32 rounding of the other engines (the tolerance of the engine tests).
    a, _ = ngi.render(s, p)
    b, _ = ngi.render(s, tuned)
    if tuned.engine == ngi.ENGINE_PERSISTENT:
        _same(b, a)
    else:
        np.testing.assert_allclose(b.mean(), a.mean(), rtol=2e-3)


def test_autotune_fp64_shading_skips_the_wavefront_engine():
    s = scenes.furnace()
    p = _p("pt", 16, 16, 4, 6, precision=ngi.PRECISION_FP64_ALL)
    tuned, trials = ngi.autotune(s, p, tune_paths=1 << 14)
    assert len(trials) == 4 and ngi.ENGINE_WAVEFRONT not in {t["engine"] for t in trials}
    assert tuned.engine in (ngi.ENGINE_PERSISTENT, ngi.ENGINE_MEGAKERNEL)


def test_autotune_reference_uniform_sampling():
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    p = ngi.default_params(renderer="pt", width=32, height=24, num_samples=50000, max_num_vertices=6)
    tuned, trials = ngi.autotune(s, p, tune_paths=1 << 16)
    assert tuned.num_samples == 50000 and len(trials) == 8
    f, st = ngi.render(s, tuned)
    assert st["samples"] == 50000 and np.isfinite(f).all() and f.mean() > 0


@pytest.mark.parametrize("renderer", ["ptdirect", "ltdirect", "bdpt"])
def test_scratch_grows_between_queued_launches(renderer):
    """A one-block launch, then a full-device launch queued behind it without
    a synchronisation: the session's shadow queues / subpath records grow in
    between, and the film equals one launch of all the samples."""
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    p = _p(renderer, 16, 16, 1, 6)
    a = ngi.Session(s, p)
    a.render(1)       # 256 paths: one block of lanes
    a.render(255)     # 65,280 paths: every resident block
    fa = a.film()
    a.close()
    b = ngi.Session(s, p)
    b.render(256)
    fb = b.film()
    b.close()
    assert np.abs(fb).max() > 0
    _same(fa, fb)


def test_wavefront_default_queues_grow_with_the_render():
    """With paths_in_flight = 0 the wavefront queues are sized from the render
    (2^16 paths for the first, 16,384-path render here) and reallocated for
    a larger later one (245,760 paths -> 2^18); the film equals one render
    of all the samples."""
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    p = _p("ptdirect", 128, 128, 1, 6)
    p.engine = ngi.ENGINE_WAVEFRONT
    a = ngi.Session(s, p)
    a.render(1)
    a.render(15)
    fa = a.film()
    a.close()
    b = ngi.Session(s, p)
    b.render(16)
    fb = b.film()
    b.close()
    assert np.abs(fb).max() > 0
    _same(fa, fb)
