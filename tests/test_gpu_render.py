"""GPU renders against the CPU FP64 oracle.

Both consume the same counter-based random numbers for a given (pixel,
sample, vertex, stream), so the GPU image equals the oracle image path by
path up to floating-point rounding; paths diverge only where an FP32 ray
hits different geometry than the FP64 one (edges, grazing angles).  The
bar, written per test: at least 97-98% of pixels within a relative 2e-4 to
5e-4 of the oracle, and the image means within a relative 2e-3 (5e-3 for
the glossy and large scenes).  An independent-seed block z-test (chi-square
over block means) checks the estimate for bias.
"""
import numpy as np
import pytest

import nanogi_amd as ngi
import oracle
from tests import scenes

pytestmark = pytest.mark.gpu

ENGINES = [ngi.ENGINE_MEGAKERNEL, ngi.ENGINE_WAVEFRONT, ngi.ENGINE_PERSISTENT]


def _params(renderer, w, h, spp, m, rr_depth=-1, seed=1, precision=ngi.PRECISION_FP32_FP64ACC,
            engine=ngi.ENGINE_MEGAKERNEL, sampling=ngi.SAMPLING_PER_PIXEL):
    return ngi.default_params(renderer=renderer, width=w, height=h, spp=spp, max_num_vertices=m, rr_depth=rr_depth,
                              seed=seed, precision=precision, engine=engine, sampling=sampling)


def _crn_close(gpu, ref, pix_rtol, frac=0.98, mean_rtol=2e-3):
    err = np.abs(gpu - ref).max(axis=2)
    scale = np.abs(ref).max(axis=2) + 1e-3 * np.abs(ref).max()
    ok = err <= pix_rtol * scale
    assert ok.mean() >= frac, f"only {ok.mean():.4f} of pixels agree (worst {np.max(err / scale):.3g})"
    np.testing.assert_allclose(gpu.reshape(-1, 3).mean(0), ref.reshape(-1, 3).mean(0), rtol=mean_rtol)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("renderer", ["pt", "ptdirect"])
@pytest.mark.parametrize("width", [ngi.BVH_2, ngi.BVH_8])
def test_cornell_common_random_numbers(engine, renderer, width):
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    s.build(width)
    p = _params(renderer, 48, 48, 8, 9, engine=engine)
    g, _ = ngi.render(s, p)
    o = oracle.render(s, p)
    _crn_close(g, o, 2e-4)


# FP64 shading still starts every bounce from an FP32 hit distance (the
# traversal is FP32, as Embree's is in the reference), so it agrees with the
# oracle to the same per-pixel tolerance as FP32 shading.
@pytest.mark.parametrize("engine", [ngi.ENGINE_MEGAKERNEL, ngi.ENGINE_PERSISTENT])
@pytest.mark.parametrize("renderer", ["pt", "ptdirect"])
@pytest.mark.parametrize("precision,tol", [(ngi.PRECISION_FP64_ALL, 2e-4), (ngi.PRECISION_FP32_ALL, 3e-4)])
def test_precision_variants(precision, tol, renderer, engine):
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    p = _params(renderer, 32, 32, 8, 9, rr_depth=2, precision=precision, engine=engine)
    g, _ = ngi.render(s, p)
    _crn_close(g, oracle.render(s, p), tol)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("renderer", ["pt", "ptdirect"])
def test_russian_roulette_and_reference_sampling(engine, renderer):
    p = _params(renderer, 32, 24, 1, -1, rr_depth=1, sampling=ngi.SAMPLING_REFERENCE_UNIFORM, engine=engine)
    p.num_samples = 32 * 24 * 16
    s2 = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 32 / 24)
    g, st = ngi.render(s2, p)
    assert st["samples"] == p.num_samples
    _crn_close(g, oracle.render(s2, p), 2e-4, frac=0.97)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("renderer", ["pt", "ptdirect"])
def test_glossy_sphere_field(engine, renderer):
    s = ngi.Scene.generate(ngi.GEN_SPHERE_FIELD, 1, 1.0, detail=4)
    p = _params(renderer, 32, 32, 4, 6, rr_depth=3, engine=engine)
    g, _ = ngi.render(s, p)
    _crn_close(g, oracle.render(s, p), 5e-4, frac=0.97, mean_rtol=5e-3)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("renderer", ["pt", "ptdirect"])
def test_yaml_scene_with_specular(engine, renderer):
    """The sample scene's Fresnel glass, mirror and glossy spheres
    (rt.hpp:804-903, 1055-1137) on every engine, with and without
    next-event estimation."""
    s = ngi.Scene.load_yaml("tests/golden/runc_data/scene.yml", 1.0)
    prims = s.desc()["primitives"]
    assert np.any(prims["type"] & ngi.PRIM_S)
    p = _params(renderer, 32, 32, 8, 8, rr_depth=2, engine=engine)
    g, _ = ngi.render(s, p)
    _crn_close(g, oracle.render(s, p), 5e-4, frac=0.97, mean_rtol=5e-3)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("renderer", ["pt", "ptdirect"])
@pytest.mark.parametrize("spec", [ngi.SPEC_REFRACTION, ngi.SPEC_FRESNEL])
def test_glass_room(engine, renderer, spec):
    """Camera and floor inside a glass box, the light outside: every
    lit path crosses a refracting interface (eta^2 radiance scaling on
    leaving the glass, total internal reflection at grazing angles).  The
    glass blocks every shadow ray, so ptdirect (which adds no emission at
    hits) renders black, as the reference's estimator does."""
    s = scenes.glass_room(spec)
    p = _params(renderer, 32, 32, 16, 10, rr_depth=3, engine=engine)
    g, _ = ngi.render(s, p)
    o = oracle.render(s, p)
    if renderer == "ptdirect":
        assert not o.any() and not g.any()
        return
    assert (o.sum(axis=2) > 0).mean() > 0.2
    _crn_close(g, o, 5e-4, frac=0.97, mean_rtol=5e-3)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("renderer", ["pt", "ptdirect"])
@pytest.mark.parametrize("kind", ["point", "directional"])
def test_point_and_directional_light_on_a_plane(engine, renderer, kind):
    """A diffuse plane under a point light or a downward directional light.
    ptdirect reaches the point light by next-event estimation
    (rt.hpp:542-547, 929-932); pt cannot hit it, and the directional light
    contributes nothing to either (rt.hpp:934-937), as in the reference."""
    s = scenes.plane_with_light(kind)
    p = _params(renderer, 24, 24, 8, 5, engine=engine)
    g, _ = ngi.render(s, p)
    o = oracle.render(s, p)
    if renderer == "ptdirect" and kind == "point":
        assert o.min() > 0
        _crn_close(g, o, 2e-4, frac=0.98)
    else:
        assert not o.any() and not g.any()


@pytest.mark.parametrize("precision,tol", [(ngi.PRECISION_FP64_ALL, 5e-4), (ngi.PRECISION_FP32_FP64ACC, 5e-4)])
def test_sphere_field_persistent_64spp(precision, tol):
    """The benchmark's scene family (displaced icosphere with
    Beckmann-glossy patches, heightfield, 17 lights) at 4 subdivisions,
    64 spp, up to 33 vertices with roulette from 4, on the default
    persistent engine: FP64 shading (the precision of the RMSE reference)
    and FP32 shading against the oracle path by path."""
    s = ngi.Scene.generate(ngi.GEN_SPHERE_FIELD, 1, 1.0, 4)
    p = _params("ptdirect", 48, 48, 64, 33, rr_depth=4, engine=ngi.ENGINE_PERSISTENT, precision=precision)
    g, st = ngi.render(s, p)
    assert st["samples"] == 48 * 48 * 64
    _crn_close(g, oracle.render(s, p), tol, frac=0.97, mean_rtol=2e-3)


@pytest.mark.parametrize("engine", ENGINES)
def test_nave_ptdirect(engine):
    s = ngi.Scene.generate(ngi.GEN_NAVE, 1, 1.0)
    p = _params("ptdirect", 32, 32, 2, 8, rr_depth=4, engine=engine)
    g, _ = ngi.render(s, p)
    _crn_close(g, oracle.render(s, p), 5e-4, frac=0.97, mean_rtol=5e-3)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("m", [2, 3, 6])
def test_furnace_closed_form(engine, m):
    Le, rho = (1.0, 0.5, 0.25), (0.5, 0.6, 0.7)
    s = scenes.furnace(Le, rho)
    g, _ = ngi.render(s, _params("pt", 16, 16, 16, m, engine=engine))
    expect = scenes.furnace_radiance(Le, rho, m)
    close = np.isclose(g, expect, rtol=1e-5).all(axis=2)
    assert close.mean() > 0.97
    np.testing.assert_allclose(g.reshape(-1, 3).mean(0), expect, rtol=2e-3)


@pytest.mark.parametrize("renderer", ["pt", "ptdirect"])
def test_independent_seed_block_z_test(renderer):
    """Unbiasedness against an independent oracle estimate, calibrated: the
    GPU image (default engine, 64 batches of 64 spp) and the oracle's (32
    independent seeds of 512 spp) are averaged over 8 x 8-pixel blocks
    (12 block-channel means of 262,144 and 1,048,576 paths, so the central
    limit theorem holds where single pixels are heavy-tailed).  With each
    side's variance from its own batch means, z = (G - O) / sqrt(VG + VO)
    is close to N(0, 1) under the null.  Bounds: sum(z^2) < 48 (chi-square,
    12 degrees of freedom: P ~ 1e-5), max |z| < 4.5, and the z-score of the
    image mean, sum(G - O) / sqrt(sum(VG + VO)), within 4.5 (a bias common
    to all blocks; a 1% bias gives about 20 for ptdirect here).  Twelve
    oracle-against-oracle trials of this recipe gave at most 34 / 3.2 / 3.0."""
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    W = H = 16
    p = _params(renderer, W, H, 1, 9, rr_depth=3, engine=ngi.ENGINE_PERSISTENT)
    sess = ngi.Session(s, p)
    for _ in range(64):
        sess.render_batch(64)
    g, var = sess.film(variance=True)
    blocks = lambda x: x.reshape(H // 8, 8, W // 8, 8, 3).mean(axis=(1, 3))
    G = blocks(g)
    VG = blocks(var) / 64.0          # pixels are independent: block variance = mean of variances / 64
    ob = np.stack([blocks(oracle.render(s, _params(renderer, W, H, 512, 9, rr_depth=3, seed=1000 + k))) for k in range(32)])
    O = ob.mean(0)
    VO = ob.var(0, ddof=1) / len(ob)
    assert np.all(VG > 0) and np.all(VO > 0)
    z = ((G - O) / np.sqrt(VG + VO)).ravel()
    chi2 = float(np.sum(z * z))
    assert 1.0 < chi2 < 48.0, f"chi-square {chi2:.1f} over {z.size} block means"
    assert np.abs(z).max() < 4.5, np.abs(z).max()
    z_img = float(np.sum(G - O) / np.sqrt(np.sum(VG + VO)))
    assert abs(z_img) < 4.5, z_img


def test_wavefront_small_queue_many_iterations():
    """A queue far smaller than the path count forces refills every
    iteration; the image must not depend on the queue capacity."""
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    p = _params("ptdirect", 24, 24, 4, 9, rr_depth=2, engine=ngi.ENGINE_WAVEFRONT)
    p.paths_in_flight = 300
    g, st = ngi.render(s, p)
    assert st["samples"] == 24 * 24 * 4
    p2 = _params("ptdirect", 24, 24, 4, 9, rr_depth=2, engine=ngi.ENGINE_MEGAKERNEL)
    m, st2 = ngi.render(s, p2)
    # Same paths, same random numbers; the two engines differ only in FP32
    # instruction selection, so a handful of paths may diverge.
    assert abs(st2["extension_rays"] - st["extension_rays"]) <= 1e-3 * st["extension_rays"]
    _crn_close(g, m, 2e-4)


def test_wavefront_rejects_fp64_shading():
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    p = _params("pt", 8, 8, 1, 4, precision=ngi.PRECISION_FP64_ALL, engine=ngi.ENGINE_WAVEFRONT)
    with pytest.raises(ngi.NgiError) as e:
        ngi.render(s, p)
    assert e.value.status == ngi.ERR_UNSUPPORTED


def test_session_batches_accumulate():
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    p = _params("pt", 16, 16, 1, 6, rr_depth=2, engine=ngi.ENGINE_WAVEFRONT)
    one, _ = ngi.render(s, _params("pt", 16, 16, 4, 6, rr_depth=2, engine=ngi.ENGINE_WAVEFRONT))
    sess = ngi.Session(s, p)
    for _ in range(4):
        sess.render(1)
    np.testing.assert_allclose(sess.film(), one, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("renderer", ["pt", "ptdirect"])
def test_persistent_engine_matches_megakernel(renderer):
    """Path regeneration changes which lane runs a path, not the path."""
    s = ngi.Scene.generate(ngi.GEN_SPHERE_FIELD, 1, 1.0, detail=4)
    out = {}
    for e in (ngi.ENGINE_MEGAKERNEL, ngi.ENGINE_PERSISTENT):
        out[e] = ngi.render(s, _params(renderer, 40, 30, 8, 12, rr_depth=3, engine=e))
    (g, st), (m, st2) = out[ngi.ENGINE_PERSISTENT], out[ngi.ENGINE_MEGAKERNEL]
    assert st["samples"] == st2["samples"] == 40 * 30 * 8
    assert abs(st["extension_rays"] - st2["extension_rays"]) <= 1e-3 * st2["extension_rays"]
    _crn_close(g, m, 2e-4)


@pytest.mark.parametrize("renderer", ["pt", "ptdirect"])
def test_fp64_persistent_engine_matches_megakernel(renderer):
    """FP64 shading on the persistent engine (resumable traversals, NEE shadow
    rays traced in place) against the FP64 megakernel: the same paths with the
    same FP64 arithmetic, so only the order of the film's FP64 additions
    differs."""
    s = ngi.Scene.generate(ngi.GEN_SPHERE_FIELD, 1, 1.0, detail=4)
    s.build(ngi.BVH_8)
    out = {}
    for e in (ngi.ENGINE_MEGAKERNEL, ngi.ENGINE_PERSISTENT):
        out[e] = ngi.render(s, _params(renderer, 40, 30, 8, 12, rr_depth=3, engine=e, precision=ngi.PRECISION_FP64_ALL))
    (g, st), (m, st2) = out[ngi.ENGINE_PERSISTENT], out[ngi.ENGINE_MEGAKERNEL]
    assert st["samples"] == st2["samples"] == 40 * 30 * 8
    for k in ("extension_rays", "shadow_rays"):
        assert abs(st[k] - st2[k]) <= 1e-4 * st2[k], k
    _crn_close(g, m, 1e-9, frac=0.999, mean_rtol=1e-4)


def test_deep_wide_bvh_resumable_at_four_blocks_and_fallback(monkeypatch):
    """A wide BVH 10 levels deep (5.8M triangles) needs 33 KiB of LDS per
    block for resumable traversals (two stacks of 11 entries and the parked
    state per lane): 4 blocks per CU fit, 5 do not.  By default the
    resumable kernels then run at 4 waves/SIMD; with NGI_RESUME_MIN_BLOCKS=5
    the launchers fall back to the kernels that run every traversal to its
    end.  ptdirect must agree with the megakernel path by path either way,
    and ltdirect (persistent only) between the two kernels."""
    s = ngi.Scene.generate(ngi.GEN_SPHERE_FIELD, 1, 1.0, detail=9)
    s.build(ngi.BVH_8)
    assert 10 <= s.bvh_info()["max_depth"] <= 13
    m, st_m = ngi.render(s, _params("ptdirect", 48, 32, 8, 12, rr_depth=3, engine=ngi.ENGINE_MEGAKERNEL))
    lt = {}
    for min_blocks in ("4", "5"):
        monkeypatch.setenv("NGI_RESUME_MIN_BLOCKS", min_blocks)
        g, st = ngi.render(s, _params("ptdirect", 48, 32, 8, 12, rr_depth=3, engine=ngi.ENGINE_PERSISTENT))
        assert st["samples"] == st_m["samples"] == 48 * 32 * 8
        assert abs(st["extension_rays"] - st_m["extension_rays"]) <= 1e-3 * st_m["extension_rays"], min_blocks
        _crn_close(g, m, 2e-4)
        lt[min_blocks] = ngi.render(s, _params("ltdirect", 48, 32, 8, 12, rr_depth=3, engine=ngi.ENGINE_PERSISTENT))
    (a, sa), (b, sb) = lt["4"], lt["5"]
    assert sa["samples"] == sb["samples"] and abs(sa["extension_rays"] - sb["extension_rays"]) <= 1e-3 * sb["extension_rays"]
    _crn_close(a, b, 2e-4)


@pytest.mark.parametrize("width", [ngi.BVH_2, ngi.BVH_8])
def test_scene_without_triangles(width):
    """A camera and a point light and nothing else: every ray misses (an
    Embree scene with no geometry), so pt and lt see nothing and the
    next-event estimators splat the light where it projects."""
    B = scenes.SceneBuilder()
    B.prim(type=ngi.PRIM_L, light_type=ngi.LIGHT_POINT, Le=(2, 1, 1), light_position=(0.1, 0.2, -2.0))
    B.camera((0, 0, 0), (0, 0, -1))
    s = ngi.Scene.create(np.zeros((0, 3)), np.zeros((0, 3), dtype=np.int64), np.zeros(0, dtype=np.int64),
                         np.array(B.prims), normals=np.zeros((0, 3)))
    s.build(width)
    for r in ("pt", "ptdirect", "lt", "ltdirect", "bdpt"):
        for eng in ENGINES:
            p = ngi.default_params(renderer=r, width=16, height=16, sampling=ngi.SAMPLING_PER_PIXEL, spp=4,
                                   max_num_vertices=4, seed=3, engine=eng)
            g, st = ngi.render(s, p)
            o = oracle.render(s, p)
            assert st["extension_rays"] >= 0 and np.isfinite(g).all()
            np.testing.assert_allclose(g, o, rtol=1e-5, atol=1e-9, err_msg=f"{r} {eng}")
            assert (np.count_nonzero(o.sum(axis=2)) == 1) == (r not in ("pt", "lt")), r


def test_fp64_shading_edge_cases():
    """FP64 shading (resumable persistent kernels with FP64 parked state and
    FP64-contribution queues) on the edge cases: no triangles at all, paths of
    0 to 3 vertices, and images smaller than a wavefront or ragged, against
    the oracle (pt / ptdirect also on the FP64 megakernel)."""
    P = ngi.PRECISION_FP64_ALL
    B = scenes.SceneBuilder()
    B.prim(type=ngi.PRIM_L, light_type=ngi.LIGHT_POINT, Le=(2, 1, 1), light_position=(0.1, 0.2, -2.0))
    B.camera((0, 0, 0), (0, 0, -1))
    empty = ngi.Scene.create(np.zeros((0, 3)), np.zeros((0, 3), dtype=np.int64), np.zeros(0, dtype=np.int64),
                             np.array(B.prims), normals=np.zeros((0, 3)))
    empty.build(ngi.BVH_8)
    corn = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    corn.build(ngi.BVH_8)
    for r in ("pt", "ptdirect", "lt", "ltdirect"):
        engines = [ngi.ENGINE_MEGAKERNEL, ngi.ENGINE_PERSISTENT] if r in ("pt", "ptdirect") else [ngi.ENGINE_PERSISTENT]
        for eng in engines:
            p = ngi.default_params(renderer=r, width=16, height=16, sampling=ngi.SAMPLING_PER_PIXEL, spp=4,
                                   max_num_vertices=4, seed=3, engine=eng, precision=P)
            g, _ = ngi.render(empty, p)
            np.testing.assert_allclose(g, oracle.render(empty, p), rtol=1e-5, atol=1e-9, err_msg=f"{r} {eng}")
            for m in (0, 1, 2, 3):
                p = ngi.default_params(renderer=r, width=16, height=16, sampling=ngi.SAMPLING_PER_PIXEL, spp=8,
                                       max_num_vertices=m, seed=2, engine=eng, precision=P)
                g, _ = ngi.render(corn, p)
                o = oracle.render(corn, p)
                if m <= 1 or not o.any():
                    assert not g.any(), (r, m, eng)
                else:
                    _crn_close(g, o, 2e-4, frac=0.97)
            for w, h in ((1, 1), (7, 3), (65, 17)):
                p = ngi.default_params(renderer=r, width=w, height=h, sampling=ngi.SAMPLING_PER_PIXEL, spp=3,
                                       max_num_vertices=5, seed=4, engine=eng, precision=P)
                g, st = ngi.render(corn, p)
                assert st["samples"] == w * h * 3
                if r in ("pt", "ptdirect"):
                    _crn_close(g, oracle.render(corn, p), 2e-4, frac=0.97)


@pytest.mark.parametrize("renderer", ["pt", "ptdirect", "lt", "ltdirect", "bdpt"])
def test_shortest_path_limits(renderer):
    """-m 0 and -m 1 render nothing; -m 2 and -m 3 are the direct-view and
    one-bounce cases, where the GPU matches the oracle to the usual bar."""
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    engines = ENGINES if renderer in ("pt", "ptdirect") else [ngi.ENGINE_PERSISTENT]
    for m in (0, 1, 2, 3):
        for eng in engines:
            p = ngi.default_params(renderer=renderer, width=16, height=16, sampling=ngi.SAMPLING_PER_PIXEL, spp=8,
                                   max_num_vertices=m, seed=2, engine=eng)
            g, _ = ngi.render(s, p)
            o = oracle.render(s, p)
            if m <= 1 or not o.any():
                assert not g.any(), (m, eng)
                continue
            _crn_close(g, o, 2e-4, frac=0.97)


@pytest.mark.parametrize("engine", ENGINES)
def test_every_pixel_gets_its_samples_at_odd_sizes(engine):
    """Camera rays only (-m 2) in a closed box of emitters with reflectance 0:
    every sample adds exactly Le to its own pixel, so any pixel the sample
    order (8x8 tiles per wavefront) skipped or visited twice would stand out."""
    Le = (1.0, 0.5, 0.25)
    s = scenes.furnace(Le, (0.0, 0.0, 0.0))
    for w, h in ((1, 1), (7, 3), (13, 9), (65, 17)):
        g, st = ngi.render(s, _params("pt", w, h, 3, 2, engine=engine))
        assert st["samples"] == w * h * 3
        assert (g <= np.array(Le) * (1 + 1e-6)).all(), (w, h)
        assert np.isclose(g, Le, rtol=1e-6).all(axis=2).mean() >= 0.99, (w, h)


@pytest.mark.parametrize("engine", ENGINES)
def test_vertex_cap_in_a_lossless_furnace(engine):
    """-m -1 without roulette in a closed box of reflectance 1: a path runs to
    the 1024-vertex cap and sees the walls' emission 1023 times.  (About one
    bounce in 10^4 leaves the box: a hit within t_min = 1e-4 of an edge can
    start its next ray past the neighbouring wall, as in the reference, so
    some paths end early; none may go past the cap.)"""
    Le = (1.0, 0.5, 0.25)
    s = scenes.furnace(Le, (1.0, 1.0, 1.0))
    g, st = ngi.render(s, _params("pt", 4, 4, 2, -1, engine=engine))
    expect = scenes.furnace_radiance(Le, (1.0, 1.0, 1.0), 1024)
    assert (g <= expect * (1 + 1e-6)).all()
    np.testing.assert_allclose(g.reshape(-1, 3).max(axis=0), expect, rtol=1e-5)
    assert np.isclose(g, expect, rtol=1e-5).all(axis=2).mean() >= 0.5
    assert 4 * 4 * 2 * 512 <= st["extension_rays"] <= 4 * 4 * 2 * 1023


@pytest.mark.parametrize("renderer", ["pt", "ptdirect", "lt", "ltdirect", "bdpt"])
def test_reference_sampling_with_few_paths(renderer):
    """Reference-uniform sampling with 1, 63 and 1001 paths (less than one
    wavefront, one short of a wavefront, a ragged launch): every renderer
    draws the oracle's paths and normalises by W*H/N."""
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    for n in (1, 63, 1001):
        p = ngi.default_params(renderer=renderer, width=16, height=16, num_samples=n, max_num_vertices=5, seed=9,
                               engine=ngi.ENGINE_PERSISTENT)
        g, st = ngi.render(s, p)
        o = oracle.render(s, p)
        assert st["samples"] == n
        err = np.abs(g - o).max(axis=2)
        bad = err > 2e-4 * (np.abs(o).max(axis=2) + 1e-3 * max(np.abs(o).max(), 1e-30))
        assert bad.sum() <= max(1, n // 200), (n, int(bad.sum()))


def test_full_size_sphere_field_fp32_against_fp64_shading():
    """The benchmark's scene at its full size (1,833,314 triangles, 1024x1024,
    up to 33 vertices, roulette from 4), 4 spp: FP32 shading with the shadow
    queue against FP64 shading on the same paths.  Size-independent bar: the
    usual per-pixel agreement and image means within 1e-3."""
    s = ngi.Scene.generate(ngi.GEN_SPHERE_FIELD, 1, 1.0, 8)
    info = s.build(ngi.BVH_8)
    assert s.num_triangles == 1833314 and info["max_depth"] >= 6
    out = {}
    for prec in (ngi.PRECISION_FP32_FP64ACC, ngi.PRECISION_FP64_ALL):
        p = ngi.default_params(renderer="ptdirect", width=1024, height=1024, sampling=ngi.SAMPLING_PER_PIXEL, spp=4,
                               max_num_vertices=33, rr_depth=4, seed=21, precision=prec)
        out[prec], st = ngi.render(s, p)
        assert st["samples"] == 1024 * 1024 * 4 and st["shadow_rays"] > 0
    _crn_close(out[ngi.PRECISION_FP32_FP64ACC], out[ngi.PRECISION_FP64_ALL], 2e-4, frac=0.97, mean_rtol=1e-3)
