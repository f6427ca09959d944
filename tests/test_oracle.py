"""Pins the CPU FP64 oracle against closed forms, invariants and known answers.

The reference ships no tests, golden images or vectors for this path, so
every oracle function is pinned here by mathematics that a plausible mistake
(dropped term, wrong sign or index, missing Jacobian) would break.
"""
import numpy as np
import pytest

import nanogi_amd as ngi
import oracle
from tests import scenes

# Random123 known-answer vectors for philox4x32-10 (Salmon et al. 2011).
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_known_answers(ctr, key, expect):
    assert tuple(int(x) for x in oracle.philox(ctr, key)) == expect


def _stat_close(run, expect, k=6, z=4.5, floor=1e-6):
    """Mean over k independent seeds within z standard errors of `expect`."""
    runs = np.array([run(seed) for seed in range(1, k + 1)])
    mean = runs.mean(0)
    se = runs.std(0, ddof=1) / np.sqrt(k)
    assert np.all(np.abs(mean - expect) <= z * se + floor * np.abs(expect) + 1e-12), (mean, expect, se)
    return mean, se


def _params(renderer, w, h, spp, m, rr_depth=-1, seed=1, sampling=ngi.SAMPLING_PER_PIXEL):
    return ngi.default_params(renderer=renderer, width=w, height=h, spp=spp, max_num_vertices=m,
                              rr_depth=rr_depth, seed=seed, sampling=sampling)


@pytest.mark.parametrize("m", [2, 3, 5])
def test_furnace_pt_exact(m):
    """pt in a closed emissive box: every path contributes Le*sum rho^k.

    Exact except for paths that leave a wall within 1e-4 of a corner: the
    adjacent wall is then closer than the absolute ray epsilon tnear = 1e-4
    (basic.hpp:85, rt.hpp:2248) and the path escapes, as in the reference."""
    Le, rho = (1.0, 0.5, 0.25), (0.5, 0.6, 0.7)
    s = scenes.furnace(Le, rho)
    film = oracle.render(s, _params("pt", 8, 8, 4, m))
    expect = np.broadcast_to(scenes.furnace_radiance(Le, rho, m), film.shape)
    close = np.isclose(film, expect, rtol=1e-12).all(axis=2)
    assert close.mean() > 0.97
    np.testing.assert_allclose(film.reshape(-1, 3).mean(0), expect[0, 0], rtol=2e-3)


@pytest.mark.parametrize("m", [2, 4])
def test_furnace_ptdirect_statistical(m):
    Le, rho = (1.0, 0.5, 0.25), (0.5, 0.6, 0.7)
    s = scenes.furnace(Le, rho)
    expect = scenes.furnace_radiance(Le, rho, m)
    _, se = _stat_close(lambda k: oracle.render(s, _params("ptdirect", 8, 8, 512, m, seed=k)).reshape(-1, 3).mean(0), expect)
    assert np.all(se < 0.02 * expect)


def test_furnace_with_russian_roulette_unbiased():
    Le, rho = (1.0, 1.0, 1.0), (0.5, 0.5, 0.5)
    s = scenes.furnace(Le, rho)
    _stat_close(lambda k: oracle.render(s, _params("pt", 8, 8, 256, 6, rr_depth=1, seed=k)).reshape(-1, 3).mean(0),
                scenes.furnace_radiance(Le, rho, 6))


@pytest.mark.parametrize("renderer", ["pt", "ptdirect"])
def test_plane_under_rectangular_light(renderer):
    """One bounce (-m 3): radiance rho * Le * F(point -> rectangle)."""
    a, b, h, Le, rho = 0.5, 0.3, 1.0, np.array([4.0, 3.0, 2.0]), np.array([0.8, 0.5, 0.3])
    s = scenes.plane_under_light(a, b, h, tuple(Le), tuple(rho))
    spp = 2048 if renderer == "ptdirect" else 16384
    expect = rho * Le * scenes.rect_form_factor(a, b, h)
    _, se = _stat_close(lambda k: oracle.render(s, _params(renderer, 2, 2, spp, 3, seed=k)).reshape(-1, 3).mean(0), expect)
    assert np.all(se < 0.03 * expect)


def test_directly_viewed_light_pt_and_splat():
    """-m 2: pt sees Le inside the light's footprint; ptdirect's camera-vertex
    NEE splat has the same expectation (src/nanogi.cpp:694-707)."""
    Le = np.array([3.0, 2.0, 1.0])
    s = scenes.light_in_view(tuple(Le))
    pt = oracle.render(s, _params("pt", 32, 32, 16, 2))
    frac = (0.5 / (3.0 * np.tan(np.radians(20.0)))) ** 2
    np.testing.assert_allclose(pt.reshape(-1, 3).mean(0), Le * frac, rtol=0.01)
    _stat_close(lambda k: oracle.render(s, _params("ptdirect", 16, 16, 64, 2, seed=k)).reshape(-1, 3).mean(0), Le * frac)
    # Pixels fully inside the footprint see exactly Le with pt.
    np.testing.assert_allclose(pt[16, 16], Le, rtol=1e-12)


def test_pt_and_ptdirect_agree_on_cornell():
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    means = {}
    for r in ("pt", "ptdirect"):
        runs = [oracle.render(s, _params(r, 8, 8, 1024, 9, seed=k)).reshape(-1, 3).mean(0) for k in range(1, 5)]
        means[r] = (np.mean(runs, 0), np.std(runs, 0, ddof=1) / 2)
    d = means["pt"][0] - means["ptdirect"][0]
    se = np.hypot(means["pt"][1], means["ptdirect"][1])
    assert np.all(np.abs(d) < 4 * se + 1e-4), (d, se)


def test_reference_uniform_matches_per_pixel_expectation():
    """Raster uniform over the image with W*H/N scaling (src/nanogi.cpp:434-437)
    has the per-pixel estimator's expectation."""
    Le, rho = (1.0, 1.0, 1.0), (0.5, 0.5, 0.5)
    s = scenes.furnace(Le, rho)
    p = _params("pt", 4, 4, 1, 4, sampling=ngi.SAMPLING_REFERENCE_UNIFORM)
    p.num_samples = 4096
    film = oracle.render(s, p)
    # Each path contributes the same value; only corner escapes (see above) lose some.
    np.testing.assert_allclose(film.reshape(-1, 3).mean(0), scenes.furnace_radiance(Le, rho, 4), rtol=1e-3)


# ---------------------------------------------------------------------------
# BSDF sampling versus evaluation (rt.hpp:692-1336)
# ---------------------------------------------------------------------------
def _glossy(alpha, eta=(1.657, 0.880, 0.521), k=(9.224, 6.270, 4.837), R=(1, 1, 1)):
    return ngi.make_primitive(type=ngi.PRIM_G, G_R=R, G_eta=eta, G_k=k, G_roughness=alpha)


def _conductor_normal_incidence(eta, k):
    eta, k = np.asarray(eta), np.asarray(k)
    return ((eta - 1) ** 2 + k ** 2) / ((eta + 1) ** 2 + k ** 2)


@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.5])
def test_conductor_fresnel_at_normal_incidence(alpha):
    """fs(z, z) = R D(z) G F(1) / 4 with D(z) = 1/(pi alpha^2), G = 1."""
    prim = _glossy(alpha)
    _, _, fs = oracle.bsdf(prim, ngi.PRIM_G, (0, 0, 1), wo=(0, 0, 1))
    F = fs * 4.0 * np.pi * alpha ** 2
    np.testing.assert_allclose(F, _conductor_normal_incidence(prim["G_eta"], prim["G_k"]), rtol=1e-12)


def test_dielectric_fresnel_at_normal_incidence():
    prim = ngi.make_primitive(type=ngi.PRIM_S, spec_type=ngi.SPEC_FRESNEL, S_R=(1, 1, 1), S_eta1=1.0, S_eta2=1.5)
    _, pdf, fs = oracle.bsdf(prim, ngi.PRIM_S, (0, 0, 1), wo=(0, 0, 1))
    np.testing.assert_allclose(pdf, ((1.0 - 1.5) / (1.0 + 1.5)) ** 2, rtol=1e-12)
    np.testing.assert_allclose(fs, pdf, rtol=1e-12)


@pytest.mark.parametrize("alpha,theta_i", [(0.1, 10.0), (0.3, 45.0), (0.6, 70.0)])
def test_beckmann_sampling_matches_pdf(alpha, theta_i):
    """Acceptance rate of the sampler == integral of its projected-solid-angle
    pdf over the hemisphere (a missing Jacobian or wrong D fails this)."""
    rng = np.random.default_rng(7)
    prim = _glossy(alpha)
    ti = np.radians(theta_i)
    wi = (np.sin(ti), 0.0, np.cos(ti))
    n = 20000
    acc = 0
    for u in rng.random((n, 3)):
        wo, pdf, _ = oracle.bsdf(prim, ngi.PRIM_G, wi, u=u)
        acc += wo[2] > 0
    # Integral of pdf_proj * cos over the hemisphere by cosine-weighted sampling:
    # E[pdf_proj * cos / (cos / pi)] = pi * E[pdf_proj].
    tot = 0.0
    for u in rng.random((n, 2)):
        r, ph = np.sqrt(u[0]), 2 * np.pi * u[1]
        wo = (r * np.cos(ph), r * np.sin(ph), np.sqrt(max(0.0, 1 - u[0])))
        _, pdf, _ = oracle.bsdf(prim, ngi.PRIM_G, wi, wo=wo)
        tot += np.pi * pdf
    assert abs(acc / n - tot / n) < 0.03, (acc / n, tot / n)


def _hemisphere_pairs(rng, n):
    v = rng.normal(size=(n, 2, 3))
    v[..., 2] = np.abs(v[..., 2])
    return v / np.linalg.norm(v, axis=2, keepdims=True)


@pytest.mark.parametrize("alpha", [0.1, 0.4])
def test_glossy_brdf_is_reciprocal(alpha):
    """f(wi, wo) = f(wo, wi) for the Beckmann conductor (rt.hpp:1032-1047;
    both cosines in the denominator, F and G of the half vector)."""
    prim = _glossy(alpha)
    for wi, wo in _hemisphere_pairs(np.random.default_rng(2), 50):
        _, _, f1 = oracle.bsdf(prim, ngi.PRIM_G, wi, wo=wo)
        _, _, f2 = oracle.bsdf(prim, ngi.PRIM_G, wo, wo=wi)
        np.testing.assert_allclose(f1, f2, rtol=1e-12, atol=1e-200)   # far-off-peak values are subnormal


@pytest.mark.parametrize("alpha,theta_i", [(0.3, 0.0), (0.3, 50.0), (0.6, 30.0)])
def test_glossy_albedo_by_bsdf_and_cosine_sampling(alpha, theta_i):
    """The albedo (integral of f over projected solid angle) estimated two
    ways: with the BSDF's own sampler (weight f / pdf, rejected samples 0)
    and with cosine-weighted directions (weight pi f).  A pdf that is off by
    a cosine or a Jacobian, or a D whose normalisation differs between f and
    pdf, moves one estimate and not the other."""
    prim = _glossy(alpha)
    ti = np.radians(theta_i)
    wi = (np.sin(ti), 0.0, np.cos(ti))
    rng = np.random.default_rng(11)
    n = 20000
    a = np.zeros(n)
    for i, u in enumerate(rng.random((n, 3))):
        wo, pdf, fs = oracle.bsdf(prim, ngi.PRIM_G, wi, u=u)
        if wo[2] > 0 and pdf > 0:
            a[i] = fs[0] / pdf
    b = np.zeros(n)
    for i, u in enumerate(rng.random((n, 2))):
        r, ph = np.sqrt(u[0]), 2 * np.pi * u[1]
        _, _, fs = oracle.bsdf(prim, ngi.PRIM_G, wi, wo=(r * np.cos(ph), r * np.sin(ph), np.sqrt(max(0.0, 1 - u[0]))))
        b[i] = np.pi * fs[0]
    se = np.sqrt(a.var() / n + b.var() / n)
    assert 0.5 < a.mean() < 1.0
    assert abs(a.mean() - b.mean()) < 4 * se, (a.mean(), b.mean(), se)


def test_lambert_sampling_is_cosine_weighted():
    rng = np.random.default_rng(3)
    prim = ngi.make_primitive(type=ngi.PRIM_D, D_R=(0.5, 0.5, 0.5))
    zs = []
    for u in rng.random((20000, 3)):
        wo, pdf, fs = oracle.bsdf(prim, ngi.PRIM_D, (0.3, 0.1, 0.9), u=u)
        assert pdf == pytest.approx(1 / np.pi)
        zs.append(wo[2])
        np.testing.assert_allclose(fs, 0.5 / np.pi)
    assert np.mean(zs) == pytest.approx(2.0 / 3.0, abs=0.01)   # E[cos] under cos/pi


def test_fresnel_branch_frequency():
    prim = ngi.make_primitive(type=ngi.PRIM_S, spec_type=ngi.SPEC_FRESNEL, S_R=(1, 1, 1), S_eta1=1.0, S_eta2=1.5)
    wi = (np.sin(1.0), 0, np.cos(1.0))
    rng = np.random.default_rng(5)
    refl, Fr = 0, None
    for u in rng.random((4000, 3)):
        wo, pdf, _ = oracle.bsdf(prim, ngi.PRIM_S, wi, u=u)
        if wo[2] > 0:
            refl += 1
            Fr = pdf
    assert abs(refl / 4000 - Fr) < 0.03


# ---------------------------------------------------------------------------
# Intersection
# ---------------------------------------------------------------------------
def test_oracle_bvh_matches_brute_force():
    s = ngi.Scene.load_yaml("tests/golden/runc_data/scene.yml", 1.0)
    rng = np.random.default_rng(11)
    o = rng.uniform(-3, 3, (2000, 3)) + np.array([0, 1, 0])
    d = rng.normal(size=(2000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = ngi.make_rays(o, d)
    a = oracle.trace(s, rays, use_bvh=False)
    b = oracle.trace(s, rays, use_bvh=True)
    assert np.array_equal(a["tri"], b["tri"])
    hit = a["tri"] != ngi.MISS
    assert hit.sum() > 200
    np.testing.assert_allclose(a["t"][hit], b["t"][hit], rtol=1e-6)


def test_oracle_hit_distance_closed_form():
    B = scenes.SceneBuilder()
    p = B.prim(type=ngi.PRIM_D, D_R=(1, 1, 1))
    B.rect(p, (0, 0, -2.0), (0, 0, 1), (1, 0, 0), 1.0, 1.0)
    B.camera((0, 0, 0), (0, 0, -1))
    s = B.scene()
    rays = ngi.make_rays([[0.1, 0.2, 0.0], [0.1, 0.2, 0.0], [5, 5, 0]], [[0, 0, -1], [0, 0, 1], [0, 0, -1]])
    h = oracle.trace(s, rays, use_bvh=False)
    assert h["tri"][0] != ngi.MISS and h["t"][0] == pytest.approx(2.0)
    assert h["tri"][1] == ngi.MISS and h["tri"][2] == ngi.MISS
    # tmax excludes hits at or beyond it (Scene::Visible range, rt.hpp:2258).
    rays["tmax"][0] = 2.0
    assert oracle.trace(s, rays[:1], use_bvh=False)["tri"][0] == ngi.MISS


# ---------------------------------------------------------------------------
# Light tracing: ProcessSample_LT / _LTDirect (src/nanogi.cpp:804-1131)
# ---------------------------------------------------------------------------
def _lt_params(renderer, w, h, n, m, seed=1, rr_depth=-1):
    p = _params(renderer, w, h, 1, m, rr_depth=rr_depth, seed=seed, sampling=ngi.SAMPLING_REFERENCE_UNIFORM)
    p.num_samples = n
    return p


def test_lt_with_a_pinhole_adds_nothing():
    """A light path that hits the sensor adds We * EvaluatePosition(geom,
    false), which is zero for a pinhole (rt.hpp:590-593): lt renders black
    while still tracing its paths."""
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    film, counts = oracle.render_paths(s, _lt_params("lt", 8, 8, 20000, 6), 0, 20000)
    assert np.all(film == 0) and counts[0] > 20000 and counts[1] == 0


def test_ltdirect_directly_viewed_light_closed_form():
    """-m 2: only the light vertex connects to the camera.  The footprint of
    the 2h x 2h light gets radiance Le scaled by the light's area, because the
    reference divides by the light's position pdf where the sensor's belongs
    (src/nanogi.cpp:1045; the pinhole's is 1)."""
    Le, half = np.array([3.0, 2.0, 1.0]), 0.3
    s = scenes.light_in_view(tuple(Le), half)
    area = (2 * half) ** 2
    frac = (half / (3.0 * np.tan(np.radians(20.0)))) ** 2
    _stat_close(lambda k: oracle.render(s, _lt_params("ltdirect", 16, 16, 20000, 2, seed=k)).reshape(-1, 3).mean(0),
                area * Le * frac)
    # Every splat lands inside the footprint, and one path carries the
    # contribution n_L A Le G We A of its light point exactly.
    img = oracle.render(s, _lt_params("ltdirect", 16, 16, 5000, 2, seed=3))
    assert img[0, 0].sum() == 0 and img[8, 8].sum() > 0


def test_ltdirect_plane_under_light_closed_form():
    """-m 3: light -> diffuse plane -> camera.  Pixel expectation
    A * rho * Le * F(x) with F the point-to-rectangle form factor."""
    a, b, h, Le, rho = 0.5, 0.3, 1.0, np.array([4.0, 3.0, 2.0]), np.array([0.8, 0.5, 0.3])
    s = scenes.plane_under_light_wide(a, b, h, tuple(Le), tuple(rho), fov=30.0)
    W = 8
    # Plane points seen through each pixel centre (camera at h/2 looking down, up = -z).
    t = np.tan(np.radians(15.0))
    c = (np.arange(W) + 0.5) / W * 2 - 1
    X, Z = np.meshgrid(c * t * 0.5 * h, c * t * 0.5 * h)
    F = np.mean([scenes.rect_form_factor_at(x, z, a, b, h) for x, z in zip(X.ravel(), Z.ravel())])
    expect = (2 * a) * (2 * b) * rho * Le * F
    _, se = _stat_close(lambda k: oracle.render(s, _lt_params("ltdirect", W, W, 400000, 3, seed=k)).reshape(-1, 3).mean(0),
                        expect)
    assert np.all(se < 0.03 * expect)


@pytest.mark.parametrize("scene", ["tilted", "room"])
def test_ltdirect_equals_area_times_ptdirect(scene):
    """Light tracing with LE shading-normal corrections estimates the same
    image as path tracing with EL ones; with one light, ltdirect's images
    are that image times the light's area.  4x4 block means, -m 4."""
    if scene == "tilted":
        s, area = scenes.tilted_normals_box(), 0.16
    else:
        s, area = ngi.Scene.generate(ngi.GEN_BATCH_ROOM, 5, 1.0, 1500), 0.09
    W = 8

    def blocks(img):
        return img.reshape(4, 2, 4, 2, 3).mean(axis=(1, 3)).ravel()

    pt = [blocks(oracle.render(s, _params("ptdirect", W, W, 2048, 4, seed=k))) for k in range(1, 5)]
    lt = [blocks(oracle.render(s, _lt_params("ltdirect", W, W, 600000, 4, seed=k))) / area for k in range(1, 5)]
    d = np.mean(pt, 0) - np.mean(lt, 0)
    se = np.hypot(np.std(pt, 0, ddof=1), np.std(lt, 0, ddof=1)) / 2
    assert np.mean(pt) > 0
    assert np.all(np.abs(d) <= 4.5 * se + 1e-3 * np.mean(pt)), (d, se)


def test_point_light_closed_form_both_directions():
    """Point light of intensity Le at height h over a diffuse plane:
    radiance rho/pi Le h / d^3.  ptdirect and ltdirect agree (a point light's
    forced position pdf is 1, so ltdirect has no area factor)."""
    h, Le, rho, W, fov = 1.0, np.array([2.0, 1.5, 1.0]), np.array([0.8, 0.5, 0.3]), 8, 20.0
    s = scenes.plane_with_light("point", h, tuple(Le), tuple(rho), fov)
    X, Z = scenes.pixel_plane_points(W, h, fov)
    expect = rho / np.pi * Le * np.mean(h / (X * X + Z * Z + h * h) ** 1.5)
    pt = oracle.render(s, _params("ptdirect", W, W, 64, 3))
    np.testing.assert_allclose(pt.reshape(-1, 3).mean(0), expect, rtol=2e-3)
    _stat_close(lambda k: oracle.render(s, _lt_params("ltdirect", W, W, 300000, 3, seed=k)).reshape(-1, 3).mean(0), expect)


def test_directional_light_closed_form():
    """Directional light straight down onto a diffuse plane.  ptdirect gets
    nothing from it (EvaluateDirection(.., false) is zero, rt.hpp:934-937).
    ltdirect starts on the reference's disk of radius r (rt.hpp:2067-2073,
    position pdf 1 / (2 pi r^2)) and divides by that pdf again at the sensor
    (src/nanogi.cpp:1045): radiance 2 pi r^2 * 2 * rho/pi * Le, the factor 2
    being the disk's true area pi r^2 against the 2 pi r^2 the pdf uses."""
    Le, rho, W, half = np.array([2.0, 1.5, 1.0]), np.array([0.8, 0.5, 0.3]), 8, 20.0
    s = scenes.plane_with_light("directional", 1.0, tuple(Le), tuple(rho), 20.0, half)
    pt = oracle.render(s, _params("ptdirect", W, W, 4, 3))
    assert np.all(pt == 0)
    r = np.sqrt(2) * half * 1.01   # |centre - bound.max| * 1.01, bound = plane
    expect = 2 * np.pi * r * r * 2 * rho / np.pi * Le
    _stat_close(lambda k: oracle.render(s, _lt_params("ltdirect", W, W, 400000, 3, seed=k)).reshape(-1, 3).mean(0), expect)


# ---------------------------------------------------------------------------
# Bidirectional path tracing: ProcessSample_BDPT (src/nanogi.cpp:1133-1189,
# include/nanogi/bdpt.hpp).  Unlike ltdirect it has no area factor: every
# strategy's contribution is divided by the pdf of that strategy, and the
# power-heuristic weights sum to one over the strategies that can make a path.
# ---------------------------------------------------------------------------
def test_bdpt_furnace():
    Le, rho, m = (1.0, 0.5, 0.25), (0.5, 0.6, 0.7), 4
    s = scenes.furnace(Le, rho)
    _stat_close(lambda k: oracle.render(s, _params("bdpt", 4, 4, 256, m, rr_depth=-1, seed=k)).reshape(-1, 3).mean(0),
                scenes.furnace_radiance(Le, rho, m))


def test_bdpt_plane_under_rectangular_light():
    a, b, h, Le, rho = 0.5, 0.3, 1.0, np.array([4.0, 3.0, 2.0]), np.array([0.8, 0.5, 0.3])
    s = scenes.plane_under_light(a, b, h, tuple(Le), tuple(rho))
    _, se = _stat_close(lambda k: oracle.render(s, _params("bdpt", 2, 2, 4096, 3, seed=k)).reshape(-1, 3).mean(0),
                        rho * Le * scenes.rect_form_factor(a, b, h))
    assert np.all(se < 0.03 * rho * Le * scenes.rect_form_factor(a, b, h))


def test_bdpt_point_and_directional_lights():
    h, Le, rho, W, fov = 1.0, np.array([2.0, 1.5, 1.0]), np.array([0.8, 0.5, 0.3]), 8, 20.0
    s = scenes.plane_with_light("point", h, tuple(Le), tuple(rho), fov)
    X, Z = scenes.pixel_plane_points(W, h, fov, sub=32)
    expect = rho / np.pi * Le * np.mean(h / (X * X + Z * Z + h * h) ** 1.5)
    _stat_close(lambda k: oracle.render(s, _params("bdpt", W, W, 64, 3, seed=k)).reshape(-1, 3).mean(0), expect)
    # Directional: the light subpath starts on the reference's disk, whose
    # position pdf is 1 / (2 pi r^2) for an area of pi r^2 (rt.hpp:2072).  A
    # 2 x 2 plane keeps the disk small enough for light subpaths to reach the view.
    s = scenes.plane_with_light("directional", 1.0, tuple(Le), tuple(rho), 60.0, half=1.0)
    _stat_close(lambda k: oracle.render(s, _params("bdpt", W, W, 1024, 3, seed=k)).reshape(-1, 3).mean(0), 2 * rho / np.pi * Le)


@pytest.mark.parametrize("scene", ["cornell", "tilted", "room"])
def test_bdpt_equals_ptdirect(scene):
    """Same expectation as path tracing with NEE, block by block (-m 5, RR from the first bounce)."""
    s = {"cornell": lambda: ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0), "tilted": scenes.tilted_normals_box,
         "room": lambda: ngi.Scene.generate(ngi.GEN_BATCH_ROOM, 5, 1.0, 1500)}[scene]()
    W = 8

    def blocks(img):
        return img.reshape(4, 2, 4, 2, 3).mean(axis=(1, 3)).ravel()

    pt = [blocks(oracle.render(s, _params("ptdirect", W, W, 2048, 5, rr_depth=1, seed=k))) for k in range(1, 5)]
    bd = [blocks(oracle.render(s, _params("bdpt", W, W, 512, 5, rr_depth=1, seed=k))) for k in range(1, 5)]
    d = np.mean(pt, 0) - np.mean(bd, 0)
    se = np.hypot(np.std(pt, 0, ddof=1), np.std(bd, 0, ddof=1)) / 2
    assert np.mean(pt) > 0
    assert np.all(np.abs(d) <= 4.5 * se + 1e-3 * np.mean(pt)), (d, se)


# ---------------------------------------------------------------------------
# Pinhole sensor and light sampling
# ---------------------------------------------------------------------------
def test_pinhole_raster_round_trip_and_importance_over_pdf():
    """RasterPosition(SampleDirection(u)) = u, We / pdf = 1 (rt.hpp:733-740,
    955-978, 1189-1212, 1348-1378); the vertical field of view is `fov`, the
    horizontal one follows the aspect W/H, and raster row 0 is the bottom."""
    B = scenes.SceneBuilder()
    p = B.prim(type=ngi.PRIM_D, D_R=(0.5, 0.5, 0.5))
    B.tri(p, (-1, -1, 0), (1, -1, 0), (0, 1, 0))
    eye, center, fov, aspect = np.array([1.0, 2.0, 3.0]), np.array([0.0, 0.2, 0.0]), 50.0, 1.5
    B.camera(tuple(eye), tuple(center), up=(0, 1, 0), fov=fov)
    ls = oracle.LoadedScene(B.scene(aspect=aspect))
    rng = np.random.default_rng(4)
    for u0, u1 in rng.random((200, 2)):
        wo, (rx, ry), We, pdf = ls.camera(u0, u1)
        assert np.linalg.norm(wo) == pytest.approx(1.0, abs=1e-12)
        assert rx == pytest.approx(u0, abs=1e-12) and ry == pytest.approx(u1, abs=1e-12)
        assert We == pytest.approx(pdf, rel=1e-12) and pdf > 0
    axis = (center - eye) / np.linalg.norm(center - eye)
    wo, _, _, _ = ls.camera(0.5, 0.5)
    np.testing.assert_allclose(wo, axis, atol=1e-12)
    up, _, _, _ = ls.camera(0.5, 1.0)
    right, _, _, _ = ls.camera(1.0, 0.5)
    assert np.degrees(np.arccos(np.dot(up, axis))) == pytest.approx(fov / 2, abs=1e-9)
    assert np.arctan(np.tan(np.arccos(np.dot(right, axis)))) == pytest.approx(
        np.arctan(aspect * np.tan(np.radians(fov / 2))), abs=1e-12)
    bottom, _, _, _ = ls.camera(0.5, 0.0)
    assert bottom[1] < axis[1]   # row 0 looks below the axis (y up)


def test_area_light_samples_triangles_by_area_and_uniformly():
    """SampleTriangleMesh (rt.hpp:488-540): a triangle is chosen with
    probability proportional to its area by the CDF search and the point is
    uniform in it; the position pdf is 1 / (total area) (rt.hpp:649-652) and
    the emitter is chosen uniformly among the lights (rt.hpp:2321-2352)."""
    B = scenes.SceneBuilder()
    li = B.prim(type=ngi.PRIM_L, light_type=ngi.LIGHT_AREA, Le=(1, 1, 1))
    B.tri(li, (0, 0, 0), (1, 0, 0), (0, 2, 0))          # area 1
    B.tri(li, (5, 0, 0), (8, 0, 0), (5, 2, 0))          # area 3
    other = B.prim(type=ngi.PRIM_L, light_type=ngi.LIGHT_AREA, Le=(1, 1, 1))
    B.tri(other, (0, 0, 5), (1, 0, 5), (0, 1, 5))
    B.camera((0, 0, 10), (0, 0, 0))
    ls = oracle.LoadedScene(B.scene())
    rng = np.random.default_rng(9)
    n = 20000
    P = np.zeros((n, 3))
    for i, (ux, uy) in enumerate(rng.random((n, 2))):
        P[i], gn, pdf_pos, pdf_sel = ls.light(0, ux, uy)
        assert pdf_pos == pytest.approx(0.25, rel=1e-12) and pdf_sel == 0.5
        np.testing.assert_allclose(gn, (0, 0, 1), atol=1e-12)
    in_a = P[:, 0] < 3.0
    assert abs(in_a.mean() - 0.25) < 4 * np.sqrt(0.25 * 0.75 / n)
    for sel, tri in ((in_a, np.array([[0, 0, 0], [1, 0, 0], [0, 2, 0]])), (~in_a, np.array([[5, 0, 0], [8, 0, 0], [5, 2, 0]]))):
        q = P[sel]
        # uniform in the triangle: the mean is its centroid, and every point is inside it
        se = q.std(axis=0) / np.sqrt(len(q)) + 1e-15
        assert np.all(np.abs(q.mean(axis=0) - tri.mean(axis=0)) < 4 * se)
        lo, hi = tri.min(axis=0), tri.max(axis=0)
        assert np.all(q >= lo - 1e-12) and np.all(q <= hi + 1e-12)
