"""Pins the oracle's Fresnel terms, refraction and radiance scaling.

Covers the conductor Fresnel term of the glossy BSDF (rt.hpp:1433-1442), the
dielectric Fresnel term (rt.hpp:1450-1466), the specular sampling branches
(rt.hpp:804-903), their evaluation with the eta^2 radiance scaling of
eye-to-light transport (rt.hpp:1055-1137) and their pdfs (rt.hpp:1259-1328).

Each check comes from optics rather than from the oracle's own formulas:
complex-amplitude Fresnel equations evaluated with numpy, Stokes
reciprocity, the Brewster angle, total internal reflection, Snell's law, the
radiance theorem (L / n^2 is invariant across an interface), and scenes whose
images follow from those laws exactly.
"""
import numpy as np
import pytest

import nanogi_amd as ngi
import oracle
from tests import scenes


# ---------------------------------------------------------------------------
# Independent optics (numpy complex arithmetic)
# ---------------------------------------------------------------------------
def fresnel_exact(n1, n2, cos_i):
    """Unpolarised reflectance of light arriving from index n1 at an
    interface with (complex) index n2: complex Snell's law gives cos(theta_t),
    then the s and p amplitude ratios.  Total internal reflection falls out
    as |r| = 1 when cos(theta_t) is imaginary."""
    n1 = complex(n1)
    n2 = complex(n2)
    ci = np.asarray(cos_i, dtype=np.complex128)
    st2 = (n1 / n2) ** 2 * (1 - ci * ci)
    ct = np.sqrt(1 - st2)
    ct = np.where(ct.imag < 0, -ct, ct)    # the evanescent branch (decaying wave)
    rs = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct)
    rp = (n2 * ci - n1 * ct) / (n2 * ci + n1 * ct)
    return 0.5 * (np.abs(rs) ** 2 + np.abs(rp) ** 2)


def fresnel_normal_transmission(n, cos_i):
    """Conductor reflectance from air with the transmitted wave taken along
    the normal (cos(theta_t) = 1), the |n| >> 1 limit of the Fresnel
    equations: r_s = (cos_i - n) / (cos_i + n), r_p = (n cos_i - 1) / (n cos_i + 1)."""
    c = np.asarray(cos_i, dtype=np.complex128)
    rs = (c - n) / (c + n)
    rp = (n * c - 1) / (n * c + 1)
    return 0.5 * (np.abs(rs) ** 2 + np.abs(rp) ** 2)


def _dielectric(spec_type, eta1, eta2, R=(1.0, 1.0, 1.0)):
    return ngi.make_primitive(type=ngi.PRIM_S, spec_type=spec_type, S_R=R, S_eta1=eta1, S_eta2=eta2)


def _oracle_dielectric_F(eta1, eta2, theta, inside):
    """Fr of the oracle's Fresnel surface: the pdf of the mirror direction
    (rt.hpp:1309-1320) for a ray arriving from outside (normal side) or inside."""
    prim = _dielectric(ngi.SPEC_FRESNEL, eta1, eta2)
    z = -np.cos(theta) if inside else np.cos(theta)
    wi = (np.sin(theta), 0.0, z)
    wo = (-np.sin(theta), 0.0, z)
    _, pdf, fs = oracle.bsdf(prim, ngi.PRIM_S, wi, wo=wo)
    np.testing.assert_allclose(fs, pdf, rtol=1e-14)     # reflection: f = R Fr with R = 1
    return pdf


# ---------------------------------------------------------------------------
# Dielectric Fresnel term (rt.hpp:1450-1466)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n1,n2", [(1.0, 1.5), (1.0, 2.4), (1.33, 1.0), (1.5, 1.0)])
def test_dielectric_fresnel_matches_complex_amplitudes(n1, n2):
    """Across incidence angles, including beyond the critical angle when
    n1 > n2.  The oracle's Eta1 is the index on the normal side; a ray from
    inside sees the indices swapped."""
    thetas = np.radians(np.linspace(0.0, 89.5, 90))
    for th in thetas:
        got_out = _oracle_dielectric_F(n1, n2, th, inside=False)
        np.testing.assert_allclose(got_out, fresnel_exact(n1, n2, np.cos(th)), rtol=1e-12, atol=1e-15)
        got_in = _oracle_dielectric_F(n1, n2, th, inside=True)
        np.testing.assert_allclose(got_in, fresnel_exact(n2, n1, np.cos(th)), rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("n1,n2", [(1.0, 1.5), (1.0, 1.33), (1.2, 2.0)])
def test_dielectric_fresnel_stokes_reciprocity(n1, n2):
    """F(theta_i; n1 -> n2) = F(theta_t; n2 -> n1) with n1 sin(theta_i) =
    n2 sin(theta_t): the interface reflects the same fraction both ways."""
    for ti in np.radians(np.linspace(1.0, 88.0, 30)):
        tt = np.arcsin(n1 / n2 * np.sin(ti))
        a = _oracle_dielectric_F(n1, n2, ti, inside=False)
        b = _oracle_dielectric_F(n1, n2, tt, inside=True)
        np.testing.assert_allclose(a, b, rtol=1e-10)


@pytest.mark.parametrize("n1,n2", [(1.0, 1.5), (1.0, 2.4)])
def test_dielectric_fresnel_brewster_angle(n1, n2):
    """At tan(theta_B) = n2 / n1 the p-polarised reflectance vanishes and the
    s part is ((n1^2 - n2^2) / (n1^2 + n2^2))^2, so F is half of that."""
    tb = np.arctan(n2 / n1)
    expect = 0.5 * ((n1 * n1 - n2 * n2) / (n1 * n1 + n2 * n2)) ** 2
    np.testing.assert_allclose(_oracle_dielectric_F(n1, n2, tb, inside=False), expect, rtol=1e-12)
    # from inside, the Brewster angle is arctan(n1 / n2) with the same value
    np.testing.assert_allclose(_oracle_dielectric_F(n1, n2, np.arctan(n1 / n2), inside=True), expect, rtol=1e-12)


def test_dielectric_total_internal_reflection_and_grazing():
    """From glass (1.5) into air, F = 1 beyond the critical angle asin(1/1.5)
    and below 1 just inside it; from air F -> 1 at grazing incidence.  Under
    TIR the Fresnel sampler always reflects and the refraction sampler
    reflects too (rt.hpp:843-849, 870-887)."""
    tc = np.arcsin(1.0 / 1.5)
    for th in (tc + 1e-6, tc + 0.1, np.radians(80.0)):
        assert _oracle_dielectric_F(1.0, 1.5, th, inside=True) == 1.0
    assert _oracle_dielectric_F(1.0, 1.5, tc - 0.01, inside=True) < 0.5
    assert _oracle_dielectric_F(1.0, 1.5, np.radians(89.999), inside=False) > 0.999
    assert _oracle_dielectric_F(1.0, 1.5, np.radians(80.0), inside=False) < _oracle_dielectric_F(
        1.0, 1.5, np.radians(89.0), inside=False)
    th = tc + 0.05
    wi = (np.sin(th), 0.0, -np.cos(th))
    for spec in (ngi.SPEC_FRESNEL, ngi.SPEC_REFRACTION):
        for uc in (0.0, 0.5, 0.999999):
            wo, pdf, fs = oracle.bsdf(_dielectric(spec, 1.0, 1.5), ngi.PRIM_S, wi, u=(0.3, 0.6, uc))
            np.testing.assert_allclose(wo, (-wi[0], 0.0, wi[2]), atol=1e-15)   # mirror, same (inner) side
            assert pdf == 1.0


# ---------------------------------------------------------------------------
# Conductor Fresnel term of the glossy BSDF (rt.hpp:1433-1442)
# ---------------------------------------------------------------------------
METALS = {   # (eta, k) per channel
    "gold": ((0.143, 0.374, 1.442), (3.983, 2.385, 1.603)),
    "aluminium": ((1.657, 0.880, 0.521), (9.224, 6.270, 4.837)),
    "copper": ((0.200, 0.924, 1.102), (3.912, 2.452, 2.142)),
}


def _oracle_conductor_F(eta, k, theta, alpha=0.3):
    """F(cos theta) of the glossy BSDF, read from f at wi and wo mirrored about
    the normal: H = n, D(n) = 1 / (pi alpha^2), G = 1, so
    f = R F D G / (4 cos_i cos_o) = F / (4 pi alpha^2 cos^2 theta)."""
    prim = ngi.make_primitive(type=ngi.PRIM_G, G_R=(1, 1, 1), G_eta=eta, G_k=k, G_roughness=alpha)
    wi = (np.sin(theta), 0.0, np.cos(theta))
    wo = (-np.sin(theta), 0.0, np.cos(theta))
    _, _, fs = oracle.bsdf(prim, ngi.PRIM_G, wi, wo=wo)
    return fs * 4.0 * np.pi * alpha * alpha * np.cos(theta) ** 2


@pytest.mark.parametrize("metal", sorted(METALS))
def test_conductor_fresnel_is_the_normal_transmission_limit(metal):
    """The reference's conductor term is the Fresnel equations for n = eta +
    i k with the transmitted wave along the normal; per channel and angle it
    equals the complex-amplitude form of that limit."""
    eta, k = METALS[metal]
    for th in np.radians(np.linspace(0.0, 88.0, 45)):
        got = _oracle_conductor_F(eta, k, th)
        expect = [fresnel_normal_transmission(complex(e, kk), np.cos(th)) for e, kk in zip(eta, k)]
        np.testing.assert_allclose(got, np.real(expect), rtol=1e-12)


@pytest.mark.parametrize("metal", sorted(METALS))
def test_conductor_fresnel_approaches_exact_fresnel(metal):
    """Against the exact Fresnel equations (complex Snell's law): within
    0.01 for real metals, with the gap shrinking like 1 / |n| as |n| grows
    (the limit the approximation takes), and F -> 1 at grazing incidence."""
    eta, k = METALS[metal]
    thetas = np.radians(np.linspace(0.0, 89.0, 60))
    gaps = []
    for scale in (1.0, 4.0):
        gap = 0.0
        for th in thetas:
            got = _oracle_conductor_F(tuple(scale * e for e in eta), tuple(scale * x for x in k), th)
            ex = [fresnel_exact(1.0, complex(scale * e, scale * kk), np.cos(th)) for e, kk in zip(eta, k)]
            gap = max(gap, float(np.max(np.abs(got - np.real(ex)))))
        gaps.append(gap)
    assert gaps[0] < 0.01, gaps
    assert gaps[1] < gaps[0] / 2.5, gaps
    assert np.all(_oracle_conductor_F(eta, k, np.radians(89.99)) > 0.99)


# ---------------------------------------------------------------------------
# Specular sampling: reflected and refracted directions (rt.hpp:804-903)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("inside", [False, True])
@pytest.mark.parametrize("spec", [ngi.SPEC_REFRACTION, ngi.SPEC_FRESNEL])
def test_refracted_direction_obeys_snell(spec, inside):
    """Unit length, the opposite side of the surface, n_i sin(theta_i) =
    n_t sin(theta_t) and the tangential component reversed with respect to
    wi (the ray goes on through the surface)."""
    n_out, n_in = 1.0, 1.6
    n_i, n_t = (n_in, n_out) if inside else (n_out, n_in)
    prim = _dielectric(spec, n_out, n_in)
    rng = np.random.default_rng(21)
    crit = np.arcsin(n_t / n_i) if n_i > n_t else np.pi / 2
    for _ in range(200):
        th = rng.uniform(0.0, min(crit - 1e-3, np.radians(89.0)))
        ph = rng.uniform(0.0, 2 * np.pi)
        s = np.sin(th)
        wi = np.array([s * np.cos(ph), s * np.sin(ph), -np.cos(th) if inside else np.cos(th)])
        wo, pdf, fs = oracle.bsdf(prim, ngi.PRIM_S, wi, u=(0.1, 0.2, 1.0))   # uc = 1: transmission
        assert abs(np.linalg.norm(wo) - 1.0) < 1e-12
        assert wo[2] * wi[2] < 0
        sin_t = np.hypot(wo[0], wo[1])
        np.testing.assert_allclose(n_i * np.hypot(wi[0], wi[1]), n_t * sin_t, rtol=1e-12)
        # tangential parts antiparallel: wi points back to where the light goes
        np.testing.assert_allclose(np.cross(wi[:2].tolist() + [0], wo[:2].tolist() + [0])[2], 0.0, atol=1e-12)
        assert np.dot(wi[:2], wo[:2]) <= 0


def test_mirror_direction_and_weight():
    """Reflection type: the mirror image of wi, weight R, nothing from below."""
    prim = _dielectric(ngi.SPEC_REFLECTION, 1.0, 1.0, R=(0.9, 0.5, 0.2))
    wi = np.array([0.3, -0.4, np.sqrt(1 - 0.25)])
    wo, pdf, fs = oracle.bsdf(prim, ngi.PRIM_S, wi, u=(0.5, 0.5, 0.5))
    np.testing.assert_allclose(wo, (-0.3, 0.4, wi[2]), atol=1e-15)
    assert pdf == 1.0
    np.testing.assert_allclose(fs, (0.9, 0.5, 0.2), rtol=1e-15)
    wo, pdf, fs = oracle.bsdf(prim, ngi.PRIM_S, -wi, u=(0.5, 0.5, 0.5))
    assert not np.any(wo) and not np.any(fs)


# ---------------------------------------------------------------------------
# Radiance scaling across an interface (rt.hpp:1086-1098, 1120-1131)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("inside", [False, True])
def test_refraction_weight_is_the_radiance_theorem(inside):
    """Eye-to-light transport carries radiance, and L / n^2 is invariant
    across an interface: the transmitted sample's weight f / pdf is
    (n_wi / n_wo)^2, with n_wi the index on wi's side.  Light-to-eye
    transport carries importance, which is not scaled (Veach 1997, 5.2).
    The Fresnel surface's reflection weight is 1 (its pdf is Fr)."""
    n_out, n_in = 1.0, 1.5
    n_wi, n_wo = (n_in, n_out) if inside else (n_out, n_in)
    th = np.radians(20.0)
    wi = (np.sin(th), 0.0, -np.cos(th) if inside else np.cos(th))
    R = np.array([0.9, 0.6, 0.3])
    for spec in (ngi.SPEC_REFRACTION, ngi.SPEC_FRESNEL):
        prim = _dielectric(spec, n_out, n_in, R=tuple(R))
        wo, pdf, fs_el = oracle.bsdf(prim, ngi.PRIM_S, wi, u=(0.0, 0.0, 1.0))
        _, _, fs_le = oracle.bsdf(prim, ngi.PRIM_S, wi, wo=wo, le=True)
        assert wo[2] * wi[2] < 0
        np.testing.assert_allclose(fs_el / pdf, R * (n_wi / n_wo) ** 2, rtol=1e-12)
        np.testing.assert_allclose(fs_le / pdf, R, rtol=1e-12)
    prim = _dielectric(ngi.SPEC_FRESNEL, n_out, n_in, R=tuple(R))
    wo, pdf, fs = oracle.bsdf(prim, ngi.PRIM_S, wi, u=(0.0, 0.0, 0.0))   # uc = 0 <= Fr: reflection
    assert wo[2] * wi[2] > 0
    np.testing.assert_allclose(fs / pdf, R, rtol=1e-12)
    np.testing.assert_allclose(pdf, fresnel_exact(n_wi, n_wo, np.cos(th)), rtol=1e-12)


def test_specular_surfaces_have_no_nee_term():
    """A delta BSDF evaluated for a connection (forceDegenerated = false) is
    zero (rt.hpp:1057-1060, 1261-1264)."""
    for spec in (ngi.SPEC_REFLECTION, ngi.SPEC_REFRACTION, ngi.SPEC_FRESNEL):
        prim = _dielectric(spec, 1.0, 1.5)
        wi = (0.0, 0.6, 0.8)
        for wo in ((0.0, -0.6, 0.8), (0.0, -0.4, -np.sqrt(0.84))):
            _, pdf, fs = oracle.bsdf(prim, ngi.PRIM_S, wi, wo=wo, force_degenerated=False)
            assert not np.any(fs)


def _params(renderer, w, h, spp, m, seed=1, rr_depth=-1):
    return ngi.default_params(renderer=renderer, width=w, height=h, spp=spp, max_num_vertices=m,
                              rr_depth=rr_depth, seed=seed, sampling=ngi.SAMPLING_PER_PIXEL)


def test_camera_inside_glass_sees_eta_squared_le():
    """Camera inside a glass cube (n = 1.5) looking out at an emitting wall,
    -m 3: every camera ray leaves the glass at under 21 degrees (no total
    internal reflection) and arrives with weight (1.5 / 1)^2 on the wall,
    so every pixel is exactly 2.25 Le.  ptdirect cannot connect through the
    glass and sees nothing (Visible is blocked, delta BSDFs have no NEE)."""
    Le = np.array([2.0, 1.0, 0.5])
    s = scenes.glass_cube_in_light_box(ngi.SPEC_REFRACTION, 1.5, tuple(Le))
    img = oracle.render(s, _params("pt", 8, 8, 4, 3))
    np.testing.assert_allclose(img, np.broadcast_to(2.25 * Le, img.shape), rtol=1e-12)
    assert not np.any(oracle.render(s, _params("ptdirect", 8, 8, 4, 3)))


def test_camera_inside_glass_fresnel_transmission():
    """Fresnel glass, camera tilted so its narrow view meets the face at 30
    degrees: a path is transmitted with probability 1 - F(30 deg; 1.5 -> 1)
    and then carries 2.25 Le; a reflected path meets the glass again at
    -m 3 and carries nothing.  Every pixel's mean is (1 - F) 2.25 Le."""
    Le = np.array([2.0, 1.0, 0.5])
    th = np.radians(30.0)
    s = scenes.glass_cube_in_light_box(ngi.SPEC_FRESNEL, 1.5, tuple(Le), look=(np.sin(th), 0.0, -np.cos(th)), fov=0.2)
    F = float(fresnel_exact(1.5, 1.0, np.cos(th)))
    img = oracle.render(s, _params("pt", 4, 4, 8192, 3))
    frac = img[..., 0] / (2.25 * Le[0])
    # each path carries 0 or 2.25 Le exactly
    np.testing.assert_allclose(img / (2.25 * Le), np.repeat(frac[..., None], 3, axis=2), rtol=1e-12)
    n = 4 * 4 * 8192
    se = np.sqrt(F * (1 - F) / n)
    assert abs(frac.mean() - (1 - F)) < 4.5 * se, (frac.mean(), 1 - F, se)


def test_slab_furnace_carries_exactly_le():
    """Closed emitting box around a Fresnel glass slab: entering and leaving
    scale by (1/1.5)^2 and 1.5^2 and reflections weigh 1, so each pt path
    carries exactly Le (or 0 if -m cuts it off inside the slab).  Every
    pixel is therefore Le times a whole number of paths over spp."""
    Le = np.array([1.0, 0.7, 0.4])
    spp = 64
    s = scenes.slab_furnace(1.5, tuple(Le))
    img = oracle.render(s, _params("pt", 8, 8, spp, 12))
    k = img / Le * spp
    np.testing.assert_allclose(k, np.round(k), atol=1e-9)
    assert k.mean() > 0.99 * spp


def test_glass_room_path_and_light_tracing_agree():
    """Light reaches the floor of the glass room only through the glass.
    Path tracing scales radiance by eta^2 when leaving the glass, light
    tracing scales nothing; both estimate the same image (ltdirect times the
    light's area, see the light-tracing tests), block by block.  eta in place
    of eta^2, or the scaling on the wrong transport direction, splits them."""
    s = scenes.glass_room(ngi.SPEC_REFRACTION)
    area, W = 0.16, 8

    def blocks(img):
        return img.reshape(2, 4, 2, 4, 3).mean(axis=(1, 3)).ravel()

    pt = [blocks(oracle.render(s, _params("pt", W, W, 4096, 4, seed=k))) for k in range(1, 5)]
    lt = []
    for k in range(1, 5):
        p = _params("ltdirect", W, W, 1, 4, seed=k)
        p.sampling = ngi.SAMPLING_REFERENCE_UNIFORM
        p.num_samples = 600000
        lt.append(blocks(oracle.render(s, p)) / area)
    d = np.mean(pt, 0) - np.mean(lt, 0)
    se = np.hypot(np.std(pt, 0, ddof=1), np.std(lt, 0, ddof=1)) / 2
    assert np.mean(pt) > 0.01
    assert np.all(np.abs(d) <= 4.5 * se + 1e-3 * np.mean(pt)), (d, se, np.mean(pt, 0), np.mean(lt, 0))
