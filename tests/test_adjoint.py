"""Adjoint (path-replay gradients) against finite differences and the FP64 oracle.

Path sampling does not depend on reflectance or emitter radiance, so with
common random numbers every path contribution is a polynomial in the
parameters and central differences (h = 1e-3) of L = sum dL/dI . I agree
with the replayed gradient up to O(h^2).
"""
import numpy as np
import pytest

import nanogi_amd as ngi
import oracle


def _params(renderer, w=12, h=12, spp=16, m=7, rr=3, seed=5):
    return ngi.default_params(renderer=renderer, width=w, height=h, sampling=ngi.SAMPLING_PER_PIXEL, spp=spp,
                              max_num_vertices=m, rr_depth=rr, seed=seed)


def _loss(scene, p, g, render):
    return float(np.sum(g * render(scene, p)))


def _fd(scene, p, g, render, which, prim, ch, h=1e-3):
    d = scene.desc()["primitives"]
    refl = np.where((d["type"] & ngi.PRIM_D)[:, None] > 0, d["D_R"], d["G_R"]).copy()
    Le = d["Le"].copy()
    base = (refl.copy(), Le.copy())
    arr = refl if which == "R" else Le
    x0 = arr[prim, ch]
    arr[prim, ch] = x0 + h
    scene.set_params(refl, Le)
    lp = _loss(scene, p, g, render)
    arr[prim, ch] = x0 - h
    scene.set_params(refl, Le)
    lm = _loss(scene, p, g, render)
    scene.set_params(*base)
    return (lp - lm) / (2 * h)


@pytest.mark.parametrize("renderer", ["pt", "ptdirect"])
def test_oracle_adjoint_matches_finite_differences(renderer):
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    p = _params(renderer)
    g = np.random.default_rng(1).normal(size=(p.height, p.width, 3))
    gR, gL = oracle.render_adjoint(s, p, g)
    prims = s.desc()["primitives"]
    light = int(np.nonzero(prims["type"] & ngi.PRIM_L)[0][0])
    for which, prim, ch in [("R", 3, 0), ("R", 4, 1), ("R", 0, 2), ("R", 5, 0), ("Le", light, 0), ("Le", light, 2)]:
        fd = _fd(s, p, g, oracle.render, which, prim, ch)
        an = (gR if which == "R" else gL)[prim, ch]
        assert abs(an - fd) <= 1e-6 * max(1.0, abs(fd)), (which, prim, ch, an, fd)
    # Only parameters that exist get gradients.
    assert np.all(gL[(prims["type"] & ngi.PRIM_L) == 0] == 0)


def test_oracle_adjoint_glossy_and_reference_sampling():
    s = ngi.Scene.generate(ngi.GEN_SPHERE_FIELD, 2, 1.0, 3)
    p = _params("ptdirect", 10, 10, 8, 5, 2)
    p.sampling = ngi.SAMPLING_REFERENCE_UNIFORM
    p.num_samples = 1000
    g = np.random.default_rng(2).normal(size=(10, 10, 3))
    gR, gL = oracle.render_adjoint(s, p, g)
    prims = s.desc()["primitives"]
    glossy = np.nonzero(prims["type"] == ngi.PRIM_G)[0]
    touched = glossy[np.abs(gR[glossy]).sum(1) > 0]
    assert len(touched) > 0
    for prim in touched[:3]:
        fd = _fd(s, p, g, oracle.render, "R", int(prim), 1)
        assert abs(gR[prim, 1] - fd) <= 1e-6 * max(1.0, abs(fd))


def _zero_some_parameters(scene, refl_zero, le_zero):
    """Set the given (primitive, channel) reflectances and radiances to 0
    (channel None = all three)."""
    d = scene.desc()["primitives"]
    refl = np.where((d["type"] & ngi.PRIM_D)[:, None] > 0, d["D_R"], d["G_R"]).copy()
    Le = d["Le"].copy()
    for prim, ch in refl_zero:
        refl[prim, slice(None) if ch is None else ch] = 0.0
    for prim, ch in le_zero:
        Le[prim, slice(None) if ch is None else ch] = 0.0
    scene.set_params(refl, Le)


@pytest.mark.parametrize("renderer", ["pt", "ptdirect"])
def test_oracle_adjoint_at_zero_parameters(renderer):
    """A zero reflectance or radiance channel still has a derivative (each
    contribution is linear in it): reflectance channel 0 of primitive 3, all
    channels of primitive 5, and one radiance channel of the light at 0.
    Central differences straddle 0 (the estimator is a polynomial in the
    parameters and accepts negative values)."""
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    prims = s.desc()["primitives"]
    light = int(np.nonzero(prims["type"] & ngi.PRIM_L)[0][0])
    _zero_some_parameters(s, [(3, 0), (5, None)], [(light, 1)])
    p = _params(renderer, spp=24)
    g = np.random.default_rng(4).normal(size=(p.height, p.width, 3))
    gR, gL = oracle.render_adjoint(s, p, g)
    cases = [("R", 3, 0), ("R", 3, 1), ("R", 5, 0), ("R", 5, 2), ("R", 4, 1), ("Le", light, 1), ("Le", light, 0)]
    for which, prim, ch in cases:
        fd = _fd(s, p, g, oracle.render, which, prim, ch)
        an = (gR if which == "R" else gL)[prim, ch]
        assert abs(an - fd) <= 1e-6 * max(1.0, abs(fd)), (which, prim, ch, an, fd)
    # The zero channels do get gradients.
    assert abs(gR[3, 0]) > 1e-3 and abs(gR[5, 2]) > 1e-3 and abs(gL[light, 1]) > 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("renderer", ["pt", "ptdirect"])
def test_gpu_adjoint_at_zero_parameters(renderer):
    """The GPU replay at zero-valued parameters against its own central
    differences (common random numbers) and the FP64 oracle's gradient."""
    s = ngi.Scene.generate(ngi.GEN_NAVE, 1, 1.0)
    _zero_some_parameters(s, [(0, 1), (3, None)], [(7, 2)])
    p = _params(renderer, 24, 24, 256, 8, 4)
    p.engine = ngi.ENGINE_PERSISTENT
    g = np.random.default_rng(5).normal(size=(24, 24, 3))
    gR, gL, _ = ngi.render_adjoint(s, p, g)
    oR, oL = oracle.render_adjoint(s, p, g)
    render = lambda sc, q: ngi.render(sc, q)[0]
    big = np.abs(np.concatenate([gR.ravel(), gL.ravel()])).max()
    for which, prim, ch in [("R", 0, 1), ("R", 3, 0), ("R", 3, 2), ("R", 4, 1), ("Le", 7, 2), ("Le", 8, 0)]:
        fd = _fd(s, p, g, render, which, prim, ch)
        an = (gR if which == "R" else gL)[prim, ch]
        assert abs(an - fd) <= 2e-3 * max(abs(fd), 1e-2 * big), (which, prim, ch, an, fd)
    assert abs(gR[0, 1]) > 1e-3 * big and abs(gL[7, 2]) > 1e-3 * big
    ref = np.concatenate([oR.ravel(), oL.ravel()])
    mine = np.concatenate([gR.ravel(), gL.ravel()])
    assert np.linalg.norm(mine - ref) <= 2e-2 * np.linalg.norm(ref)


@pytest.mark.gpu
@pytest.mark.parametrize("renderer", ["pt", "ptdirect"])
def test_gpu_adjoint_matches_oracle_and_fd(renderer):
    s = ngi.Scene.generate(ngi.GEN_NAVE, 1, 1.0)
    # 256 spp: the few paths whose FP32 rays hit other geometry than the FP64
    # ones weigh ~1/sqrt(paths) in the gradient; at 16 spp they reach 3.5%.
    p = _params(renderer, 24, 24, 256, 8, 4)
    p.engine = ngi.ENGINE_PERSISTENT
    g = np.random.default_rng(3).normal(size=(24, 24, 3))
    gR, gL, _ = ngi.render_adjoint(s, p, g)
    oR, oL = oracle.render_adjoint(s, p, g)
    ref = np.concatenate([oR.ravel(), oL.ravel()])
    mine = np.concatenate([gR.ravel(), gL.ravel()])
    # Central differences of the GPU renderer itself (common random numbers):
    # the replayed gradient must reproduce them (the 2% bar of the adjoint).
    render = lambda sc, q: ngi.render(sc, q)[0]
    big = np.abs(mine).max()
    for which, prim, ch in [("R", 0, 0), ("R", 3, 1), ("R", 4, 2), ("Le", 7, 2), ("Le", 8, 0)]:
        fd = _fd(s, p, g, render, which, prim, ch)
        an = (gR if which == "R" else gL)[prim, ch]
        assert abs(an - fd) <= 2e-3 * max(abs(fd), 1e-2 * big), (which, prim, ch, an, fd)
    # FP32 replay against the FP64 oracle on the same paths: the gradient
    # vectors agree to 2% in L2 (paths whose FP32 rays hit different
    # geometry than the FP64 ones carry the difference).
    assert np.linalg.norm(mine - ref) <= 2e-2 * np.linalg.norm(ref)
    assert np.dot(mine, ref) / (np.linalg.norm(mine) * np.linalg.norm(ref)) > 0.999


@pytest.mark.gpu
def test_gpu_adjoint_binary_and_wide_bvh_agree():
    """The binary BVH replays paths with traversals run to the end by every
    lane, the wide BVH with resumable traversals; the same paths give the
    same gradients (both within the 2% of the FP64 oracle)."""
    g = np.random.default_rng(6).normal(size=(24, 24, 3))
    out = []
    for width in (ngi.BVH_2, ngi.BVH_8):
        s = ngi.Scene.generate(ngi.GEN_NAVE, 1, 1.0)
        s.build(width)
        p = _params("ptdirect", 24, 24, 64, 8, 4)
        p.engine = ngi.ENGINE_PERSISTENT
        gR, gL, _ = ngi.render_adjoint(s, p, g)
        out.append(np.concatenate([gR.ravel(), gL.ravel()]))
    oR, oL = oracle.render_adjoint(s, p, g)
    ref = np.concatenate([oR.ravel(), oL.ravel()])
    assert np.linalg.norm(out[1] - out[0]) <= 1e-3 * np.linalg.norm(out[0])
    for mine in out:
        assert np.linalg.norm(mine - ref) <= 2e-2 * np.linalg.norm(ref)


@pytest.mark.gpu
def test_autograd_harness_and_inverse_rendering_descends():
    torch = pytest.importorskip("torch")
    from nanogi_amd import autograd as nga
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    refl0, Le0, rmask, lmask = nga.scene_parameters(s)
    p = _params("ptdirect", 24, 24, 64, 6, 3, seed=11)
    target, _ = ngi.render(s, p)
    target = torch.from_numpy(target)
    refl = torch.where(rmask[:, None], torch.full_like(refl0, 0.5), refl0).requires_grad_(True)
    Le = Le0.clone().requires_grad_(True)
    opt = torch.optim.Adam([refl], lr=0.03)
    losses = []
    for it in range(30):
        p.seed = 100 + 2 * it
        fn = nga.render_fn(s, p, adjoint_seed=101 + 2 * it)
        img = fn(refl, Le)
        loss = ((img - target) ** 2).mean()
        opt.zero_grad()
        loss.backward()
        refl.grad[~rmask] = 0
        opt.step()
        with torch.no_grad():
            refl.clamp_(0.01, 1.0)
        losses.append(float(loss.detach()))
    assert losses[-1] < 0.5 * losses[0], losses
    assert float((refl.detach() - refl0)[rmask].abs().mean()) < 0.5 * float((0.5 - refl0[rmask]).abs().mean())


@pytest.mark.gpu
def test_gpu_adjoint_beyond_the_lds_accumulators():
    """324 floor tiles, each its own primitive: the first 256 primitives'
    gradients are summed in LDS, the others go straight to global FP64
    atomics; both must agree with the oracle's gradient."""
    from tests import scenes
    B = scenes.SceneBuilder()
    rng = np.random.default_rng(11)
    n = 18
    for i in range(n):
        for j in range(n):
            pid = B.prim(type=ngi.PRIM_D, D_R=tuple(rng.uniform(0.2, 0.9, 3)))
            B.rect(pid, (-1 + (2 * i + 1) / n, 0, -1 + (2 * j + 1) / n), (0, 1, 0), (1, 0, 0), 1 / n, 1 / n)
    lp = B.prim(type=ngi.PRIM_L, light_type=ngi.LIGHT_AREA, Le=(6.0, 5.0, 4.0))
    B.rect(lp, (0, 1.5, 0), (0, -1, 0), (1, 0, 0), 0.3, 0.3)
    B.camera((0, 1.2, 0.01), (0, 0, 0), up=(0, 0, -1), fov=90.0)
    s = B.scene()
    assert len(B.prims) > 300
    p = _params("ptdirect", 24, 24, 128, 3, -1)
    p.engine = ngi.ENGINE_PERSISTENT
    g = np.random.default_rng(4).normal(size=(24, 24, 3))
    gR, gL, _ = ngi.render_adjoint(s, p, g)
    oR, oL = oracle.render_adjoint(s, p, g)
    for mine, ref in ((gR[:256], oR[:256]), (gR[256:], oR[256:]), (gL, oL)):
        assert np.linalg.norm(ref) > 0
        assert np.linalg.norm(mine - ref) <= 2e-2 * np.linalg.norm(ref)
