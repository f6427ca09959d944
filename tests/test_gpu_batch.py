"""Batched rendering (ngi_render_batch) against one render per scene.

A scene's paths in a batch are the paths ngi_render draws for the same
parameters (same ids, seeds and stream ids), traced by the same vertex loop,
so each batch film equals the single-scene film path by path, up to FP32
rounding and the order of the FP64 film additions.  The single-scene renders are themselves pinned against the
oracle by test_gpu_render.py; one small scene is also checked against the
oracle directly here.
"""
import numpy as np
import pytest

import nanogi_amd as ngi
import oracle

pytestmark = pytest.mark.gpu


def _p(renderer, w, h, spp, m=8, rr=3, seed=1, stream_id=0, sampling=ngi.SAMPLING_PER_PIXEL, num_samples=0):
    p = ngi.default_params(renderer=renderer, width=w, height=h, spp=spp, max_num_vertices=m, rr_depth=rr, seed=seed,
                           sampling=sampling, engine=ngi.ENGINE_PERSISTENT)
    p.stream_id = stream_id
    if num_samples:
        p.num_samples = num_samples
    return p


def _same(a, b):
    """Same paths and random numbers as the single-scene render.  The two
    kernels are compiled separately, so the compiler may fuse FP32
    multiply-adds differently and a path that grazes an edge can take
    another branch: the bar is the one used between the GPU and the oracle,
    98% of pixels within a relative 2e-4 and image means within 5e-3."""
    err = np.abs(a - b).max(axis=2)
    scale = np.abs(b).max(axis=2) + 1e-3 * np.abs(b).max()
    frac = (err <= 2e-4 * scale).mean()
    assert frac >= 0.98, frac
    np.testing.assert_allclose(a.reshape(-1, 3).mean(0), b.reshape(-1, 3).mean(0), rtol=5e-3)


@pytest.mark.parametrize("renderer", ["pt", "ptdirect"])
def test_batch_films_equal_single_renders(renderer):
    scenes = [ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0),
              ngi.Scene.generate(ngi.GEN_BATCH_ROOM, 7, 1.5, 3000),
              ngi.Scene.generate(ngi.GEN_BATCH_ROOM, 8, 0.75, 40),
              ngi.Scene.generate(ngi.GEN_SPHERE_FIELD, 3, 1.0, 4)]
    params = [_p(renderer, 33, 29, 5, seed=3, stream_id=0),
              _p(renderer, 48, 32, 2, m=12, rr=4, seed=4, stream_id=1),
              _p(renderer, 16, 24, 17, m=-1, rr=1, seed=5, stream_id=2),
              _p(renderer, 40, 40, 1, seed=6, stream_id=3, sampling=ngi.SAMPLING_REFERENCE_UNIFORM, num_samples=4000)]
    films, ms, st = ngi.render_batch(scenes, params)
    assert st["samples"] == 33 * 29 * 5 + 48 * 32 * 2 + 16 * 24 * 17 + 4000
    assert np.all(ms > 0) and np.all(ms <= st["ms_total"] + 1.0)
    for s, p, f in zip(scenes, params, films):
        ref, _ = ngi.render(s, p)
        assert f.shape == ref.shape and np.abs(ref).max() > 0
        _same(f, ref)


def test_batch_matches_oracle():
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    p = _p("ptdirect", 24, 24, 8, m=9, rr=-1, seed=9)
    films, _, _ = ngi.render_batch([s, s], [p, p])
    o = oracle.render(s, p)
    # Two copies of one scene in one kernel: identical paths, so only the
    # order of the FP64 film additions differs.
    np.testing.assert_allclose(films[0], films[1], rtol=1e-9, atol=1e-12 * np.abs(films[1]).max())
    err = np.abs(films[0] - o).max(axis=2)
    scale = np.abs(o).max(axis=2) + 1e-3 * np.abs(o).max()
    assert (err <= 2e-4 * scale).mean() >= 0.98
    np.testing.assert_allclose(films[0].reshape(-1, 3).mean(0), o.reshape(-1, 3).mean(0), rtol=2e-3)


@pytest.mark.parametrize("renderer", ["pt", "ptdirect"])
def test_batch_of_many_small_scenes_and_binary_bvh(renderer):
    """Tiny scenes (a wave crosses many of them, draining its shadow queue
    each time it leaves one) on the binary BVH."""
    scenes, params = [], []
    rng = np.random.default_rng(0)
    for i in range(37):
        s = ngi.Scene.generate(ngi.GEN_BATCH_ROOM, 100 + i, 1.0, int(rng.integers(12, 600)))
        s.build(ngi.BVH_2)
        scenes.append(s)
        params.append(_p(renderer, int(rng.integers(4, 20)), int(rng.integers(4, 20)), int(rng.integers(1, 6)), seed=i,
                         stream_id=i))
    films, ms, _ = ngi.render_batch(scenes, params)
    for k in (0, 17, 36):
        _same(films[k], ngi.render(scenes[k], params[k])[0])
    assert np.all(ms > 0)


def test_batch_rejects_mixed_bvh_widths():
    a = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    b = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    a.build(ngi.BVH_2)
    b.build(ngi.BVH_8)
    with pytest.raises(ngi.NgiError) as e:
        ngi.render_batch([a, b], [_p("pt", 8, 8, 1)] * 2)
    assert e.value.status == ngi.ERR_INVALID_ARG


def test_batch_graph_replays_frames_and_follows_scene_updates():
    """Each frame replays the captured graph (fill kernel + persistent launch);
    eager renders run between replays, and every replay is checked."""
    scenes = [ngi.Scene.generate(ngi.GEN_BATCH_ROOM, 20 + i, 1.0, n) for i, n in enumerate((30, 900, 4000))]
    params = [_p("ptdirect", 24 + 8 * i, 20, 3 + i, seed=10 + i, stream_id=i) for i in range(3)]
    b = ngi.Batch(scenes, params)
    cfg = b.config()
    assert cfg["block_size"] in (64, 128, 256) and cfg["blocks_per_cu"] >= 1 and cfg["use_graph"] == 1
    assert cfg["chunk_paths"] == 256   # a few thousand paths: the smallest chunk
    for frame in (0, 1, 0, 2):
        b.render(frame)
        b.sync()
        for i, (s, p) in enumerate(zip(scenes, params)):
            q = ngi.default_params()
            ctypes_copy(q, p)
            q.seed = p.seed + frame
            _same(b.film(i), ngi.render(s, q)[0])
        ms, st = b.times()
        assert np.all(ms > 0) and st["samples"] == sum(p.width * p.height * p.spp for p in params)
        assert st["extension_rays"] >= st["samples"]
    # Changed reflectances are picked up by the next frame.
    d = scenes[1].desc()["primitives"]
    refl = np.where((d["type"] & ngi.PRIM_D)[:, None] > 0, d["D_R"], d["G_R"]) * 0.5
    scenes[1].set_params(refl, d["Le"])
    b.render(0)
    b.sync()
    _same(b.film(1), ngi.render(scenes[1], params[1])[0])
    b.close()


@pytest.mark.parametrize("block", [64, 128, 256])
def test_batch_fixed_launch_shapes_and_plain_launches(block):
    scenes = [ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0), ngi.Scene.generate(ngi.GEN_BATCH_ROOM, 3, 1.0, 500)]
    params = [_p("pt", 20, 20, 4, seed=1), _p("pt", 30, 10, 6, seed=2, stream_id=1)]
    b = ngi.Batch(scenes, params, ngi.batch_opts(block_size=block, blocks_per_cu=2, use_graph=0, chunk_paths=100))
    assert b.config()["block_size"] == block and b.config()["blocks_per_cu"] <= 2 and b.config()["chunk_paths"] == 100
    b.render(0)
    b.sync()
    for i in range(2):
        _same(b.film(i), ngi.render(scenes[i], params[i])[0])


def ctypes_copy(dst, src):
    for name, _ in src._fields_:
        setattr(dst, name, getattr(src, name))


def test_batch_with_an_empty_scene_and_single_path_scenes():
    """A scene without triangles (a camera and a point light), scenes of one
    path and a one-pixel image between ordinary ones: each film equals its
    single-scene render, and the empty scene shows only the light's splat."""
    from tests import scenes as sc
    B = sc.SceneBuilder()
    B.prim(type=ngi.PRIM_L, light_type=ngi.LIGHT_POINT, Le=(2, 1, 1), light_position=(0.1, 0.2, -2.0))
    B.camera((0, 0, 0), (0, 0, -1))
    empty = ngi.Scene.create(np.zeros((0, 3)), np.zeros((0, 3), dtype=np.int64), np.zeros(0, dtype=np.int64),
                             np.array(B.prims), normals=np.zeros((0, 3)))
    empty.build(ngi.BVH_8)
    room = ngi.Scene.generate(ngi.GEN_BATCH_ROOM, 9, 1.0, 500)
    room.build(ngi.BVH_8)
    scenes = [room, empty, room, room, empty]
    params = [_p("ptdirect", 24, 24, 4, seed=1, stream_id=0),
              _p("ptdirect", 16, 16, 8, seed=2, stream_id=1),
              _p("ptdirect", 1, 1, 1, seed=3, stream_id=2),
              _p("ptdirect", 8, 8, 1, seed=4, stream_id=3, sampling=ngi.SAMPLING_REFERENCE_UNIFORM, num_samples=1),
              _p("ptdirect", 4, 4, 2, seed=5, stream_id=4)]
    films, ms, st = ngi.render_batch(scenes, params)
    assert st["samples"] == 24 * 24 * 4 + 16 * 16 * 8 + 1 + 1 + 4 * 4 * 2
    for s, p, f in zip(scenes, params, films):
        ref, _ = ngi.render(s, p)
        np.testing.assert_allclose(f, ref, rtol=2e-4, atol=1e-6 * max(np.abs(ref).max(), 1.0))
    assert np.count_nonzero(films[1].sum(axis=2)) == 1


def test_batch_chunk_size_follows_the_paths_per_wave():
    """Without a fixed chunk size a batch takes about 16 chunks per resident
    wave (a power of two in [256, 8192]): 2 x 512^2 x 128 spp = 67M paths get
    chunks above 256.  The chunk size changes only which wave renders which
    paths, so the films equal those of chunks of 256."""
    scenes = [ngi.Scene.generate(ngi.GEN_BATCH_ROOM, 40 + i, 1.0, n) for i, n in enumerate((200, 3000))]
    params = [_p("ptdirect", 512, 512, 128, seed=3 + i, stream_id=i) for i in range(2)]
    auto = ngi.Batch(scenes, params, ngi.batch_opts(block_size=64, use_graph=0))
    c = auto.config()["chunk_paths"]
    assert 256 < c <= 8192 and c & (c - 1) == 0
    auto.render(0)
    auto.sync()
    fa = [auto.film(i) for i in range(2)]
    auto.close()
    small = ngi.Batch(scenes, params, ngi.batch_opts(block_size=64, use_graph=0, chunk_paths=256))
    small.render(0)
    small.sync()
    for i in range(2):
        fb = small.film(i)
        assert np.abs(fb).max() > 0
        np.testing.assert_allclose(fa[i], fb, rtol=1e-9, atol=1e-12 * np.abs(fb).max())
    small.close()


def test_set_params_waits_for_a_running_frame():
    """Two batches over one scene: frame 0 of the first is still running when
    the parameters change and the second batch renders (its refresh rewrites
    the scene's parameter tables).  The rewrite waits for every batch stream
    that reads the scene, so the first frame renders the old parameters
    throughout and the second the new ones."""
    def fresh():
        return ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    s = fresh()
    p = _p("ptdirect", 256, 256, 64, m=9, rr=-1, seed=12)
    a = ngi.Batch([s], [p])
    b = ngi.Batch([s], [p])
    d = s.desc()["primitives"]
    refl = np.where((d["type"] & ngi.PRIM_D)[:, None] > 0, d["D_R"], d["G_R"]).copy()
    new_refl = np.where(refl > 0, 0.95, 0.0)
    new_le = d["Le"] * 0.25
    a.render(0)                       # enqueued, not waited for
    s.set_params(new_refl, new_le)    # host copy only; uploaded by the next refresh
    b.render(0)
    a.sync()
    b.sync()
    old_ref, _ = ngi.render(fresh(), p)
    s_new = fresh()
    s_new.set_params(new_refl, new_le)
    new_ref, _ = ngi.render(s_new, p)
    _same(a.film(0), old_ref)
    _same(b.film(0), new_ref)
    a.close()
    b.close()


def test_session_survives_a_scene_rebuild():
    """A scene rebuilt with another BVH releases its device copy; a session
    of it keeps its film and renders on after uploading the scene again."""
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    s.build(ngi.BVH_8)
    p = _p("pt", 32, 32, 4, m=6, rr=-1, seed=2)
    sess = ngi.Session(s, p)
    sess.render(4)
    before = sess.film()
    s.build(ngi.BVH_2)
    np.testing.assert_array_equal(sess.film(), before)
    st = sess.stats()
    assert st["samples"] == 32 * 32 * 4
    sess.render(4)
    both = sess.film()
    q = _p("pt", 32, 32, 8, m=6, rr=-1, seed=2)
    ref, _ = ngi.render(s, q)
    _same(both, ref)
    sess.reset()
    assert not sess.film().any()
    sess.close()


def test_reduced_room_batch_against_the_oracle():
    """The many-room workload's recipe (seeded log-uniform draws: 10-5,000
    triangles, 64-512 px, 16-512 spp; ptdirect, at most 16 segments,
    roulette from depth 4), first 16 rooms, images capped at 64 x 64 and
    16 spp so the oracle finishes in seconds.  One captured-graph frame;
    every film against oracle.render of its own room with the same seed and
    stream id (common random numbers)."""
    rng = np.random.default_rng(2024)
    lu = lambda lo, hi: float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
    scenes, params = [], []
    for i in range(16):
        tris = int(round(lu(10, 5000)))
        side = min(int(round(lu(64, 512))), 64)
        spp = min(int(round(lu(16, 512))), 16)
        s = ngi.Scene.generate(ngi.GEN_BATCH_ROOM, 1000 + i, 1.0, max(tris, 12))
        s.build(8)
        scenes.append(s)
        p = ngi.default_params(renderer="ptdirect", width=side, height=side, sampling=ngi.SAMPLING_PER_PIXEL, spp=spp,
                               max_num_vertices=17, rr_depth=4, rr_prob=0.5, seed=77, engine=ngi.ENGINE_PERSISTENT)
        p.stream_id = i
        params.append(p)
    assert max(sc.num_triangles for sc in scenes) > 1000 and min(sc.num_triangles for sc in scenes) < 200
    b = ngi.Batch(scenes, params)
    assert b.config()["use_graph"]
    b.render(0)
    b.sync()
    for i, (s, p) in enumerate(zip(scenes, params)):
        f = b.film(i)
        o = oracle.render(s, p)
        err = np.abs(f - o).max(axis=2)
        scale = np.abs(o).max(axis=2) + 1e-3 * np.abs(o).max()
        frac = (err <= 2e-4 * scale).mean()
        assert frac >= 0.98, (i, s.num_triangles, frac)
        np.testing.assert_allclose(f.reshape(-1, 3).mean(0), o.reshape(-1, 3).mean(0), rtol=2e-3, err_msg=str(i))
    b.close()
