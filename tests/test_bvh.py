"""Structure of the flattened BVHs (decoded with numpy from ngi_scene_bvh_export)."""
import numpy as np
import pytest

import nanogi_amd as ngi

EMPTY = 0x7FFFFFFF


def _leaf(ref):
    v = (~np.int64(ref)) & 0xFFFFFFFF
    return int(v >> 4), int((v & 15) + 1)


def _tri_bounds(scene):
    d = scene.desc()
    P = d["positions"].astype(np.float32)[d["faces"]]
    return P.min(axis=1), P.max(axis=1)


def _decode2(nodes):
    rec = nodes.view(np.float32).reshape(-1, 16)
    refs = rec[:, 12:14].view(np.int32)
    lo = np.stack([np.stack([rec[:, 0], rec[:, 2], rec[:, 8]], 1), np.stack([rec[:, 4], rec[:, 6], rec[:, 10]], 1)], 1)
    hi = np.stack([np.stack([rec[:, 1], rec[:, 3], rec[:, 9]], 1), np.stack([rec[:, 5], rec[:, 7], rec[:, 11]], 1)], 1)
    return refs, lo, hi


def _decode8(nodes):
    rec = nodes.reshape(-1, 128)
    org = rec[:, 0:12].copy().view(np.float32).reshape(-1, 3)
    ex = rec[:, 12:15].astype(np.int32) - 127
    valid = rec[:, 15]
    q = rec[:, 16:64].reshape(-1, 6, 8).astype(np.float64)
    scale = np.ldexp(1.0, ex)                                   # (n, 3)
    lo = org[:, None, :] + q[:, 0:3, :].transpose(0, 2, 1) * scale[:, None, :]
    hi = org[:, None, :] + q[:, 3:6, :].transpose(0, 2, 1) * scale[:, None, :]
    refs = rec[:, 64:96].copy().view(np.int32).reshape(-1, 8)
    return refs, lo.astype(np.float32), hi.astype(np.float32), valid


def _check(scene, width):
    info = scene.build(width)
    nodes, perm = scene.bvh_export()
    nt = scene.num_triangles
    assert sorted(perm.tolist()) == list(range(nt))
    tlo, thi = _tri_bounds(scene)
    tlo, thi = tlo[perm], thi[perm]   # leaf order
    if width == 2:
        refs, lo, hi = _decode2(nodes)
        kids = [[(refs[i, k], lo[i, k], hi[i, k]) for k in range(2)] for i in range(len(refs))]
    else:
        refs, lo, hi, valid = _decode8(nodes)
        kids = []
        for i in range(len(refs)):
            row = []
            for k in range(8):
                has = bool((valid[i] >> k) & 1)
                assert has == (refs[i, k] != EMPTY)
                if has:
                    row.append((refs[i, k], lo[i, k], hi[i, k]))
            kids.append(row)
    covered = np.zeros(nt, dtype=np.int32)
    depth = 0
    parent_of = {}

    def visit(i, d):
        nonlocal depth
        depth = max(depth, d)
        bl, bh = np.full(3, np.inf, np.float32), np.full(3, -np.inf, np.float32)
        for ref, clo, chi in kids[i]:
            if ref >= 0:
                assert ref not in parent_of or width == 2 and len(kids) == 1
                parent_of[ref] = i
                sl, sh = visit(ref, d + 1)
            else:
                first, count = _leaf(ref)
                covered[first:first + count] += 1
                sl, sh = tlo[first:first + count].min(0), thi[first:first + count].max(0)
            # conservative: the stored child box contains its subtree
            assert np.all(clo <= sl) and np.all(chi >= sh), (i, clo, sl, chi, sh)
            bl, bh = np.minimum(bl, sl), np.maximum(bh, sh)
        return bl, bh

    import sys
    sys.setrecursionlimit(10000)
    visit(0, 0)
    # A one-leaf binary tree stores that leaf in both child slots of the root.
    root_is_leaf_pair = width == 2 and len(kids) == 1 and kids[0][0][0] < 0 and kids[0][0][0] == kids[0][1][0]
    if root_is_leaf_pair:
        covered //= 2
    assert np.all(covered == 1), "every triangle in exactly one leaf"
    assert depth == info["max_depth"] or width == 2
    return info


@pytest.mark.parametrize("width", [2, 8])
@pytest.mark.parametrize("kind", [ngi.GEN_CORNELL, ngi.GEN_NAVE, ngi.GEN_BATCH_ROOM])
def test_bvh_structure(kind, width):
    s = ngi.Scene.generate(kind, 3, 1.0, 700 if kind == ngi.GEN_BATCH_ROOM else 0)
    info = _check(s, width)
    assert info["node_bytes"] == (64 if width == 2 else 128)
    if width == 8:
        assert info["max_depth"] < 32


@pytest.mark.parametrize("width", [2, 8])
def test_single_and_tiny_scenes(width):
    from tests import scenes
    for ntri in (1, 2, 3, 5):
        B = scenes.SceneBuilder()
        p = B.prim(type=ngi.PRIM_D, D_R=(1, 1, 1))
        for k in range(ntri):
            B.tri(p, (k, 0, 0), (k + 1, 0, 0), (k, 1, 0))
        B.camera((0, 0, 5), (0, 0, 0))
        _check(B.scene(), width)


def test_build_is_deterministic_across_thread_counts():
    s = ngi.Scene.generate(ngi.GEN_SPHERE_FIELD, 1, 1.0, 6)
    out = []
    for threads in (1, 3, 8):
        s.build(8, num_threads=threads)
        out.append(s.bvh_export())
    for nodes, perm in out[1:]:
        assert np.array_equal(nodes, out[0][0]) and np.array_equal(perm, out[0][1])


def test_wide_slots_follow_child_octants():
    """Children sit in the slot of their centroid's octant when free."""
    s = ngi.Scene.generate(ngi.GEN_NAVE, 1, 1.0)
    s.build(8)
    nodes, _ = s.bvh_export()
    refs, lo, hi, valid = _decode8(nodes)
    agree = total = 0
    for i in range(len(refs)):
        ks = [k for k in range(8) if (valid[i] >> k) & 1]
        if len(ks) < 2:
            continue
        c = 0.5 * (lo[i, ks] + hi[i, ks])
        pc = 0.5 * (lo[i, ks].min(0) + hi[i, ks].max(0))
        for j, k in enumerate(ks):
            octant = int(c[j, 0] > pc[0]) | int(c[j, 1] > pc[1]) << 1 | int(c[j, 2] > pc[2]) << 2
            agree += octant == k
            total += 1
    assert agree / total > 0.5


def test_traversal_cost_option_shapes_the_tree():
    """A larger SAH node-visit cost makes fewer, larger leaves (up to the leaf cap)."""
    s = ngi.Scene.generate(ngi.GEN_SPHERE_FIELD, 1, 1.0, 3)
    cheap = s.build(ngi.BVH_8, max_leaf_size=8, traversal_cost=0.5)
    dear = s.build(ngi.BVH_8, max_leaf_size=8, traversal_cost=4.0)
    assert dear["num_leaves"] < cheap["num_leaves"]
    assert dear["num_triangles"] == cheap["num_triangles"]
    with pytest.raises(ngi.NgiError) as e:
        s.build(ngi.BVH_8, traversal_cost=-1.0)
    assert e.value.status == ngi.ERR_INVALID_ARG
