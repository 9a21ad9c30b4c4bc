"""Host kd-tree builder (product) vs the oracle's restatement, and kd traversal vs brute force."""
import json
import os

import numpy as np
import pytest

import yar_amd
from tests import oracle_lib as O
from tests import raygen, scenes


def _meshes(scene):
    d = scene.desc
    for i in range(d.mesh_count):
        m = d.meshes[i]
        v = np.ctypeslib.as_array(m.vertices, shape=(m.vertex_count * 3,)).copy()
        idx = np.ctypeslib.as_array(m.indices, shape=(m.triangle_count * 3,)).copy()
        yield v, idx


@pytest.mark.parametrize("which", ["c1", "c2_small"])
def test_builder_byte_identical_to_oracle(which, c1_path, c2_small_path):
    scene = yar_amd.Scene(c1_path if which == "c1" else c2_small_path)
    n = 0
    for v, idx in _meshes(scene):
        t = yar_amd.build_kdtree(v, idx)
        ow, ob, od = O.kdtree_mesh(v, idx)
        assert np.array_equal(t.words, ow)
        assert np.array_equal(t.bounds, ob)
        assert t.depth == od
        n += 1
    assert n == scene.desc.mesh_count


def test_builder_regression_meshes_identical():
    data = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kdtree_regression_meshes.json")))
    for m in data["meshes"]:
        v = np.array(m["vertices"], np.float32)
        idx = np.array(m["indices"], np.int32)
        t = yar_amd.build_kdtree(v, idx)
        ow, ob, od = O.kdtree_mesh(v, idx)
        assert np.array_equal(t.words, ow)


def test_box_builder_identical_to_oracle():
    rng = np.random.default_rng(11)
    lo = rng.uniform(-10, 10, (500, 3)).astype(np.float32)
    hi = lo + rng.uniform(0, 2, (500, 3)).astype(np.float32)
    hi[::7, 1] = lo[::7, 1]                          # planar boxes exercise the planar-primitive rule
    boxes = np.concatenate([lo, hi], 1)
    t = yar_amd.build_box_kdtree(boxes)
    ow, ob, od = O.kdtree_boxes(boxes)
    assert np.array_equal(t.words, ow)


def _leaf_lists(words):
    """Walk a tree: returns (max depth, list of leaf id lists) and checks the encoding."""
    leaves, deepest = [], 0
    stack = [(0, 0)]
    while stack:
        n, lvl = stack.pop()
        w0, w1 = int(words[2 * n]), int(words[2 * n + 1])
        if (w0 & 3) == 3:
            cnt = w0 >> 2
            ids = [int(words[2 * n + 1 + i]) for i in range(cnt)]
            assert ids == sorted(ids)
            leaves.append(ids)
            deepest = max(deepest, lvl)
        else:
            stack.append((n + 1, lvl + 1))
            stack.append((w0 >> 2, lvl + 1))
    return deepest, leaves


def _leaf_boxes(words, lo, hi):
    """Leaf id lists with the box of each leaf (the root box split at every interior node's plane)."""
    out = []
    stack = [(0, np.array(lo, np.float64), np.array(hi, np.float64))]
    while stack:
        n, blo, bhi = stack.pop()
        w0, w1 = int(words[2 * n]), int(words[2 * n + 1])
        if (w0 & 3) == 3:
            out.append(([int(words[2 * n + 1 + i]) for i in range(w0 >> 2)], blo, bhi))
            continue
        axis = w0 & 3
        split = float(np.array([w1], np.uint32).view(np.float32)[0])
        below_hi, above_lo = bhi.copy(), blo.copy()
        below_hi[axis] = split
        above_lo[axis] = split
        stack.append((n + 1, blo, below_hi))
        stack.append((w0 >> 2, above_lo, bhi))
    return out


def _tris_overlap_open_box(tri, lo, hi):
    """Separating-axis test (13 axes) of triangles (n, 3, 3) against the open box (lo, hi), float64."""
    c, h = 0.5 * (lo + hi), 0.5 * (hi - lo)
    v = tri - c
    ok = (v.min(1) < h).all(1) & (v.max(1) > -h).all(1)
    edges = [v[:, 1] - v[:, 0], v[:, 2] - v[:, 1], v[:, 0] - v[:, 2]]
    nrm = np.cross(edges[0], edges[1])
    ok &= np.abs((nrm * v[:, 0]).sum(1)) < (np.abs(nrm) * h).sum(1)
    for e in edges:
        for k in range(3):
            axis = np.cross(e, np.eye(3)[k])
            p = np.einsum("nij,nj->ni", v, axis)
            r = (np.abs(axis) * h).sum(1)
            ok &= (p.min(1) < r) & (p.max(1) > -r)
    return ok


@pytest.mark.parametrize("which", ["c1", "c2_small"])
def test_builder_structure_invariants(which, c1_path, c2_small_path):
    # depth <= get_max_depth_limit (kdtree.cpp:199-203); every triangle that overlaps the interior of
    # a leaf's box is listed in that leaf (the builder clips triangles to node boxes and keeps every
    # primitive whose clipped bounds reach a child, kdtree_builder.cpp:207-301)
    scene = yar_amd.Scene(c1_path if which == "c1" else c2_small_path)
    checked = 0
    for mi, (v, idx) in enumerate(_meshes(scene)):
        if which == "c2_small" and mi > 3:
            break
        t = yar_amd.build_kdtree(v, idx)
        deepest, leaves = _leaf_lists(t.words)
        ntri = idx.size // 3
        limit = min(int(np.round(8 + 1.3 * np.floor(np.log2(ntri)))), 40)
        assert deepest <= limit
        # every triangle is referenced by some leaf
        seen = set(i for lst in leaves for i in lst)
        assert seen == set(range(ntri))
        tri = v.reshape(-1, 3).astype(np.float64)[idx.reshape(-1, 3)]
        tlo, thi = tri.min(1), tri.max(1)
        b = t.bounds.astype(np.float64)
        # float32 clipping decides overlaps to a few ulps of the coordinates: shrink every leaf box by
        # 1e-4 of its extent plus 1e-6 of the largest coordinate before demanding the listing
        eps = 1e-6 * np.abs(b).max()
        for ids, lo, hi in _leaf_boxes(t.words, b[:3], b[3:]):
            ext = hi - lo
            slo, shi = lo + 1e-4 * ext + eps, hi - 1e-4 * ext - eps
            cand = np.nonzero((tlo < shi).all(1) & (thi > slo).all(1))[0]
            if cand.size == 0 or np.any(shi <= slo):
                continue
            inside = cand[_tris_overlap_open_box(tri[cand], slo, shi)]
            missing = set(inside.tolist()) - set(ids)
            assert not missing, (mi, sorted(missing)[:5])
            checked += inside.size
    assert checked > 0


def test_kdtree_save_layout(tmp_path):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    idx = np.array([0, 1, 2, 0, 2, 3, 0, 3, 1], np.int32)
    t = yar_amd.build_kdtree(v, idx)
    p = tmp_path / "0.kdtree"
    t.save(str(p))
    raw = p.read_bytes()
    assert len(raw) == 24 + 8 + 4 + 8 * (t.words.size // 2)          # kdtree.cpp:140-160
    assert np.array_equal(np.frombuffer(raw[:24], np.float32), t.bounds)
    assert np.frombuffer(raw[24:32], np.uint64)[0] == 0
    assert np.frombuffer(raw[32:36], np.uint32)[0] == t.words.size // 2
    assert np.array_equal(np.frombuffer(raw[36:], np.uint32), t.words)


@pytest.mark.parametrize("which", ["c1", "reg0", "reg1"])
def test_oracle_kd_traversal_equals_brute_force(which, c1_path):
    # validate_triangle_mesh_kdtree (test_kdtree.cpp:138-194): kd-tree t == brute-force t exactly
    if which == "c1":
        scene = yar_amd.Scene(c1_path)
        rays = raygen.random_rays(scene, 40_000, seed=2)
    else:
        scene = scenes.regression_mesh_scene(int(which[-1]))
        m = scene.desc.meshes[0]
        v = np.ctypeslib.as_array(m.vertices, shape=(m.vertex_count * 3,))
        idx = np.ctypeslib.as_array(m.indices, shape=(m.triangle_count * 3,))
        rays = raygen.rays_at_mesh(v, idx, 100_000, seed=2)
    os_ = O.OracleScene(scene)
    a = os_.trace(rays)
    b = os_.trace(rays, brute=True)
    assert np.array_equal(a["t"], b["t"])
    hit = b["object"] >= 0
    same = (a["object"] == b["object"]) & (a["triangle"] == b["triangle"])
    # identical hit ids except exact ties between coincident triangles
    assert same[hit].mean() > 0.9999
    assert hit.mean() > 0.05


def _sah_cost(words, bounds):
    """Expected cost of a ray query against the tree: surface-area weighted interior visits plus
    leaf primitive tests, relative to the root box."""
    def area(lo, hi):
        d = np.maximum(hi - lo, 0)
        return 2 * (d[0] * d[1] + d[0] * d[2] + d[1] * d[2])
    b = bounds.astype(np.float64)
    root = area(b[:3], b[3:])
    cost = 0.0
    stack = [(0, b[:3].copy(), b[3:].copy())]
    while stack:
        n, lo, hi = stack.pop()
        w0, w1 = int(words[2 * n]), int(words[2 * n + 1])
        a = area(lo, hi) / root
        if (w0 & 3) == 3:
            cost += a * (w0 >> 2)
            continue
        cost += a
        axis = w0 & 3
        split = float(np.array([w1], np.uint32).view(np.float32)[0])
        bh, al = hi.copy(), lo.copy()
        bh[axis] = split
        al[axis] = split
        stack.append((n + 1, lo, bh))
        stack.append((w0 >> 2, al, hi))
    return cost


@pytest.mark.parametrize("which", ["c1", "c2_ground", "c5_sphere"])
def test_binned_builder_invariants_and_quality(which, c1_path, c2_small_path, tmp_path):
    # the binned fast path (nodes above 1024 triangles) keeps the kd-tree invariants that make
    # traversal exact, and its SAH cost stays close to the reference builder's
    if which == "c1":
        v, idx = list(_meshes(yar_amd.Scene(c1_path)))[-1]
    elif which == "c2_ground":
        meshes = list(_meshes(yar_amd.Scene(c2_small_path)))
        v, idx = max(meshes, key=lambda m: m[1].size)
    else:
        p = yar_amd.generate_scene(5, str(tmp_path / "c5.pbrt"), seed=1, scale=0.1)
        v, idx = list(_meshes(yar_amd.Scene(p)))[0]
    ntri = idx.size // 3
    assert ntri > 1024
    exact = yar_amd.build_kdtree(v, idx)
    binned = yar_amd.build_kdtree(v, idx, binned=True)
    assert not np.array_equal(exact.words, binned.words)
    assert np.array_equal(exact.bounds, binned.bounds)
    deepest, leaves = _leaf_lists(binned.words)
    assert deepest <= min(int(np.round(8 + 1.3 * np.floor(np.log2(ntri)))), 40)
    assert set(i for lst in leaves for i in lst) == set(range(ntri))
    tri = v.reshape(-1, 3).astype(np.float64)[idx.reshape(-1, 3)]
    tlo, thi = tri.min(1), tri.max(1)
    b = binned.bounds.astype(np.float64)
    eps = 1e-6 * np.abs(b).max()
    for ids, lo, hi in _leaf_boxes(binned.words, b[:3], b[3:]):
        ext = hi - lo
        slo, shi = lo + 1e-4 * ext + eps, hi - 1e-4 * ext - eps
        cand = np.nonzero((tlo < shi).all(1) & (thi > slo).all(1))[0]
        if cand.size == 0 or np.any(shi <= slo):
            continue
        inside = cand[_tris_overlap_open_box(tri[cand], slo, shi)]
        assert not (set(inside.tolist()) - set(ids))
    ce, cb = _sah_cost(exact.words, exact.bounds), _sah_cost(binned.words, binned.bounds)
    print(f"{which}: {ntri} tris, SAH cost exact {ce:.3f} binned {cb:.3f} (ratio {cb / ce:.3f})")
    assert cb <= 1.25 * ce
