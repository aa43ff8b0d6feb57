"""The oracle's LBVH/TLAS against brute force over every triangle, plus
structural invariants of the built hierarchy."""
import numpy as np
import pytest

from luz_amd import api
from oracle import oracle as orc
from tests.helpers import cube_mesh, oracle_scene, random_rays, scaled_instances


def _check_invariants(nodes_u8: np.ndarray, n_prims: int):
    f = nodes_u8.view(np.float32).reshape(-1, 16)
    u = nodes_u8.view(np.uint32).reshape(-1, 16)
    ch = u[:, 12:14]
    if n_prims <= 1:
        assert len(f) == 1
        return
    assert len(f) == n_prims - 1
    leaves = ch[(ch >> 31) == 1] & 0x7FFFFFFF
    assert sorted(leaves.tolist()) == list(range(n_prims))
    inner = ch[(ch >> 31) == 0]
    assert sorted(inner.tolist()) == list(range(1, n_prims - 1))
    assert u[0, 14] == 0xFFFFFFFF
    for i in range(n_prims - 1):
        for side in range(2):
            c = int(ch[i, side])
            if c >> 31:
                continue
            assert u[c, 14] == i  # parent link
            lo = f[i][[0 + side, 4 + side, 8 + side]]
            hi = f[i][[2 + side, 6 + side, 10 + side]]
            clo = np.minimum(f[c][[0, 4, 8]], f[c][[1, 5, 9]])
            chi = np.maximum(f[c][[2, 6, 10]], f[c][[3, 7, 11]])
            assert np.all(lo <= clo) and np.all(hi >= chi)


def test_cube_scene_bvh_invariants_and_brute_force():
    s = api.HostScene(api.SCENE_CUBE, 0)
    o = oracle_scene(s)
    for m in range(s.mesh_count):
        n, t, p = o.blas_arrays(m)
        _check_invariants(n, len(p))
        assert sorted(p.tolist()) == list(range(len(p)))
    rng = np.random.default_rng(11)
    rays = np.concatenate([random_rays(rng, 20000, (0, 0.5, 0), 3.0, 30.0),
                           random_rays(rng, 20000, (0, 0.5, 0), 1.0, 2.0, short=True)])
    h1, st = o.any_hit(rays)
    h2, _ = o.any_hit(rays, brute=True)
    assert 1000 < h1.sum() < len(rays) - 1000
    np.testing.assert_array_equal(h1, h2)
    c1 = o.closest_hit(rays)
    c2 = o.closest_hit(rays, brute=True)
    np.testing.assert_array_equal(c1[0].view(np.uint32), c2[0].view(np.uint32))
    np.testing.assert_array_equal(c1[1], c2[1])
    np.testing.assert_array_equal(c1[2], c2[2])


def test_sphere_mesh_bvh_against_brute_force():
    s = api.HostScene(api.SCENE_SPHERES, 0)
    v, idx = s.mesh(0)
    assert len(idx) // 3 == 20480 and len(v) == 10242
    o = orc.OracleScene()
    o.add_mesh(v, idx)
    xf = np.array([[2, 0, 0, 1], [0, 2, 0, -1], [0, 0, 2, 0.5]], np.float32)
    o.set_instances(xf.reshape(1, 12), [0])
    n, t, p = o.blas_arrays(0)
    _check_invariants(n, len(p))
    rng = np.random.default_rng(12)
    rays = np.concatenate([random_rays(rng, 1500, (1, -1, 0.5), 3.0, 10.0),
                           random_rays(rng, 1500, (1, -1, 0.5), 2.4, 1.0, short=True)])
    h1, _ = o.any_hit(rays)
    h2, _ = o.any_hit(rays, brute=True)
    assert 200 < h1.sum() < len(rays) - 200
    np.testing.assert_array_equal(h1, h2)
    c1 = o.closest_hit(rays)
    c2 = o.closest_hit(rays, brute=True)
    np.testing.assert_array_equal(c1[0].view(np.uint32), c2[0].view(np.uint32))
    np.testing.assert_array_equal(c1[2], c2[2])


def test_tlas_world_to_object_is_inverse_and_masks_drop_instances():
    s = api.HostScene(api.SCENE_SPHERES, 0)
    o2w, mesh, _ = s.nodes()
    o = oracle_scene(s)
    n, w2o, order = o.tlas_arrays()
    assert len(order) == len(mesh) and sorted(order.tolist()) == list(range(len(mesh)))
    _check_invariants(n, len(order))
    for slot, i in enumerate(order):
        a = np.eye(4); a[:3] = o2w[i].reshape(3, 4)
        b = np.eye(4); b[:3] = w2o[slot].reshape(3, 4)
        np.testing.assert_allclose(b @ a, np.eye(4), atol=2e-5)
    mask = np.full(len(mesh), 0xFF); mask[3] = 0
    o.set_instances(o2w, mesh, mask=mask)
    _, _, order2 = o.tlas_arrays()
    assert 3 not in order2.tolist() and len(order2) == len(mesh) - 1


def test_degenerate_sizes():
    o = orc.OracleScene()
    v = np.zeros((3, 12), np.float32)
    v[1, 0] = 1.0
    v[2, 1] = 1.0
    one = o.add_mesh(v, np.array([0, 1, 2], np.uint32))
    empty = o.add_mesh(v, np.zeros(0, np.uint32))
    o.set_instances(np.tile(np.eye(3, 4, dtype=np.float32).reshape(1, 12), (2, 1)), [one, empty])
    n, t, p = o.blas_arrays(one)
    assert len(n) == 1 and len(p) == 1
    n, t, p = o.blas_arrays(empty)
    assert len(n) == 1 and len(p) == 0
    _, _, order = o.tlas_arrays()
    assert order.tolist() == [0]  # the empty mesh's instance can never be hit
    rays = np.array([[0.2, 0.2, 1, 0, 0, 0, -1, 5], [2, 2, 1, 0, 0, 0, -1, 5]], np.float32)
    h, _ = o.any_hit(rays)
    assert h.tolist() == [1, 0]


def _leaf_sets(nodes_u8):
    """Frozen set of leaf slots under every internal node, and the leaf boxes."""
    u = nodes_u8.view(np.uint32).reshape(-1, 16)
    f = nodes_u8.view(np.float32).reshape(-1, 16)
    boxes = {}

    def under(i):
        out = set()
        for side in range(2):
            c = int(u[i, 12 + side])
            if c >> 31:
                out.add(c & 0x7FFFFFFF)
                boxes[c & 0x7FFFFFFF] = (f[i][[0 + side, 4 + side, 8 + side]].astype(np.float64),
                                         f[i][[2 + side, 6 + side, 10 + side]].astype(np.float64))
            else:
                out |= under(c)
        sets.add(frozenset(out))
        return out

    sets = set()
    under(0)
    return sets, boxes


def test_tlas_clustering_closed_form():
    """Two tight pairs far apart: the pairs merge first (nodes 2 and 1, in
    slot order), then the root joins them."""
    v, idx = cube_mesh()
    o = orc.OracleScene()
    m = o.add_mesh(v, idx)
    o.set_instances(scaled_instances([(11.5, 0, 0), (0, 0, 0), (10, 0, 0), (1.5, 0, 0)], [1] * 4), [m] * 4)
    n, _, order = o.tlas_arrays()
    u = n.view(np.uint32).reshape(-1, 16)
    assert order.tolist() == [1, 3, 2, 0]                 # Morton order along x
    assert u[0, 12:15].tolist() == [2, 1, 0xFFFFFFFF]
    assert u[2, 12:15].tolist() == [0x80000000, 0x80000001, 0]
    assert u[1, 12:15].tolist() == [0x80000002, 0x80000003, 0]
    _check_invariants(n, 4)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_tlas_clustering_equals_greedy_agglomerative(seed):
    """With at most 17 instances the clustering radius (16) spans every
    cluster, and merging all mutual nearest neighbours yields the same tree as
    greedy agglomerative clustering (repeatedly merge the pair of least union
    area), because the union area never shrinks under a merge (Walter et al.
    2008).  The greedy tree is built here by exhaustive search in FP64."""
    rng = np.random.default_rng(seed)
    k = 14
    v, idx = cube_mesh()
    o = orc.OracleScene()
    m = o.add_mesh(v, idx)
    o.set_instances(scaled_instances(rng.uniform(-20, 20, (k, 3)), rng.uniform(0.3, 4.0, k)), [m] * k)
    n, _, order = o.tlas_arrays()
    _check_invariants(n, k)
    sets, boxes = _leaf_sets(n)
    clusters = [(frozenset([s]), boxes[s][0], boxes[s][1]) for s in range(k)]
    greedy = set()

    def area(lo, hi):
        e = hi - lo
        return e[0] * e[1] + e[1] * e[2] + e[2] * e[0]

    while len(clusters) > 1:
        best = None
        for a in range(len(clusters)):
            for b in range(a + 1, len(clusters)):
                c = area(np.minimum(clusters[a][1], clusters[b][1]), np.maximum(clusters[a][2], clusters[b][2]))
                if best is None or c < best[0]:
                    best = (c, a, b)
        _, a, b = best
        merged = (clusters[a][0] | clusters[b][0], np.minimum(clusters[a][1], clusters[b][1]),
                  np.maximum(clusters[a][2], clusters[b][2]))
        greedy.add(merged[0])
        clusters = [c for i, c in enumerate(clusters) if i not in (a, b)] + [merged]
    assert sets == greedy


def test_tlas_puts_the_ground_next_to_the_root():
    """The sphere-field scene's ground quad spans the whole scene: clustering
    merges it last, so it is a child of the root and no other subtree's box is
    inflated by it (a Morton-order hierarchy buries it among the spheres)."""
    s = api.HostScene(api.SCENE_SPHERES, 0)
    o = oracle_scene(s)
    n, _, order = o.tlas_arrays()
    f = n.view(np.float32).reshape(-1, 16)
    u = n.view(np.uint32).reshape(-1, 16)
    ext = [f[0][2 + c] - f[0][0 + c] for c in range(2)]                 # x extent of each root child
    big = int(np.argmax(ext))
    assert u[0, 12 + big] >> 31 == 1                                    # a leaf: the ground instance
    assert ext[big] > 4 * ext[1 - big]
