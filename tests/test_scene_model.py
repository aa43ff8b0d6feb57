"""The host scene model (AssetManager.hpp conventions) and the workload
generators: closed forms for transforms, camera and light packing."""
import math

import numpy as np
import pytest

from luz_amd import api


def test_cube_scene_content():
    s = api.HostScene(api.SCENE_CUBE, 0)
    assert s.mesh_count == 2
    v0, i0 = s.mesh(0)
    v1, i1 = s.mesh(1)
    assert len(i0) // 3 == 2 and len(i1) // 3 == 12
    np.testing.assert_allclose(v0[:, 0:3].min(0), [-10, 0, -10])
    np.testing.assert_allclose(v0[:, 0:3].max(0), [10, 0, 10])
    np.testing.assert_allclose(v1[:, 0:3].min(0), [-0.5, 0, -0.5])
    np.testing.assert_allclose(v1[:, 0:3].max(0), [0.5, 1, 0.5])
    # cube normals are face normals of unit length
    np.testing.assert_allclose(np.linalg.norm(v1[:, 3:6], axis=1), 1.0)
    o2w, mesh, mats = s.nodes()
    np.testing.assert_array_equal(o2w, np.tile(np.eye(3, 4, dtype=np.float32).reshape(1, 12), (2, 1)))
    for m in mats:
        assert (m.material >> 16) & 0xFF == 255   # occlusion 1 (opaque.frag:57)
    sb = s.pack()
    assert sb.num_lights == 1 and sb.ao_num_samples == 4 and sb.lights[0].num_shadow_samples == 1
    assert sb.lights[0].type == api.LIGHT_DIRECTIONAL
    d = np.array(sb.lights[0].direction[:])
    want = np.array([-0.3, -1.0, -0.2]) / np.linalg.norm([-0.3, -1.0, -0.2])
    np.testing.assert_allclose(d, want, atol=1e-6)
    np.testing.assert_allclose(np.array(sb.lights[0].position[:]), -1000 * want, rtol=1e-6)
    assert abs(sb.ao_min - 1e-4) < 1e-9 and sb.ao_max == 1.0
    np.testing.assert_allclose(sb.cam_pos[:], [3, 2, 3])


def test_camera_conventions():
    s = api.HostScene(api.SCENE_CUBE, 0)
    W, H, cam = s.camera()
    assert (W, H) == (256, 256)
    f = np.array(cam.forward[:], np.float64)
    r = np.array(cam.right[:], np.float64)
    u = np.array(cam.up[:], np.float64)
    np.testing.assert_allclose(f, -np.array([3, 2, 3]) / np.linalg.norm([3, 2, 3]), atol=1e-7)
    t = math.tan(math.radians(30))
    assert abs(np.linalg.norm(r) - t) < 1e-6 and abs(np.linalg.norm(u) - t) < 1e-6
    assert abs(r @ f) < 1e-7 and abs(u @ f) < 1e-7 and abs(r[1]) < 1e-7
    assert u[1] < 0  # image rows grow downwards (Vulkan y flip)
    s2 = api.HostScene(api.SCENE_SPHERES, 0)
    W, H, cam = s2.camera()
    assert (W, H) == (1920, 1080)
    assert abs(np.linalg.norm(cam.right[:]) / np.linalg.norm(cam.up[:]) - 1920 / 1080) < 1e-5


def test_sphere_field_generator():
    s = api.HostScene(api.SCENE_SPHERES, 0)
    assert s.mesh_count == 6
    assert s.instanced_triangles == 50 * 20480 + 8192 == 1032192
    for k in range(5):
        v, idx = s.mesh(k)
        assert len(v) == 10242 and len(idx) == 3 * 20480
        # closed 2-manifold: every undirected edge is shared by exactly two triangles
        t = idx.reshape(-1, 3)
        e = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1)
        _, counts = np.unique(e[:, 0].astype(np.int64) * 20000 + e[:, 1], return_counts=True)
        assert np.all(counts == 2)
        r = np.linalg.norm(v[:, 0:3], axis=1)
        assert 0.85 < r.min() and r.max() < 1.15      # radial fBm, amplitude 0.1
    o2w, mesh, mats = s.nodes()
    assert len(mesh) == 51 and mesh[-1] == 5
    for i in range(50):
        m = o2w[i].reshape(3, 4).astype(np.float64)
        R = m[:, :3]
        sc = np.linalg.norm(R[:, 0])
        assert 0.5 <= sc <= 3.0
        np.testing.assert_allclose(R.T @ R, sc * sc * np.eye(3), atol=1e-5 * sc * sc)
        assert np.linalg.det(R) > 0
        assert np.all(np.abs(m[:, 3]) <= 50.0)
    sb = s.pack()
    assert sb.num_lights == 9 and sb.ao_num_samples == 16
    types = [sb.lights[i].type for i in range(9)]
    assert types.count(api.LIGHT_DIRECTIONAL) == 1 and types.count(api.LIGHT_POINT) == 8


def test_generators_are_deterministic():
    a = api.HostScene(api.SCENE_SPHERES, 0)
    b = api.HostScene(api.SCENE_SPHERES, 0)
    c = api.HostScene(api.SCENE_SPHERES, 1)
    np.testing.assert_array_equal(a.mesh(2)[0], b.mesh(2)[0])
    np.testing.assert_array_equal(a.nodes()[0], b.nodes()[0])
    assert not np.array_equal(a.nodes()[0], c.nodes()[0])


def test_animation_moves_instances_rigidly():
    s = api.HostScene(api.SCENE_SPHERES, 0)
    o0 = s.nodes()[0]
    s.animate(7)
    o7 = s.nodes()[0]
    assert not np.array_equal(o0[:50], o7[:50])
    np.testing.assert_array_equal(o0[50], o7[50])   # the ground does not move
    for i in range(50):
        a = o0[i].reshape(3, 4)[:, :3].astype(np.float64)
        b = o7[i].reshape(3, 4)[:, :3].astype(np.float64)
        np.testing.assert_allclose(np.linalg.norm(a, axis=0), np.linalg.norm(b, axis=0), rtol=1e-5)
