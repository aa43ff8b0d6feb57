"""Scene ingest on the CPU: the reference's own assets (glTF binary, OBJ
meshes, the .luz project without its blob) load into the scene model with
the geometry the files hold, the orbit camera of default.luz reproduces the
eye the reference saved, and a generated scene survives a .luz/.luzbin
round trip bit for bit."""
import json
import time
from pathlib import Path

import numpy as np
import pytest

from luz_amd import api

ASSETS = Path(__file__).resolve().parent / "golden" / "assets"


def _geometry(scene):
    out = []
    for i in range(scene.mesh_count):
        v, idx = scene.mesh(i)
        out.append((v, idx))
    return out


def test_cube_glb_is_a_closed_cube():
    s = api.HostScene.load(ASSETS / "cube.glb")
    geo = _geometry(s)
    assert len(geo) == 1
    v, idx = geo[0]
    assert len(v) == 24 and len(idx) == 36
    pos = v[:, 0:3]
    np.testing.assert_allclose(np.abs(pos), 1.0, atol=1e-6)          # the corners of [-1, 1]^3
    # welded by position, every edge is shared by exactly two triangles
    key = {tuple(p): k for k, p in enumerate(np.unique(pos.round(6), axis=0))}
    tri = np.array([[key[tuple(pos[i].round(6))] for i in idx[t:t + 3]] for t in range(0, 36, 3)])
    edges = {}
    for a, b, c in tri:
        for e in ((a, b), (b, c), (c, a)):
            edges[tuple(sorted(e))] = edges.get(tuple(sorted(e)), 0) + 1
    assert len(key) == 8 and set(edges.values()) == {2}
    # the stored vertex normals agree with the face orientation
    nrm = v[:, 3:6]
    for t in range(0, 36, 3):
        a, b, c = pos[idx[t:t + 3]]
        fn = np.cross(b - a, c - a)
        assert np.dot(fn, nrm[idx[t]]) > 0
    o2w, mesh, mats = s.nodes()
    assert len(mesh) == 1


def _obj_reference(path):
    """Independent reading of an OBJ: triangle count after fan triangulation and
    the bounds of the referenced positions."""
    verts, tris, used = [], 0, set()
    for line in path.read_text().splitlines():
        parts = line.split()
        if not parts:
            continue
        if parts[0] == "v":
            verts.append([float(x) for x in parts[1:4]])
        elif parts[0] == "f":
            tris += len(parts) - 3
            for p in parts[1:]:
                i = int(p.split("/")[0])
                used.add(i - 1 if i > 0 else len(verts) + i)
    P = np.array(verts)[sorted(used)]
    return tris, P.min(0), P.max(0)


@pytest.mark.parametrize("name", ["point.obj", "spot.obj", "directional.obj"])
def test_obj_meshes_match_an_independent_reading(name):
    s = api.HostScene.load(ASSETS / name)
    geo = _geometry(s)
    tris, lo, hi = _obj_reference(ASSETS / name)
    assert sum(len(i) // 3 for _, i in geo) == tris
    allp = np.concatenate([v[:, 0:3] for v, _ in geo])
    np.testing.assert_allclose(allp.min(0), lo, atol=1e-6)
    np.testing.assert_allclose(allp.max(0), hi, atol=1e-6)
    for v, idx in geo:
        assert idx.max() < len(v) and len(idx) % 3 == 0
        n = np.linalg.norm(v[:, 3:6], axis=1)
        assert (n > 0.99).all() and (n < 1.01).all()


def test_default_luz_scene_settings_and_orbit_camera():
    s = api.HostScene.load(ASSETS / "default.luz")    # the project's .luzbin is not shipped
    proj = json.loads((ASSETS / "default.luz").read_text())
    scene = proj["scenes"][str(proj["initialScene"])]
    mesh_nodes = []

    def walk(nodes):
        for n in nodes:
            if n["type"] == 6:
                mesh_nodes.append(n)
            walk(n["children"])
    walk(scene["nodes"])
    assert s.mesh_count == 0 and s.missing_meshes == len(mesh_nodes) > 0
    sb = s.pack()
    assert sb.ao_num_samples == scene["aoSamples"] == 4
    assert sb.lights[0].num_shadow_samples == 0 or sb.num_lights > 0
    np.testing.assert_allclose(sb.ambient_intensity, scene["ambientLight"], rtol=1e-7)
    np.testing.assert_allclose([sb.ao_min, sb.ao_max], [scene["aoMin"], scene["aoMax"]], rtol=1e-7)
    cam = next(n for n in scene["nodes"] if n["type"] == 8)
    # the orbit camera's eye, recomputed from rotation/zoom/center, is the eye the reference saved
    np.testing.assert_allclose(sb.cam_pos[:], cam["eye"], rtol=2e-6, atol=2e-6)


def test_luz_round_trip_is_exact(tmp_path):
    s = api.HostScene(api.SCENE_SPHERES, 3)
    p = tmp_path / "spheres.luz"
    t0 = time.perf_counter()
    s.save(p)
    t1 = time.perf_counter()
    r = api.HostScene.load(p)
    t2 = time.perf_counter()
    mb = (p.stat().st_size + (tmp_path / "spheres.luzbin").stat().st_size) / 1e6
    print(f"{mb:.1f} MB: save {mb / (t1 - t0):.0f} MB/s, load {mb / (t2 - t1):.0f} MB/s", flush=True)
    assert r.mesh_count == s.mesh_count
    for (va, ia), (vb, ib) in zip(_geometry(s), _geometry(r)):
        np.testing.assert_array_equal(va.view(np.uint32), vb.view(np.uint32))
        np.testing.assert_array_equal(ia, ib)
    oa, ma, mata = s.nodes()
    ob, mb_, matb = r.nodes()
    np.testing.assert_array_equal(oa.view(np.uint32), ob.view(np.uint32))
    np.testing.assert_array_equal(ma, mb_)
    assert [bytes(m) for m in mata] == [bytes(m) for m in matb]
    assert bytes(s.pack()) == bytes(r.pack())


def test_load_errors_are_reported(tmp_path):
    with pytest.raises(api.LuzError, match="cannot read"):
        api.HostScene.load(tmp_path / "missing.glb")
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nv 1 0 0\nf 1 2 7\n")
    with pytest.raises(api.LuzError, match="out of range"):
        api.HostScene.load(bad)
    j = tmp_path / "bad.luz"
    j.write_text("{\"assets\": [1, 2,}")
    with pytest.raises(api.LuzError, match="JSON"):
        api.HostScene.load(j)
