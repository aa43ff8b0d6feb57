"""Shared set-up for parity tests: the same scene on the oracle and the HIP path."""
from __future__ import annotations

import ctypes as C

import numpy as np

from luz_amd import api
from oracle import oracle as orc


def oracle_scene(scene: api.HostScene) -> orc.OracleScene:
    o = orc.OracleScene()
    for i in range(scene.mesh_count):
        v, idx = scene.mesh(i)
        o.add_mesh(v, idx)
    o2w, mesh, _ = scene.nodes()
    o.set_instances(o2w, mesh)
    return o


def random_rays(rng: np.random.Generator, n: int, center, extent: float, tmax: float = 50.0,
                short: bool = False) -> np.ndarray:
    """Rays from points around `center` in random directions (8 floats each)."""
    o = np.asarray(center, np.float32) + rng.uniform(-extent, extent, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    if short:
        d *= rng.uniform(0.2, 2.0, (n, 1)).astype(np.float32)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = rng.uniform(0.0, 0.01, n).astype(np.float32)
    r[:, 4:7] = d
    r[:, 7] = rng.uniform(0.5, tmax, n).astype(np.float32)
    return r


def set_light_direction(sb: api.SceneBlock, index: int, d) -> None:
    d = np.asarray(d, np.float64)
    d = d / np.linalg.norm(d)
    for k in range(3):
        sb.lights[index].direction[k] = float(d[k])
        sb.lights[index].position[k] = float(-d[k] * 1000.0)


def denoise_rmse(ctx, dn, fr) -> dict:
    """RMSE over foreground pixels against the frame's multi-spp targets of the
    raw 1-spp estimate, the 5x5 box filter (non-learned) and the network."""
    from oracle import denoise as odn
    n = fr.width * fr.height
    L = fr.num_lights
    ao_o, vis_o = ctx.buffer(n * 4), ctx.buffer(max(L, 1) * n * 4)
    dn.denoise(fr, ao_o, vis_o)
    ao_d = ao_o.download(np.float32, n).astype(np.float64)
    vis_d = vis_o.download(np.float32, (max(L, 1), n)).astype(np.float64)[:L]
    host = fr.host()
    fg = fr.foreground
    ao_raw = host["ao"].astype(np.float64) / host["ao_n"]
    vis_raw = 1.0 - host["shadow"].astype(np.float64) / max(host["shadow_n"], 1)
    box = [odn.box_filter(host, l) for l in range(max(L, 1))]
    box_ao = np.mean([b[0] for b in box], 0)
    box_vis = np.array([b[1] for b in box[:L]])

    def r(a, b):
        return float(np.sqrt(np.mean((a[..., fg] - b[..., fg]) ** 2)))

    ta, tv = fr.target_ao_host.astype(np.float64), fr.target_vis_host.astype(np.float64)
    return {"ao": {"raw": r(ao_raw, ta), "box": r(box_ao, ta), "mlp": r(ao_d, ta)},
            "vis": {"raw": r(vis_raw, tv), "box": r(box_vis, tv), "mlp": r(vis_d, tv)}}


def cube_mesh():
    """Unit cube [0,1]^3: 8 vertices (48-byte MeshVertex rows), 12 triangles."""
    v = np.zeros((8, 12), np.float32)
    for k in range(8):
        v[k, :3] = (k & 1, (k >> 1) & 1, (k >> 2) & 1)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    idx = [i for a, b, c, d in quads for i in (a, b, c, a, c, d)]
    return v, np.array(idx, np.uint32)


def scaled_instances(xyz, scale):
    """Object-to-world rows (12 floats) of uniform scale `scale[i]` and translation `xyz[i]`."""
    m = np.zeros((len(xyz), 12), np.float32)
    for i, ((x, y, z), s) in enumerate(zip(xyz, scale)):
        m[i] = [s, 0, 0, x, 0, s, 0, y, 0, 0, s, z]
    return m
