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
