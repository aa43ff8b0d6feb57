"""Writes the golden fixtures of the cube workload (256x256) with the CPU oracle.

The reference has no tests or golden images (its shaders run in the Vulkan
driver); these fixtures freeze the oracle's outputs so that any change of the
arithmetic specification shows up as a diff, and the HIP path is compared
against the same bytes.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from luz_amd import api  # noqa: E402
from oracle import oracle as orc  # noqa: E402

LIGHT_VARIANTS = {"config": None, "towards_camera": (0.8, -1.0, 0.6)}


def cube_case(variant: str):
    s = api.HostScene(api.SCENE_CUBE, 0)
    o = orc.OracleScene()
    for i in range(s.mesh_count):
        v, idx = s.mesh(i)
        o.add_mesh(v, idx)
    o2w, mesh, mats = s.nodes()
    o.set_instances(o2w, mesh)
    W, H, cam = s.camera()
    g = o.primary_gbuffer(cam, mats, W, H, threads=1)
    sb = s.pack()
    d = LIGHT_VARIANTS[variant]
    if d is not None:
        d = np.asarray(d, np.float64) / np.linalg.norm(d)
        for k in range(3):
            sb.lights[0].direction[k] = float(d[k])
            sb.lights[0].position[k] = float(-1000.0 * d[k])
    sh, ao, st = o.trace_visibility(g, W, H, sb, threads=1)
    return s, g, sb, sh, ao, st


def gbuffer_digest(g: dict) -> str:
    h = hashlib.sha256()
    for k in ("position", "normal", "albedo", "material", "emission"):
        h.update(np.ascontiguousarray(g[k]).tobytes())
    return h.hexdigest()


def main() -> None:
    for variant in LIGHT_VARIANTS:
        _, g, _, sh, ao, st = cube_case(variant)
        np.savez_compressed(HERE / f"cube_{variant}.npz", shadow=sh, ao=ao,
                            gbuffer_sha256=np.array(gbuffer_digest(g)),
                            rays=np.array(st.rays), node_visits=np.array(st.node_visits),
                            tri_tests=np.array(st.tri_tests), inst_visits=np.array(st.inst_visits))
        print(variant, "rays", st.rays, "shadowed", int(sh.sum()), "ao misses", int(ao.sum()))


if __name__ == "__main__":
    main()
