"""Ingested scenes through the GPU path: the reference's cube.glb and its
OBJ light gizmos, placed on a ground plane with a camera and lights added
through the scene API, render G-buffers and visibility bit-exact with the
oracle."""
from pathlib import Path

import numpy as np
import pytest

from luz_amd import api
from tests.helpers import oracle_scene

pytestmark = pytest.mark.gpu

ASSETS = Path(__file__).resolve().parent / "golden" / "assets"


@pytest.mark.parametrize("name", ["cube.glb", "point.obj", "spot.obj"])
def test_ingested_scene_bit_exact(ctx, name):
    s = api.HostScene.load(ASSETS / name)
    W, H = 320, 240
    s.set_camera((4.0, 3.0, 5.0), (0.0, 0.0, 0.0), 45.0, W, H)
    s.add_light(api.LIGHT_DIRECTIONAL, (10.0, 30.0, 20.0), direction=(-0.3, -1.0, -0.5), intensity=3.0)
    s.add_light(api.LIGHT_POINT, (3.0, 4.0, -2.0), intensity=50.0, radius=0.5)
    s.set_samples(2, 8, 1e-4, 1.0)
    g = api.GPUScene(ctx, s)
    o = oracle_scene(s)
    _, _, cam = s.camera()
    _, _, mats = s.nodes()
    gd = api.GBufferDevice(ctx, W, H)
    ctx.primary_gbuffer(g.tlas, cam, mats, gd.desc())
    got = gd.download()
    want = o.primary_gbuffer(cam, mats, W, H)
    for k in ("position", "normal"):
        np.testing.assert_array_equal(got[k].view(np.uint32), want[k].view(np.uint32), err_msg=k)
    fg = int((np.abs(got["normal"]).sum(1) != 0).sum())
    sb = s.pack()
    nl = sb.num_lights
    dsh, dao = ctx.buffer(nl * W * H), ctx.buffer(W * H)
    ctx.trace_visibility(g.tlas, gd.desc(), sb, dsh, dao)
    wsh, wao, _ = o.trace_visibility(got, W, H, sb)
    print(f"{name}: {fg} foreground pixels, {int(wsh.sum())} occluded shadow samples, "
          f"{int(wao.sum())} AO misses", flush=True)
    assert fg > 1000
    np.testing.assert_array_equal(dsh.download(np.uint8, (nl, W * H)), wsh)
    np.testing.assert_array_equal(dao.download(np.uint8, W * H), wao)
    g.close()
