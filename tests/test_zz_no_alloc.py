"""The allocation-free visibility entry point, luzrt_trace_visibility_static,
called through the C ABI: its counts equal luzrt_trace_visibility's on two
calls in a row (the second on a smaller image, after the first left its
records and hits in the shared scratch), and sizes over its limits are
rejected before anything runs."""
import ctypes as C

import numpy as np
import pytest

from luz_amd import api
from tests.helpers import set_light_direction

pytestmark = pytest.mark.gpu

E_LIMIT = -5


def _frame(ctx, kind, W, H):
    s = api.HostScene(kind, 0)
    g = api.GPUScene(ctx, s)
    _, _, cam = s.camera()
    _, _, mats = s.nodes()
    gd = api.GBufferDevice(ctx, W, H)
    ctx.primary_gbuffer(g.tlas, cam, mats, gd.desc())
    return s, g, gd


def _both(ctx, g, gd, sb, W, H):
    nl = max(sb.num_lights, 1)
    out = []
    for fn in (api.lib.luzrt_trace_visibility_static, api.lib.luzrt_trace_visibility):
        sh, ao = ctx.buffer(nl * W * H), ctx.buffer(W * H)
        desc = gd.desc()
        rc = fn(ctx.handle, g.tlas.handle, C.byref(desc), C.byref(sb), sh.ptr, ao.ptr)
        assert rc == 0, api.lib.luzrt_last_error(ctx.handle)
        ctx.sync()
        out.append((sh.download(np.uint8, (nl, W * H)), ao.download(np.uint8, W * H), ctx.timings().rays))
    return out


def test_static_visibility_equals_the_allocating_call(ctx):
    # first call: the 1080p sphere field at 720x540 (388,800 pixels, 25 ray slots)
    s, g, gd = _frame(ctx, api.SCENE_SPHERES, 720, 540)
    sb = s.pack()
    (ssh, sao, srays), (wsh, wao, wrays) = _both(ctx, g, gd, sb, 720, 540)
    print(f"sphere field 720x540: {srays} rays, shadowed {int(wsh.sum())}", flush=True)
    assert srays == wrays and wsh.sum() > 0
    np.testing.assert_array_equal(ssh, wsh)
    np.testing.assert_array_equal(sao, wao)
    g.close()
    # second call, smaller: the cube frame at 256x256 with its shadow towards the camera
    s2, g2, gd2 = _frame(ctx, api.SCENE_CUBE, 256, 256)
    sb2 = s2.pack()
    set_light_direction(sb2, 0, (0.8, -1.0, 0.6))
    (ssh, sao, srays), (wsh, wao, wrays) = _both(ctx, g2, gd2, sb2, 256, 256)
    print(f"cube 256x256: {srays} rays, shadowed {int(wsh.sum())}", flush=True)
    assert srays == wrays and wsh.sum() > 0
    np.testing.assert_array_equal(ssh, wsh)
    np.testing.assert_array_equal(sao, wao)
    g2.close()


def test_static_visibility_rejects_sizes_over_its_limits(ctx):
    s, g, gd = _frame(ctx, api.SCENE_CUBE, 256, 256)
    sb = s.pack()
    desc = gd.desc()
    desc.width, desc.height = 1024, 1024          # 2^20 pixels (checked before any launch)
    sh, ao = ctx.buffer(1024 * 1024), ctx.buffer(1024 * 1024)
    assert api.lib.luzrt_trace_visibility_static(ctx.handle, g.tlas.handle, C.byref(desc), C.byref(sb),
                                                 sh.ptr, ao.ptr) == E_LIMIT
    assert b"pixels" in api.lib.luzrt_last_error(ctx.handle)
    sb.ao_num_samples = 255
    sb.lights[0].num_shadow_samples = 255          # 256x256 x 510 slots > 2^24 rays
    assert api.lib.luzrt_trace_visibility_static(ctx.handle, g.tlas.handle, C.byref(gd.desc()), C.byref(sb),
                                                 sh.ptr, ao.ptr) == E_LIMIT
    g.close()
