"""Error behaviour of the C ABI on a GPU: arguments over the limits or
malformed are rejected with the documented status and a message, before any
kernel runs, and the context stays usable afterwards (the reference would
overflow its fixed-size buffers or raise Vulkan validation errors instead,
GPUScene.cpp:51,240)."""
import ctypes as C

import numpy as np
import pytest

from luz_amd import api
from tests.helpers import set_light_direction

pytestmark = pytest.mark.gpu

E_INVALID_ARG, E_IN_USE, E_LIMIT = -1, -4, -5


def _status(fn):
    with pytest.raises(api.LuzError) as e:
        fn()
    return e.value.status


def _cube_scene(ctx):
    s = api.HostScene(api.SCENE_CUBE, 0)
    g = api.GPUScene(ctx, s)
    W, H, cam = s.camera()
    _, _, mats = s.nodes()
    gd = api.GBufferDevice(ctx, W, H)
    ctx.primary_gbuffer(g.tlas, cam, mats, gd.desc())
    return s, g, gd, W, H


def test_visibility_rejects_bad_scene_blocks(ctx):
    s, g, gd, W, H = _cube_scene(ctx)
    sh, ao = ctx.buffer(W * H * 64), ctx.buffer(W * H)

    def run(mutate):
        sb = s.pack()
        mutate(sb)
        return _status(lambda: ctx.trace_visibility(g.tlas, gd.desc(), sb, sh, ao))

    assert run(lambda sb: setattr(sb, "num_lights", 65)) == E_LIMIT
    assert b"num_lights 65" in api.lib.luzrt_last_error(ctx.handle)
    assert run(lambda sb: setattr(sb, "ao_num_samples", 256)) == E_LIMIT
    assert run(lambda sb: setattr(sb.lights[0], "num_shadow_samples", 256)) == E_LIMIT
    assert run(lambda sb: setattr(sb.lights[0], "type", 7)) == E_INVALID_ARG
    desc = gd.desc()
    desc.width = 0
    assert _status(lambda: ctx.trace_visibility(g.tlas, desc, s.pack(), sh, ao)) == E_LIMIT
    desc = gd.desc()
    desc.normal = None
    assert _status(lambda: ctx.trace_visibility(g.tlas, desc, s.pack(), sh, ao)) == E_INVALID_ARG
    # the context still works: a valid call after the rejected ones
    sb = s.pack()
    set_light_direction(sb, 0, (0.3, -1.0, 0.2))
    ctx.trace_visibility(g.tlas, gd.desc(), sb, sh, ao)
    assert ao.download(np.uint8, W * H).max() <= sb.ao_num_samples
    g.close()


def test_builds_reject_malformed_geometry(ctx):
    v = np.zeros((3, 12), np.float32)
    v[1, 0] = 1.0
    v[2, 2] = 1.0
    dv = ctx.to_device(v)
    di = ctx.to_device(np.array([0, 1, 2], np.uint32))
    assert _status(lambda: ctx.create_blas(dv, 48, 3, di, 2)) == E_INVALID_ARG        # not triangles
    assert _status(lambda: ctx.create_blas(dv, 10, 3, di, 3)) == E_INVALID_ARG        # bad stride
    assert _status(lambda: ctx.create_blas(dv, 48, 0, di, 3)) == E_INVALID_ARG        # no vertices
    bad = ctx.to_device(np.array([0, 1, 7], np.uint32))
    assert _status(lambda: ctx.create_blas(dv, 48, 3, bad, 3)) == E_INVALID_ARG       # index >= count
    blas = ctx.create_blas(dv, 48, 3, di, 3)
    tlas = ctx.create_tlas(2)
    eye = np.eye(3, 4, dtype=np.float32)
    three = [api.make_instance(blas, eye, i) for i in range(3)]
    assert _status(lambda: tlas.build(three)) == E_LIMIT                              # over capacity
    singular = np.zeros((3, 4), np.float32)
    assert _status(lambda: tlas.build([api.make_instance(blas, singular, 0)])) == E_INVALID_ARG
    nan = eye.copy()
    nan[0, 3] = np.nan
    assert _status(lambda: tlas.build([api.make_instance(blas, nan, 0)])) == E_INVALID_ARG
    tlas.build([api.make_instance(blas, eye, 0)])
    assert _status(blas.destroy) == E_IN_USE                                          # still referenced
    tlas.destroy()
    blas.destroy()


def test_denoiser_rejects_bad_parameters(ctx):
    from luz_amd import denoiser as dnm
    dn = dnm.Denoiser(ctx, 0)
    p = np.zeros(dnm.N_PARAMS, np.float32)
    p[5] = np.inf
    assert _status(lambda: setattr(dn, "params", p)) == E_INVALID_ARG
    short = np.zeros(dnm.N_PARAMS - 1, np.float32)
    rc = api.lib.luzrt_denoiser_set_params(ctx.handle, dn.handle,
                                           short.ctypes.data_as(C.POINTER(C.c_float)), len(short))
    assert rc == E_INVALID_ARG
    dn.close()


def test_traversal_overflow_is_reported_by_the_next_sync(ctx):
    """A ray that exceeds its step limit (forced here from the host with a
    limit of 3 steps; kernels still end normally) is written as a miss and
    latched on the device: the next luzrt_sync returns LUZRT_E_LIMIT once,
    although a further call ran in between, and the condition is then clear."""
    s, g, gd, W, H = _cube_scene(ctx)
    sb = s.pack()
    set_light_direction(sb, 0, (0.8, -1.0, 0.6))
    sh, ao = ctx.buffer(W * H), ctx.buffer(W * H)
    ctx.sync()
    ctx.set_step_limit(3)
    try:
        ctx.trace_visibility(g.tlas, gd.desc(), sb, sh, ao)
        # a second call before the check does not clear the latch
        ctx.trace_visibility(g.tlas, gd.desc(), sb, sh, ao)
        assert _status(ctx.sync) == E_LIMIT
        assert b"step limit" in api.lib.luzrt_last_error(ctx.handle)
        ctx.sync()
    finally:
        ctx.set_step_limit(0)
    ctx.trace_visibility(g.tlas, gd.desc(), sb, sh, ao)
    ctx.sync()
    from tests.helpers import oracle_scene
    host = gd.download()
    wsh, wao, _ = oracle_scene(s).trace_visibility(host, W, H, sb)
    np.testing.assert_array_equal(sh.download(np.uint8, (1, W * H)), wsh)
    np.testing.assert_array_equal(ao.download(np.uint8, W * H), wao)
    g.close()


def test_batch_rejects_shared_tlas_and_outputs(ctx):
    s, g, gd, W, H = _cube_scene(ctx)
    W2, H2, cam = s.camera()
    _, _, mats = s.nodes()
    sb = s.pack()
    other = api.GPUScene(ctx, s)
    gd2 = api.GBufferDevice(ctx, W, H)

    def entry(tlas, gbuf, bufs, camera=None):
        return {"tlas": tlas, "gbuffer": gbuf.desc(), "scene": sb, "camera": camera, "materials": mats,
                "shadow": bufs[0], "ao": bufs[1], "out": bufs[2]}

    def bufs():
        return ctx.buffer(W * H), ctx.buffer(W * H), ctx.buffer(W * H * 16)

    a, b = bufs(), bufs()
    assert _status(lambda: api.Batch(ctx, [entry(g.tlas, gd, a), entry(g.tlas, gd2, b)])) == E_INVALID_ARG
    assert b"same TLAS" in api.lib.luzrt_last_error(ctx.handle)
    assert _status(lambda: api.Batch(ctx, [entry(g.tlas, gd, a), entry(other.tlas, gd2, a)])) == E_INVALID_ARG
    assert _status(lambda: api.Batch(ctx, [entry(g.tlas, gd, a, cam), entry(other.tlas, gd, b)])) == E_INVALID_ARG
    # G-buffers that no camera writes may be shared
    ok = api.Batch(ctx, [entry(g.tlas, gd, a), entry(other.tlas, gd, b)])
    ok.close()
    other.close()
    g.close()
