"""GPU parity of the precision form: 8-wide BLASes with 8-bit child boxes
(bytes identical to the oracle's), FP32 watertight visibility (identical
counts), packed-FP16 shading (radiance error against the FP64 oracle)."""
import numpy as np
import pytest

from luz_amd import api
from oracle import oracle as orc
from tests.helpers import oracle_scene, random_rays

pytestmark = pytest.mark.gpu


@pytest.fixture()
def compact(ctx):
    ctx.set_precision(api.PRECISION_COMPACT)
    yield ctx
    ctx.set_precision(api.PRECISION_PLAIN)


@pytest.mark.parametrize("kind", [api.SCENE_CUBE, api.SCENE_SPHERES])
def test_wide_bytes_match_oracle(compact, kind):
    s = api.HostScene(kind, 0)
    g = api.GPUScene(compact, s)
    o = oracle_scene(s)
    o.build_wide()
    for m in range(s.mesh_count):
        gn, gt, gp = g.blas(m).wide_arrays()
        on, ot, op = o.blas_wide_arrays(m)
        print(f"kind {kind} mesh {m}: {len(on)} wide nodes", flush=True)
        np.testing.assert_array_equal(gp, op)
        np.testing.assert_array_equal(gt.view(np.uint32), ot.view(np.uint32))
        np.testing.assert_array_equal(gn, on)
    g.close()


def test_wide_trace_rays_match_oracle(compact):
    s = api.HostScene(api.SCENE_SPHERES, 0)
    g = api.GPUScene(compact, s)
    o = oracle_scene(s)
    o.build_wide()
    rng = np.random.default_rng(17)
    rays = np.concatenate([random_rays(rng, 40000, (0, 0, 0), 60.0, 200.0),
                           random_rays(rng, 40000, (0, 0, 0), 50.0, 3.0, short=True)])
    want, _ = o.any_hit_wide(rays)
    dr = compact.to_device(rays)
    dh = compact.buffer(len(rays))
    compact.trace_rays(g.tlas, dr, len(rays), dh)
    np.testing.assert_array_equal(dh.download(np.uint8, len(rays)), want)
    g.close()


def test_compact_visibility_bit_exact_and_fp16_radiance(compact):
    s = api.HostScene(api.SCENE_SPHERES, 0)
    g = api.GPUScene(compact, s)
    o = oracle_scene(s)
    o.build_wide()
    W, H = 480, 270
    _, _, cam = s.camera()
    _, _, mats = s.nodes()
    gd = api.GBufferDevice(compact, W, H)
    compact.primary_gbuffer(g.tlas, cam, mats, gd.desc())
    host = gd.download()
    sb = s.pack()
    dsh, dao = compact.buffer(sb.num_lights * W * H), compact.buffer(W * H)
    compact.trace_visibility(g.tlas, gd.desc(), sb, dsh, dao)
    gsh = dsh.download(np.uint8, (sb.num_lights, W * H))
    gao = dao.download(np.uint8, W * H)
    wsh, wao, st = o.trace_visibility(host, W, H, sb, wide=True)
    np.testing.assert_array_equal(gsh, wsh)
    np.testing.assert_array_equal(gao, wao)
    out = compact.buffer(W * H * 16)
    compact.shade(gd.desc(), sb, dsh, dao, out)
    rad = out.download(np.float32, (W * H, 4))[:, :3].astype(np.float64)
    ref = orc.shade(host, W, H, sb, wsh, wao, f64=True)[:, :3]
    act = np.abs(host["normal"]).sum(1) != 0
    rmse = np.sqrt(np.mean((rad[act] - ref[act]) ** 2))
    scale = np.sqrt(np.mean(ref[act] ** 2))
    print(f"FP16 shading: radiance RMSE {rmse:.3g} (relative {rmse / scale:.3g})", flush=True)
    assert np.isfinite(rad).all()
    assert rmse / scale < 5e-3
    g.close()


@pytest.mark.parametrize("ao_max", [None, 16.0])
def test_wide_mode_equals_plain(ctx, ao_max):
    """PRECISION_WIDE traces the 8-wide BVH and shades in FP32.  On the sphere
    field, with the scene's AO radius and with radius 16 (where AO rays meet
    occluders), its counts equal the oracle's binary FP32 traversal and its
    light-pass radiance equals PLAIN's bit for bit, from the same G-buffer."""
    s = api.HostScene(api.SCENE_SPHERES, 0)
    g = api.GPUScene(ctx, s)
    o = oracle_scene(s)
    W, H = 480, 270
    _, _, cam = s.camera()
    _, _, mats = s.nodes()
    gd = api.GBufferDevice(ctx, W, H)
    ctx.primary_gbuffer(g.tlas, cam, mats, gd.desc())
    host = gd.download()
    sb = s.pack()
    if ao_max is not None:
        sb.ao_max = ao_max
    nl = sb.num_lights
    out = ctx.buffer(W * H * 16)
    ctx.light_pass(g.tlas, gd.desc(), sb, out)
    want = out.download(np.float32, (W * H, 4))
    try:
        ctx.set_precision(api.PRECISION_WIDE)
        g.update_resources_gpu(s)                 # the TLAS over the 8-wide BLASes
        dsh, dao = ctx.buffer(nl * W * H), ctx.buffer(W * H)
        ctx.trace_visibility(g.tlas, gd.desc(), sb, dsh, dao)
        gsh = dsh.download(np.uint8, (nl, W * H))
        gao = dao.download(np.uint8, W * H)
        out2 = ctx.buffer(W * H * 16)
        ctx.light_pass(g.tlas, gd.desc(), sb, out2)
        got = out2.download(np.float32, (W * H, 4))
    finally:
        ctx.set_precision(api.PRECISION_PLAIN)
    wsh, wao, _ = o.trace_visibility(host, W, H, sb)
    print(f"ao_max {ao_max}: shadowed {int(wsh.sum())}, AO misses {int(wao.sum())}", flush=True)
    np.testing.assert_array_equal(gsh, wsh)
    np.testing.assert_array_equal(gao, wao)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    g.close()
