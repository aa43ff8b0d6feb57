"""GPU parity of the HIP path against the CPU oracle.

Integer results (BVH bytes, G-buffer bytes, shadow/AO counts, hit flags) must
be identical; radiance within 1e-3 relative (|gpu - cpu| <= 1e-3 * max(|cpu|, 1e-3)).
"""
import numpy as np
import pytest

from luz_amd import api
from oracle import oracle as orc
from tests.helpers import cube_mesh, oracle_scene, random_rays, scaled_instances, set_light_direction

pytestmark = pytest.mark.gpu

RADIANCE_RTOL = 1e-3


def _gpu_scene(ctx, kind, seed=0):
    s = api.HostScene(kind, seed)
    g = api.GPUScene(ctx, s)
    return s, g


def test_blas_bytes_match_oracle(ctx):
    for kind in (api.SCENE_CUBE, api.SCENE_SPHERES):
        s, g = _gpu_scene(ctx, kind)
        o = oracle_scene(s)
        for m in range(s.mesh_count):
            gn, gt, gp = g.blas(m).arrays()
            on, ot, op = o.blas_arrays(m)
            print(f"scene {kind} mesh {m}: {len(op)} tris, {len(on)} nodes", flush=True)
            assert gn.shape == on.shape
            np.testing.assert_array_equal(gp, op)
            np.testing.assert_array_equal(gt.view(np.uint32), ot.view(np.uint32))
            np.testing.assert_array_equal(gn, on)
        g.close()


def test_tlas_bytes_match_oracle(ctx):
    s, g = _gpu_scene(ctx, api.SCENE_SPHERES)
    o = oracle_scene(s)
    gn, gw, go = g.tlas.arrays()
    on, ow, oo = o.tlas_arrays()
    np.testing.assert_array_equal(go, oo)
    np.testing.assert_array_equal(gw.view(np.uint32), ow.view(np.uint32))
    np.testing.assert_array_equal(gn, on)
    g.close()


@pytest.mark.parametrize("count,ties", [(2, False), (3, False), (17, False), (40, True), (1500, False),
                                         (4100, False), (8192, False)])
def test_tlas_clustering_bytes_match_oracle(ctx, count, ties):
    """TLAS hierarchy by clustering rounds: one tile of clusters up to 1024,
    several above, the 8192-instance maximum, and all-equal boxes (every merge
    decided by the index tie-break)."""
    rng = np.random.default_rng(count)
    v, idx = cube_mesh()
    blas = ctx.create_blas(ctx.to_device(v), 48, len(v), ctx.to_device(idx), len(idx))
    if ties:
        xf = scaled_instances(np.zeros((count, 3)), np.ones(count))
    else:
        xf = scaled_instances(rng.uniform(-200, 200, (count, 3)), rng.uniform(0.2, 5.0, count))
    tlas = ctx.create_tlas(count)
    print(f"TLAS of {count} instances (ties {ties})", flush=True)
    tlas.build([api.make_instance(blas, xf[i], i) for i in range(count)])
    gn, gw, go = tlas.arrays()
    o = orc.OracleScene()
    m = o.add_mesh(v, idx)
    o.set_instances(xf, [m] * count)
    on, ow, oo = o.tlas_arrays()
    np.testing.assert_array_equal(go, oo)
    np.testing.assert_array_equal(gw.view(np.uint32), ow.view(np.uint32))
    np.testing.assert_array_equal(gn, on)
    tlas.destroy()
    blas.destroy()


def test_tlas_rebuilt_between_calls_without_waiting(ctx):
    """TLAS builds run on a stream of their own.  Four rebuilds of one TLAS
    with different instance placements, each followed at once by a ray call
    into its own buffer, nothing waited for in between: every call sees the
    instances of the build before it and none of the build after it."""
    rng = np.random.default_rng(5)
    v, idx = cube_mesh()
    blas = ctx.create_blas(ctx.to_device(v), 48, len(v), ctx.to_device(idx), len(idx))
    count = 40
    tlas = ctx.create_tlas(count)
    rays = random_rays(rng, 30000, (0, 0, 0), 30.0, 60.0)
    dr = ctx.to_device(rays)
    placements, outs = [], []
    for k in range(4):
        xf = scaled_instances(rng.uniform(-25, 25, (count, 3)), rng.uniform(0.5, 3.0, count))
        tlas.build([api.make_instance(blas, xf[i], i) for i in range(count)])
        dh = ctx.buffer(len(rays))
        ctx.trace_rays(tlas, dr, len(rays), dh)
        placements.append(xf)
        outs.append(dh)
    got = [dh.download(np.uint8, len(rays)) for dh in outs]
    for k, xf in enumerate(placements):
        o = orc.OracleScene()
        m = o.add_mesh(v, idx)
        o.set_instances(xf, [m] * count)
        want, _ = o.any_hit(rays)
        print(f"build {k}: {int(want.sum())} of {len(rays)} rays hit", flush=True)
        assert 0 < want.sum() < len(rays)
        np.testing.assert_array_equal(got[k], want)
    tlas.destroy()
    blas.destroy()


@pytest.mark.parametrize("kind", [api.SCENE_CUBE, api.SCENE_SPHERES])
def test_trace_rays_match_oracle(ctx, kind):
    s, g = _gpu_scene(ctx, kind)
    o = oracle_scene(s)
    rng = np.random.default_rng(7)
    if kind == api.SCENE_CUBE:
        rays = np.concatenate([random_rays(rng, 20000, (0, 0.5, 0), 2.0, 20.0),
                               random_rays(rng, 20000, (0, 0.5, 0), 1.0, 2.0, short=True)])
    else:
        rays = np.concatenate([random_rays(rng, 40000, (0, 0, 0), 60.0, 200.0),
                               random_rays(rng, 40000, (0, 0, 0), 50.0, 3.0, short=True)])
    want, _ = o.any_hit(rays)
    dr = ctx.to_device(rays)
    dh = ctx.buffer(len(rays))
    ctx.trace_rays(g.tlas, dr, len(rays), dh)
    got = dh.download(np.uint8, len(rays))
    print(f"kind {kind}: {int(want.sum())} of {len(rays)} rays hit", flush=True)
    assert 0 < want.sum() < len(rays)
    np.testing.assert_array_equal(got, want)
    g.close()


def _gbuffers(ctx, s, g, o, width=None, height=None):
    W, H, cam = s.camera()
    if width:
        W, H = width, height
    _, _, mats = s.nodes()
    gd = api.GBufferDevice(ctx, W, H)
    ctx.primary_gbuffer(g.tlas, cam, mats, gd.desc())
    got = gd.download()
    want = o.primary_gbuffer(cam, mats, W, H)
    return W, H, gd, got, want


@pytest.mark.parametrize("kind,res", [(api.SCENE_CUBE, None), (api.SCENE_SPHERES, (480, 270))])
def test_primary_gbuffer_bit_exact(ctx, kind, res):
    s, g = _gpu_scene(ctx, kind)
    o = oracle_scene(s)
    W, H, gd, got, want = _gbuffers(ctx, s, g, o, *(res or (None, None)))
    for k in ("position", "normal"):
        np.testing.assert_array_equal(got[k].view(np.uint32), want[k].view(np.uint32), err_msg=k)
    for k in ("albedo", "material", "emission"):
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    g.close()


@pytest.mark.parametrize("kind,res,light_dir", [
    (api.SCENE_CUBE, None, None),
    (api.SCENE_CUBE, None, (0.8, -1.0, 0.6)),      # shadow falls towards the camera
    (api.SCENE_SPHERES, (480, 270), None),
])
def test_visibility_bit_exact(ctx, kind, res, light_dir):
    s, g = _gpu_scene(ctx, kind)
    o = oracle_scene(s)
    W, H, gd, got, _ = _gbuffers(ctx, s, g, o, *(res or (None, None)))
    sb = s.pack()
    if light_dir is not None:
        set_light_direction(sb, 0, light_dir)
    nl = sb.num_lights
    dsh = ctx.buffer(nl * W * H)
    dao = ctx.buffer(W * H)
    ctx.trace_visibility(g.tlas, gd.desc(), sb, dsh, dao)
    gsh = dsh.download(np.uint8, (nl, W * H))
    gao = dao.download(np.uint8, W * H)
    t = ctx.timings()
    wsh, wao, st = o.trace_visibility(got, W, H, sb)
    print(f"kind {kind} {W}x{H}: {t.rays} rays (oracle {st.rays}), shadowed {int(wsh.sum())}, "
          f"AO misses {int(wao.sum())}", flush=True)
    assert t.rays == st.rays
    np.testing.assert_array_equal(gsh, wsh)
    np.testing.assert_array_equal(gao, wao)
    if light_dir is not None:
        assert wsh.sum() > 100
    g.close()


@pytest.mark.parametrize("order", [api.RAY_ORDER_SLOT, api.RAY_ORDER_GROUPED, api.RAY_ORDER_AUTO])
def test_visibility_same_in_every_ray_order(ctx, order):
    """The order in which a visibility call claims its rays (all pixels' first
    sample first, or the samples of 64 neighbouring pixels together) changes
    no count: both against the oracle, on an image whose active-pixel count
    leaves a partial last group of pixels."""
    s, g = _gpu_scene(ctx, api.SCENE_SPHERES)
    o = oracle_scene(s)
    W, H, gd, got, _ = _gbuffers(ctx, s, g, o, 331, 187)
    active = int((np.abs(got["normal"]).sum(1) != 0).sum())
    assert active % 64 != 0
    sb = s.pack()
    nl = sb.num_lights
    dsh, dao = ctx.buffer(nl * W * H), ctx.buffer(W * H)
    ctx.set_ray_order(order)
    try:
        print(f"ray order {order}: {active} active pixels", flush=True)
        ctx.trace_visibility(g.tlas, gd.desc(), sb, dsh, dao)
    finally:
        ctx.set_ray_order(api.RAY_ORDER_AUTO)
    wsh, wao, _ = o.trace_visibility(got, W, H, sb)
    np.testing.assert_array_equal(dsh.download(np.uint8, (nl, W * H)), wsh)
    np.testing.assert_array_equal(dao.download(np.uint8, W * H), wao)
    g.close()


@pytest.mark.parametrize("ao_max,radius", [(1.0, None), (16.0, None), (16.0, 3.0)])
def test_group_culling_changes_no_count(ctx, ao_max, radius):
    """Ray-group culling (on by default) against the oracle and against the
    same call without it, on the sphere field with a partial last group of
    pixels: AO radius 1 (beyond the reach of every group's AO rays) and 16
    (occluders within reach), and point lights of radius 3 (wide shadow
    shafts).  At AO radius 1 the culled call traces fewer rays than it casts
    (at 16 every group of this small image can reach an instance)."""
    s, g = _gpu_scene(ctx, api.SCENE_SPHERES)
    o = oracle_scene(s)
    W, H, gd, got, _ = _gbuffers(ctx, s, g, o, 331, 187)
    sb = s.pack()
    sb.ao_max = ao_max
    if radius is not None:
        for l in range(sb.num_lights):
            if sb.lights[l].type == api.LIGHT_POINT:
                sb.lights[l].radius = radius
    nl = sb.num_lights
    out = {}
    for cull in (True, False):
        dsh, dao = ctx.buffer(nl * W * H), ctx.buffer(W * H)
        ctx.set_group_culling(cull)
        try:
            print(f"ao_max {ao_max} radius {radius} culling {cull}", flush=True)
            ctx.trace_visibility(g.tlas, gd.desc(), sb, dsh, dao)
        finally:
            ctx.set_group_culling(True)
        t = ctx.timings()
        out[cull] = (dsh.download(np.uint8, (nl, W * H)), dao.download(np.uint8, W * H), t.traced_rays,
                     t.rays)
    wsh, wao, _ = o.trace_visibility(got, W, H, sb)
    print(f"rays {out[True][3]}, traced with culling {out[True][2]}, shadowed {int(wsh.sum())}, "
          f"AO hits {int((sb.ao_num_samples - wao.astype(int)).sum())}", flush=True)
    for cull in (True, False):
        np.testing.assert_array_equal(out[cull][0], wsh, err_msg=f"shadow, culling {cull}")
        np.testing.assert_array_equal(out[cull][1], wao, err_msg=f"AO, culling {cull}")
    assert out[False][2] == out[False][3]
    if ao_max == 1.0:
        assert out[True][2] < out[True][3]
    g.close()


def test_visibility_more_than_64_rays_per_pixel(ctx):
    """8 shadow samples per light and 24 AO samples (96 rays per pixel): past
    the 64 ray slots that ray-group culling and the resolve's hit mask cover,
    so the call traces every ray and resolves each pixel's counts slot by
    slot.  Counts against the oracle."""
    s, g = _gpu_scene(ctx, api.SCENE_SPHERES)
    o = oracle_scene(s)
    W, H, gd, got, _ = _gbuffers(ctx, s, g, o, 160, 96)
    sb = s.pack()
    sb.ao_max = 16.0
    for l in range(sb.num_lights):
        sb.lights[l].num_shadow_samples = 8
    sb.ao_num_samples = 24
    nl = sb.num_lights
    dsh, dao = ctx.buffer(nl * W * H), ctx.buffer(W * H)
    ctx.trace_visibility(g.tlas, gd.desc(), sb, dsh, dao)
    t = ctx.timings()
    print(f"{t.rays} rays, {t.traced_rays} traced", flush=True)
    assert t.traced_rays == t.rays
    wsh, wao, _ = o.trace_visibility(got, W, H, sb)
    np.testing.assert_array_equal(dsh.download(np.uint8, (nl, W * H)), wsh)
    np.testing.assert_array_equal(dao.download(np.uint8, W * H), wao)
    g.close()


@pytest.mark.parametrize("kind,res", [(api.SCENE_CUBE, None), (api.SCENE_SPHERES, (480, 270))])
def test_light_pass_radiance(ctx, kind, res):
    s, g = _gpu_scene(ctx, kind)
    o = oracle_scene(s)
    W, H, gd, got, _ = _gbuffers(ctx, s, g, o, *(res or (None, None)))
    sb = s.pack()
    out = ctx.buffer(W * H * 16)
    ctx.light_pass(g.tlas, gd.desc(), sb, out)
    rad = out.download(np.float32, (W * H, 4))
    wsh, wao, _ = o.trace_visibility(got, W, H, sb)
    want = orc.shade(got, W, H, sb, wsh, wao)
    err = np.abs(rad - want)
    tol = RADIANCE_RTOL * np.maximum(np.abs(want), 1e-3)
    print(f"max rel err {float((err / np.maximum(np.abs(want), 1e-3)).max()):.3g}", flush=True)
    assert np.all(err <= tol)
    g.close()


def test_light_pass_without_stage_timing(ctx):
    """With stage timing off the light pass writes the same radiance, bit for
    bit, the timings report 0 ms for every stage and still count the rays;
    turned back on, the stages are timed again."""
    s, g = _gpu_scene(ctx, api.SCENE_SPHERES)
    o = oracle_scene(s)
    W, H, gd, _, _ = _gbuffers(ctx, s, g, o, 480, 270)
    sb = s.pack()
    out = ctx.buffer(W * H * 16)
    ctx.light_pass(g.tlas, gd.desc(), sb, out)
    on = ctx.timings()
    want = out.download(np.float32, (W * H, 4))
    try:
        ctx.set_stage_timing(False)
        out2 = ctx.buffer(W * H * 16)
        ctx.light_pass(g.tlas, gd.desc(), sb, out2)
        off = ctx.timings()
        got = out2.download(np.float32, (W * H, 4))
    finally:
        ctx.set_stage_timing(True)
    print(f"timed: trace {on.trace_ms:.4f} ms, {on.rays} rays; untimed: {off.rays} rays", flush=True)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    assert on.trace_ms > 0 and on.shade_ms > 0
    assert (off.build_ms, off.compact_ms, off.raygen_ms, off.trace_ms, off.resolve_ms, off.shade_ms) == (0,) * 6
    assert off.rays == on.rays > 0 and off.traced_rays == on.traced_rays
    ctx.light_pass(g.tlas, gd.desc(), sb, out2)
    assert ctx.timings().trace_ms > 0
    g.close()


@pytest.mark.parametrize("shadow_type",[api.SHADOW_RAYTRACED, api.SHADOW_DISABLED, api.SHADOW_MAP])
def test_spot_light_and_shadow_types(ctx, shadow_type):
    """The cube scene with a spot light and a point light beside its
    directional light, two soft-shadow samples per light and 8 AO samples,
    in each shadow type: counts bit-exact and the light pass within 1e-3 of
    the oracle.  The spot's cone (light.frag:199-204, angles in radians)
    lights part of the frame; without ray-traced shadows no shadow ray is
    cast and every light is fully shadowed (light.frag:141-168), so only
    ambient and emission remain."""
    s = api.HostScene(api.SCENE_CUBE, 0)
    s.add_light(api.LIGHT_SPOT, (1.5, 3.0, 1.0), direction=(-0.45, -0.85, -0.3), color=(1.0, 0.8, 0.6),
                intensity=40.0, radius=0.3)
    s.add_light(api.LIGHT_POINT, (-2.0, 2.5, 1.5), color=(0.3, 0.5, 1.0), intensity=25.0, radius=0.2)
    s.set_samples(2, 8)
    s.set_shadow_type(shadow_type)
    g = api.GPUScene(ctx, s)
    o = oracle_scene(s)
    W, H, gd, got, _ = _gbuffers(ctx, s, g, o)
    sb = s.pack()
    assert sb.num_lights == 3 and sb.lights[1].type == api.LIGHT_SPOT
    nl = sb.num_lights
    dsh, dao = ctx.buffer(nl * W * H), ctx.buffer(W * H)
    ctx.trace_visibility(g.tlas, gd.desc(), sb, dsh, dao)
    gsh, gao = dsh.download(np.uint8, (nl, W * H)), dao.download(np.uint8, W * H)
    rays = ctx.timings().rays
    wsh, wao, st = o.trace_visibility(got, W, H, sb)
    active = int((np.abs(got["normal"]).sum(1) != 0).sum())
    per_px = 8 + (2 * nl if shadow_type == api.SHADOW_RAYTRACED else 0)
    print(f"shadow type {shadow_type}: {rays} rays, shadowed samples per light {wsh.sum(1)}", flush=True)
    assert rays == st.rays == active * per_px
    np.testing.assert_array_equal(gsh, wsh)
    np.testing.assert_array_equal(gao, wao)
    out = ctx.buffer(W * H * 16)
    ctx.light_pass(g.tlas, gd.desc(), sb, out)
    rad = out.download(np.float32, (W * H, 4))
    want = orc.shade(got, W, H, sb, wsh, wao)
    err = np.abs(rad - want) / np.maximum(np.abs(want), 1e-3)
    print(f"max rel err {float(err.max()):.3g}", flush=True)
    assert err.max() <= RADIANCE_RTOL
    # the spot light's share of the frame, from the FP64 oracle with and without it
    dark = api.SceneBlock()
    import ctypes as C
    C.memmove(C.byref(dark), C.byref(sb), C.sizeof(api.SceneBlock))
    dark.lights[1].intensity = 0.0
    spot = np.abs(orc.shade(got, W, H, sb, wsh, wao, f64=True) - orc.shade(got, W, H, dark, wsh, wao, f64=True))
    lit_by_spot = int((spot[:, :3].max(1) > 1e-4).sum())
    if shadow_type == api.SHADOW_RAYTRACED:
        assert wsh.sum() > 100 and lit_by_spot > 1000
    else:
        assert wsh.sum() == 0 and lit_by_spot == 0
        for l in range(nl):
            dark.lights[l].intensity = 0.0
        np.testing.assert_allclose(rad, orc.shade(got, W, H, dark, wsh, wao), rtol=1e-6, atol=1e-7)
    g.close()


def _engine_images(ctx, s, got, W, H, pad):
    """Luz's own G-buffer images for a frame (DeferredRenderer.cpp:175-248):
    RGBA32F normal, D32F depth (the camera's (proj view P).z / w, 1 on the
    background), RGBA8 albedo/material/emission, each with `pad` bytes of
    filler (NaN / 0xAB) after every row; and the descriptor that reads them in
    place (depth mode)."""
    m = s.camera_matrices()
    V = m["view"].astype(np.float64).reshape(4, 4).T
    Pm = m["proj"].astype(np.float64).reshape(4, 4).T
    fg = np.abs(got["normal"]).sum(1) != 0
    P = np.concatenate([got["position"].astype(np.float64), np.ones((W * H, 1))], 1)
    clip = P @ (Pm @ V).T
    depth = np.where(fg, clip[:, 2] / clip[:, 3], 1.0).astype(np.float32).reshape(H, W)
    fpad = pad // 4
    dimg = np.full((H, W + fpad), np.nan, np.float32)
    dimg[:, :W] = depth
    nimg = np.full((H, W * 4 + fpad), np.nan, np.float32)
    nimg[:, :W * 4] = np.concatenate([got["normal"], np.zeros((W * H, 1), np.float32)], 1).reshape(H, W * 4)
    words = {}
    for k in ("albedo", "material", "emission"):
        a = np.full((H, W + fpad), 0xABABABAB, np.uint32)
        a[:, :W] = got[k].reshape(H, W)
        words[k] = a
    bufs = {"depth": ctx.to_device(dimg), "normal": ctx.to_device(nimg)}
    bufs.update({k: ctx.to_device(v) for k, v in words.items()})
    desc = api.GBuffer()
    desc.width, desc.height = W, H
    desc.position = None
    desc.normal, desc.albedo = bufs["normal"].ptr, bufs["albedo"].ptr
    desc.material, desc.emission = bufs["material"].ptr, bufs["emission"].ptr
    desc.depth = bufs["depth"].ptr
    desc.normal_stride = 16
    desc.normal_pitch = W * 16 + pad
    desc.albedo_pitch = desc.material_pitch = desc.emission_pitch = W * 4 + pad
    desc.depth_pitch = W * 4 + pad
    for k in range(16):
        desc.inverse_proj[k] = float(m["inverse_proj"][k])
        desc.inverse_view[k] = float(m["inverse_view"][k])
    # what the oracle sees: packed planes with DepthToWorld positions
    host = {"position": orc.depth_to_world(dimg, W, H, m["inverse_proj"], m["inverse_view"]),
            "normal": got["normal"].copy(), "albedo": got["albedo"].copy(),
            "material": got["material"].copy(), "emission": got["emission"].copy()}
    return desc, bufs, host


@pytest.mark.parametrize("kind,res", [(api.SCENE_CUBE, None), (api.SCENE_SPHERES, (480, 270))])
def test_engine_gbuffer_images_read_in_place(ctx, kind, res):
    """The engine's own attachments passed unchanged: RGBA32F normals (16-byte
    stride), the D32F depth with DepthToWorld on the device, RGBA8 planes,
    every row padded.  Visibility equals the oracle's on the DepthToWorld
    positions bit for bit, and the light pass is within 1e-3."""
    s, g = _gpu_scene(ctx, kind)
    o = oracle_scene(s)
    W, H, gd, got, _ = _gbuffers(ctx, s, g, o, *(res or (None, None)))
    desc, bufs, host = _engine_images(ctx, s, got, W, H, pad=64)
    sb = s.pack()
    if kind == api.SCENE_CUBE:
        set_light_direction(sb, 0, (0.8, -1.0, 0.6))
    nl = sb.num_lights
    dsh, dao = ctx.buffer(nl * W * H), ctx.buffer(W * H)
    ctx.trace_visibility(g.tlas, desc, sb, dsh, dao)
    wsh, wao, st = o.trace_visibility(host, W, H, sb)
    print(f"kind {kind} {W}x{H} depth mode: {st.rays} rays, shadowed {int(wsh.sum())}", flush=True)
    np.testing.assert_array_equal(dsh.download(np.uint8, (nl, W * H)), wsh)
    np.testing.assert_array_equal(dao.download(np.uint8, W * H), wao)
    out = ctx.buffer(W * H * 16)
    ctx.light_pass(g.tlas, desc, sb, out)
    rad = out.download(np.float32, (W * H, 4))
    want = orc.shade(host, W, H, sb, wsh, wao)
    err = np.abs(rad - want) / np.maximum(np.abs(want), 1e-3)
    print(f"max rel err {float(err.max()):.3g}", flush=True)
    assert err.max() <= RADIANCE_RTOL
    # a position plane with a 16-byte stride and padded rows reads as the packed one
    pimg = np.full((H, W * 4 + 16), np.nan, np.float32)
    pimg[:, :W * 4] = np.concatenate([got["position"], np.ones((W * H, 1), np.float32)], 1).reshape(H, W * 4)
    dpos = ctx.to_device(pimg)
    desc.position = dpos.ptr
    desc.position_stride = 16
    desc.position_pitch = W * 16 + 64
    ctx.trace_visibility(g.tlas, desc, sb, dsh, dao)
    psh, pao, _ = o.trace_visibility(got, W, H, sb)
    np.testing.assert_array_equal(dsh.download(np.uint8, (nl, W * H)), psh)
    np.testing.assert_array_equal(dao.download(np.uint8, W * H), pao)
    g.close()
