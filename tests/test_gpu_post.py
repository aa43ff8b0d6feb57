"""GPU parity of the passes after the light pass: TAA bit-exact with the
oracle on rendered light buffers (jittered camera, a moving camera, with and
without the Mitchell reconstruction), and the ACES compose within the
tolerance of the device pow."""
import numpy as np
import pytest

from luz_amd import api
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def frames(ctx):
    s = api.HostScene(api.SCENE_SPHERES, 0)
    g = api.GPUScene(ctx, s)
    W, H = 480, 270
    _, _, cam = s.camera()
    from luz_amd.denoiser import crop_camera
    cam = crop_camera(cam, 1920, 1080, 600, 300, W, H)
    _, _, mats = s.nodes()
    gd = api.GBufferDevice(ctx, W, H)
    ctx.primary_gbuffer(g.tlas, cam, mats, gd.desc())
    sb = s.pack()
    lights = []
    for f in (0, 1):
        sb.frame = f
        out = ctx.buffer(W * H * 16)
        ctx.light_pass(g.tlas, gd.desc(), sb, out)
        lights.append(out)
    yield {"scene": s, "gd": gd, "lights": lights, "W": W, "H": H}
    g.close()


def _taa(ctx, fr, vp, pvp, ivp, j, pj, reconstruct):
    W, H = fr["W"], fr["H"]
    gd = fr["gd"]
    out = ctx.buffer(W * H * 16)
    d = api.TaaDesc()
    d.width, d.height = W, H
    d.light, d.history, d.out = fr["lights"][1].ptr, fr["lights"][0].ptr, out.ptr
    d.position, d.normal = gd.position.ptr, gd.normal.ptr
    d.view_proj[:] = [float(v) for v in vp]
    d.prev_view_proj[:] = [float(v) for v in pvp]
    d.inv_view_proj[:] = [float(v) for v in ivp]
    d.jitter[:] = [float(v) for v in j]
    d.prev_jitter[:] = [float(v) for v in pj]
    d.reconstruct = reconstruct
    ctx.taa(d)
    got = out.download(np.float32, (W * H, 4))
    host = gd.download()
    want = orc.taa(fr["lights"][1].download(np.float32, (W * H, 4)), fr["lights"][0].download(np.float32, (W * H, 4)),
                   host["position"], host["normal"], W, H, vp, pvp, ivp, j, pj, reconstruct)
    return got, want


@pytest.mark.parametrize("reconstruct", [0, 1])
def test_taa_bit_exact_with_jitter(ctx, frames, reconstruct):
    s, W, H = frames["scene"], frames["W"], frames["H"]
    j0, j1 = api.jitter(3, W, H), api.jitter(4, W, H)
    vp, ivp = s.view_proj(j1)
    pvp, _ = s.view_proj(j0)
    got, want = _taa(ctx, frames, vp, pvp, ivp, j1, j0, reconstruct)
    changed = float(np.mean(np.any(got != frames["lights"][1].download(np.float32, (W * H, 4)), 1)))
    print(f"reconstruct={reconstruct}: {changed:.1%} of pixels blended", flush=True)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_taa_bit_exact_with_camera_motion(ctx, frames):
    s, W, H = frames["scene"], frames["W"], frames["H"]
    vp, ivp = s.view_proj()
    T = np.eye(4, dtype=np.float64)
    T[:3, 3] = [0.4, -0.2, 0.3]                  # previous camera displaced (world translate)
    M = vp.astype(np.float64).reshape(4, 4).T @ T
    pvp = M.T.reshape(16).astype(np.float32)
    got, want = _taa(ctx, frames, vp, pvp, ivp, (0.0, 0.0), (0.0, 0.0), 0)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_compose_matches_oracle(ctx, frames):
    W, H = frames["W"], frames["H"]
    n = W * H
    rgba8, rgb = ctx.buffer(n * 4), ctx.buffer(n * 16)
    ctx.compose(W, H, frames["lights"][0], rgba8, rgb)
    w8, wf = orc.compose(frames["lights"][0].download(np.float32, (n, 4)))
    g8, gf = rgba8.download(np.uint32, n), rgb.download(np.float32, (n, 4))
    np.testing.assert_allclose(gf, wf, rtol=2e-6, atol=1e-7)
    diff = np.abs(g8.view(np.uint8).astype(int) - w8.view(np.uint8).astype(int))
    assert diff.max() <= 1 and (diff > 0).mean() < 1e-3
