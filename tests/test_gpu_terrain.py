"""GPU parity of the precision form on its own workload: the fBm terrain
(2047x1023 quads, about 4.2 M triangles) plus 40 displaced icospheres
(about 5.0 M instanced triangles in all), 4 directional and 32 point lights,
8 AO rays per pixel, through the 3840x2160 workload's camera at 960x540.

The 8-wide quantised BVH gives counts identical to the oracle's 8-wide
traversal and to its FP64 traversal (every 16th row); the packed-FP16
shading stays within a relative radiance RMSE of 5e-3 of the FP64 oracle
(light.frag:171-235)."""
import numpy as np
import pytest

from luz_amd import api
from oracle import oracle as orc
from tests.helpers import oracle_scene

pytestmark = pytest.mark.gpu

W, H = 960, 540
FP64_ROW_STEP = 16


def test_terrain_precision_form(ctx):
    ctx.set_precision(api.PRECISION_COMPACT)
    try:
        s = api.HostScene(api.SCENE_TERRAIN, 0)
        print(f"terrain scene: {s.instanced_triangles} instanced triangles", flush=True)
        g = api.GPUScene(ctx, s)
        o = oracle_scene(s)
        o.build_wide()
        _, _, cam = s.camera()
        _, _, mats = s.nodes()
        gd = api.GBufferDevice(ctx, W, H)
        ctx.primary_gbuffer(g.tlas, cam, mats, gd.desc())
        host = gd.download()
        want_g = o.primary_gbuffer(cam, mats, W, H)
        for k in ("position", "normal"):
            np.testing.assert_array_equal(host[k].view(np.uint32), want_g[k].view(np.uint32), err_msg=k)
        sb = s.pack()
        nl = sb.num_lights
        kinds = [sb.lights[l].type for l in range(nl)]
        assert nl == 36 and kinds.count(api.LIGHT_DIRECTIONAL) == 4 and sb.ao_num_samples == 8
        dsh, dao = ctx.buffer(nl * W * H), ctx.buffer(W * H)
        ctx.trace_visibility(g.tlas, gd.desc(), sb, dsh, dao)
        gsh = dsh.download(np.uint8, (nl, W * H))
        gao = dao.download(np.uint8, W * H)
        rays = ctx.timings().rays
        wsh, wao, st = o.trace_visibility(host, W, H, sb, wide=True)
        print(f"{rays} rays (oracle {st.rays}), occluded shadow samples {int(wsh.sum())}, "
              f"mean AO misses {float(wao.mean()):.3f}", flush=True)
        assert rays == st.rays
        np.testing.assert_array_equal(gsh, wsh)
        np.testing.assert_array_equal(gao, wao)
        assert 0 < wsh.sum() and 0 < wao.sum() < W * H * 8
        # FP64 traversal on sampled rows: hit/miss mismatches of the FP32 counts
        sh64, ao64, _ = o.trace_visibility(host, W, H, sb, f64=True, row_step=FP64_ROW_STEP)
        rows = np.zeros((H, W), bool)
        rows[::FP64_ROW_STEP] = True
        rows = rows.ravel()
        mism = int((gsh[:, rows] != sh64[:, rows]).sum()) + int((gao[rows] != ao64[rows]).sum())
        print(f"count mismatches against FP64 on every {FP64_ROW_STEP}th row: {mism}", flush=True)
        assert mism == 0
        # packed-FP16 shading against the FP64 oracle
        out = ctx.buffer(W * H * 16)
        ctx.shade(gd.desc(), sb, dsh, dao, out)
        rad = out.download(np.float32, (W * H, 4))[:, :3].astype(np.float64)
        ref = orc.shade(host, W, H, sb, wsh, wao, f64=True)[:, :3]
        act = np.abs(host["normal"]).sum(1) != 0
        rmse = np.sqrt(np.mean((rad[act] - ref[act]) ** 2))
        scale = np.sqrt(np.mean(ref[act] ** 2))
        print(f"FP16 shading: radiance RMSE {rmse:.3g} (relative {rmse / scale:.3g})", flush=True)
        assert np.isfinite(rad).all() and rmse / scale < 5e-3
        g.close()
    finally:
        ctx.set_precision(api.PRECISION_PLAIN)
