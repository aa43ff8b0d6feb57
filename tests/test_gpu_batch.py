"""GPU parity of batched frames: every scene of a batch (refit, TLAS rebuild,
primary G-buffer, one multi-scene traversal launch, shading; eager and
hipGraph replays) reproduces the single-scene calls bit for bit, and its
visibility counts equal the oracle's on the same G-buffer."""
import numpy as np
import pytest

from luz_amd import api
from luz_amd.batching import BatchWorkload
from tests.helpers import oracle_scene

pytestmark = pytest.mark.gpu

# uneven scenes: (seed, triangle budget, width, height)
SCENES = [(3, 1500, 96, 64), (11, 30000, 160, 120), (5, 6000, 64, 64), (8, 90000, 200, 150),
          (21, 1000, 40, 30)]


def single_scene(ctx, wl, i, frame):
    """The single-scene path on scene i at animation `frame`."""
    h, g = wl.hosts[i], wl.gpus[i]
    h.animate(frame)
    g.update_resources_gpu(h)
    W, H, cam = h.camera()
    _, _, mats = h.nodes()
    gd = api.GBufferDevice(ctx, W, H)
    ctx.primary_gbuffer(g.tlas, cam, mats, gd.desc())
    sb = h.pack()
    sb.frame = frame
    n = W * H
    dsh, dao = ctx.buffer(max(sb.num_lights, 1) * n), ctx.buffer(n)
    out = ctx.buffer(n * 16)
    ctx.trace_visibility(g.tlas, gd.desc(), sb, dsh, dao)
    ctx.shade(gd.desc(), sb, dsh, dao, out)
    return (gd.download(), dsh.download(np.uint8, (sb.num_lights, n)), dao.download(np.uint8, n),
            out.download(np.float32, (n, 4)))


def assert_same(got, want, tag):
    gg, gs, ga, go = got
    wg, ws, wa, wo = want
    for k in gg:
        np.testing.assert_array_equal(gg[k].view(np.uint32), wg[k].view(np.uint32), err_msg=f"{tag} {k}")
    np.testing.assert_array_equal(gs, ws, err_msg=f"{tag} shadow")
    np.testing.assert_array_equal(ga, wa, err_msg=f"{tag} ao")
    np.testing.assert_array_equal(go.view(np.uint32), wo.view(np.uint32), err_msg=f"{tag} radiance")


@pytest.fixture()
def workload(ctx):
    wl = BatchWorkload(ctx, SCENES)
    yield wl
    wl.close()
    ctx.set_trace_config(6, 48, 16, 8)   # the library defaults


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_matches_single_scene_calls(ctx, workload, use_graph):
    # several frames (graph replays) with eager single-scene work in between
    for frame in (0, 3, 7):
        rays = workload.frame(frame, use_graph)
        print(f"graph={use_graph} frame {frame}: {rays} rays", flush=True)
        got = [workload.outputs(i) for i in range(len(SCENES))]
        total = 0
        for i in range(len(SCENES)):
            want = single_scene(ctx, workload, i, frame)
            assert_same(got[i], want, f"scene {i} frame {frame}")
            active = int((np.abs(want[0]["normal"]).sum(1) != 0).sum())
            sb = workload.blocks[i]
            per_px = sum(sb.lights[l].num_shadow_samples for l in range(sb.num_lights)) + sb.ao_num_samples
            total += active * per_px
        assert rays == total


def test_batch_visibility_matches_oracle(ctx, workload):
    frame = 5
    workload.frame(frame, True)
    for i in (0, 1, 3):
        h = workload.hosts[i]
        h.animate(frame)
        gbuf, sh, ao, _ = workload.outputs(i)
        sb = h.pack()
        sb.frame = frame
        o = oracle_scene(h)
        W, H, _ = h.camera()
        wsh, wao, _ = o.trace_visibility(gbuf, W, H, sb)
        print(f"scene {i}: {W}x{H}, oracle shadow sum {int(wsh.sum())}, ao sum {int(wao.sum())}", flush=True)
        np.testing.assert_array_equal(sh, wsh)
        np.testing.assert_array_equal(ao, wao)


def test_batch_grouped_ray_order_matches_oracle(ctx):
    """A batch created under the grouped ray order (the samples of 64
    neighbouring pixels claimed together in every scene) gives the oracle's
    counts, like the default order."""
    ctx.set_ray_order(api.RAY_ORDER_GROUPED)
    try:
        wl = BatchWorkload(ctx, SCENES)
    finally:
        ctx.set_ray_order(api.RAY_ORDER_AUTO)
    try:
        frame = 4
        wl.frame(frame, True)
        for i in (0, 3, 4):
            h = wl.hosts[i]
            h.animate(frame)
            gbuf, sh, ao, _ = wl.outputs(i)
            sb = h.pack()
            sb.frame = frame
            W, H, _ = h.camera()
            wsh, wao, _ = oracle_scene(h).trace_visibility(gbuf, W, H, sb)
            print(f"grouped order, scene {i}: {W}x{H}", flush=True)
            np.testing.assert_array_equal(sh, wsh)
            np.testing.assert_array_equal(ao, wao)
    finally:
        wl.close()


def test_batch_mixes_cameras_and_caller_gbuffers(ctx):
    """Scenes without a camera keep the G-buffer their caller wrote (no primary
    tiles of their own in the batch's tile space, between scenes that have
    them) and still match the single-scene calls bit for bit."""
    from types import SimpleNamespace
    specs = [SCENES[0], SCENES[1], SCENES[2], SCENES[4]]
    frame = 2
    hosts, gpus, gbufs, outs, blocks, entries = [], [], [], [], [], []
    for i, (seed, budget, w, hgt) in enumerate(specs):
        h = api.HostScene(api.SCENE_RANDOM, seed, budget, w, hgt)
        g = api.GPUScene(ctx, h)
        h.animate(frame)
        g.update_resources_gpu(h)
        W, H, cam = h.camera()
        _, _, mats = h.nodes()
        sb = h.pack()
        gb = api.GBufferDevice(ctx, W, H)
        caller = i in (0, 2)   # the first scene and one in the middle
        if caller:
            ctx.primary_gbuffer(g.tlas, cam, mats, gb.desc())
        bufs = (ctx.buffer(max(sb.num_lights, 1) * W * H), ctx.buffer(W * H), ctx.buffer(W * H * 16))
        hosts.append(h); gpus.append(g); gbufs.append(gb); outs.append(bufs); blocks.append(sb)
        entries.append({"tlas": g.tlas, "gbuffer": gb.desc(), "scene": sb,
                        "camera": None if caller else cam, "materials": [] if caller else mats,
                        "shadow": bufs[0], "ao": bufs[1], "out": bufs[2]})
    batch = api.Batch(ctx, entries)
    try:
        for i, (h, g) in enumerate(zip(hosts, gpus)):
            arr, n = g.instances(h)
            batch.set_instances(i, arr, n)
        batch.frame(frame, True)
        wl = SimpleNamespace(hosts=hosts, gpus=gpus)
        for i in range(len(specs)):
            n = gbufs[i].width * gbufs[i].height
            L = blocks[i].num_lights
            got = (gbufs[i].download(), outs[i][0].download(np.uint8, (L, n)),
                   outs[i][1].download(np.uint8, n), outs[i][2].download(np.float32, (n, 4)))
            print(f"mixed batch scene {i}: camera={entries[i]['camera'] is not None}", flush=True)
            assert_same(got, single_scene(ctx, wl, i, frame), f"mixed scene {i}")
    finally:
        batch.close()


def test_batch_recaptures_on_config_change(ctx, workload):
    want = None
    for cfg in [(6, 48, 16, 32), (6, 16, 8, 16), (8, 32, 32, 8)]:
        ctx.set_trace_config(*cfg)
        workload.frame(2, True)
        got = [workload.outputs(i) for i in range(len(SCENES))]
        if want is None:
            want = got
            continue
        for i in range(len(SCENES)):
            assert_same(got[i], want[i], f"config {cfg} scene {i}")


def test_batch_rejects_compact_precision(ctx, workload):
    ctx.set_precision(api.PRECISION_COMPACT)
    try:
        with pytest.raises(api.LuzError):
            workload.batch.frame(0, True)
    finally:
        ctx.set_precision(api.PRECISION_PLAIN)
