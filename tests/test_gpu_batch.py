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


def test_batch_without_culling_matches_oracle(ctx):
    """A batch created with ray-group culling off (every ray traced, slot-major
    order) gives the oracle's counts, as the default culled batch does."""
    ctx.set_group_culling(False)
    try:
        wl = BatchWorkload(ctx, SCENES)
    finally:
        ctx.set_group_culling(True)
    try:
        frame = 6
        wl.frame(frame, True)
        for i in (1, 2, 3):
            h = wl.hosts[i]
            h.animate(frame)
            gbuf, sh, ao, _ = wl.outputs(i)
            sb = h.pack()
            sb.frame = frame
            W, H, _ = h.camera()
            wsh, wao, _ = oracle_scene(h).trace_visibility(gbuf, W, H, sb)
            print(f"unculled batch, scene {i}: {W}x{H}", flush=True)
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


def test_batch_graph_follows_a_reallocated_spill_buffer(ctx, workload):
    """A graph captured at 6 groups per CU holds the context's traversal spill
    buffer; a single-scene call at 10 groups per CU reallocates that buffer,
    so the replay after the configuration is set back must recapture rather
    than write the freed one: every scene still equals the single-scene calls."""
    ctx.set_trace_config(6, 48, 16, 8)
    workload.frame(1, True)
    ctx.set_trace_config(10, 48, 16, 8)
    single_scene(ctx, workload, 0, 1)
    ctx.set_trace_config(6, 48, 16, 8)
    frame = 3
    workload.frame(frame, True)
    got = [workload.outputs(i) for i in range(len(SCENES))]
    for i in range(len(SCENES)):
        assert_same(got[i], single_scene(ctx, workload, i, frame), f"after reallocation, scene {i}")


def test_batch_refits_a_single_triangle_blas(ctx):
    """A one-triangle mesh registered for refit moves with its vertices in
    every batched frame (its single root is rebuilt), as luzrt_blas_refit
    moves it in the single-scene path."""
    s = api.HostScene(api.SCENE_CUBE, 0)
    g = api.GPUScene(ctx, s)
    W, H, cam = s.camera()
    _, _, mats = s.nodes()
    sb = s.pack()

    def tri_at(frame):
        v = np.zeros((3, 12), np.float32)
        y = 1.3 + 0.2 * frame
        v[:, :3] = [[-0.8, y, -0.8], [0.8, y, -0.6], [0.0, y + 0.6, 0.9]]
        v[:, 4] = 1.0
        return v

    dv = ctx.to_device(tri_at(0))
    tblas = ctx.create_blas(dv, 48, 3, ctx.to_device(np.array([0, 1, 2], np.uint32)), 3)
    arr, n = g.instances(s)
    insts = [arr[i] for i in range(n)] + [api.make_instance(tblas, np.eye(3, 4), n)]
    tlas, tlas1 = ctx.create_tlas(16), ctx.create_tlas(16)
    gb, gb1 = api.GBufferDevice(ctx, W, H), api.GBufferDevice(ctx, W, H)
    bufs = [ctx.buffer(W * H), ctx.buffer(W * H), ctx.buffer(W * H * 16)]
    batch = api.Batch(ctx, [{"tlas": tlas, "gbuffer": gb.desc(), "scene": sb, "camera": cam, "materials": mats,
                             "shadow": bufs[0], "ao": bufs[1], "out": bufs[2]}])
    try:
        batch.add_refit(tblas, dv.ptr)
        prev = None
        for frame in (1, 2):
            dv.upload(tri_at(frame))
            batch.set_instances(0, (api.Instance * len(insts))(*insts), len(insts))
            batch.frame(frame, True)
            batch.rays()
            got = gb.download()
            # the single-scene path: refit, TLAS rebuild, primary rays
            tblas.refit(dv)
            tlas1.build(insts)
            ctx.primary_gbuffer(tlas1, cam, mats, gb1.desc())
            want = gb1.download()
            for k in got:
                np.testing.assert_array_equal(got[k].view(np.uint32), want[k].view(np.uint32), err_msg=k)
            above = int((got["position"][:, 1] > 1.2).sum())
            print(f"frame {frame}: {above} pixels see the moving triangle", flush=True)
            assert above > 200
            if prev is not None:
                assert not np.array_equal(prev, got["position"])
            prev = got["position"]
    finally:
        batch.close()
        tlas.destroy()
        tlas1.destroy()
        tblas.destroy()
        g.close()


def test_batch_reports_traversal_overflow_once(ctx, workload):
    """A frame whose rays exceed the step limit (3 steps, set from the host)
    makes luzrt_batch_rays return LUZRT_E_LIMIT once; frames at the default
    limit report nothing."""
    ctx.set_step_limit(3)
    try:
        workload.stage(1)
        workload.batch.frame(1, True)
        with pytest.raises(api.LuzError) as e:
            workload.batch.rays()
        assert e.value.status == -5
        workload.batch.rays()
    finally:
        ctx.set_step_limit(0)
    workload.frame(2, True)
