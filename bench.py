"""Benchmark of the ray-traced lighting path on one MI355X.

A step is one frame of the 1920x1080 sphere-field workload (5 fBm-displaced
level-5 icospheres x 50 instances + 64x64 ground grid, 1,032,192 instanced
triangles; 1 directional + 8 point lights, 1 shadow ray per light and 16 AO
rays per pixel): TLAS rebuild over the 51 instances, shadow + AO visibility
(compaction, ray generation, persistent traversal, resolve) and GGX lighting.
The G-buffer is produced once by primary rays before timing and stays
resident in HBM, as do the BLASes.

Prints one JSON line.  `value` is shadow+AO rays traced per second over the
whole frame (Mrays/s).  `roofline` refers to the traversal kernel: its
algorithmic bytes per launch (per-ray node, triangle and instance counts of
the CPU oracle's canonical traversal of the identical BVH, plus the G-buffer
reads and count writes per pixel) over its HIP-event time, the profiled HBM
traffic, and the profiled VALU issue fraction that binds it.  `cpu_baseline`
is the oracle (FP32, same arithmetic) on the whole frame on this host's
cores.  `parity` compares every pixel's counts with the FP32 oracle and a
row sample with the FP64 oracle.  `ao_radius16` repeats the frame with an AO
radius of 16, where occluders fall within range of the lifted AO origins.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "shadow+AO Mrays/s & ms/frame @1080p, 1M tris, 1 GPU; AO RMSE vs FP64 oracle"
HBM_PEAK_GBS = 8000.0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-row-step", type=int, default=8, help="oracle traces every k-th row")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and parity sample")
    ap.add_argument("--workload", default="spheres1080p", choices=["spheres1080p", "terrain4k", "batch64", "denoise1080p"],
                    help="spheres1080p: the headline 1080p sphere field (plain FP32 path); "
                         "terrain4k: 3840x2160 terrain + spheres, ~5.2M tris, precision form; "
                         "batch64: 64 animated scenes of uneven size and resolution per frame; "
                         "denoise1080p: 1-spp + learned denoiser against the plain multi-spp tracer")
    ap.add_argument("--bvh", default="wide", choices=["wide", "binary"],
                    help="spheres1080p: trace the 8-wide BVH with FP32 shading (PRECISION_WIDE, the "
                         "default) or the binary BVH (PRECISION_PLAIN); counts and radiance are the same")
    ap.add_argument("--scenes", type=int, default=64, help="batch64: scenes per frame")
    ap.add_argument("--autotune", action="store_true",
                    help="search the traversal launch shape on this workload before timing")
    ap.add_argument("--groups-per-cu", type=int, default=0)
    ap.add_argument("--refill", type=int, default=0)
    ap.add_argument("--stack-entries", type=int, default=0, help="LDS stack entries per lane (8, 16 or 32)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (PMC counter calibration)")
    ap.add_argument("--traffic", default=None,
                    help="HBM bytes and VALU instructions per traversal launch measured with rocprofv3 "
                         "--pmc (JSON; default: the workload's file under profiles/)")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus}); "
              "the lighting path does not shard across GPUs here", file=sys.stderr)
        return 2

    import numpy as np

    from luz_amd import api

    ctx = api.Context(0)
    if args.groups_per_cu or args.refill or args.stack_entries:
        ctx.set_trace_config(args.groups_per_cu, args.refill, args.stack_entries)
    if args.workload == "batch64":
        return bench_batch(ctx, args)
    if args.workload == "denoise1080p":
        return bench_denoise(ctx, args)
    compact = args.workload == "terrain4k"
    wide = compact or args.bvh == "wide"
    if compact:
        ctx.set_precision(api.PRECISION_COMPACT)
    elif wide:
        ctx.set_precision(api.PRECISION_WIDE)
    scene = api.HostScene(api.SCENE_TERRAIN if compact else api.SCENE_SPHERES, 0)
    gs = api.GPUScene(ctx, scene)
    W, H, cam = scene.camera()
    _, _, mats = scene.nodes()
    gbuf = api.GBufferDevice(ctx, W, H)
    ctx.primary_gbuffer(gs.tlas, cam, mats, gbuf.desc())
    sb = scene.pack()
    out = ctx.buffer(W * H * 16)
    nl = sb.num_lights
    tuned = None
    if args.autotune:
        from luz_amd import autotune
        tuned = autotune.tune_visibility(ctx, gs.tlas, gbuf.desc(), sb)

    def frame():
        gs.update_resources_gpu(scene)            # TLAS rebuild (GPUScene.cpp:354-365)
        ctx.light_pass(gs.tlas, gbuf.desc(), sb, out)

    # Frames are submitted back to back, as a renderer does: the host packs
    # frame k+1's instances while the GPU runs frame k (the library does not
    # wait for the GPU inside a frame).  Stage events are off meanwhile: each
    # one holds the next kernel back by about 8 us.
    ctx.set_stage_timing(False)
    for _ in range(args.warmup):
        frame()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame()
    ctx.sync()
    elapsed = time.perf_counter() - t0
    ctx.set_stage_timing(True)
    # per-stage GPU times (HIP events) from frames that are each waited for
    trace_ms, build_ms, shade_ms, vis_ms = [], [], [], []
    for _ in range(min(args.steps, 10)):
        frame()
        t = ctx.timings()
        trace_ms.append(t.trace_ms)
        build_ms.append(t.build_ms)
        shade_ms.append(t.shade_ms)
        vis_ms.append(t.compact_ms + t.raygen_ms + t.trace_ms + t.resolve_ms)
    rays = int(t.rays)
    active = int(t.active_pixels)
    traced = int(t.traced_rays)
    ms_frame = elapsed / args.steps * 1e3
    value = rays * args.steps / elapsed / 1e6
    trace_avg = float(np.mean(trace_ms))

    record = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_frame, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp16" if compact else "fp32",
        "data": "synthetic",
        "config": {
            "workload": ("3840x2160 fBm terrain (2047x1023 quads) + 40 fBm icospheres, 8-wide BVH with "
                         "8-bit child boxes, packed-FP16 GGX, 4 dir + 32 point lights, 1 shadow ray/light "
                         "+ 8 AO rays/px") if compact else
                        ("1920x1080 sphere field: 50 instances of 5 fBm icospheres (20,480 tris) + "
                         "64x64 ground grid, 1,032,192 instanced tris, 1 dir + 8 point lights, "
                         "1 shadow ray/light + 16 AO rays/px, " +
                         ("8-wide BVH with 8-bit child boxes" if wide else "binary BVH") +
                         ", FP32 shading"),
            "bvh": "8-wide" if wide else "binary",
            "width": W, "height": H, "instanced_tris": int(scene.instanced_triangles),
            "rays_per_frame": rays, "active_pixels": active, "traced_rays_per_frame": traced,
            "frame_ms_breakdown": {"tlas_build": round(float(np.mean(build_ms)), 4),
                                   "visibility": round(float(np.mean(vis_ms)), 4),
                                   "trace_kernel": round(trace_avg, 4),
                                   "shade": round(float(np.mean(shade_ms)), 4)},
            "visibility_mrays_per_s": round(rays / (float(np.mean(vis_ms)) * 1e-3) / 1e6, 2),
        },
    }
    if tuned is not None:
        record["config"]["autotune"] = {k: tuned[k] for k in
                                        ("groups_per_cu", "refill", "stack_entries", "tile_width", "table")}

    prof = {}
    tpath = Path(args.traffic) if args.traffic else (
        ROOT / "profiles" / ("k_trace_any_terrain4k_traffic.json" if compact else
                             "k_trace_any_traffic.json" if wide else "k_trace_any_binary_traffic.json"))
    if tpath.exists():
        try:
            prof = json.loads(tpath.read_text())
        except (ValueError, OSError):
            prof = {}
    traffic = prof.get("hbm_bytes_per_launch")

    if not args.no_cpu and compact:
        record.update(compact_parity(ctx, gs, scene, gbuf, sb, out, W, H, trace_avg, rays, prof))
    elif not args.no_cpu:
        from oracle import oracle as orc
        host = gbuf.download()
        o = orc.OracleScene()
        for i in range(scene.mesh_count):
            v, idx = scene.mesh(i)
            o.add_mesh(v, idx)
        o2w, mesh_ids, _ = scene.nodes()
        o.set_instances(o2w, mesh_ids)
        if wide:
            o.build_wide()
        threads, cores_avail = cpu_threads()
        par, st = visibility_parity(ctx, gs, gbuf, o, host, sb, W, H, threads, args.cpu_row_step * 4, wide)
        # CPU baseline: the oracle's FP32 visibility of the whole frame, median of 5
        o.trace_visibility(host, W, H, sb, threads=threads)      # warm-up frame
        times = []
        for _ in range(5):
            c0 = time.perf_counter()
            o.trace_visibility(host, W, H, sb, threads=threads)
            times.append(time.perf_counter() - c0)
        cpu_s = float(np.median(times))
        c0 = time.perf_counter()
        _, _, st1 = o.trace_visibility(host, W, H, sb, threads=1, row_step=args.cpu_row_step * 2)
        cpu1_s = time.perf_counter() - c0
        record["roofline"] = roofline(ctx, st, rays, traced, W * H, nl, trace_avg, traffic, prof, wide)
        record["cpu_baseline"] = {
            "value": round(st.rays / cpu_s / 1e6, 2), "unit": "Mrays/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle FP32 visibility of the whole frame ({st.rays} rays), median of 5 "
                      f"after a warm-up frame: {cpu_s * 1e3:.0f} ms/frame on {threads} threads "
                      f"({cores_avail} CPUs in this process's affinity mask, {cpu_model()}); "
                      f"single thread on every {args.cpu_row_step * 2}th row: "
                      f"{st1.rays / cpu1_s / 1e6:.2f} Mrays/s",
            "ms_per_frame": round(cpu_s * 1e3, 1),
            "single_thread_mrays_per_s": round(st1.rays / cpu1_s / 1e6, 3),
            "cpus_available": cores_avail,
            "cpu_model": cpu_model(),
        }
        record["parity"] = par
        # the same frame with an AO radius at which occluders fall in range (16):
        # the reference lifts the AO origin by 0.01 x the camera distance (1-2.5
        # here), so with its radius of 1 every AO ray of this scene misses
        sb16 = api.SceneBlock()
        import ctypes as C
        C.memmove(C.byref(sb16), C.byref(sb), C.sizeof(api.SceneBlock))
        sb16.ao_max = 16.0
        record["ao_radius16"] = ao_radius_frame(ctx, gs, scene, gbuf, o, host, sb16, out, W, H, threads,
                                                args, wide)
    print(json.dumps(record), flush=True)
    gs.close()
    ctx.close()
    return 0


def cpu_threads():
    """Oracle threads: every CPU of this process's affinity mask, at most 16
    (the host CPU share of one GPU on the measurement boxes)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    return max(1, min(avail, 16)), avail


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def visibility_parity(ctx, gs, gbuf, o, host, sb, W, H, threads, f64_step, wide=False):
    """GPU counts against the FP32 oracle (binary BVH) on every pixel and the
    FP64 oracle on every f64_step-th row; the oracle's stats (canonical
    traversal counts) of the BVH the GPU traced: the 8-wide one when wide."""
    import numpy as np

    nl = sb.num_lights
    gpu_sh = ctx.buffer(max(nl, 1) * W * H)
    gpu_ao = ctx.buffer(W * H)
    ctx.trace_visibility(gs.tlas, gbuf.desc(), sb, gpu_sh, gpu_ao)
    gsh = gpu_sh.download(np.uint8, (max(nl, 1), W * H))[:nl]
    gao = gpu_ao.download(np.uint8, W * H)
    wsh, wao, st = o.trace_visibility(host, W, H, sb, threads=threads)
    rows64 = np.zeros((H, W), bool)
    rows64[::f64_step] = True
    act = np.abs(host["normal"]).sum(1) != 0
    m64 = rows64.reshape(-1) & act
    dsh, dao, _ = o.trace_visibility(host, W, H, sb, threads=threads, row_step=f64_step, f64=True)
    if wide:
        _, _, st = o.trace_visibility(host, W, H, sb, threads=threads, wide=True)
    na = max(sb.ao_num_samples, 1)
    ao_rmse = float(np.sqrt(np.mean((gao[m64].astype(np.float64) / na - dao[m64] / na) ** 2))) if m64.any() else 0.0
    par = {
        "compared_pixels": W * H,
        "shadow_mismatch": int((gsh != wsh).sum()),
        "ao_mismatch": int((gao != wao).sum()),
        "ao_rmse_vs_fp64": round(ao_rmse, 6),
        "count_mismatch_vs_fp64": int((gsh[:, m64] != dsh[:, m64]).sum() + (gao[m64] != dao[m64]).sum()),
        "fp64_sample_rows_every": f64_step,
        "ao_fraction_below_1": round(float(np.mean(gao[act] < sb.ao_num_samples)), 6),
        "mean_ao": round(float(np.mean(gao[act].astype(np.float64) / na)), 6),
    }
    return par, st


def roofline(ctx, st, rays, traced, pixels, nl, trace_ms, traffic, prof, wide=False) -> dict:
    """The traversal kernel against its rooflines.  Algorithmic bytes per launch
    follow the visibility formula (96 n_wide + 64 n_node + 36 n_tri + 64 n_inst)
    per ray + 24 P + (L + 1) P, with per-ray counts from the oracle's canonical
    traversal of the identical BVH (n_wide: 8-wide BLAS nodes, n_node: binary
    nodes), over the rays the launch traces (ray-group culling
    resolves the others as misses before it); the VALU issue fraction is the
    profiled SQ_INSTS_VALU per launch over CUs x 4 SIMDs x clock / 4 (one
    wave64 VALU instruction per SIMD every 4 cycles) over the kernel time."""
    nr = max(st.rays, 1)
    per_ray = (96.0 * st.wide_node_visits + 64.0 * st.node_visits + 36.0 * st.tri_tests +
               64.0 * st.inst_visits) / nr
    b_vis = traced * per_ray + 24.0 * pixels + (nl + 1) * pixels
    t = trace_ms * 1e-3
    achieved = b_vis / t / 1e9
    cus, clk_khz = ctx.device_info()
    clock_hz = max(clk_khz, 1) * 1e3
    valu = prof.get("valu_insts_per_launch")
    valu_frac = valu / (cus * 4 * clock_hz / 4) / t if valu else None
    hbm_frac = traffic / t / 1e9 / HBM_PEAK_GBS if traffic else None
    # the binding resource is the one whose measured fraction is larger: VALU
    # issue against the HBM traffic actually drawn (the algorithmic bytes are
    # mostly served from the caches and bound nothing)
    bound = "valu" if valu_frac is not None and valu_frac > (hbm_frac or 0.0) else "hbm"
    return {
        "bound": bound, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        "kernel": "k_trace_any<wide>" if wide else "k_trace_any",
        "algorithmic_bytes_per_launch": round(b_vis), "rays": rays, "traced_rays": traced,
        "per_ray_counts": {"wide_node": round(st.wide_node_visits / nr, 3),
                           "node": round(st.node_visits / nr, 3), "tri": round(st.tri_tests / nr, 3),
                           "inst": round(st.inst_visits / nr, 3)},
        # the BVH (about 11 MB) is served from L2 and the Infinity Cache: the
        # measured HBM bytes per launch over the same time give the HBM fraction drawn
        "traffic_frac": round(hbm_frac, 4) if hbm_frac else None,
        "valu_issue_frac": round(valu_frac, 4) if valu_frac else None,
        "valu_insts_per_launch": valu,
        "compute_units": cus, "clock_mhz": round(clock_hz / 1e6),
        "counters_from": prof.get("source"),
        # node/triangle/instance bytes plus the 32-byte pixel record each ray is
        # generated from and its hit byte, per ray
        "bytes_per_ray_with_pixel_records": round(per_ray + 33.0, 2),
    }


def ao_radius_frame(ctx, gs, scene, gbuf, o, host, sb, out, W, H, threads, args, wide) -> dict:
    """The headline frame with AO radius 16: timing, parity and traversal counts."""
    import numpy as np

    def frame():
        gs.update_resources_gpu(scene)
        ctx.light_pass(gs.tlas, gbuf.desc(), sb, out)

    ctx.set_stage_timing(False)
    for _ in range(max(args.warmup, 1)):
        frame()
    ctx.sync()
    steps = max(args.steps // 2, 5)
    t0 = time.perf_counter()
    for _ in range(steps):
        frame()
    ctx.sync()
    elapsed = time.perf_counter() - t0
    ctx.set_stage_timing(True)
    trace_ms = []
    for _ in range(5):
        frame()
        t = ctx.timings()
        trace_ms.append(t.trace_ms)
    rays = int(t.rays)
    traced = int(t.traced_rays)
    par, st = visibility_parity(ctx, gs, gbuf, o, host, sb, W, H, threads, args.cpu_row_step * 4, wide)
    nr = max(st.rays, 1)
    return {
        "ao_max": float(sb.ao_max), "steps": steps, "ms_per_frame": round(elapsed / steps * 1e3, 4),
        "mrays_per_s": round(rays * steps / elapsed / 1e6, 2), "trace_kernel_ms": round(float(np.median(trace_ms)), 4),
        "rays_per_frame": rays, "traced_rays_per_frame": traced,
        "per_ray_counts": {"wide_node": round(st.wide_node_visits / nr, 3),
                           "node": round(st.node_visits / nr, 3), "tri": round(st.tri_tests / nr, 3),
                           "inst": round(st.inst_visits / nr, 3)},
        "parity": par,
    }


def bench_batch(ctx, args) -> int:
    """Batched serving frames: per-frame p50/p99 latency and aggregate Mrays/s."""
    import numpy as np

    from luz_amd import autotune
    from luz_amd.batching import BatchWorkload

    t0 = time.perf_counter()
    wl = BatchWorkload(ctx, list(range(args.scenes)))
    setup_s = time.perf_counter() - t0
    print(f"batch: {args.scenes} scenes, {wl.pixels} px, {wl.triangles} instanced tris, "
          f"set-up {setup_s:.1f} s", file=sys.stderr, flush=True)
    tuned = None
    if args.autotune:
        tuned = autotune.tune_batch(ctx, wl)
        print(f"batch autotune: {({k: tuned[k] for k in autotune.KEYS})} {tuned['ms']:.2f} ms",
              file=sys.stderr, flush=True)
    res = wl.run(args.steps, warmup=args.warmup, first=1000)
    # one replayed frame checked scene by scene against the oracle on the smallest scenes
    parity = None
    if not args.no_cpu:
        from oracle import oracle as orc
        frame = 1000 + args.warmup + args.steps
        wl.frame(frame, True)
        order = sorted(range(args.scenes), key=lambda i: wl.gbufs[i].width * wl.gbufs[i].height)[:4]
        mism = 0
        checked = 0
        for i in order:
            h = wl.hosts[i]
            gbuf, sh, ao, _ = wl.outputs(i)
            sb = h.pack()
            sb.frame = frame
            o = orc.OracleScene()
            for m in range(h.mesh_count):
                v, idx = h.mesh(m)
                o.add_mesh(v, idx)
            o2w, mesh_ids, _ = h.nodes()
            o.set_instances(o2w, mesh_ids)
            W, H, _ = h.camera()
            wsh, wao, _ = o.trace_visibility(gbuf, W, H, sb, threads=cpu_threads()[0])
            mism += int((sh != wsh).sum() + (ao != wao).sum())
            checked += W * H
        parity = {"scenes_checked": len(order), "pixels_checked": checked, "count_mismatch": mism}
    res_set = sorted({(g.width, g.height) for g in wl.gbufs})
    record = {
        "metric": "batched frames: per-frame p50/p99 latency and aggregate shadow+AO Mrays/s",
        "value": round(res["mrays_per_s"], 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["mean_ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.scenes} seeded sphere-field scenes from the 1080p workload's generator, "
                         "tris log-uniform in [1e3, 5e5], resolutions from {320x240, 640x480, 1280x720, "
                         "1920x1080}; per frame: rigid animation, BLAS refit, TLAS rebuild, primary "
                         "G-buffer, 1 dir + 8 point lights, 9 shadow + 16 AO rays/px, GGX shading; one "
                         "hipGraph replay per frame"),
            "scenes": args.scenes, "pixels": wl.pixels, "instanced_tris": wl.triangles,
            "resolutions": [f"{w}x{h}" for w, h in res_set],
            "p50_ms": round(res["p50_ms"], 3), "p99_ms": round(res["p99_ms"], 3),
            "rays_per_frame": int(res["rays_per_frame"]),
        },
    }
    if tuned is not None:
        record["config"]["autotune"] = {k: tuned[k] for k in autotune.KEYS}
        record["config"]["autotune"]["table"] = tuned["table"]
    if parity is not None:
        record["parity"] = parity
    print(json.dumps(record), flush=True)
    wl.close()
    ctx.close()
    return 0


def box5(plane, fg, W: int, H: int):
    """Non-learned comparison: the mean of a 1-spp plane over the foreground
    pixels of the clamped 5x5 window (the base the denoiser's outputs are added to)."""
    import numpy as np

    v = np.where(fg, plane, 0.0).reshape(H, W)
    m = fg.reshape(H, W).astype(np.float64)
    acc, cnt = np.zeros((H, W)), np.zeros((H, W))
    for dy in range(-2, 3):
        ys = np.clip(np.arange(H) + dy, 0, H - 1)
        for dx in range(-2, 3):
            xs = np.clip(np.arange(W) + dx, 0, W - 1)
            acc += v[np.ix_(ys, xs)]
            cnt += m[np.ix_(ys, xs)]
    return (acc / np.maximum(cnt, 1.0)).reshape(-1)


def bench_denoise(ctx, args) -> int:
    """1-spp visibility + learned denoiser + lighting against the plain k-spp
    tracer on a held-out 1080p sphere-field scene: AO / visibility RMSE
    against 256-spp references (GPU FP32 on every pixel; FP64 oracle on a row
    sample) and GPU frame time of each; MFMA fraction of the denoiser."""
    import ctypes as C

    import numpy as np

    from luz_amd import api, denoiser as dnm

    if not dnm.WEIGHTS.exists():
        print(f"bench.py: no trained denoiser weights at {dnm.WEIGHTS}", file=sys.stderr)
        return 2
    seed = 900100
    scene = api.HostScene(api.SCENE_SPHERES, seed)
    gs = api.GPUScene(ctx, scene)
    W, H, cam = scene.camera()
    _, _, mats = scene.nodes()
    gd = api.GBufferDevice(ctx, W, H)
    ctx.primary_gbuffer(gs.tlas, cam, mats, gd.desc())
    sb = scene.pack()
    sb.ao_max = dnm.AO_RADIUS   # AO out to 16: with the reference's radius of 1 AO is 1 everywhere here
    L, n = sb.num_lights, W * H

    def block(spp, frame):
        b = api.SceneBlock()
        C.memmove(C.byref(b), C.byref(sb), C.sizeof(api.SceneBlock))
        for l in range(b.num_lights):
            b.lights[l].num_shadow_samples = spp
        b.ao_num_samples = spp
        b.frame = frame
        return b

    sh = ctx.buffer(max(L, 1) * n)
    ao = ctx.buffer(n)
    out = ctx.buffer(n * 16)
    vis_o, ao_o = ctx.buffer(max(L, 1) * n * 4), ctx.buffer(n * 4)
    dn = dnm.Denoiser(ctx, 0)
    dn.load(dnm.WEIGHTS)
    fr = dnm.Frame(ctx, W, H, gd.position, gd.normal, tuple(sb.cam_pos), L, sh, 1, ao, 1)
    reps = max(args.steps // 5, 3)

    def vis_ms(t):
        return t.compact_ms + t.raygen_ms + t.trace_ms + t.resolve_ms

    # 256-spp reference on the GPU (two independent 128-spp passes)
    ref_sh = np.zeros((max(L, 1), n))
    ref_ao = np.zeros(n)
    for frame in (1, 2):
        ctx.trace_visibility(gs.tlas, gd.desc(), block(128, frame), sh, ao)
        ref_sh += sh.download(np.uint8, (max(L, 1), n))
        ref_ao += ao.download(np.uint8, n)
    ref_ao /= 256.0
    ref_vis = 1.0 - ref_sh[:L] / 256.0
    host = gd.download()
    fg = np.abs(host["normal"]).sum(1) != 0

    def rmse(a, b, m):
        return float(np.sqrt(np.mean((a[..., m] - b[..., m]) ** 2)))

    # the denoised frame: 1-spp visibility + denoiser + lighting
    b1 = block(1, 0)
    times = []
    for r in range(reps + 1):
        ctx.trace_visibility(gs.tlas, gd.desc(), b1, sh, ao)
        t_v = ctx.timings()
        dn.denoise(fr, ao_o, vis_o)
        t_d = ctx.timings()
        ctx.shade_fractions(gd.desc(), sb, vis_o, ao_o, out)
        t_s = ctx.timings()
        if r:
            times.append((vis_ms(t_v), t_d.denoise_ms, t_s.shade_ms))
    tv, td, ts = (float(np.median([t[i] for t in times])) for i in range(3))
    dn_ao = ao_o.download(np.float32, n).astype(np.float64)
    dn_vis = vis_o.download(np.float32, (max(L, 1), n)).astype(np.float64)[:L]
    raw_ao = ao.download(np.uint8, n).astype(np.float64)
    raw_vis = 1.0 - sh.download(np.uint8, (max(L, 1), n))[:L].astype(np.float64)
    box_ao, box_vis = box5(raw_ao, fg, W, H), np.array([box5(v, fg, W, H) for v in raw_vis])
    rows = []
    for k in (1, 2, 4, 8, 16, 32, 64):
        bk = block(k, 0)
        tt = []
        for r in range(reps + 1):
            ctx.trace_visibility(gs.tlas, gd.desc(), bk, sh, ao)
            t_v = ctx.timings()
            ctx.shade(gd.desc(), bk, sh, ao, out)
            t_s = ctx.timings()
            if r:
                tt.append(vis_ms(t_v) + t_s.shade_ms)
        a_k = ao.download(np.uint8, n).astype(np.float64) / k
        v_k = 1.0 - sh.download(np.uint8, (max(L, 1), n))[:L].astype(np.float64) / k
        rows.append({"spp": k, "frame_ms": round(float(np.median(tt)), 4),
                     "ao_rmse": round(rmse(a_k, ref_ao, fg), 6), "vis_rmse": round(rmse(v_k, ref_vis, fg), 6),
                     "_ao": a_k})
    den = {"frame_ms": round(tv + td + ts, 4), "visibility_ms": round(tv, 4), "denoise_ms": round(td, 4),
           "shade_ms": round(ts, 4), "ao_rmse": round(rmse(dn_ao, ref_ao, fg), 6),
           "vis_rmse": round(rmse(dn_vis, ref_vis, fg), 6),
           "raw_1spp_ao_rmse": round(rmse(raw_ao, ref_ao, fg), 6),
           "raw_1spp_vis_rmse": round(rmse(raw_vis, ref_vis, fg), 6),
           "box5_1spp_ao_rmse": round(rmse(box_ao, ref_ao, fg), 6),
           "box5_1spp_vis_rmse": round(rmse(box_vis, ref_vis, fg), 6)}
    # equal accuracy: the first plain sample count whose RMSE against the 256-spp reference is
    # at or below the denoised frame's, for AO (the metric) and for the soft-shadow visibility
    match = next((r for r in rows if r["ao_rmse"] <= den["ao_rmse"]), None)
    match_vis = next((r for r in rows if r["vis_rmse"] <= den["vis_rmse"]), None)
    spp64 = next(r for r in rows if r["spp"] == 64)
    # algorithmic FLOPs of k_dn_infer per foreground pixel: the 125 light-independent
    # inputs of layer 1 once, then per light the 25 visibility inputs of layer 1,
    # layers 2..4 (64x64) and layer 5 (64x2)
    flops_px = 2.0 * 125 * 64 + max(L, 1) * 2.0 * (25 * 64 + 3 * 64 * 64 + 64 * 2)
    flops = float(fg.sum()) * flops_px
    achieved = flops / (td * 1e-3) / 1e12
    dn_traffic = None
    tpath = ROOT / "profiles" / "k_dn_infer_traffic.json"
    if tpath.exists():
        try:
            dn_traffic = json.loads(tpath.read_text()).get("hbm_bytes_per_launch")
        except (ValueError, OSError):
            dn_traffic = None
    record = {
        "metric": "AO RMSE vs FP64 256-spp oracle and ms/frame: 1-spp + learned denoiser vs plain multi-spp",
        "value": den["frame_ms"], "unit": "ms/frame", "n_gpus": 1, "steps": reps, "warmup": 1,
        "ms_per_step": den["frame_ms"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 (MFMA, FP32 accumulate)", "data": "synthetic",
        "config": {"workload": "1920x1080 sphere field (held-out scene seed 900100), 1 shadow ray per "
                               "light (9 lights) + 1 AO ray (radius 16) per pixel, per-pixel MLP 150-64x4-2 on 9 (pixel, "
                               "light) samples per pixel",
                   "denoised": den,
                   "plain": [{k: v for k, v in r.items() if k != "_ao"} for r in rows],
                   "ao_radius": dnm.AO_RADIUS,
                   "equal_accuracy_spp": match["spp"] if match else None,
                   "equal_accuracy_on": "AO RMSE vs the 256-spp reference",
                   "ao_fraction_below_1_ref256": round(float(np.mean(ref_ao[fg] < 1.0)), 6),
                   "speedup_vs_equal_accuracy": round(match["frame_ms"] / den["frame_ms"], 2) if match else None,
                   "visibility_equal_accuracy_spp": match_vis["spp"] if match_vis else None,
                   "visibility_speedup_vs_equal_accuracy":
                       round(match_vis["frame_ms"] / den["frame_ms"], 2) if match_vis else None,
                   "speedup_vs_64spp": round(spp64["frame_ms"] / den["frame_ms"], 2)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 1), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(achieved / 2500.0, 4), "traffic": dn_traffic, "kernel": "k_dn_infer",
                     "flops_per_pixel": flops_px, "timed": "denoise stage: per-pixel prepass + k_dn_infer"},
    }
    if not args.no_cpu:
        from oracle import oracle as orc
        o = orc.OracleScene()
        for i in range(scene.mesh_count):
            v, idx = scene.mesh(i)
            o.add_mesh(v, idx)
        o2w, mesh_ids, _ = scene.nodes()
        o.set_instances(o2w, mesh_ids)
        step = 16
        threads = cpu_threads()[0]
        f_sh = np.zeros((max(L, 1), n))
        f_ao = np.zeros(n)
        c0 = time.perf_counter()
        for frame in (1, 2):
            dsh, dao, _ = o.trace_visibility(host, W, H, block(128, frame), threads=threads, row_step=step, f64=True)
            f_sh += dsh
            f_ao += dao
        cpu_s = time.perf_counter() - c0
        rows_m = np.zeros((H, W), bool)
        rows_m[::step] = True
        m = rows_m.reshape(-1) & fg
        f_ao /= 256.0
        f_vis = 1.0 - f_sh[:L] / 256.0
        record["config"]["fp64_oracle"] = {
            "rows_every": step, "pixels": int(m.sum()), "cpu_s": round(cpu_s, 1), "threads": threads,
            "denoised_ao_rmse": round(rmse(dn_ao, f_ao, m), 6),
            "denoised_vis_rmse": round(rmse(dn_vis, f_vis, m), 6),
            "gpu_ref256_vs_fp64_ao_rmse": round(rmse(ref_ao, f_ao, m), 6),
            "plain_ao_rmse": {r["spp"]: round(rmse(r["_ao"], f_ao, m), 6) for r in rows},
        }
    print(json.dumps(record), flush=True)
    dn.close()
    gs.close()
    ctx.close()
    return 0


def compact_parity(ctx, gs, scene, gbuf, sb, out, W, H, trace_avg, rays, prof) -> dict:
    """Precision form: hit/miss mismatch and radiance RMSE against the FP64 oracle,
    exactness against the FP32 oracle, on row samples of the 4K frame; the
    8-wide traversal kernel against its rooflines as roofline() rates the
    binary one (algorithmic bytes with 96 bytes per 8-wide node, the bound from
    the profiled HBM traffic and VALU instructions per launch)."""
    import numpy as np

    from oracle import oracle as orc
    nl = sb.num_lights
    host = gbuf.download()
    dsh, dao = ctx.buffer(nl * W * H), ctx.buffer(W * H)
    ctx.trace_visibility(gs.tlas, gbuf.desc(), sb, dsh, dao)
    ctx.shade(gbuf.desc(), sb, dsh, dao, out)
    gsh = dsh.download(np.uint8, (nl, W * H))
    gao = dao.download(np.uint8, W * H)
    rad = out.download(np.float32, (W * H, 4))[:, :3].astype(np.float64)
    o = orc.OracleScene()
    for i in range(scene.mesh_count):
        v, idx = scene.mesh(i)
        o.add_mesh(v, idx)
    o2w, mesh_ids, _ = scene.nodes()
    o.set_instances(o2w, mesh_ids)
    o.build_wide()
    threads = cpu_threads()[0]
    step32, step64 = 16, 64
    c0 = time.perf_counter()
    wsh, wao, st = o.trace_visibility(host, W, H, sb, threads=threads, row_step=step32, wide=True)
    cpu_s = time.perf_counter() - c0
    dsh64, dao64, _ = o.trace_visibility(host, W, H, sb, threads=threads, row_step=step64, f64=True)
    act = np.abs(host["normal"]).sum(1) != 0
    rows32 = np.zeros((H, W), bool); rows32[::step32] = True; m32 = rows32.reshape(-1)
    rows64 = np.zeros((H, W), bool); rows64[::step64] = True; m64 = rows64.reshape(-1) & act
    ref64 = orc.shade(host, W, H, sb, dsh64, dao64, f64=True)[:, :3]
    rmse = float(np.sqrt(np.mean((rad[m64] - ref64[m64]) ** 2)))
    scale = float(np.sqrt(np.mean(ref64[m64] ** 2)))
    na = max(sb.ao_num_samples, 1)
    per_ray = (96.0 * st.wide_node_visits + 64.0 * st.node_visits + 36.0 * st.tri_tests +
               64.0 * st.inst_visits) / max(st.rays, 1)
    pixels = int(act.sum())
    t = trace_avg * 1e-3
    achieved = (rays * per_ray + 24.0 * pixels + (nl + 1) * pixels) / t / 1e9
    cus, clk_khz = ctx.device_info()
    clock_hz = max(clk_khz, 1) * 1e3
    traffic, valu = prof.get("hbm_bytes_per_launch"), prof.get("valu_insts_per_launch")
    valu_frac = valu / (cus * 4 * clock_hz / 4) / t if valu else None
    hbm_frac = traffic / t / 1e9 / HBM_PEAK_GBS if traffic else None
    bound = "valu" if valu_frac is not None and valu_frac > (hbm_frac or 0.0) else "hbm"
    return {
        "roofline": {"bound": bound, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": "k_trace_any<wide>",
                     "algorithmic_bytes_per_ray": round(per_ray, 2),
                     "traffic_frac": round(hbm_frac, 4) if hbm_frac else None,
                     "valu_issue_frac": round(valu_frac, 4) if valu_frac else None,
                     "valu_insts_per_launch": valu, "counters_from": prof.get("source"),
                     "bytes_per_ray_with_pixel_records": round(per_ray + 33.0, 2)},
        "cpu_baseline": {"value": round(st.rays / cpu_s / 1e6, 2), "unit": "Mrays/s", "cores": threads,
                         "kind": "port",
                         "sample": f"oracle FP32 8-wide visibility on every {step32}th row ({st.rays} rays)"},
        "parity": {
            "fp32_oracle_rows_every": step32,
            "shadow_mismatch_vs_fp32": int((gsh[:, m32] != wsh[:, m32]).sum()),
            "ao_mismatch_vs_fp32": int((gao[m32] != wao[m32]).sum()),
            "fp64_oracle_rows_every": step64,
            "count_mismatch_vs_fp64": int((gsh[:, m64] != dsh64[:, m64]).sum() + (gao[m64] != dao64[m64]).sum()),
            "ao_rmse_vs_fp64": round(float(np.sqrt(np.mean((gao[m64] / na - dao64[m64] / na) ** 2))), 6),
            "radiance_rmse_vs_fp64": round(rmse, 6),
            "radiance_rmse_relative": round(rmse / max(scale, 1e-30), 6),
        },
    }


def calibrate_copy(mib: int) -> int:
    """A plain device-to-device copy of known size: its FETCH_SIZE / WRITE_SIZE
    readings scale the traversal kernel's counters to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": n * 4, "launches": 3}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
