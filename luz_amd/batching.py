"""Batched serving workload: many independent scenes rendered per frame.

Each scene is a seeded LUZSCENE_RANDOM sphere field, drawn by the 1080p
sphere field's generator (triangle budget log-uniform in [1e3, 5e5],
resolution drawn from {320x240, 640x480, 1280x720, 1920x1080}, 1 directional
+ 8 point lights, 1 shadow ray per light + 16 AO rays per pixel) with its
own BLASes, TLAS, G-buffer and light buffer.
Per frame every scene's mesh nodes move (seeded rigid motion,
luzscene_animate), their instances are staged, and one luzrt_batch_frame
refits the BLASes, rebuilds all TLASes, renders the G-buffers by primary
rays, traces every scene's shadow and AO rays in one persistent launch and
shades, replayed from a hipGraph.
"""
from __future__ import annotations

import time

import numpy as np

from . import api


class BatchWorkload:
    def __init__(self, ctx: api.Context, specs, refit: bool = True):
        """specs: seeds (budget and resolution drawn from the seed) or
        (seed, triangle budget, width, height) tuples."""
        self.ctx = ctx
        self.hosts, self.gpus, self.gbufs, self.outs = [], [], [], []
        self.blocks = []
        entries = []
        for spec in specs:
            seed, budget, width, height = spec if isinstance(spec, tuple) else (spec, 0, 0, 0)
            h = api.HostScene(api.SCENE_RANDOM, seed, budget, width, height)
            g = api.GPUScene(ctx, h)
            W, H, cam = h.camera()
            _, _, mats = h.nodes()
            sb = h.pack()
            gb = api.GBufferDevice(ctx, W, H)
            sh = ctx.buffer(max(sb.num_lights, 1) * W * H)
            ao = ctx.buffer(W * H)
            out = ctx.buffer(W * H * 16)
            self.hosts.append(h)
            self.gpus.append(g)
            self.gbufs.append(gb)
            self.outs.append((sh, ao, out))
            self.blocks.append(sb)
            entries.append({"tlas": g.tlas, "gbuffer": gb.desc(), "scene": sb, "camera": cam,
                            "materials": mats, "shadow": sh, "ao": ao, "out": out})
        self.batch = api.Batch(ctx, entries)
        if refit:
            for h, g in zip(self.hosts, self.gpus):
                for m in range(h.mesh_count):
                    self.batch.add_refit(g.blas(m), g.vertices(m))

    @property
    def pixels(self) -> int:
        return sum(g.width * g.height for g in self.gbufs)

    @property
    def triangles(self) -> int:
        return sum(h.instanced_triangles for h in self.hosts)

    def stage(self, frame: int) -> None:
        """Animate every scene to `frame` and stage its instances."""
        for i, (h, g) in enumerate(zip(self.hosts, self.gpus)):
            h.animate(frame)
            arr, n = g.instances(h)
            self.batch.set_instances(i, arr, n)

    def frame(self, frame: int, use_graph: bool = True) -> int:
        """One synchronous frame; returns the rays traced."""
        self.stage(frame)
        self.batch.frame(frame, use_graph)
        return self.batch.rays()

    def run(self, frames: int, warmup: int = 2, first: int = 0, use_graph: bool = True) -> dict:
        """Per-frame wall-clock latency (staging through completion) and rays."""
        for f in range(warmup):
            self.frame(first + f, use_graph)
        lat, rays = [], []
        for f in range(frames):
            t0 = time.perf_counter()
            r = self.frame(first + warmup + f, use_graph)
            lat.append((time.perf_counter() - t0) * 1e3)
            rays.append(r)
        lat = np.array(lat)
        return {"frames": frames, "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
                "rays_per_frame": float(np.mean(rays)),
                "mrays_per_s": float(np.sum(rays) / (lat.sum() * 1e-3) / 1e6)}

    def outputs(self, i: int):
        """Scene i's G-buffer planes, shadow counts, AO counts and radiance."""
        gb = self.gbufs[i]
        n = gb.width * gb.height
        sh, ao, out = self.outs[i]
        L = self.blocks[i].num_lights
        return (gb.download(), sh.download(np.uint8, (L, n)), ao.download(np.uint8, n),
                out.download(np.float32, (n, 4)))

    def close(self) -> None:
        self.batch.close()
        for g in self.gpus:
            g.close()
        for h in self.hosts:
            h.close()
