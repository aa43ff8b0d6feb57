"""Autotuner for the traversal launch parameters.

Searches, on the live workload, the launch shape of the visibility pipeline:
  groups_per_cu  resident workgroups per CU of the persistent traversal
                 (wave count: 4 waves per group),
  refill         idle lanes of a wave that trigger a refill from the wave's
                 ray chunk,
  stack_entries  LDS stack entries per lane (8, 16, 32) before spilling,
  tile_width     ray-tile shape of the pixel compaction (8x8, 16x4, 32x2
                 pixels per 64-lane wave),
  ray_order      order of the rays of a visibility call (single scene only):
                 slot-major or 64-pixel groups (luzrt_set_ray_order).
The search is a coordinate descent: the (groups_per_cu, refill) grid first,
then the stack depth, then the tile shape, then the ray order, each with the
others fixed at the best so far.  Single-scene candidates are timed with the library's own HIP
events around the visibility stages; batched frames by the wall-clock
latency of whole frames (hipGraph replays).  The best median wins.

    from luz_amd import autotune
    best = autotune.tune_visibility(ctx, tlas, gbuffer_desc, scene_block)
    best = autotune.tune_batch(ctx, batch_workload)
"""
from __future__ import annotations

import itertools
import json
import statistics
import time
from pathlib import Path

from . import api

GROUPS_PER_CU = (4, 6, 8, 10)
REFILL = (8, 16, 32, 48)
STACK_ENTRIES = (8, 16, 32)
TILE_WIDTH = (8, 16, 32)
RAY_ORDER = (api.RAY_ORDER_SLOT, api.RAY_ORDER_GROUPED)
KEYS = ("groups_per_cu", "refill", "stack_entries", "tile_width", "ray_order")


def _apply(ctx: api.Context, cfg: dict) -> None:
    ctx.set_trace_config(cfg["groups_per_cu"], cfg["refill"], cfg["stack_entries"], cfg["tile_width"])
    ctx.set_ray_order(cfg["ray_order"])


def _descend(ctx: api.Context, measure, groups, refills, stacks, tiles, start: dict, orders=()):
    table = []
    best = dict(start)

    def run(cands):
        nonlocal best
        for c in cands:
            cfg = dict(best, **c)
            _apply(ctx, cfg)
            ms = measure()
            table.append(dict(cfg, ms=ms))
        top = min((r for r in table), key=lambda r: r["ms"])
        best = {k: top[k] for k in KEYS}

    run([{"groups_per_cu": g, "refill": r} for g, r in itertools.product(groups, refills)])
    run([{"stack_entries": s} for s in stacks])
    run([{"tile_width": t} for t in tiles])
    if orders:
        run([{"ray_order": o} for o in orders])
    _apply(ctx, best)
    top = min(table, key=lambda r: r["ms"])
    return dict(best, ms=top["ms"], table=table)


def _cached(ctx, cache, key):
    if cache is not None and key and cache.exists():
        known = json.loads(cache.read_text())
        if key in known:
            cfg = dict(START, **known[key])   # entries stored before a key existed take its start value
            _apply(ctx, cfg)
            return cfg
    return None


def _store(cache, key, result):
    if cache is not None and key:
        known = json.loads(cache.read_text()) if cache.exists() else {}
        known[key] = {k: result[k] for k in KEYS + ("ms",)}
        cache.write_text(json.dumps(known, indent=1))


START = {"groups_per_cu": 6, "refill": 48, "stack_entries": 16, "tile_width": 8,
         "ray_order": api.RAY_ORDER_AUTO}


def tune_visibility(ctx: api.Context, tlas: api.Tlas, g: api.GBuffer, sb: api.SceneBlock,
                    reps: int = 3, groups=GROUPS_PER_CU, refills=REFILL, stacks=STACK_ENTRIES,
                    tiles=TILE_WIDTH, orders=RAY_ORDER, cache: Path | None = None,
                    key: str | None = None) -> dict:
    """Returns the best configuration (KEYS, "ms" = visibility ms, "table") and applies it."""
    hit = _cached(ctx, cache, key)
    if hit is not None:
        return hit
    nl = max(sb.num_lights, 1)
    shadow = ctx.buffer(nl * g.width * g.height)
    ao = ctx.buffer(g.width * g.height)

    def measure():
        ctx.trace_visibility(tlas, g, sb, shadow, ao)   # warm-up (allocations, caches)
        ctx.timings()
        times = []
        for _ in range(reps):
            ctx.trace_visibility(tlas, g, sb, shadow, ao)
            t = ctx.timings()
            times.append(t.compact_ms + t.raygen_ms + t.trace_ms + t.resolve_ms)
        return statistics.median(times)

    result = _descend(ctx, measure, groups, refills, stacks, tiles, START, orders)
    _store(cache, key, result)
    return result


def tune_batch(ctx: api.Context, workload, reps: int = 3, groups=GROUPS_PER_CU, refills=REFILL,
               stacks=STACK_ENTRIES, tiles=TILE_WIDTH, frame: int = 0) -> dict:
    """Tunes batched frames (luz_amd.batching.BatchWorkload) by median frame latency (the ray
    order of a batch is fixed when it is created and is not searched here)."""

    def measure():
        workload.frame(frame, True)                       # capture + first replay
        times = []
        for i in range(reps):
            t0 = time.perf_counter()
            workload.frame(frame + 1 + i, True)
            times.append((time.perf_counter() - t0) * 1e3)
        return statistics.median(times)

    return _descend(ctx, measure, groups, refills, stacks, tiles, START)
