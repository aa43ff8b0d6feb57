"""Autotuner for the traversal launch parameters.

Searches the persistent traversal's launch shape on the live workload: the
number of resident workgroups per CU (wave occupancy: 4 waves per group) and
the refill threshold (idle lanes of a wave that trigger a refill from the
wave's ray chunk).  Each candidate is timed with the library's own HIP events
around the traversal kernel; the best median wins.

    from luz_amd import autotune
    best = autotune.tune_visibility(ctx, tlas, gbuffer_desc, scene_block)
"""
from __future__ import annotations

import itertools
import json
import statistics
from pathlib import Path

from . import api

GROUPS_PER_CU = (4, 6, 8, 10)
REFILL = (8, 16, 32, 48)


def tune_visibility(ctx: api.Context, tlas: api.Tlas, g: api.GBuffer, sb: api.SceneBlock,
                    reps: int = 3, groups=GROUPS_PER_CU, refills=REFILL, cache: Path | None = None,
                    key: str | None = None) -> dict:
    """Returns {"groups_per_cu", "refill", "trace_ms", "table"} and applies the best."""
    if cache is not None and key and cache.exists():
        known = json.loads(cache.read_text())
        if key in known:
            best = known[key]
            ctx.set_trace_config(best["groups_per_cu"], best["refill"])
            return best
    nl = max(sb.num_lights, 1)
    shadow = ctx.buffer(nl * g.width * g.height)
    ao = ctx.buffer(g.width * g.height)
    table = []
    for gpc, refill in itertools.product(groups, refills):
        ctx.set_trace_config(gpc, refill)
        ctx.trace_visibility(tlas, g, sb, shadow, ao)   # warm-up (allocations, caches)
        ctx.timings()
        times = []
        for _ in range(reps):
            ctx.trace_visibility(tlas, g, sb, shadow, ao)
            times.append(ctx.timings().trace_ms)
        table.append({"groups_per_cu": gpc, "refill": refill, "trace_ms": statistics.median(times)})
    best = min(table, key=lambda r: r["trace_ms"])
    ctx.set_trace_config(best["groups_per_cu"], best["refill"])
    result = dict(best, table=table)
    if cache is not None and key:
        known = json.loads(cache.read_text()) if cache.exists() else {}
        known[key] = {k: result[k] for k in ("groups_per_cu", "refill", "trace_ms")}
        cache.write_text(json.dumps(known, indent=1))
    return result
