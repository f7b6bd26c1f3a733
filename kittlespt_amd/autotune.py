"""Autotuner of the wavefront launch shapes.

Knobs (kpt_set_tuning): lane refill on/off, queue entries claimed per wave
atomic (with or without refill), refill threshold, persistent workgroups per compute unit, material
sort on/off, per-frame hipGraph replay on/off, spatial ray sort on/off, the
bounce from which the remaining paths finish in one path-per-lane launch, the
traversal stack entries per lane held in LDS (the rest in private memory) and
the workgroup size of the trace, shadow and tail kernels.
The tuner renders real frames of the caller's scene and camera
and keeps the fastest setting by coordinate descent (each knob swept with the
others fixed, two passes).  Results are cached in a JSON file keyed by the
workload so later runs reuse them.
"""
from __future__ import annotations

import json
import os
import time

SPACE = {
    "refill": [0, 1],
    "blocks_per_cu": [4, 8, 12, 16],
    "chunk": [64, 128, 256, 512],
    "refill_min": [8, 16, 32],
    "sort": [0, 1, 2],
    "graph": [0, 1],
    "ray_sort": [0, 1],
    "tail": [0, 1, 2, 3, 4],
    "stack": [8, 16, 24, 32],
    "block": [64, 128, 256],
    "suspend": [0, 8, 16, 24],
    "overlap": [0, 1],
}
ORDER = ["overlap", "tail", "suspend", "ray_sort", "refill", "blocks_per_cu", "stack", "block", "sort", "graph", "chunk",
         "refill_min"]


def measure(ctx, cfg, spp=1, repeats=2):
    ctx.set_tuning(**cfg)
    ctx.render(spp)                       # warm the setting
    best = float("inf")
    for _ in range(repeats):
        ctx.render(spp)
        best = min(best, ctx.timing()[0])
    return best


def tune(ctx, spp=1, repeats=2, passes=2, log=None, skip=()):
    cur = ctx.tuning()
    cur_t = measure(ctx, cur, spp, repeats)
    history = [(dict(cur), cur_t)]
    for _ in range(passes):
        improved = False
        for knob in ORDER:
            if knob in skip or (knob == "refill_min" and not cur["refill"]):
                continue
            for v in SPACE[knob]:
                if v == cur[knob]:
                    continue
                cand = dict(cur)
                cand[knob] = v
                t = measure(ctx, cand, spp, repeats)
                history.append((cand, t))
                if log:
                    log(f"[tune] {cand} -> {t:.2f} ms")
                if t < cur_t * 0.98:
                    cur, cur_t, improved = cand, t, True
        if not improved:
            break
    ctx.set_tuning(**cur)
    ctx.reset()
    return cur, cur_t, history


def cached_tune(ctx, key, path, spp=1, log=None):
    db = {}
    if os.path.exists(path):
        try:
            db = json.load(open(path))
        except (OSError, ValueError):
            db = {}
    if key in db:
        ctx.set_tuning(**db[key]["tuning"])
        return db[key]["tuning"], db[key]["ms"], True
    t0 = time.time()
    best, ms, _ = tune(ctx, spp=spp, log=log)
    db[key] = {"tuning": best, "ms": ms, "tuned_in_s": round(time.time() - t0, 2)}
    try:
        with open(path, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
    except OSError:
        pass
    return best, ms, False
