#!/usr/bin/env python3
"""Benchmark: Msamples/s of the path tracer on the 1080p instanced workload.

Workload (one step = one pass of the hot path over one batch): 1920x1080,
8 samples per pixel (8 reference frames of 1 spp with static accumulation),
max depth 8, 1,000 instances of 8 seeded BLAS (50k-200k triangles each),
16 textured GGX materials, Nishita sky.  Synthetic seeded data; scene and
buffers resident in HBM before timing starts.

Prints ONE JSON line with the metric, a roofline object for the dominant
kernel (closest-hit traversal, HBM-bound) and a CPU baseline (the oracle on
the host cores, bounded pixel sample).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from kittlespt_amd import abi, autotune, scenes  # noqa: E402
from kittlespt_amd.context import Context  # noqa: E402

METRIC = "Msamples/s @1080p + time-to-target-RMSE vs FP64, cached vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0

# algorithmic bytes per traversal event, reference node/triangle layout
# (BVH2Node 32 B, Triangle positions 36 B + 4 B index, BLAS2 matrix + ids 72 B)
BYTES_PER_EVENT = {"blas_interior_pops": 64, "blas_leaf_pops": 32, "tri_tests": 40, "blas_entries": 72,
                   "tlas_pops": 32}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_settings(depth, integrator):
    s = abi.KptSettings()
    s.max_depth = depth
    s.bloom_final_blend = 0.175
    s.bloom_internal_blend = 0.65
    s.sun_angular_diameter_rad = 0.0087
    s.sun_phi_rad = 4.18879
    s.sun_theta_rad = math.radians(30.0)
    s.sun_emission_factor = 1.0
    s.seed_mode = abi.SEED_XOR
    s.integrator = integrator
    return s


def cpu_baseline(scene, st, cam, W, H, budget_s):
    import oracle
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    t0 = time.time()
    orc = oracle.OracleScene(scene)
    log(f"[cpu] oracle scene built in {time.time() - t0:.1f}s")
    rng = np.random.default_rng(1234)
    n = 2048
    spent = 0.0
    total_samples = 0
    while True:
        pix = rng.choice(W * H, size=n, replace=False).astype(np.int32)
        t0 = time.time()
        orc.render_pixels(st, cam, W, H, pix, 0, 1, threads=cores)
        dt = time.time() - t0
        spent += dt
        total_samples += n
        log(f"[cpu] {n} pixel-samples in {dt:.2f}s")
        if spent >= budget_s * 0.5 or n >= 1 << 20:
            break
        n = int(min(1 << 20, n * max(2.0, min(8.0, (budget_s * 0.5) / max(dt, 1e-3)))))
    orc.close()
    return {"value": total_samples / spent / 1e6, "unit": "Msamples/s", "cores": cores, "kind": "port",
            "sample": f"{total_samples} random pixels of the 1920x1080 frame, 1 sample each, depth "
                      f"{st.max_depth}, same scene/camera/seeds, oracle C++ (IEEE float) on {cores} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=8)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--integrator", choices=["wavefront", "megakernel"], default="wavefront")
    ap.add_argument("--instances", type=int, default=1000)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tuning", default="", help="fixed knobs, e.g. refill=0,blocks_per_cu=8,sort=1")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="only run one device-to-device copy of this many bytes (HBM counter calibration)")
    args = ap.parse_args()

    if args.calibrate_copy:
        from kittlespt_amd.context import DeviceBuffer, hip
        a, b = DeviceBuffer(args.calibrate_copy), DeviceBuffer(args.calibrate_copy)
        rc = hip().hipMemcpy(b.ptr, a.ptr, args.calibrate_copy, 3)   # device to device
        hip().hipDeviceSynchronize()
        print(json.dumps({"calibrate_copy_bytes": args.calibrate_copy, "rc": rc}), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build measures one GPU "
              f"(multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    t0 = time.time()
    scene = scenes.scene_instanced(num_instances=args.instances)
    log(f"[bench] scene generated in {time.time() - t0:.1f}s: {len(scene.positions)} triangles, "
        f"{scene.num_instances} instances")
    integ = abi.INTEGRATOR_WAVEFRONT if args.integrator == "wavefront" else abi.INTEGRATOR_MEGAKERNEL
    W, H = args.width, args.height
    ctx = Context(0)
    t0 = time.time()
    ctx.upload(scene)
    log(f"[bench] scene built and uploaded in {time.time() - t0:.1f}s: info {ctx.info()}")
    st = make_settings(args.depth, integ)
    ctx.set_settings(st)
    ctx.resize(W, H)
    proj = scenes.perspective_fov_lh(scene.camera_fov, W, H, 1.0, 100.0)
    view = scenes.look_at_view(scene.camera_eye, scene.camera_target)
    ctx.set_view(proj, view)
    cam = ctx.camera()

    if integ == abi.INTEGRATOR_WAVEFRONT and args.tuning:
        kv = dict(x.split("=") for x in args.tuning.split(","))
        ctx.set_tuning(**{k: int(v) for k, v in kv.items()})
    elif integ == abi.INTEGRATOR_WAVEFRONT and not args.no_autotune:
        t0 = time.time()
        best, ms, _ = autotune.tune(ctx, spp=1, log=log)
        log(f"[bench] autotuned in {time.time() - t0:.1f}s: {best} ({ms:.2f} ms per 1-spp frame)")
    tuning = ctx.tuning() if integ == abi.INTEGRATOR_WAVEFRONT else None

    for i in range(args.warmup):
        ctx.render(args.spp)
        log(f"[bench] warmup step {i} {ctx.timing()[0]:.1f} ms")

    # traversal event counts of one step (measurement pass, outside the timed region)
    counted = None
    if integ == abi.INTEGRATOR_WAVEFRONT:
        ctx.set_trace_counting(True)
        ctx.render(args.spp)
        counted = ctx.render_counters()
        count_launches = ctx.timing()[2]
        ctx.set_trace_counting(False)
        log(f"[bench] counted step: {counted} over {count_launches:.0f} trace launches")

    ctx.reset()
    render_ms, trace_ms, launches = 0.0, 0.0, 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        ctx.render(args.spp)
        tm = ctx.timing()
        render_ms += tm[0]
        trace_ms += tm[1]
        launches += tm[2]
    wall = time.perf_counter() - t0
    samples = W * H * args.spp * args.steps
    value = samples / wall / 1e6
    log(f"[bench] {args.steps} steps in {wall:.3f}s (event time {render_ms:.1f} ms), trace {trace_ms:.1f} ms "
        f"over {launches:.0f} launches")

    roof = None
    if counted is not None and launches > 0:
        bytes_step = sum(BYTES_PER_EVENT[k] * counted[k] for k in BYTES_PER_EVENT)
        bytes_per_launch = bytes_step / count_launches
        avg_launch_s = (trace_ms / launches) / 1e3
        achieved = bytes_per_launch / avg_launch_s / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "kernel": "k_trace (wavefront closest hit)",
                "bytes_per_launch": int(bytes_per_launch), "avg_launch_ms": round(avg_launch_s * 1e3, 4)}

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(scene, st, cam, W, H, args.cpu_budget)
        except Exception as e:  # the baseline is reported, never fatal
            log(f"[cpu] baseline failed: {e}")

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded scene generators)",
        "config": {"workload": f"instanced {W}x{H}, {args.spp} spp, depth {args.depth}, "
                               f"{scene.num_instances} instances of 8 BLAS, 16 textured materials, Nishita sky",
                   "integrator": args.integrator, "width": W, "height": H, "spp": args.spp,
                   "max_depth": args.depth, "triangles": int(len(scene.positions)),
                   "instances": int(scene.num_instances), "parallelism": "single GPU", "tuning": tuning},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
