#!/usr/bin/env python3
"""Benchmark: Msamples/s of the path tracer on the 1080p instanced workload.

Workload (one step = one pass of the hot path over one batch): 1920x1080,
8 samples per pixel (8 reference frames of 1 spp with static accumulation),
max depth 8, 1,000 instances of 8 seeded BLAS (50k-200k triangles each),
16 textured GGX materials, Nishita sky.  Synthetic seeded data; scene and
buffers resident in HBM before timing starts.

Prints ONE JSON line with the metric, a roofline object for the dominant
kernel (closest-hit traversal, HBM-bound), a CPU baseline (the oracle on
the host cores, bounded pixel sample) and the time-to-target-RMSE record:
the plain wavefront tracer's RMSE at --ttt-spp against an FP64 converged
render of a random pixel subset is the target, and the wall-clock (HIP event
frame times) each method needs to reach it is reported, the cached method
training its network online from random weights with its warm-up counted.
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
# path-per-lane tail, per path it finishes: its queue entry (4 B) and state read
# (ray origin, direction, throughput, radiance, previous vertex: 5 x 16 B),
# its radiance written (16 B); shading reads (attributes, textures) not counted
TAIL_BYTES_PER_PATH = 100
STAGE_KERNELS = {"trace": "k_trace (wavefront closest hit)",
                 "tail": "k_tail (path-per-lane tail: closest and any-hit traversal, shading)"}
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "wavefront_instanced1080p_traffic.json")
# VALU-bound stages: FP32 vector peak of the MI355X
VALU_PEAK_TFLOPS = 157.3
VALU_STAGES = {"shade": "k_shade (primary vertex: surface, BSDF, next-event estimate, primary misses' sky)",
               "sky": "k_sky (sky of the tail's escaped paths)"}
# radiance-cache MLP on the matrix cores: BF16 dense peak
MFMA_PEAK_TFLOPS = 2500.0
# 32 -> 64 -> 64 -> 64 -> 64 -> 64 -> 3, no biases
MLP_MACS = 32 * 64 + 4 * 64 * 64 + 64 * 3                 # 18,624 per query
MLP_FWD_FLOP = 2 * MLP_MACS                               # 37,248
# training sample: forward, input gradients of layers 1..5 (64x64 x 4, 3x64), weight gradients
MLP_TRAIN_FLOP = MLP_FWD_FLOP + 2 * (4 * 64 * 64 + 3 * 64) + 2 * MLP_MACS   # 107,648


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


_ORACLE = {}


def oracle_scene(scene):
    """The host oracle's copy of the scene (built once, shared by the CPU
    baseline and the FP64 reference)."""
    import oracle
    if id(scene) not in _ORACLE:
        t0 = time.time()
        _ORACLE[id(scene)] = oracle.OracleScene(scene)
        log(f"[cpu] oracle scene built in {time.time() - t0:.1f}s")
    return _ORACLE[id(scene)]


def cpu_baseline(scene, st, cam, W, H, budget_s):
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    orc = oracle_scene(scene)
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
        if spent >= budget_s * 0.5:
            break
        n = int(min(1 << 20, n * max(2.0, min(8.0, (budget_s * 0.5 - spent) / max(dt, 1e-3)))))
    # the same on one core, over a smaller sample
    n1 = max(2048, int(total_samples / spent / cores * budget_s * 0.1))
    pix = rng.choice(W * H, size=min(n1, W * H), replace=False).astype(np.int32)
    t0 = time.time()
    orc.render_pixels(st, cam, W, H, pix, 0, 1, threads=1)
    one = len(pix) / (time.time() - t0) / 1e6
    log(f"[cpu] 1 core: {len(pix)} pixel-samples, {one:.3f} Msamples/s")
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return {"value": total_samples / spent / 1e6, "unit": "Msamples/s", "cores": cores, "kind": "port",
            "sample": f"{total_samples} random pixels of the 1920x1080 frame, 1 sample each, depth "
                      f"{st.max_depth}, same scene/camera/seeds, oracle C++ (IEEE float) on {cores} threads",
            "single_core_value": round(one, 4), "single_core_sample": f"{len(pix)} random pixels, 1 sample each",
            "cpu_model": model}


def fp64_reference(scene, st, cam, W, H, pixels, spp):
    """Converged FP64 render of a pixel subset on the host cores (the oracle's
    double instantiation): the yardstick of the RMSE, never the thing measured."""
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    orc = oracle_scene(scene)
    t0 = time.time()
    # frames from 2^20 on: seeds the GPU's first frames never use
    ref = orc.render_pixels(st, cam, W, H, pixels, 1 << 20, spp, double=True, threads=cores)
    log(f"[ref] FP64 reference: {len(pixels)} pixels x {spp} spp in {time.time() - t0:.1f}s on {cores} threads")
    return ref.astype(np.float64)


def op_counts(scene, st, cam, W, H, n):
    """Algorithmic FP32 ops per sample of the frame, by phase, over a random
    pixel sample (the oracle instantiated with a counting float; the
    yardstick of the VALU rooflines, never the thing timed)."""
    import oracle
    orc = oracle_scene(scene)
    pix = np.random.default_rng(4321).choice(W * H, size=n, replace=False).astype(np.int32)
    t0 = time.time()
    ops = oracle.count_ops(orc, st, cam, W, H, pix, frame=0)
    log(f"[ops] counted {n} samples in {time.time() - t0:.1f}s")
    return ops


def bytes_per_sample(scene, st, cam, W, H, n):
    """Algorithmic traversal bytes of one sample (all its closest-hit and
    shadow rays, reference layout: BYTES_PER_EVENT per event) over a random
    pixel sample, counted by the oracle: mean, p50, p99 and rays per sample
    (the yardstick of the traversal roofline's per-unit figure, never the
    thing timed)."""
    import oracle
    orc = oracle_scene(scene)
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    pix = np.random.default_rng(2468).choice(W * H, size=n, replace=False).astype(np.int32)
    t0 = time.time()
    ev = oracle.sample_events(orc, st, cam, W, H, pix, frame=0, threads=cores)
    log(f"[bytes] counted {n} samples in {time.time() - t0:.1f}s")
    w = np.array([BYTES_PER_EVENT[k] for k in ("blas_interior_pops", "blas_leaf_pops", "tri_tests", "blas_entries",
                                               "tlas_pops")], np.float64)
    b = ev[:, :5].astype(np.float64) @ w
    return {"mean": round(float(b.mean()), 1), "p50": round(float(np.percentile(b, 50)), 1),
            "p99": round(float(np.percentile(b, 99)), 1), "rays_per_sample": round(float(ev[:, 5].mean()), 4),
            "sample": f"{n} random pixels of the frame, 1 sample each, frame 0, the reference's traversal",
            "per_event_bytes": dict(BYTES_PER_EVENT)}


def measured_traffic(stage, W, H, instances, depth, spp, tail):
    """HBM bytes per launch of a stage's kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes of this workload (scaled by the
    copy-kernel calibration), or (None, None) for another workload or another
    split of the bounces (the `tail` knob decides which bounces each kernel
    covers)."""
    try:
        with open(TRAFFIC_JSON) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None, None
    for e in t.get("entries", []):
        w = e.get("workload", {})
        if e.get("stage") == stage and (
                (w.get("width"), w.get("height"), w.get("instances"), w.get("depth"), w.get("spp"), w.get("tail")) ==
                (W, H, instances, depth, spp, tail)):
            b = 1024 * (e["fetch_kib_per_dispatch"] * t["fetch_scale"] +
                        e["write_kib_per_dispatch"] * t["write_scale"])
            return int(b), os.path.relpath(TRAFFIC_JSON, ROOT)
    return None, None


def rel_rmse(est, ref):
    """RMSE relative to the reference, per channel, with a floor of 10 % of the
    mean so dark pixels do not dominate."""
    floor = 0.1 * ref.mean()
    return float(np.sqrt(np.mean((est - ref) ** 2 / (ref ** 2 + floor ** 2))))


def run_curve(ctx, pixels, ref, frames, checkpoints, warmup_frames=0, label="", cache_acc=None):
    """Renders `frames` frames of 1 spp (after `warmup_frames` frames whose
    time counts but whose samples are discarded); returns [(seconds, spp,
    rmse)] at the checkpoints.  Seconds are HIP-event times of the render
    calls; the frames between two checkpoints are one call, so the tracer
    runs with its frames in flight as in the headline.
    cache_acc (a dict, or None): per accumulated frame, the cache stages'
    HIP-event ms and launches and the frame's queries and training records
    are added to it (a call's last frame's counts stand for each of its
    frames)."""
    seconds = 0.0
    if warmup_frames:
        ctx.render(warmup_frames)
        seconds += ctx.timing()[0] / 1e3
        ctx.reset()
    out = []
    done = 0
    for f in sorted(c for c in set(checkpoints) | {frames} if 0 < c <= frames):
        n = f - done
        ctx.render(n)
        done = f
        seconds += ctx.timing()[0] / 1e3
        if cache_acc is not None:
            for stage in ("cache_infer", "cache_train"):
                sst = ctx.stage_stats(stage)
                cache_acc[stage + "_ms"] = cache_acc.get(stage + "_ms", 0.0) + sst["ms"]
                cache_acc[stage + "_launches"] = cache_acc.get(stage + "_launches", 0) + sst["launches"]
            cs = ctx.cache_stats()
            cache_acc["queries"] = cache_acc.get("queries", 0) + n * int(cs["queries"])
            cache_acc["train_records"] = cache_acc.get("train_records", 0) + n * int(cs["train_records"])
            # the frame's Adam steps consume at most train_samples of its records
            cache_acc["trained"] = cache_acc.get("trained", 0) + n * min(int(cs["train_records"]),
                                                                       cache_acc.get("train_samples", 1 << 62))
            cache_acc["frames"] = cache_acc.get("frames", 0) + n
        if f in checkpoints:
            est = ctx.estimate().reshape(-1, 3)[pixels].astype(np.float64)
            out.append((seconds, f, rel_rmse(est, ref)))
            extra = ""
            if ctx.cache_settings().enabled:
                cs = ctx.cache_stats()
                extra = f"  (cache loss {cs['loss']:.4f}, {int(cs['train_records'])} records)"
            log(f"[ttt] {label} {f:4d} spp  {seconds:7.3f} s  rel RMSE {out[-1][2]:.4f}{extra}")
    return out


def batched_views(ctx, scene, n_views, spp, seed=64):
    """The multi-view workload: n_views cameras around the instance box with
    square resolutions log-uniform in [128, 1024], all traced as one set of
    paths per sample (kpt_render_views), with the configuration's material
    sort and a hipGraph per sample (the eager, unsorted figure beside it)."""
    views = scenes.multi_view_cameras(scene, n_views, seed)
    pixels = sum(w * h for _, _, w, h in views)
    keep = ctx.tuning()
    res = {}
    for name, kn in (("graph_sorted", {"graph": 1, "sort": 1}), ("eager", {"graph": 0})):
        ctx.set_tuning(**kn)
        ctx.render_views(views, spp=1)                  # allocation, capture and warm-up
        ctx.render_views(views, spp=spp)
        res[name] = ctx.timing()[0]
        ctx.set_tuning(**keep)
    ms = res["graph_sorted"]
    out = {"views": n_views, "pixels": int(pixels), "spp": spp, "ms": round(ms, 3),
           "Msamples_per_s": round(pixels * spp / (ms / 1e3) / 1e6, 3),
           "tuning": "hipGraph per sample, material sort at the bounces the stage kernels run",
           "eager_unsorted_ms": round(res["eager"], 3),
           "resolutions": "square, side log-uniform in [128, 1024]"}
    log(f"[views] {out}")
    return out


def axis_angle(axis, ang):
    x, y, z = axis / np.linalg.norm(axis)
    c, s = math.cos(ang), math.sin(ang)
    return np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s, 0],
                     [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s, 0],
                     [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c), 0],
                     [0, 0, 0, 1]], np.float64)


def full_frame_4k(ctx, scene, frames, seed=100, tlas_device=2):
    """The animated full frame: 3840x2160, the first 100 instances moving with
    seeded linear + angular velocities, the TLAS refitted on the device every
    frame (tlas_device 2; 0 rebuilds it on the host as the reference does),
    motion vectors, temporal accumulation, 256-bin auto exposure (ISO 100,
    f/2.8, 1/60 s), 6-level bloom, AgX, radiance cache on (training online)."""
    keep = ctx.tuning()["tlas_device"]
    ctx.set_tuning(tlas_device=tlas_device)
    W, H = 3840, 2160
    rng = np.random.default_rng(seed)
    base = np.array(scene.instance_model[:100], np.float64).reshape(100, 4, 4).transpose(0, 2, 1)  # row-major
    vel = rng.uniform(-3.0, 3.0, (100, 3))
    axes = rng.normal(size=(100, 3))
    omega = rng.uniform(0.2, 1.5, 100)
    st = make_settings(8, abi.INTEGRATOR_WAVEFRONT)
    st.use_temporal_accumulation, st.enable_auto_exposure, st.bloom_generate_bloom = 1, 1, 1
    ctx.set_settings(st)
    ctx.resize(W, H)
    proj = scenes.perspective_fov_lh(scene.camera_fov, W, H, 1.0, 100.0)
    ctx.set_view(proj, scenes.look_at_view(scene.camera_eye, scene.camera_target))
    ctx.set_exposure(abi.KptExposure.camera(aperture=2.8, iso=100, shutter=1 / 60))
    ctx.set_frame_delta(1000.0 / 60.0)
    ctx.cache_configure(enabled=True, train=True, seed=13)
    times, host = [], []
    for f in range(frames):
        t = f / 60.0
        models = np.empty((100, 16), np.float32)
        for i in range(100):
            m = base[i].copy()
            m[:3, 3] += vel[i] * t
            m = m @ axis_angle(axes[i], omega[i] * t)
            models[i] = m.T.reshape(-1)                 # column-major
        h0 = time.perf_counter()
        ctx.set_instance_transforms(0, models)
        ctx.render(1)
        host.append(time.perf_counter() - h0)
        times.append(ctx.timing()[0])
    out = ctx.readback(4)
    ctx.cache_configure(enabled=False, train=False, seed=13)
    ctx.set_instance_transforms(0, scene.instance_model[:100])   # the later legs render the scene at rest
    ctx.set_tuning(tlas_device=keep)
    steady = times[2:] if len(times) > 4 else times
    rec = {"width": W, "height": H, "frames": frames, "animated_instances": 100,
           "gpu_ms_per_frame": round(float(np.mean(steady)), 3),
           "wall_ms_per_frame": round(1e3 * float(np.mean(host[2:] if len(host) > 4 else host)), 3),
           "Msamples_per_s": round(W * H / (float(np.mean(steady)) / 1e3) / 1e6, 3),
           "output_finite": bool(np.isfinite(out).all()),
           "tlas": "device refit per frame" if tlas_device == 2 else "host rebuild per frame",
           "stages": "TLAS per frame, motion vectors, temporal, AE, bloom, AgX, cache"}
    log(f"[4k] {rec}")
    return rec


def time_at(curve, target):
    """First time the curve reaches the target RMSE (log-log interpolation
    between checkpoints), or None."""
    prev = None
    for s, spp, e in curve:
        if e <= target:
            if prev is None or prev[2] <= e:
                return s
            ps, _, pe = prev
            t = (math.log(pe) - math.log(target)) / (math.log(pe) - math.log(e))
            return math.exp(math.log(ps) + t * (math.log(s) - math.log(ps)))
        prev = (s, spp, e)
    return None


def blas_build_leg(scene, host_ctx=None, host_ms=None):
    """Wall time of the scene's BLAS builds on host threads and on the device
    (tuning knob blas_device; the second device build, the first also loads
    the kernels), and whether the trees are equal node for node."""
    own = host_ctx is None
    if own:
        host_ctx = Context(0)
        host_ctx.set_tuning(blas_device=0)
        host_ctx.upload(scene)
        host_ms = host_ctx.build_timing()[0]
    bctx = Context(0)
    bctx.set_tuning(blas_device=1)
    dev_ms = []
    for _ in range(2):
        bctx.upload(scene)
        dev_ms.append(bctx.build_timing()[0])
    same = all(bctx.blas(m)[0].tobytes() == host_ctx.blas(m)[0].tobytes() and
               np.array_equal(bctx.blas(m)[1], host_ctx.blas(m)[1]) for m in range(len(scene.mesh_count)))
    bctx.close()
    if own:
        host_ctx.close()
    return {"triangles": int(len(scene.positions)), "meshes": int(len(scene.mesh_count)),
            "host_threads_ms": round(host_ms, 2), "device_ms": round(dev_ms[-1], 2),
            "device_first_ms": round(dev_ms[0], 2), "trees_equal": bool(same)}


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
    ap.add_argument("--op-count-pixels", type=int, default=2048,
                    help="pixel sample of the FP32 op count behind the shading stages' VALU rooflines (0: skip)")
    ap.add_argument("--bytes-pixels", type=int, default=4096,
                    help="pixel sample of the per-sample traversal bytes distribution (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tuning", default="", help="fixed knobs, e.g. refill=0,blocks_per_cu=8,sort=1")
    ap.add_argument("--views", type=int, default=64, help="cameras of the batched multi-view leg (0: skip)")
    ap.add_argument("--full4k", type=int, default=12, help="frames of the animated 4K full-frame leg (0: skip)")
    ap.add_argument("--no-graph-leg", action="store_true", help="skip the hipGraph replay leg")
    ap.add_argument("--no-megakernel-leg", action="store_true", help="skip the megakernel comparison leg")
    ap.add_argument("--ttt-spp", type=int, default=256,
                    help="time-to-target run: the plain tracer's RMSE at this many spp is the target (0: skip)")
    ap.add_argument("--ttt-pixels", type=int, default=2048)
    ap.add_argument("--ref-spp", type=int, default=8192, help="spp of the FP64 reference pixels")
    ap.add_argument("--cache-warmup", type=int, default=8, help="cache training frames before accumulating")
    ap.add_argument("--cache-nee", type=int, default=1, choices=[0, 1],
                    help="1: trace the cache vertex's next-event estimate, the network learns the rest")
    ap.add_argument("--cache-wide", type=float, default=0.0,
                    help="paths end in the cache at the first vertex reached through a diffuse sample or a lobe "
                         "at least this rough (0: always at the bounce-1 vertex)")
    ap.add_argument("--profile-cached", type=int, default=0,
                    help="only render this many cached 1-spp frames of the workload (for rocprofv3)")
    ap.add_argument("--blas-build-only", action="store_true",
                    help="only build the instanced scene's BLASes on host threads and on the device, print one JSON line")
    ap.add_argument("--calibrate-gather", type=int, default=0,
                    help="only run this many scattered 16-byte reads of a 1 GiB table (HBM counter calibration)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="only run one device-to-device copy of this many bytes (HBM counter calibration)")
    args = ap.parse_args()

    if args.calibrate_gather:
        # scattered 16-byte reads of a 1 GiB table: FETCH_SIZE per read, against
        # the copy's streaming calibration
        from kittlespt_amd.context import DeviceBuffer
        records = 1 << 26
        src, dst = DeviceBuffer(records * 16), DeviceBuffer(args.calibrate_gather * 16)
        idx = DeviceBuffer.from_array(np.random.default_rng(7).integers(0, records, args.calibrate_gather,
                                                                         dtype=np.int32))
        rc = abi.lib().kpt_measure_gather16(src.ptr, records, idx.ptr, args.calibrate_gather, dst.ptr)
        print(json.dumps({"calibrate_gather_reads": args.calibrate_gather, "table_bytes": records * 16, "rc": rc}),
              flush=True)
        return
    if args.calibrate_copy:
        from kittlespt_amd.context import DeviceBuffer, hip
        a, b = DeviceBuffer(args.calibrate_copy), DeviceBuffer(args.calibrate_copy)
        rc = hip().hipMemcpy(b.ptr, a.ptr, args.calibrate_copy, 3)   # device to device
        hip().hipDeviceSynchronize()
        print(json.dumps({"calibrate_copy_bytes": args.calibrate_copy, "rc": rc}), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this benchmark measures one GPU "
              f"(multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    if args.blas_build_only:
        scene = scenes.scene_instanced(num_instances=args.instances, tex_size=8)
        print(json.dumps(blas_build_leg(scene)), flush=True)
        return
    t0 = time.time()
    scene = scenes.scene_instanced(num_instances=args.instances)
    log(f"[bench] scene generated in {time.time() - t0:.1f}s: {len(scene.positions)} triangles, "
        f"{scene.num_instances} instances")
    integ = abi.INTEGRATOR_WAVEFRONT if args.integrator == "wavefront" else abi.INTEGRATOR_MEGAKERNEL
    W, H = args.width, args.height
    ctx = Context(0)
    ctx.set_tuning(blas_device=0)          # host threads here; the device build is timed beside it
    t0 = time.time()
    ctx.upload(scene)
    host_build_ms = ctx.build_timing()[0]
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
        # per-launch trace events need the eager launch path: graph replay is
        # measured separately below
        # four frames per trial, so that two frames in flight take part in the choice
        best, ms, _ = autotune.tune(ctx, spp=4, log=log, skip=("graph",))
        log(f"[bench] autotuned in {time.time() - t0:.1f}s: {best} ({ms / 4:.2f} ms per 1-spp frame)")
    tuning = ctx.tuning() if integ == abi.INTEGRATOR_WAVEFRONT else None

    if args.profile_cached:
        # only the cached tracer's frames (training online), for a kernel profile
        ctx.cache_configure(enabled=True, train=True, train_samples=65536, train_steps=4, learning_rate=1e-3,
                            radiance_scale=1e4, seed=11, nee_at_vertex=args.cache_nee,
                            wide_roughness=args.cache_wide)
        times = []
        for f in range(args.profile_cached):
            ctx.render(1)
            times.append(ctx.timing()[0])
            log(f"[profile] cached frame {f}: {times[-1]:.2f} ms")
        print(json.dumps({"profile_cached_frames": args.profile_cached,
                          "ms_per_frame_after_2": round(float(np.mean(times[2:] or times)), 3)}), flush=True)
        return

    for i in range(args.warmup):
        ctx.render(args.spp)
        log(f"[bench] warmup step {i} {ctx.timing()[0]:.1f} ms")

    # traversal event counts of one step (measurement pass, outside the timed region)
    counted = None
    if integ == abi.INTEGRATOR_WAVEFRONT:
        ctx.set_trace_counting(True)
        ctx.render(args.spp)
        counted = {sname: ctx.stage_stats(sname) for sname in STAGE_KERNELS}
        ctx.set_trace_counting(False)
        for sname, c in counted.items():
            log(f"[bench] counted step, {sname}: {c}")

    ctx.reset()
    render_ms = 0.0
    timed_stages = list(STAGE_KERNELS) + list(VALU_STAGES)
    stage_ms = {sname: 0.0 for sname in timed_stages}
    stage_launches = {sname: 0 for sname in timed_stages}
    t0 = time.perf_counter()
    for i in range(args.steps):
        ctx.render(args.spp)
        render_ms += ctx.timing()[0]
        if integ == abi.INTEGRATOR_WAVEFRONT:
            for sname in timed_stages:
                sst = ctx.stage_stats(sname)
                stage_ms[sname] += sst["ms"]
                stage_launches[sname] += sst["launches"]
    wall = time.perf_counter() - t0
    samples = W * H * args.spp * args.steps
    value = samples / wall / 1e6
    # With two frames in flight each stage's launches share the GPU with the
    # other frame's stages, so their HIP-event durations include that sharing;
    # one more step (untimed) with one frame in flight gives each launch's
    # duration alone on the GPU.
    iso_ms, iso_launches = None, None
    if integ == abi.INTEGRATOR_WAVEFRONT and ctx.tuning().get("overlap", 0):
        ctx.set_tuning(overlap=0)
        ctx.render(args.spp)
        iso_ms = {sname: ctx.stage_stats(sname)["ms"] for sname in timed_stages}
        iso_launches = {sname: ctx.stage_stats(sname)["launches"] for sname in timed_stages}
        ctx.set_tuning(overlap=1)
    log(f"[bench] {args.steps} steps in {wall:.3f}s (event time {render_ms:.1f} ms), " +
        ", ".join(f"{sname} {stage_ms[sname]:.1f} ms over {stage_launches[sname]} launches"
                  for sname in STAGE_KERNELS))

    # roofline of each instrumented stage kernel; the bench line's `roofline`
    # is the one with the most time in the timed steps
    roof, stage_roofs = None, {}
    if counted is not None:
        tail_knob = (tuning or {}).get("tail", 0)
        for sname, c in counted.items():
            if stage_launches[sname] <= 0 or c["launches"] <= 0:
                continue
            bytes_step = sum(BYTES_PER_EVENT[k] * c["events"][k] for k in BYTES_PER_EVENT)
            if sname == "tail":
                bytes_step += TAIL_BYTES_PER_PATH * c["paths"]
            bytes_per_launch = bytes_step / c["launches"]
            avg_launch_s = (stage_ms[sname] / stage_launches[sname]) / 1e3
            achieved = bytes_per_launch / avg_launch_s / 1e9
            traffic, traffic_src = measured_traffic(sname, W, H, args.instances, args.depth, args.spp, tail_knob)
            r = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                 "kernel": STAGE_KERNELS[sname],
                 "bytes_per_launch": int(bytes_per_launch), "avg_launch_ms": round(avg_launch_s * 1e3, 4),
                 "ms_per_step": round(stage_ms[sname] / args.steps, 3)}
            if traffic_src:
                r["traffic_source"] = traffic_src
            if iso_ms and iso_launches.get(sname, 0) > 0:
                iso_s = iso_ms[sname] / iso_launches[sname] / 1e3
                r["alone"] = {"avg_launch_ms": round(iso_s * 1e3, 4),
                              "achieved": round(bytes_per_launch / iso_s / 1e9, 2),
                              "frac": round(bytes_per_launch / iso_s / 1e9 / HBM_PEAK_GBS, 5),
                              "note": "one frame in flight (untimed step after the timed ones): the launch "
                                      "alone on the GPU; the figures above are the timed steps', two frames "
                                      "in flight"}
            stage_roofs[sname] = r
        if stage_roofs:
            dom = max(stage_roofs, key=lambda k: stage_ms[k])
            roof = dict(stage_roofs[dom])
            roof["stages"] = {k: {x: v[x] for x in ("kernel", "achieved", "frac", "avg_launch_ms", "ms_per_step",
                                                    "traffic", "alone") if x in v}
                              for k, v in stage_roofs.items()}

    # per-sample algorithmic traversal bytes (distribution) and rays per sample
    bps = None
    if integ == abi.INTEGRATOR_WAVEFRONT and args.bytes_pixels > 0:
        try:
            bps = bytes_per_sample(scene, st, cam, W, H, args.bytes_pixels)
            log(f"[bytes] per sample: {bps}")
            if roof is not None:
                roof["bytes_per_sample"] = bps
        except Exception as e:  # a reported figure, never fatal
            log(f"[bytes] count failed: {e}")

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(scene, st, cam, W, H, args.cpu_budget)
            if bps is not None:
                # rays per sample from the counted pixel sample above
                cpu["mrays_per_s"] = round(cpu["value"] * bps["rays_per_sample"], 4)
                cpu["mrays_note"] = "Msamples/s x rays per sample (closest-hit + shadow) of the counted pixel sample"
        except Exception as e:  # the baseline is reported, never fatal
            log(f"[cpu] baseline failed: {e}")

    # FP32 vector rooflines of the shading stages: the algorithmic op count
    # per sample (the oracle instantiated with a counting float over a pixel
    # sample: +, -, *, /, sqrt and transcendentals, one op each) times the
    # frame's samples, over the launch's HIP-event time
    if roof is not None and args.op_count_pixels > 0:
        try:
            ops = op_counts(scene, st, cam, W, H, args.op_count_pixels)
            flop = {k: v["arith"] + v["sqrt"] + v["transcendental"] for k, v in ops.items()}
            per_launch = {"shade": (flop["shade_primary"] + flop["sky_primary"]) * W * H,
                          "sky": flop["sky_later"] * W * H}
            for sname, kname in VALU_STAGES.items():
                if stage_launches[sname] <= 0:
                    continue
                # the launch alone on the GPU when frames overlapped in the timed steps
                # (a short stage's event bracket then also holds the other frame's work)
                alone = bool(iso_ms and iso_launches.get(sname, 0) > 0)
                avg_s = (iso_ms[sname] / iso_launches[sname] if alone
                         else stage_ms[sname] / stage_launches[sname]) / 1e3
                ach = per_launch[sname] / avg_s / 1e12
                roof["stages"][sname] = {"kernel": kname, "bound": "valu", "unit": "TFLOP/s",
                                         "achieved": round(ach, 3), "peak": VALU_PEAK_TFLOPS,
                                         "frac": round(ach / VALU_PEAK_TFLOPS, 5),
                                         "flop_per_launch": int(per_launch[sname]),
                                         "avg_launch_ms": round(avg_s * 1e3, 4),
                                         "ms_per_step": round(stage_ms[sname] / args.steps, 3)}
                if alone:
                    roof["stages"][sname]["duration"] = "the launch alone on the GPU (one frame in flight)"
            roof["op_counts_per_sample"] = {k: {x: round(y, 2) for x, y in v.items()} for k, v in ops.items()}
            roof["op_counts_sample"] = (f"{args.op_count_pixels} random pixels, 1 sample each, oracle with a "
                                        f"counting float; the sky toward the sun is a per-frame constant the "
                                        f"kernels evaluate once (not counted per sample)")
            log(f"[ops] per sample: {roof['op_counts_per_sample']}")
        except Exception as e:  # measurement aid, never fatal
            log(f"[ops] op count failed: {e}")

    # the same steps with every frame replayed as a captured hipGraph
    graph_rec = None
    if integ == abi.INTEGRATOR_WAVEFRONT and not args.no_graph_leg:
        ctx.set_tuning(graph=1)
        ctx.render(args.spp)
        ctx.reset()
        t0 = time.perf_counter()
        for i in range(args.steps):
            ctx.render(args.spp)
        gwall = time.perf_counter() - t0
        ctx.set_tuning(graph=0)
        graph_rec = {"value": round(samples / gwall / 1e6, 3), "ms_per_step": round(gwall * 1e3 / args.steps, 3),
                     "note": "same workload, each 1-spp frame captured as one hipGraph and replayed"}
        log(f"[bench] graph replay: {graph_rec}")

    # the plain megakernel (one thread runs a whole path, the reference's
    # structure) on the same workload, for comparison
    mega = None
    if integ == abi.INTEGRATOR_WAVEFRONT and not args.no_megakernel_leg:
        ctx.set_settings(make_settings(args.depth, abi.INTEGRATOR_MEGAKERNEL))
        ctx.render(args.spp)
        ctx.reset()
        msteps = max(1, min(args.steps, 2))
        t0 = time.perf_counter()
        for i in range(msteps):
            ctx.render(args.spp)
        mwall = time.perf_counter() - t0
        ctx.set_settings(st)
        mega = {"value": round(W * H * args.spp * msteps / mwall / 1e6, 3),
                "ms_per_step": round(mwall * 1e3 / msteps, 3), "steps": msteps,
                "note": "same workload, megakernel integrator (one lane per path)"}
        log(f"[bench] megakernel: {mega}")

    views = None
    if integ == abi.INTEGRATOR_WAVEFRONT and args.views > 0:
        views = batched_views(ctx, scene, args.views, 4)

    full4k = None
    if integ == abi.INTEGRATOR_WAVEFRONT and args.full4k > 0:
        full4k = full_frame_4k(ctx, scene, args.full4k, tlas_device=2)
        rebuild = full_frame_4k(ctx, scene, args.full4k, tlas_device=0)
        full4k["host_rebuild"] = {k: rebuild[k] for k in ("gpu_ms_per_frame", "wall_ms_per_frame", "Msamples_per_s")}
        # back to the headline frame for the time-to-target leg
        ctx.set_settings(st)
        ctx.resize(W, H)
        ctx.set_view(proj, view)

    # time to a target RMSE against an FP64 converged render: plain wavefront
    # vs wavefront + radiance cache (trained online from scratch, its warm-up
    # frames counted)
    ttt = None
    if integ == abi.INTEGRATOR_WAVEFRONT and args.ttt_spp > 0:
        rng = np.random.default_rng(77)
        pix = np.sort(rng.choice(W * H, size=args.ttt_pixels, replace=False)).astype(np.int32)
        ref = fp64_reference(scene, st, cam, W, H, pix, args.ref_spp)
        checkpoints = sorted({1, 2, 4, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256} | {args.ttt_spp})
        plain_cp = [c for c in checkpoints if c <= args.ttt_spp]
        ctx.cache_configure(enabled=False, train=False)
        ctx.reset()
        plain = run_curve(ctx, pix, ref, args.ttt_spp, set(plain_cp), label="plain ")
        target = plain[-1][2]
        t_plain = plain[-1][0]
        ctx.cache_configure(enabled=True, train=True, train_samples=65536, train_steps=4, learning_rate=1e-3,
                            radiance_scale=1e4, seed=11, nee_at_vertex=args.cache_nee,
                            wide_roughness=args.cache_wide)
        cache_acc = {"train_samples": 65536}
        cached = run_curve(ctx, pix, ref, args.ttt_spp, set(plain_cp), warmup_frames=args.cache_warmup,
                           label="cached", cache_acc=cache_acc)
        t_cached = time_at(cached, target)
        ctx.cache_configure(enabled=False, train=False, seed=11)
        ttt = {"target_rel_rmse": round(target, 5), "target": f"plain wavefront at {args.ttt_spp} spp",
               "plain_s": round(t_plain, 4), "cached_s": None if t_cached is None else round(t_cached, 4),
               "speedup": None if t_cached is None else round(t_plain / t_cached, 3),
               "cached_rmse_at_equal_spp": round(cached[-1][2], 5),
               # throughput of each tracer over the accumulated frames (HIP-event frame times)
               "plain_msamples_per_s": round(W * H * (plain[-1][1] - plain[0][1]) / max(plain[-1][0] - plain[0][0], 1e-9)
                                             / 1e6, 3),
               "cached_msamples_per_s": round(W * H * (cached[-1][1] - cached[0][1])
                                              / max(cached[-1][0] - cached[0][0], 1e-9) / 1e6, 3),
               "cache_warmup_frames": args.cache_warmup, "cache_nee_at_vertex": args.cache_nee,
               "cache_wide_roughness": args.cache_wide,
               "frames_in_flight": ("two, for both tracers (the frames between checkpoints are one render "
                                    "call); a cached frame waits for the previous one before its bounce-1 "
                                    "work, whose cache records and inference follow the previous training"),
               "reference": f"FP64 oracle render of {args.ttt_pixels} random pixels x {args.ref_spp} spp "
                            f"(host cores)",
               "curve_plain": [[round(s, 4), n, round(e, 5)] for s, n, e in plain],
               "curve_cached": [[round(s, 4), n, round(e, 5)] for s, n, e in cached]}
        # matrix-core rooflines of the cache's MLP kernels over the accumulated frames
        mlp = {}
        nf = max(1, cache_acc.get("frames", 0))
        if cache_acc.get("cache_infer_launches", 0) > 0:
            avg_s = cache_acc["cache_infer_ms"] / cache_acc["cache_infer_launches"] / 1e3
            fl = cache_acc["queries"] / nf * MLP_FWD_FLOP
            mlp["infer"] = {"kernel": "k_rc_infer (BF16 MFMA forward of every cache query)", "bound": "mfma",
                            "unit": "TFLOP/s", "achieved": round(fl / avg_s / 1e12, 3), "peak": MFMA_PEAK_TFLOPS,
                            "frac": round(fl / avg_s / 1e12 / MFMA_PEAK_TFLOPS, 5), "queries_per_frame":
                            int(cache_acc["queries"] / nf), "flop_per_launch": int(fl),
                            "avg_launch_ms": round(avg_s * 1e3, 4)}
        if cache_acc.get("cache_train_launches", 0) > 0:
            avg_s = cache_acc["cache_train_ms"] / cache_acc["cache_train_launches"] / 1e3
            used = cache_acc["trained"] / nf
            fl = used * MLP_TRAIN_FLOP
            mlp["train"] = {"kernel": "k_rc_train + k_rc_reduce + k_rc_adam (4 Adam steps per frame)",
                            "bound": "mfma", "unit": "TFLOP/s", "achieved": round(fl / avg_s / 1e12, 3),
                            "peak": MFMA_PEAK_TFLOPS, "frac": round(fl / avg_s / 1e12 / MFMA_PEAK_TFLOPS, 5),
                            "samples_per_frame": int(used), "flop_per_frame": int(fl),
                            "avg_ms_per_frame": round(avg_s * 1e3, 4)}
        for v in mlp.values():
            v["duration"] = ("HIP events over the cached curve's render calls, two frames in flight: the next "
                             "frame's primary stages share the GPU with these launches")
        ttt["mlp_roofline"] = mlp
        if roof is not None:
            roof.setdefault("stages", {}).update({"cache_" + k: v for k, v in mlp.items()})
        log(f"[ttt] {ttt}")

    blas_build = blas_build_leg(scene, ctx, host_build_ms)
    log(f"[bench] BLAS build {blas_build}")

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
        "roofline": roof, "cpu_baseline": cpu, "time_to_target": ttt, "batched_views": views,
        "graph_replay": graph_rec, "megakernel": mega, "full_frame_4k": full4k, "blas_build": blas_build,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
