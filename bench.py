"""Frames/s of the MI355X KinectFusion pipeline on the 512^3 precision
workload: BF16 voxels with the learned ICP warm start against the plain FP32
pipeline, with the TSDF integration kernel's HBM roofline and a CPU baseline.

Workload ("512^3-5mm-room-640x480"): 640x480 seeded synthetic room trajectory
(uint16 mm depth, sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2 m, 1 % dropouts,
RGBA8), fx = fy = 525, cx = 319.5, cy = 239.5; 512^3 voxels of 5 mm
(2.56 m, camera margin 0.5 m), truncation 0.05 m, angle-weighted TSDF with
colour, bilateral r = 3 / 4.5 px / 30 mm.  A step is one frame: pyramid,
tracking (raycast + ICP per level), integration.  Frames are resident in HBM
before the timed region.

Pipelines timed on the same frames:
  * headline: BF16 voxels + learned warm start (pose network proposal, 2
    level-0 ICP iterations, fallback to the constant-velocity seed with the
    full 10/5/4 schedule when the final RMS exceeds 1 cm or the last
    increment 1 mm / 1 mrad);
  * plain FP32: float voxels, previous-pose seed, 10/5/4 iterations
    (the reference's method);
  * BF16 plain: bf16 voxels with the plain method.
The pose network is trained first (untimed) on pairs the pipeline generates
from 64 rooms other than the evaluated one (room seeds 1-64; the timed
trajectory is room 0; the held-out error is measured on room 1001).

Beside the headline line's `roofline` (the bf16 voxel-update kernel) the JSON
carries `roofline_float_store` (the plain FP32 pipeline's float voxel-update
kernel), `roofline_raycast` (the headline's level-0 raycast, its samples and
hits counted by the CPU oracle over the GPU's own volume; with the CPU leg)
and `hashed` (the 2048^3 hashed-hall workload: frames/s, ATE and the block
kernel's roofline).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

INTR = (525.0, 525.0, 319.5, 239.5)
RES, VOXEL, MARGIN, TRUNC = 512, 0.005, 0.5, 0.05
# Learned warm start: ICP iterations per level from the network's proposal
# (finest first).  On this workload 2 level-0 iterations reach the ATE of 4
# (0.089 vs 0.088 mm, every seed kept); 1 iteration is faster but less
# accurate (0.165 mm).
LEARNED_ITERATIONS = (2, 0, 0)
HBM_PEAK_GBPS = 8000.0
METRIC = "frames/s @512³ (bf16+learned-init vs FP32 plain, ATE≤5mm); TSDF HBM GB/s"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_config(K, voxel: str, init: str, **kw):
    half = np.float32(0.5) * np.float32(RES) * np.float32(VOXEL)
    return K.PipelineConfig(
        voxel=voxel, init=init, tsdf_rule="angle-weighted", color="float",
        volume=K.VolumeGeometry((RES, RES, RES), VOXEL, (-float(half), -float(half), -MARGIN), TRUNC),
        integration=K.TsdfIntegrationOptions(depth_scale=1000.0),
        depth=K.DepthProcessingOptions(levels=3, bilateral_radius=3, bilateral_spatial_sigma=4.5,
                                       bilateral_depth_sigma=0.03),
        intrinsics=K.CameraIntrinsics(*INTR), width=640, height=480, icp_iterations=(10, 5, 4),
        learned_iterations=LEARNED_ITERATIONS, learned_max_rms=0.01, **kw)


def ate(traj: np.ndarray, gt: np.ndarray) -> float:
    from kinectfusion_amd.trajectory import ate as ate_stats

    return ate_stats(traj, gt)["rmse"]


def run_pipeline(K, torch, cfg, depth_dev, color_dev, warmup: int, steps: int, net=None):
    p = K.Pipeline(cfg)
    if net is not None:
        p.attach_posenet(net)
    for i in range(warmup):
        p.process_frame(depth_dev[i], color_dev[i], report=False, sync_inputs=False)
    p.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(warmup, warmup + steps):
        p.process_frame(depth_dev[i], color_dev[i], report=False, sync_inputs=False)
    p.synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    traj = p.trajectory()
    p.close()
    return dt, traj


def stage_pass(K, cfg, depth_dev, color_dev, first: int, count: int, net=None, keep: bool = False,
               increments: list | None = None):
    """Untimed eager replays of the same frames: one with per-stage HIP
    events on the library stream (the voxel-update kernel as it runs in the
    timed loop), one counting the fused voxels of every frame (the kernel's
    counting variant; identical poses and volumes, so identical counts).
    keep: the counting pipeline is returned open (its volume after the last
    frame), with its trajectory.  increments: receives (translation,
    rotation, status) of each kept learned run's finest level."""
    cfg.stage_timing, cfg.count_fused = True, False
    p = K.Pipeline(cfg)
    if net is not None:
        p.attach_posenet(net)
    times, learned = [], 0
    for i in range(first + count):
        res = p.process_frame(depth_dev[i], color_dev[i], report=True, sync_inputs=False)
        st = p.stage_times()
        if i >= first:
            times.append([st.pyramid_ms, st.proposal_ms, st.raycast_ms, st.icp_ms, st.integrate_ms,
                          st.integrate_kernel_ms, st.total_ms])
            learned += res.used_learned
            if res.used_learned and increments is not None:
                r0 = res.level[0]
                increments.append((r0.update_translation, r0.update_rotation, r0.status))
    p.close()
    cfg.stage_timing, cfg.count_fused = False, True
    p = K.Pipeline(cfg)
    if net is not None:
        p.attach_posenet(net)
    fused, before = [], (0, 0)
    for i in range(first + count):
        p.process_frame(depth_dev[i], color_dev[i], report=False, sync_inputs=False)
        now = p.fused_counts()
        if i >= first:
            fused.append((now[0] - before[0], now[1] - before[1]))
        before = now
    cfg.count_fused = False
    if keep:
        return np.array(times), np.array(fused, np.float64), learned, p
    p.close()
    return np.array(times), np.array(fused, np.float64), learned


def integration_roofline(times, fused, sv: int, kernel: str) -> dict:
    """Algorithmic bytes of the voxel-update kernel per launch (read and
    write of every fused voxel and colour voxel: N_fused 2 s_v + N_colour 2
    16) over its mean launch time from HIP events on the library stream."""
    alg = fused[:, 0] * 2 * sv + fused[:, 1] * 2 * 16
    kern_ms = times[:, 5]
    achieved = float(alg.sum() / (kern_ms.sum() / 1e3) / 1e9)
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4), "kernel": kernel, "bytes_per_launch": float(alg.mean()),
            "avg_launch_ms": float(kern_ms.mean()), "fused_voxels_per_launch": float(fused[:, 0].mean()),
            "colour_voxels_per_launch": float(fused[:, 1].mean())}


def train_posenet(K, cfg, pairs_n: int, epochs: int):
    from kinectfusion_amd import posenet as PN

    t0 = time.perf_counter()
    pairs, targets, stats = PN.generate_pairs(cfg, pairs_n, seed=7)
    t1 = time.perf_counter()
    net = PN.PoseNet(PN.PoseNetConfig(seed=7))
    hist = PN.train(net, pairs, targets, epochs=epochs, seed=7, log=log)
    t2 = time.perf_counter()
    held, held_t, _ = PN.generate_pairs(cfg, 512, seed=99, scene_seeds=(1001,))
    err = PN.proposal_errors(net, held, held_t)
    info = {"pairs": pairs_n, "epochs": epochs, "labels": "converged ICP", "icp_label_failures": stats["icp_failed"],
            "generate_s": round(t1 - t0, 1), "train_s": round(t2 - t1, 1), "final_loss": round(hist[-1], 5),
            "heldout_median_err_m": round(err[0], 5), "heldout_median_err_rad": round(err[1], 5)}
    log(f"posenet: {info}")
    del pairs, targets, held, held_t
    return net, info


def cpu_baseline(scene, frames: int):
    """The CPU oracle (single thread) on the first `frames` frames of the same
    workload: frame 0 integrates, frames 1.. track coarse-to-fine and integrate."""
    from oracle import kf_oracle as ko

    half = np.float32(0.5) * np.float32(RES) * np.float32(VOXEL)
    vol = ko.Volume.create(RES, VOXEL, (-float(half), -float(half), -MARGIN), TRUNC, ko.FLOAT, True)
    opts = ko.integration_options(1000.0)
    iters = {0: 10, 1: 5, 2: 4}
    pose = np.eye(4)
    t_total = 0.0
    timed = 0
    poses = [pose]
    for i in range(frames + 1):
        depth, rgba, _ = scene.frame(i)
        t0 = time.perf_counter()
        pyr = ko.pyramid(depth, INTR, 1000.0, levels=3, radius=3, sigma_s=4.5, sigma_r=0.03)
        ok = True
        if i > 0:
            tracked = pose.copy()
            for level in (2, 1, 0):
                k = pyr[level]["intrinsics"]
                s = 1 << level
                pts, nrm, _, _ = ko.raycast(vol, tracked, k, 640 // s, 480 // s)
                new_pose, r = ko.icp_track_level(pyr[level]["vertices"], pyr[level]["normals"], pts, nrm, k,
                                                 tracked, tracked, iters[level])
                if r.status not in (0, 1):
                    ok = False
                    break
                tracked = new_pose
            pose = tracked
            poses.append(pose)
        if ok:
            ko.integrate(vol, 1, depth, INTR, ko.rigid_inverse(pose), opts, color=rgba, normals=pyr[0]["normals"])
        dt = time.perf_counter() - t0
        if i > 0:
            t_total += dt
            timed += 1
        log(f"cpu oracle frame {i}: {dt:.2f} s")
    return timed / t_total, t_total, timed, np.stack(poses)


def raycast_algorithmic(p, pose, voxel: str, raycast_ms: float) -> dict:
    """Algorithmic bytes of one 640x480 raycast of the pipeline's volume at
    `pose`, with the samples and hits counted by the CPU oracle's raycast of
    the same voxels: 28 B written per pixel (point, normal, colour), 8
    corners x s_v read per sample, and per hit the normal's 6 samples (48
    corners x s_v) and the colour's 8 colour voxels (16 B each); over the
    level-0 raycast's time from HIP events."""
    from oracle import kf_oracle as ko

    vox, col = p.host_view()
    half = np.float32(0.5) * np.float32(RES) * np.float32(VOXEL)
    kind = {"float": ko.FLOAT, "quantized": ko.QUANTIZED, "bf16": ko.BF16}[voxel]
    vol = ko.Volume.create(RES, VOXEL, (-float(half), -float(half), -MARGIN), TRUNC, kind, col is not None)
    vol.voxels[:] = vox
    if col is not None:
        vol.colors[:] = col
    del vox, col
    pts, _, _, samples = ko.raycast(vol, pose, INTR, 640, 480)
    hits = int(np.isfinite(pts).all(-1).sum())
    sv = 8 if voxel == "float" else 4
    npx = 640 * 480
    alg = 28 * npx + samples * 8 * sv + hits * (48 * sv + 8 * 16)
    achieved = alg / (raycast_ms / 1e3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4), "kernel": f"raycast_kernel<{voxel}>, 640x480",
            "bytes_per_launch": float(alg), "avg_launch_ms": raycast_ms, "samples": int(samples), "hits": hits,
            "samples_per_ray": round(samples / npx, 2)}


def hashed_bench(K, torch, steps: int, warmup: int) -> dict:
    """The sparse hashed volume workload: 2048^3 voxels of 5 mm in 8^3
    blocks, 2^20 hash buckets, a 500 000-block pool, float voxels without
    colour, the seeded 6 x 3 x 4 m hall with 40 boxes and spheres, plain
    previous-pose tracking (10/5/4).  Frames rendered on the device."""
    scene = K.SyntheticScene("hall", seed=0)
    n = warmup + steps
    gt = np.stack([scene.pose(i) for i in range(n)])
    depth = scene.render_device(gt, np.arange(n))
    frames = [depth[i] for i in range(n)]
    half = np.float32(0.5) * np.float32(2048) * np.float32(VOXEL)
    cfg = K.PipelineConfig(
        voxel="float", init="previous", tsdf_rule="angle-weighted", color="none", storage="hashed",
        hash_buckets=1 << 20, block_capacity=500_000,
        volume=K.VolumeGeometry((2048, 2048, 2048), VOXEL, (-float(half), -float(half), -MARGIN), TRUNC),
        integration=K.TsdfIntegrationOptions(depth_scale=1000.0, mode="band"),
        depth=K.DepthProcessingOptions(levels=3, bilateral_radius=3, bilateral_spatial_sigma=4.5,
                                       bilateral_depth_sigma=0.03),
        intrinsics=K.CameraIntrinsics(*INTR), width=640, height=480, icp_iterations=(10, 5, 4))
    dt, traj = run_pipeline(K, torch, cfg, frames, [None] * n, warmup, steps)
    times, fused, _ = stage_pass(K, cfg, frames, [None] * n, warmup, min(steps, 40))
    p = K.Pipeline(cfg)
    for i in range(n):
        p.process_frame(frames[i], None, report=False, sync_inputs=False)
    stats = p.hashed_stats()
    p.close()
    roof = integration_roofline(times, fused, 8, "integrate_blocks_kernel<float>")
    return {"workload": "2048^3-5mm-hashed-hall-640x480", "frames": n, "fps": round(steps / dt, 2),
            "ate_m": round(ate(traj, gt[: len(traj)]), 6), "blocks_allocated": stats["allocated"],
            "overflow": stats["overflow"],
            "stage_ms": {k: round(float(np.mean(times[:, j])), 4) for j, k in
                         enumerate(["pyramid", "proposal", "raycast", "icp", "integrate", "integrate_kernel",
                                    "total"])},
            "roofline": roof}


def pose_gap(a: np.ndarray, b: np.ndarray):
    """Largest rotation angle (rad) and translation distance (m) between
    matching camera_to_world poses of two trajectories."""
    rot, tr = 0.0, 0.0
    for pa, pb in zip(a, b):
        r = pa[:3, :3].T @ pb[:3, :3]
        rot = max(rot, float(np.arccos(np.clip((np.trace(r) - 1.0) / 2.0, -1.0, 1.0))))
        tr = max(tr, float(np.linalg.norm(pa[:3, 3] - pb[:3, 3])))
    return rot, tr


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def committed_traffic(voxel: str):
    """HBM bytes per launch of the integrate kernel as measured by two rocprofv3 counter passes
    (FETCH_SIZE, then WRITE_SIZE, each scaled by what the counter reads for a plain 256 MiB copy),
    from the committed summary under profiles/; None when there is none for this voxel type."""
    path = Path(__file__).resolve().parent / "profiles" / f"integrate_kernel_512cubed_{voxel}_pmc.md"
    if not path.exists():
        return None, None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1]), f"profiles/{path.name}"
    return None, None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=90)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-frames", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--train-pairs", type=int, default=20000)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--only", choices=["headline", "fp32", "bf16_plain"], default=None,
                    help="time one pipeline only (profiling)")
    ap.add_argument("--no-hashed", action="store_true", help="skip the 2048^3 hashed-volume line")
    ap.add_argument("--hashed-steps", type=int, default=90)
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="first run one plain device copy of this many MiB (the reference dispatch for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X; --gpus {args.gpus} is not "
              "supported (multi-GPU sharding is a design note only).", file=sys.stderr)
        return 2

    import torch

    import kinectfusion_amd as K

    if args.copy_probe_mb > 0:
        n = args.copy_probe_mb << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {n} bytes read, {n} bytes written")
        del src, dst

    total = args.warmup + args.steps
    scene = K.SyntheticScene("room", seed=0)
    log(f"rendering {total} synthetic frames")
    depth_h, color_h, gt = [], [], []
    for i in range(total):
        d, c, pose = scene.frame(i)
        depth_h.append(d)
        color_h.append(c)
        gt.append(pose)
    gt = np.array(gt)
    depth_dev = [torch.from_numpy(d.view(np.int16)).cuda() for d in depth_h]
    color_dev = [torch.from_numpy(c.view(np.int32)).cuda() for c in color_h]
    torch.cuda.synchronize()

    net, net_info = None, None
    if args.only in (None, "headline"):
        log(f"training the pose network on {args.train_pairs} generated pairs, {args.epochs} epochs")
        net, net_info = train_posenet(K, make_config(K, "bf16", "learned"), args.train_pairs, args.epochs)

    runs = {"headline": ("bf16", "learned"), "fp32": ("float", "previous"), "bf16_plain": ("bf16", "previous")}
    results = {}
    for name, (voxel, init) in runs.items():
        if args.only and args.only != name:
            continue
        log(f"timing {name} pipeline ({voxel} voxels, {init} seed): {args.warmup} warmup + {args.steps} frames")
        dt, traj = run_pipeline(K, torch, make_config(K, voxel, init), depth_dev, color_dev, args.warmup, args.steps,
                                net if init == "learned" else None)
        results[name] = {"seconds": dt, "fps": args.steps / dt, "ate_m": ate(traj, gt[: len(traj)]), "traj": traj}
        log(f"{name}: {args.steps / dt:.1f} frames/s, ATE {results[name]['ate_m'] * 1000:.3f} mm")

    head = "headline" if "headline" in results else next(iter(results))
    voxel, init = runs[head]
    log("stage passes: HIP events on the library stream, then exact fused-voxel counts")
    n_stage = min(args.steps, 40)
    increments: list = []
    times, fused, learned_used, kept = stage_pass(K, make_config(K, voxel, init), depth_dev, color_dev,
                                                  args.warmup, n_stage, net if init == "learned" else None,
                                                  keep=True, increments=increments)
    sv = 4 if voxel == "bf16" else 8
    roof = integration_roofline(times, fused, sv, f"integrate_kernel<{voxel}, colour>")
    inc = np.array(increments or [(0.0, 0.0, 0)], np.float64)
    learned_increment = {"kept_runs": len(increments), "converged": int((inc[:, 2] == 1).sum()),
                         "translation_m_p50_p90_max": [float(np.percentile(inc[:, 0], q)) for q in (50, 90, 100)],
                         "rotation_rad_p50_p90_max": [float(np.percentile(inc[:, 1], q)) for q in (50, 90, 100)]}
    stage_ms = {k: round(float(np.mean(times[:, j])), 4) for j, k in
                enumerate(["pyramid", "proposal", "raycast", "icp", "integrate", "integrate_kernel", "total"])}
    roof_float = None
    if "fp32" in results and head != "fp32":
        log("stage passes of the plain FP32 pipeline (float-store roofline)")
        t32, f32, _ = stage_pass(K, make_config(K, "float", "previous"), depth_dev, color_dev, args.warmup, n_stage)
        roof_float = integration_roofline(t32, f32, 8, "integrate_kernel<float, colour>")

    traffic, traffic_src = committed_traffic(voxel)
    roof["traffic"] = traffic
    roof["traffic_source"] = traffic_src

    hashed = None
    if not args.no_hashed:
        log(f"hashed volume: 2048^3 hall, {args.hashed_steps} timed frames")
        hashed = hashed_bench(K, torch, args.hashed_steps, args.warmup)
        log(f"hashed: {hashed['fps']} frames/s, ATE {hashed['ate_m'] * 1000:.3f} mm")

    cpu = None
    raycast_roof = None
    if not args.no_cpu:
        # the level-0 raycast of the headline's volume at its last pose, its
        # samples counted by the oracle (the learned pass raycasts level 0 only,
        # so the stage's raycast time is that launch's)
        if init == "learned":
            log("raycast samples of the headline volume counted by the CPU oracle")
            raycast_roof = raycast_algorithmic(kept, kept.trajectory()[-1], voxel, stage_ms["raycast"])
        log(f"cpu baseline: oracle on {args.cpu_frames} tracked frames (single thread)")
        fps_cpu, secs, n, oracle_poses = cpu_baseline(scene, args.cpu_frames)
        cpu = {"value": fps_cpu, "unit": "frames/s", "cores": 1, "kind": "port",
               "sample": f"frames 1-{n} of the same 512^3 workload after integrating frame 0: pyramid, 3-level "
                         f"tracking and full-volume integration on one core ({secs:.1f} s; host: {cpu_model()}, "
                         f"nproc {os.cpu_count()})"}
        if "fp32" in results:
            # the plain FP32 pipeline runs the oracle's configuration: its poses for the same
            # frames, against the oracle's FP64 tracking
            rot, tr = pose_gap(results["fp32"]["traj"][1: n + 1], oracle_poses[1:])
            cpu["fp32_pose_vs_oracle"] = {"frames": n, "max_rot_rad": rot, "max_trans_m": tr,
                                          "tolerance": {"rad": 1e-4, "m": 1e-4}}

    r = results[head]
    line = {
        "metric": METRIC,
        "value": round(r["fps"], 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * r["seconds"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 voxels, fp32 arithmetic, fp64 solve; bf16 MFMA pose network" if voxel == "bf16" else "fp32",
        "data": "synthetic",
        "config": {"workload": "512^3-5mm-room-640x480", "frames": args.warmup + args.steps, "voxel": voxel,
                   "init": init, "icp_iterations": [10, 5, 4],
                   "learned_iterations": list(LEARNED_ITERATIONS) if init == "learned" else None, "color": "float"},
        "ate_m": round(r["ate_m"], 6),
        "stage_ms": stage_ms,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if roof_float is not None:
        line["roofline_float_store"] = roof_float
    if raycast_roof is not None:
        line["roofline_raycast"] = raycast_roof
    if hashed is not None:
        line["hashed"] = hashed
    if init == "learned":
        line["learned_seed_kept"] = f"{learned_used}/{n_stage}"
        line["learned_last_increment"] = learned_increment
        line["posenet"] = net_info
    if "fp32" in results:
        line["fp32_plain_fps"] = round(results["fp32"]["fps"], 2)
        line["fp32_plain_ate_m"] = round(results["fp32"]["ate_m"], 6)
        line["speedup_vs_fp32"] = round(r["fps"] / results["fp32"]["fps"], 4)
    if "bf16_plain" in results:
        line["bf16_plain_fps"] = round(results["bf16_plain"]["fps"], 2)
        line["bf16_plain_ate_m"] = round(results["bf16_plain"]["ate_m"], 6)
    print(json.dumps(line))
    kept.close()
    if net is not None:
        net.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
