"""Build recipe for the native libraries of kinectfusion_amd.

Produces, in-tree:
  kinectfusion_amd/libkf_hip.so    -- gfx950 HIP kernels + the C ABI (include/kf_hip.h)
  kinectfusion_amd/libkf_scene.so  -- seeded synthetic RGB-D scenes (host C++ renderer and the same
                                      per-pixel renderer as a gfx950 kernel)
  kinectfusion_amd/libkf_io.so     -- host image codec helper of the reconstruction app (PNG row filters)

HIP flags: gfx950 only, -O3, no FMA contraction (bit-exact parity with the
reference's float expression order), IEEE fp32 divide/sqrt (hip default),
packed-fp32 instruction selection disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["kf_api", "kf_depth", "kf_integrate", "kf_raycast", "kf_icp", "kf_mesh", "kf_posenet"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ... (exit {proc.returncode})")
    noise = [
        line for line in (proc.stdout + proc.stderr).splitlines()
        if line.strip() and "packed-fp32-ops" not in line and "is not a recognized feature" not in line
        and "(ignoring feature)" not in line
    ]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


def _hip_object(name: str) -> Path:
    src = CSRC / f"{name}.hip"
    obj = BUILD / f"{name}.o"
    _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    return obj


def build(jobs: int = 8) -> dict[str, Path]:
    BUILD.mkdir(exist_ok=True)
    sources = [s for s in HIP_SOURCES if (CSRC / f"{s}.hip").exists()]
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(sources)))) as pool:
        objects = list(pool.map(_hip_object, sources))
    lib = PKG / "libkf_hip.so"
    tmp = PKG / "libkf_hip.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objects), "-o", str(tmp)])
    os.replace(tmp, lib)
    scene = PKG / "libkf_scene.so"
    tmp = PKG / "libkf_scene.so.tmp"
    host_obj = BUILD / "scene.o"
    _run(["g++", "-std=c++20", "-O2", "-fPIC", "-fopenmp", "-c", str(CSRC / "scene.cpp"), "-o", str(host_obj)])
    dev_obj = _hip_object("scene_device")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", str(host_obj), str(dev_obj), "-lgomp", "-o", str(tmp)])
    os.replace(tmp, scene)
    io = PKG / "libkf_io.so"
    tmp = PKG / "libkf_io.so.tmp"
    _run(["g++", "-std=c++20", "-O2", "-fPIC", "-shared", str(CSRC / "kf_io.cpp"), "-o", str(tmp)])
    os.replace(tmp, io)
    return {"hip": lib, "scene": scene, "io": io}


if __name__ == "__main__":
    out = build()
    for k, v in out.items():
        print(k, v)
