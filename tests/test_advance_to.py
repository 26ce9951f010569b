"""The raycaster's closed-form runs of float steps (advance_to,
kinectfusion_amd/csrc/kf_steps.hpp), compiled from the same header for the
host with g++, against the plain loop it replaces,
    while (len < lim) { last = len; len += step; }
every addition rounded to nearest: the returned length and the last length
before the final step equal the loop's bit for bit.  Cases: the march's
own ranges (voxel sizes 1 mm - 5 cm, step scales 0.1 - 1, rays to 12 m),
limits placed on and next to binade edges and multiples of the step, and
steps that are odd multiples of half an ulp (ties, which the parity of the
length decides)."""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = r"""
#include "kf_steps.hpp"
extern "C" void check(const float* len, const float* step, const float* lim, float* out, long n) {
  for (long i = 0; i < n; ++i) {
    float a = len[i], la = -1.0F;
    while (a < lim[i]) { la = a; a += step[i]; }
    float lb = -1.0F;
    const float b = kf::advance_to(len[i], step[i], lim[i], lb);
    out[4 * i + 0] = a;
    out[4 * i + 1] = la;
    out[4 * i + 2] = b;
    out[4 * i + 3] = lb;
  }
}
"""


@pytest.fixture(scope="module")
def check(tmp_path_factory):
    d = tmp_path_factory.mktemp("steps")
    (d / "steps.cpp").write_text(SRC)
    lib = d / "libsteps.so"
    subprocess.run(["g++", "-O2", "-std=c++20", "-ffp-contract=off", "-shared", "-fPIC",
                    f"-I{ROOT / 'kinectfusion_amd' / 'csrc'}", str(d / "steps.cpp"), "-o", str(lib)], check=True)
    so = C.CDLL(str(lib))
    so.check.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long]

    def run(length, step, lim):
        length, step, lim = (np.ascontiguousarray(v, np.float32) for v in (length, step, lim))
        out = np.empty((length.size, 4), np.float32)
        so.check(length.ctypes.data, step.ctypes.data, lim.ctypes.data, out.ctypes.data, length.size)
        return out

    return run


def same(out):
    loop = out[:, :2].view(np.uint32)
    fast = out[:, 2:].view(np.uint32)
    bad = np.nonzero((loop != fast).any(1))[0]
    return bad


def test_march_ranges(check):
    rng = np.random.default_rng(11)
    n = 200_000
    voxel = rng.uniform(0.001, 0.05, n)
    step = (voxel * rng.uniform(0.1, 1.0, n)).astype(np.float32)
    length = rng.uniform(0.05, 12.0, n).astype(np.float32)
    lim = (length + rng.uniform(-0.1, 3.0, n)).astype(np.float32)
    out = check(length, step, lim)
    bad = same(out)
    assert bad.size == 0, (length[bad[:5]], step[bad[:5]], lim[bad[:5]], out[bad[:5]])
    assert (out[:, 0] > length).mean() > 0.9  # most cases take steps


def test_limits_on_binade_edges_and_step_multiples(check):
    rng = np.random.default_rng(12)
    edges = np.float32([0.25, 0.5, 1.0, 2.0, 4.0, 8.0])
    cases = []
    for e in edges:
        for _ in range(3000):
            step = np.float32(rng.uniform(0.0005, 0.03))
            length = np.float32(e - rng.uniform(0.0, 0.5) * e)
            for lim in (e, np.nextafter(e, np.float32(0)), np.nextafter(e, np.float32(10)),
                        np.float32(length + step * rng.integers(1, 400)), np.float32(e + step * rng.integers(0, 50))):
                cases.append((length, step, lim))
            # a length on the edge itself
            cases.append((e, step, np.float32(e + rng.uniform(0.0, 2.0))))
    length, step, lim = (np.array(c, np.float32) for c in zip(*cases))
    out = check(length, step, lim)
    bad = same(out)
    assert bad.size == 0, (length[bad[:5]], step[bad[:5]], lim[bad[:5]], out[bad[:5]])


def test_ties_are_decided_by_parity(check):
    """Steps of (q + 1/2) ulps of the length's binade: every addition is a
    tie, so the step alternates between q and q + 1 ulps with the parity of
    the length; the closed form takes them one at a time."""
    rng = np.random.default_rng(13)
    cases = []
    for _ in range(20000):
        e = int(rng.integers(-3, 3))
        u = np.float32(2.0 ** (e - 23))
        q = int(rng.integers(1, 40000))
        step = np.float32((q + 0.5) * float(u))
        length = np.float32(2.0 ** e * rng.uniform(1.0, 1.9))
        lim = np.float32(length + step * rng.integers(1, 60))
        cases.append((length, step, lim))
    length, step, lim = (np.array(c, np.float32) for c in zip(*cases))
    out = check(length, step, lim)
    bad = same(out)
    assert bad.size == 0, (length[bad[:5]], step[bad[:5]], lim[bad[:5]], out[bad[:5]])


def test_no_step_and_tiny_lengths(check):
    length = np.float32([0.3, 0.3, 1e-30, 0.0, 5e-39, 1.0])
    step = np.float32([0.01, 0.01, 0.01, 0.002, 0.002, 0.125])
    lim = np.float32([0.3, 0.1, 0.05, 0.01, 0.01, 3.0])
    out = check(length, step, lim)
    assert same(out).size == 0
    assert out[0, 2] == np.float32(0.3) and out[0, 3] == np.float32(-1.0)  # no step: last untouched
