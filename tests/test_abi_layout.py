"""Struct layouts of include/kf_hip.h as a C compiler sees them, against the
ctypes mirror in kinectfusion_amd/_native.py: a C translation unit that
includes the header is compiled with gcc and prints sizeof/offsetof of every
field the mirror declares (no device needed)."""
import ctypes as C
import shutil
import subprocess

import pytest

from conftest import ROOT

STRUCTS = ["KfConfig", "KfIcpOptions", "KfIcpResult", "KfLevelView", "KfMapsView", "KfDepthFrame",
           "KfStageTimes", "KfFrameResult", "KfPosenetConfig"]
C_NAME = {"KfConfig": "kf_config", "KfIcpOptions": "kf_icp_options", "KfIcpResult": "kf_icp_result",
          "KfLevelView": "kf_level_view", "KfMapsView": "kf_maps_view", "KfDepthFrame": "kf_depth_frame",
          "KfStageTimes": "kf_stage_times", "KfFrameResult": "kf_frame_result",
          "KfPosenetConfig": "kf_posenet_config"}


def c_field(name: str) -> str:
    return "lambda" if name == "lambda_" else name


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    from kinectfusion_amd import _native as N

    if shutil.which("gcc") is None:
        pytest.skip("gcc is not available")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "kf_hip.h"', "int main(void) {"]
    for s in STRUCTS:
        cn = C_NAME[s]
        lines.append(f'  printf("{s} . %zu %zu\\n", sizeof({cn}), _Alignof({cn}));')
        for name, _ in getattr(N, s)._fields_:
            f = c_field(name)
            lines.append(f'  printf("{s} {name} %zu %zu\\n", offsetof({cn}, {f}), sizeof((({cn}*)0)->{f}));')
    lines += ["  return 0;", "}"]
    d = tmp_path_factory.mktemp("abi")
    src, exe = d / "layout.c", d / "layout"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", str(ROOT / "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    table = {}
    for line in out.splitlines():
        s, f, a, b = line.split()
        table[(s, f)] = (int(a), int(b))
    return table


@pytest.mark.parametrize("struct", STRUCTS)
def test_struct_layout_matches_the_header(c_layout, struct):
    from kinectfusion_amd import _native as N

    cls = getattr(N, struct)
    size, align = c_layout[(struct, ".")]
    assert C.sizeof(cls) == size, (struct, C.sizeof(cls), size)
    assert C.alignment(cls) == align
    for name, _ in cls._fields_:
        desc = getattr(cls, name)
        assert (desc.offset, desc.size) == c_layout[(struct, name)], (struct, name)
