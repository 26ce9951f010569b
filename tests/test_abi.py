"""The C ABI library loads and exports every entry point include/kf_hip.h
declares (no device calls)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared():
    text = (ROOT / "include" / "kf_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for n in ["kf_create", "kf_destroy", "kf_pyramid_build", "kf_integrate", "kf_raycast", "kf_icp_reduce",
              "kf_icp_track_maps", "kf_track", "kf_process_frame", "kf_mesh_extract", "kf_volume_download"]:
        assert n in names


def test_integration_guide_maps_every_entry_point():
    """The integration guide's interface table names every declared entry
    point, so a maintainer wiring the reference to the library finds each."""
    guide = (ROOT / "INTEGRATION.md").read_text()
    missing = [n for n in declared() if f"`{n}`" not in guide]
    assert not missing, missing


def test_library_exports_every_declared_symbol():
    from kinectfusion_amd import _native

    lib = ctypes.CDLL(str(_native.HIP_LIB))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    bound = {name for name, _, _ in _native.HIP_SYMBOLS}
    assert set(declared()) <= bound, set(declared()) - bound


def test_config_defaults_are_the_reference_defaults():
    import ctypes as C

    from kinectfusion_amd import _native as N

    c = N.KfConfig()
    N.hip().kf_config_defaults(C.byref(c))
    assert (c.voxel_size, c.truncation_distance, c.max_weight) == (C.c_float(0.01).value, C.c_float(0.05).value,
                                                                    196.0)
    assert (c.bilateral_radius, c.levels, list(c.icp_iterations)[:3]) == (2, 3, [10, 5, 4])
    o = N.KfIcpOptions()
    N.hip().kf_icp_options_defaults(C.byref(o))
    assert o.min_correspondences == 64 and abs(o.min_normal_dot - 0.9659258) < 1e-7


def test_cpu_space_is_refused():
    import pytest

    import kinectfusion_amd as K

    with pytest.raises(NotImplementedError):
        K.PipelineConfig(space="cpu").native()


def test_device_images_are_validated_before_any_pointer_crosses():
    """Host tensors, strided views, wrong dtypes and wrong shapes are refused
    with ValueError before their data pointer reaches the library."""
    import pytest
    import torch

    from kinectfusion_amd.pipeline import _check_image

    cpu = torch.zeros((4, 6), dtype=torch.int16)
    with pytest.raises(ValueError, match="device"):
        _check_image(cpu, "depth", ("int16", "uint16"), (4, 6))

    class FakeCuda:  # a device tensor's attributes, without a device
        def __init__(self, t, contiguous=True):
            self.t, self.is_cuda, self._c = t, True, contiguous
            self.dtype, self.shape = t.dtype, t.shape

        def is_contiguous(self):
            return self._c

    with pytest.raises(ValueError, match="contiguous"):
        _check_image(FakeCuda(cpu, contiguous=False), "depth", ("int16",), (4, 6))
    with pytest.raises(ValueError, match="dtype"):
        _check_image(FakeCuda(cpu.float()), "depth", ("int16",), (4, 6))
    with pytest.raises(ValueError, match="shape"):
        _check_image(FakeCuda(cpu), "depth", ("int16",), (6, 4))
    assert _check_image(FakeCuda(cpu), "depth", ("int16",), (4, 6)) is not None
    assert _check_image(None, "color", ("int32",), (4, 6), optional=True) is None
