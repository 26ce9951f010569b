"""The reconstruction app's host side (the reference's src/app and its I/O),
without a GPU:

* PNG: files written by Pillow with its adaptive row filters (every filter
  type occurs, checked) decode to Pillow's pixels -- 16-bit depth, RGB,
  RGBA, gray, gray+alpha, palette; this module's files decode in Pillow to
  the written pixels (read_png.cpp / write_png.cpp semantics);
* PLY: the byte layout of mesh.ply and the raycast point cloud
  (frame_output.cpp:65-185) against struct-packed records;
* the --pipelines TOML file (pipeline_config.cpp:51-279): overrides,
  reference, compare interval, and every error the reference raises;
* CLI defaults equal the reference's (app_options.hpp:21-83) and validation
  rejects what validate_options rejects;
* the comparator's known answers (test/tests.cpp:837-878 and :530-560) and
  the ablation CSV (frame_output.cpp:215-242);
* a generated TUM directory reads back through the TUM sensor.
"""
import io
import struct
import zlib

import numpy as np
import pytest

from kinectfusion_amd.app import frame_output as FO
from kinectfusion_amd.app import names, options, png
from kinectfusion_amd.app.pipeline_config import parse_pipeline_set
from kinectfusion_amd.app.pipeline_set import Comparator, HostMaps, PipelineComparison, decode_voxels
from kinectfusion_amd.pipeline import CameraIntrinsics, PipelineConfig

PIL = pytest.importorskip("PIL.Image")


def pil_png(img, mode):
    buf = io.BytesIO()
    im = PIL.fromarray(img)  # the mode follows from the array's shape and type
    assert im.mode == mode
    im.save(buf, format="PNG")
    return buf.getvalue()


def filter_types(data: bytes) -> set[int]:
    px, hdr = None, None
    chunks, pos, idat = [], 8, b""
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        if kind == b"IDAT":
            idat += body
        pos += 12 + n
    w, h, depth, ctype = hdr[:4]
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    row = (w * ch * depth + 7) // 8
    raw = zlib.decompress(idat)
    return {raw[y * (row + 1)] for y in range(h)}


def smooth_image(h, w, ch, seed, dtype=np.uint8, hi=256):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    base = (np.sin(x / 7.0) + np.cos(y / 5.0) + 2.0) * (hi / 4.5)
    img = base[..., None] + rng.integers(0, max(2, hi // 64), (h, w, ch))
    return np.clip(img, 0, hi - 1).astype(dtype)


def test_png_decodes_pillow_files_with_every_filter():
    seen = set()
    for mode, ch in (("RGB", 3), ("RGBA", 4), ("L", 1), ("LA", 2)):
        img = smooth_image(37, 53, ch, seed=ch)
        data = pil_png(img if ch > 1 else img[..., 0], mode)
        seen |= filter_types(data)
        px, hdr = png.decode(data)
        assert np.array_equal(px, img), mode
        rgba = png.to_rgba8(px, hdr).view(np.uint8).reshape(37, 53, 4)
        want = {1: np.concatenate([img, img, img, np.full_like(img, 255)], 2),
                2: np.concatenate([img[..., :1]] * 3 + [img[..., 1:]], 2),
                3: np.concatenate([img, np.full_like(img[..., :1], 255)], 2), 4: img}[ch]
        assert np.array_equal(rgba, want), mode
    assert seen >= {1, 2, 4}, seen  # Sub, Up and Paeth rows decoded
    # Pillow's encoder never picks Average: a hand-worked 1-channel 2 x 3 image
    # with filter 3 on every row; predictions floor((left + up) / 2) in full
    # precision (200 + 250 does not wrap) and the sum modulo 256
    raw = bytes([3, 10, 20, 3, 5, 7, 3, 0, 250])
    body = struct.pack(">IIBBBBB", 2, 3, 8, 0, 0, 0, 0)
    chunk = lambda k, b: struct.pack(">I", len(b)) + k + b + struct.pack(">I", zlib.crc32(k + b))
    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", body) + chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b"")
    # row 0: 10, 20 + 10 // 2 = 25; row 1: 5 + 10 // 2 = 10, 7 + (10 + 25) // 2 = 24;
    # row 2: 0 + 10 // 2 = 5, (250 + (5 + 24) // 2) % 256 = 8
    assert png.decode(data)[0][..., 0].tolist() == [[10, 25], [10, 24], [5, 8]]
    raw = bytes([0, 0, 250, 3, 200, 0])  # left 200 and up 250: (200 + 250) // 2 = 225, not 450 % 256 // 2
    body = struct.pack(">IIBBBBB", 2, 2, 8, 0, 0, 0, 0)
    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", body) + chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b"")
    assert png.decode(data)[0][..., 0].tolist() == [[0, 250], [200, 225]]
    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", 2, 1, 8, 0, 0, 0, 0)) + \
        chunk(b"IDAT", zlib.compress(bytes([3, 250, 0]))) + chunk(b"IEND", b"")
    assert png.decode(data)[0][..., 0].tolist() == [[250, 125]]
    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", 1, 2, 8, 0, 0, 0, 0)) + \
        chunk(b"IDAT", zlib.compress(bytes([0, 250, 3, 200]))) + chunk(b"IEND", b"")
    assert png.decode(data)[0][..., 0].tolist() == [[250], [(200 + 125) % 256]]
    depth = smooth_image(48, 64, 1, seed=9, dtype=np.uint16, hi=40000)[..., 0]
    data = pil_png(depth, "I;16")
    px, hdr = png.decode(data)
    assert (hdr["bit_depth"], hdr["color_type"]) == (16, 0) and np.array_equal(px[..., 0], depth)
    pal = PIL.fromarray(smooth_image(20, 30, 3, seed=4)).quantize(16)
    buf = io.BytesIO()
    pal.save(buf, format="PNG")
    px, hdr = png.decode(buf.getvalue())
    assert hdr["color_type"] == 3 and np.array_equal(px, np.asarray(pal.convert("RGB")))


def test_png_files_written_here_read_in_pillow(tmp_path):
    rgba = smooth_image(40, 60, 4, seed=1)
    png.write_png(rgba.view(np.uint32)[..., 0], tmp_path / "c.png")
    back = np.asarray(PIL.open(tmp_path / "c.png"))
    assert back.shape == (40, 60, 4) and np.array_equal(back, rgba)
    assert png.read_color_png(tmp_path / "c.png").view(np.uint8).reshape(40, 60, 4).tolist() == rgba.tolist()
    depth = smooth_image(30, 40, 1, seed=2, dtype=np.uint16, hi=65535)[..., 0]
    png.write_depth_png(depth, tmp_path / "d.png")
    im = PIL.open(tmp_path / "d.png")
    assert np.array_equal(np.asarray(im).astype(np.uint16), depth)
    assert np.array_equal(png.read_depth_png(tmp_path / "d.png"), depth)
    with pytest.raises(png.PngError, match="16-bit grayscale"):
        png.read_depth_png(tmp_path / "c.png")
    bad = bytearray((tmp_path / "c.png").read_bytes())
    bad[40] ^= 0xFF
    with pytest.raises(png.PngError):
        png.decode(bytes(bad))


class _Mesh:
    positions = np.array([[0.0, 1.0, 2.0], [3.5, -1.0, 0.25], [1, 1, 1]], np.float32)
    normals = np.array([[0, 0, 1], [0, 1, 0], [1, 0, 0]], np.float32)
    colors = np.array([[300.0, -4.0, 17.9], [255.0, 0.0, 128.5], [1, 2, 3]], np.float32)
    triangles = np.array([[0, 1, 2], [2, 1, 0]], np.uint32)


def test_mesh_ply_layout(tmp_path):
    FO.write_mesh_ply(_Mesh, tmp_path / "mesh.ply")
    data = (tmp_path / "mesh.ply").read_bytes()
    header = ("ply\nformat binary_little_endian 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
              "property float z\nproperty float nx\nproperty float ny\nproperty float nz\nproperty uchar red\n"
              "property uchar green\nproperty uchar blue\nelement face 2\n"
              "property list uchar uint vertex_indices\nend_header\n").encode()
    assert data.startswith(header)
    body = data[len(header):]
    v0 = struct.pack("<6f3B", 0.0, 1.0, 2.0, 0, 0, 1, 255, 0, 17)  # clamped, truncated channels
    v1 = struct.pack("<6f3B", 3.5, -1.0, 0.25, 0, 1, 0, 255, 0, 128)
    assert body[:27] == v0 and body[27:54] == v1
    faces = body[81:]
    assert faces == struct.pack("<B3I", 3, 0, 1, 2) + struct.pack("<B3I", 3, 2, 1, 0)
    counts, verts, f = FO.read_ply(tmp_path / "mesh.ply")
    assert counts == {"vertex": 3, "face": 2} and np.array_equal(f["v"], _Mesh.triangles)


def test_point_cloud_ply_keeps_finite_pixels(tmp_path):
    pts = np.full((2, 2, 3), np.nan, np.float32)
    nrm = np.full((2, 2, 3), np.nan, np.float32)
    pts[0, 1] = [1, 2, 3]
    nrm[0, 1] = [0, 0, -1]
    pts[1, 0] = [4, 5, 6]  # normal invalid: dropped
    col = np.array([[0, 0x00332211], [0, 0]], np.uint32)
    assert FO.write_point_cloud(HostMaps(pts, nrm, col), tmp_path / "p.ply") == 1
    data = (tmp_path / "p.ply").read_bytes()
    assert b"element vertex 1\n" in data and b"element face" not in data
    assert data.endswith(struct.pack("<6f3B", 1, 2, 3, 0, 0, -1, 0x11, 0x22, 0x33))


TOML = """
reference = "bf16"
compare-every-n-frames = 5

[[pipeline]]
name = "baseline"

[[pipeline]]
name = "bf16"
voxel = "bf16"
color = "none"
tsdf-variant = "classic"
integration = "band"
max-weight = 64.0
observation-weight = 2
icp-damping = "diagonal"
icp-lambda = 0.01
icp-adaptive-damping = true

[[pipeline]]
name = "sparse"
storage = "sparse"
sparse-capacity = 1000
integration = "band"
"""


def test_pipeline_toml(tmp_path):
    base = PipelineConfig(name="baseline")
    f = tmp_path / "p.toml"
    f.write_text(TOML)
    cfg = parse_pipeline_set(f, base, 1)
    assert [p.name for p in cfg.pipelines] == ["baseline", "bf16", "sparse"]
    assert cfg.reference == "bf16" and cfg.compare_every_n_frames == 5
    b = cfg.pipelines[1]
    assert (b.voxel, b.color, b.tsdf_rule, b.integration.mode) == ("bf16", "none", "classic", "band")
    assert (b.integration.max_weight, b.integration.observation_weight) == (64.0, 2.0)
    assert cfg.icp_damping["bf16"] == ("diagonal", 0.01, True) and cfg.icp_damping["baseline"][0] == "none"
    assert cfg.pipelines[0].voxel == "float" and base.name == "baseline"  # defaults untouched
    assert (cfg.pipelines[2].storage, cfg.pipelines[2].block_capacity) == ("hashed", 1000)
    for text, msg in [("[[pipeline]]\nvoxel = 'bf16'\n", "requires a name"),
                      ("[[pipeline]]\nname = 'a,b'\n", "commas"),
                      ("[[pipeline]]\nname = 'a'\nresolution = 3\n", "Unknown pipeline config key 'resolution'"),
                      ("[[pipeline]]\nname = 'a'\nmax-weight = 'x'\n", "wrong type"),
                      ("[[pipeline]]\nname = 'a'\nvoxel = 'half'\n", "Unknown voxel store 'half'"),
                      ("frames = 3\n[[pipeline]]\nname = 'a'\n", "Unknown pipeline config key 'frames'"),
                      ("reference = 'a'\n", "at least one"),
                      ("pipeline = 3\n", "array of tables"),
                      ("[[pipeline]\n", "Failed to parse")]:
        f.write_text(text)
        with pytest.raises(ValueError, match=msg):
            parse_pipeline_set(f, base, 1)


def test_cli_defaults_and_validation():
    o = options.parse_args([])
    assert (o.max_frames, o.volume_resolution, o.voxel_size, o.truncation_distance) == (30, 512, 0.01, 0.05)
    assert (o.volume_camera_margin, o.depth_scale, o.min_depth, o.max_depth) == (2.56, 5000.0, 0.4, 8.0)
    assert (o.pyramid_levels, o.bilateral_radius, o.bilateral_spatial_sigma, o.bilateral_depth_sigma) == \
        (3, 2, 2.0, 0.08)
    assert [o.icp_iterations_for_level(level) for level in range(4)] == [10, 5, 4, 4]
    assert (o.mesh_min_weight, o.voxel_store, o.tsdf_variant, o.space, o.compare_every_n_frames) == \
        (2.0, "float", "angle-weighted", "hip", 1)
    assert o.projective_tsdf_distance and o.bilateral_filter and o.write_mesh and not o.raycast_seed_previous
    o = options.parse_args(["data", "--no-bilateral-filter", "--icp-iterations", "3", "--voxel", "bf16",
                            "--storage", "sparse", "--integration", "band", "--raycast-seed-previous",
                            "--no-write-point-clouds", "--frames", "-1"])
    assert not o.bilateral_filter and o.icp_iterations_for_level(0) == 3 and o.max_frames == -1
    cfg = o.pipeline_config(CameraIntrinsics(525, 525, 319.5, 239.5), 640, 480)
    assert (cfg.voxel, cfg.storage, cfg.integration.mode, cfg.raycast.seed_from_previous) == \
        ("bf16", "hashed", "band", True)
    assert cfg.block_capacity == 64 ** 3 // 4
    assert np.array_equal(np.float32(cfg.volume.origin), np.float32([-2.56, -2.56, -2.56]))
    for argv, msg in [(["--voxel-size", "0"], "voxel size"), (["--min-depth", "3", "--max-depth", "2"], "depth range"),
                      (["--icp-iterations", "0"], "ICP iterations"), (["--min-normal-dot", "2"], "normal dot"),
                      (["--icp-adaptive-damping"], "requires --icp-damping"),
                      (["--volume-camera-margin", "9"], "camera margin"), (["--init", "learned"], "--posenet")]:
        with pytest.raises(ValueError, match=msg):
            options.validate_options(options.parse_args(argv))
    with pytest.raises(ValueError, match="Unknown raycast backend"):
        names.raycast_backend_from_name("ray")
    with pytest.raises(NotImplementedError, match="not built"):
        options.parse_args(["--space", "cpu"]).pipeline_config(CameraIntrinsics(1, 1, 0, 0), 8, 8).native()
    with pytest.raises(NotImplementedError, match="band-march"):
        options.parse_args(["--raycast", "band-march"]).pipeline_config(CameraIntrinsics(1, 1, 0, 0), 8, 8).native()


def test_surface_comparison_known_answer():
    """test/tests.cpp:837-878."""
    nan = np.full((2, 2, 3), np.nan, np.float32)
    p = HostMaps(nan.copy(), nan.copy(), np.zeros((2, 2), np.uint32))
    r = HostMaps(nan.copy(), nan.copy(), np.zeros((2, 2), np.uint32))
    p.points[0, 0], p.normals[0, 0] = [0, 0, 1], [1, 0, 0]
    r.points[0, 0], r.normals[0, 0] = [0.1, 0, 1], [0, 1, 0]
    p.points[0, 1], p.normals[0, 1] = [0, 0, 1], [0, 0, -1]   # (x 1, y 0): primary only
    r.points[1, 0], r.normals[1, 0] = [0, 0, 1], [0, 0, -1]   # (x 0, y 1): reference only
    c = Comparator.compare_maps(p, r)
    assert (c.compared_pixels, c.only_primary, c.only_reference) == (1, 1, 1)
    assert c.max_point_distance == pytest.approx(0.1) and c.mean_point_distance == pytest.approx(0.1)
    assert c.max_normal_angle == pytest.approx(np.pi / 2) and c.mean_normal_angle == pytest.approx(np.pi / 2)
    with pytest.raises(ValueError, match="matching dimensions"):
        Comparator.compare_maps(p, HostMaps(nan[:1], nan[:1], np.zeros((1, 2), np.uint32)))


def test_volume_comparison_and_codecs():
    t = np.array([1.0, 0.5, -0.25, 0.0], np.float32)
    w = np.array([0.0, 2.0, 3.0, 1.0], np.float32)
    t2 = np.array([1.0, 0.25, -0.25, 0.0], np.float32)
    w2 = np.array([1.0, 2.0, 4.0, 0.0], np.float32)
    c = Comparator.compare_volumes((t, w), (t2, w2))
    assert (c.compared_voxels, c.only_primary, c.only_reference) == (2, 1, 1)
    assert (c.max_distance_delta, c.mean_distance_delta, c.max_weight_delta) == (0.25, 0.125, 1.0)
    same = Comparator.compare_volumes((t, w), (t.copy(), w.copy()))
    assert same.max_distance_delta == 0.0 and same.max_weight_delta == 0.0 and same.compared_voxels == 3
    # store codecs (volume.hpp:28-103): int16 32767 -> 1, bf16 0x3F80 -> 1, weight raw 512 -> 2
    q = np.array([[32767, 512], [np.uint16(-16384 & 0xFFFF), 256]], np.uint16)
    tq, wq = decode_voxels(q, "quantized")
    assert tq[0] == np.float32(32767) * np.float32(1 / 32767) and wq.tolist() == [2.0, 1.0]
    assert tq[1] == np.float32(-16384) * np.float32(1 / 32767)
    b = np.array([[0x3F80, 512], [0xBE80, 0]], np.uint16)
    tb, wb = decode_voxels(b, "bf16")
    assert tb.tolist() == [1.0, -0.25] and wb.tolist() == [2.0, 0.0]


def test_ablation_csv(tmp_path):
    from kinectfusion_amd.app.pipeline_set import SurfaceMapsComparison, VolumeComparison

    out = FO.FrameOutput(tmp_path, write_raycast_images=False, write_point_clouds=False)
    c = PipelineComparison("bf16", VolumeComparison(10, 1, 2, 0.5, 0.125, 3.0),
                           SurfaceMapsComparison(7, 0, 1, 0.1, 1e-5, 1.5707964, 0.25))
    out.append_ablation_stats(0, [c])
    out.append_ablation_stats(5, [c])
    text = (tmp_path / "ablation_stats.csv").read_text().splitlines()
    assert text[0] == FO.CSV_HEADER.strip()
    assert text[1] == "0,bf16,10,1,2,0.5,0.125,3,7,0,1,0.1,1e-05,1.5707964,0.25"
    assert text[2].startswith("5,bf16,") and len(text) == 3
    out2 = FO.FrameOutput(tmp_path, False, False)  # a new run truncates the file
    out2.append_ablation_stats(9, [c])
    assert len((tmp_path / "ablation_stats.csv").read_text().splitlines()) == 2
    out.close()
    out2.close()


def test_generated_tum_dataset_reads_back(tmp_path):
    from kinectfusion_amd import SyntheticScene
    from kinectfusion_amd.app import sensor as S
    from kinectfusion_amd.trajectory import read_tum

    S.write_tum_dataset(tmp_path / "seq", "room", seed=0, frames=3)
    sensor = S.TumSensor(tmp_path / "seq")
    scene = SyntheticScene("room", seed=0)
    assert sensor.count == 3 and sensor.intrinsics == CameraIntrinsics(525.0, 525.0, 319.5, 239.5)
    for i in range(3):
        assert sensor.process_next_frame()
        d, c, _ = scene.frame(i)
        assert np.array_equal(sensor.depth, d.astype(np.uint16) * 5) and np.array_equal(sensor.color, c)
        assert sensor.timestamp() == pytest.approx(1305031102.0 + i / 30.0)
    assert not sensor.process_next_frame()
    stamps, pos, _ = read_tum(sensor.groundtruth())
    assert len(stamps) == 3 and np.allclose(pos[2], scene.pose(2)[:3, 3])
    sensor.close()
    pre = S.TumSensor(tmp_path / "seq", preload=True)
    assert pre.process_next_frame() and pre.index == 0
    pre.close()
    assert S.tum_depth_intrinsics("x/rgbd_dataset_freiburg2_desk") == S.TUM_FREIBURG2
    assert S.tum_depth_intrinsics("x/rgbd_dataset_freiburg3_office") == S.TUM_FREIBURG3
    assert S.tum_depth_intrinsics("anything") == S.TUM_FREIBURG1
    with pytest.raises(FileNotFoundError):
        S.TumSensor(tmp_path / "missing")
