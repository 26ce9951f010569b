"""The reconstruction app end to end on the GPU (src/app: main, Reconstruction,
PipelineSet, FrameOutput, AteEvaluator):

* a generated sequence written as a TUM directory (16-bit PNG depth at 5000
  per metre, PNG colour, index files, ground truth) runs through an ablation
  set of three pipelines (float, bf16, int16 voxels) from a --pipelines
  file: every frame's raycast PNG and point cloud per pipeline, the
  ablation CSV rows, a trajectory and a mesh per pipeline, the ATE against
  groundtruth.txt;
* the reference pipeline's trajectory is bit-identical to a plain Pipeline
  driven frame by frame with the same configuration (the app adds no
  arithmetic of its own), and the float pipeline compared with itself in a
  second set reports zero deviation;
* sparse storage with band integration runs from the CLI flags alone.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ARGS = ["--volume-resolution", "128", "--voxel-size", "0.02", "--volume-camera-margin", "0.5",
        "--bilateral-radius", "3", "--bilateral-spatial-sigma", "4.5", "--bilateral-depth-sigma", "0.03"]

PIPELINES = """
compare-every-n-frames = 1

[[pipeline]]
name = "baseline"

[[pipeline]]
name = "bf16"
voxel = "bf16"

[[pipeline]]
name = "int16"
voxel = "quantized"
"""


@pytest.fixture(scope="module")
def dataset(tmp_path_factory):
    from kinectfusion_amd.app.sensor import write_tum_dataset

    return write_tum_dataset(tmp_path_factory.mktemp("tum") / "synthetic_room", "room", seed=0, frames=6)


def run_app(argv):
    from kinectfusion_amd.app import options
    from kinectfusion_amd.app.reconstruction import Reconstruction

    opts = options.parse_args(argv)
    options.validate_options(opts)
    rec = Reconstruction(opts)
    assert rec.run() == 0
    return rec


def test_ablation_set_writes_every_output(dataset, tmp_path, torch_gpu):
    from kinectfusion_amd.app.frame_output import read_ply
    from kinectfusion_amd.app.png import read_color_png
    from kinectfusion_amd.trajectory import read_tum

    (tmp_path / "p.toml").write_text(PIPELINES)
    out = tmp_path / "out"
    rec = run_app([str(dataset), *ARGS, "--frames", "6", "--pipelines", str(tmp_path / "p.toml"),
                   "--output-dir", str(out)])
    assert set(rec.ate) == {"baseline", "bf16", "int16"}
    for name, stats in rec.ate.items():
        print(f"{name}: ATE {stats.rmse * 1000:.2f} mm over {stats.pairs} frames")
        assert stats.pairs == 6 and stats.rmse < 0.01
    for sub in ("", "bf16", "int16"):
        d = out / sub if sub else out
        stamps, pos, _ = read_tum(d / "trajectory.txt")
        assert len(stamps) == 6
        for i in range(6):
            img = read_color_png(d / f"frame_{i:06d}_raycast.png")
            assert img.shape == (480, 640) and np.count_nonzero(img) > 1000
            counts, verts, _ = read_ply(d / f"frame_{i:06d}_raycast_point_cloud.ply")
            assert counts["vertex"] > 1000 and np.isfinite(verts["p"]).all()
        counts, verts, faces = read_ply(d / "mesh.ply")
        assert counts["face"] > 1000 and faces["v"].max() < counts["vertex"]
    rows = (out / "ablation_stats.csv").read_text().splitlines()
    assert len(rows) == 1 + 6 * 2
    first = rows[1].split(",")
    assert first[:2] == ["0", "bf16"] and int(first[2]) > 100_000 and int(first[8]) > 1000
    # bf16 and int16 stores deviate from float by their codecs' resolution on
    # average (single voxels of small weight, where each pipeline's own pose
    # and rounding move the average most, deviate more)
    mean = max(float(r.split(",")[6]) for r in rows[1:])
    print(f"largest mean |delta tsdf| vs float: {mean:.2e}")
    assert 0.0 < mean < 5e-3


def test_app_trajectory_is_the_pipelines(dataset, tmp_path, torch_gpu):
    import kinectfusion_amd as K
    from kinectfusion_amd.app import options
    from kinectfusion_amd.app.sensor import TumSensor

    rec = run_app([str(dataset), *ARGS, "--frames", "6", "--output-dir", str(tmp_path / "o"),
                   "--no-write-raycast-images", "--no-write-point-clouds", "--no-write-mesh"])
    app_traj = np.stack([p for _, p in rec.tracks[0].trajectory])
    sensor = TumSensor(dataset)
    opts = options.parse_args([str(dataset), *ARGS])
    cfg = opts.pipeline_config(sensor.intrinsics, 640, 480)
    cfg.integration.depth_scale = 5000.0
    p = K.Pipeline(cfg)
    while sensor.process_next_frame():
        d = torch_gpu.from_numpy(sensor.depth.view(np.int16)).cuda()
        c = torch_gpu.from_numpy(sensor.color.view(np.int32)).cuda()
        p.process_frame(d, c)
    assert np.array_equal(p.trajectory(), app_traj)
    p.close()
    sensor.close()
    # a pipeline compared with an identical one deviates nowhere
    (tmp_path / "same.toml").write_text('[[pipeline]]\nname = "a"\n[[pipeline]]\nname = "b"\n')
    run_app([str(dataset), *ARGS, "--frames", "3", "--pipelines", str(tmp_path / "same.toml"),
             "--output-dir", str(tmp_path / "s"), "--no-write-raycast-images", "--no-write-point-clouds",
             "--no-write-mesh"])
    for row in (tmp_path / "s" / "ablation_stats.csv").read_text().splitlines()[1:]:
        v = row.split(",")
        assert v[3:8] == ["0", "0", "0", "0", "0"] and v[9:] == ["0", "0", "0", "0", "0", "0"], row


def test_sparse_storage_from_flags(tmp_path, torch_gpu):
    rec = run_app(["synthetic:room:0", *ARGS, "--frames", "5", "--storage", "sparse", "--integration", "band",
                   "--sparse-capacity", "20000", "--output-dir", str(tmp_path / "sp"), "--no-write-point-clouds"])
    assert rec.ate["baseline"].rmse < 0.01
    assert (tmp_path / "sp" / "mesh.ply").stat().st_size > 10_000
    with pytest.raises(ValueError, match="band integration"):
        run_app(["synthetic:room:0", *ARGS, "--frames", "2", "--storage", "sparse",
                 "--output-dir", str(tmp_path / "bad")])
