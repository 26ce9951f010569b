"""kf_icp_reduce_static, the allocation-free form of kf_icp_reduce, through
the C ABI: on real maps of the synthetic room its sums and count equal
kf_icp_reduce's bit for bit (the same sweep and reduction kernels over the
same partition), on two calls in a row, the second with a smaller input
(level 1 after level 0), so rows left by the first call cannot leak into the
second.  Oversized maps are rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

from kf_helpers import pipeline_config, to_dev

pytestmark = pytest.mark.gpu

ROOM_VOXEL = 0.02


@pytest.fixture(scope="module")
def K():
    import kinectfusion_amd

    return kinectfusion_amd


def static_reduce(N, torch, lv, ln, mv, mn, k, cam, mod, max_dist, min_dot):
    lh, lw = lv.shape[:2]
    mh, mw = mv.shape[:2]
    sums = torch.full((29,), float("nan"), dtype=torch.float64, device="cuda")
    count = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    kk = (C.c_float * 4)(k.fx, k.fy, k.cx, k.cy)
    cc = (C.c_float * 12)(*np.asarray(cam, np.float32).reshape(12))
    mm = (C.c_float * 12)(*np.asarray(mod, np.float32).reshape(12))
    torch.cuda.synchronize()
    status = N.hip().kf_icp_reduce_static(lv.data_ptr(), ln.data_ptr(), lw, lh, mv.data_ptr(), mn.data_ptr(), mw,
                                          mh, kk, cc, mm, max_dist, min_dot, sums.data_ptr(), count.data_ptr(),
                                          None)
    N.check(status)
    torch.cuda.synchronize()
    return sums.cpu().numpy(), int(count.cpu().numpy()[0])


def test_static_reduce_equals_context_reduce(K, ko, torch_gpu):
    from kinectfusion_amd import _native as N

    scene = K.SyntheticScene("room", seed=0)
    p = K.Pipeline(pipeline_config(K, voxel=ROOM_VOXEL))
    for i in range(2):
        depth, rgba, pose = scene.frame(i)
        p.integrate(to_dev(torch_gpu, depth), ko.rigid_inverse(pose), color=to_dev(torch_gpu, rgba))
    depth, _, _ = scene.frame(2)
    src = K.PyramidSource(p)
    src.build(to_dev(torch_gpu, depth))
    model_pose = scene.pose(1)
    cam = ko.pose12(model_pose)
    mod = ko.pose12(ko.rigid_inverse(model_pose))
    tracker = K.ProjectiveIcpTracker(p)
    opt = tracker.options
    for level in (0, 1):  # the second call has the smaller input
        lv = src.level(level)
        maps = p.raycast(model_pose, level)
        p.synchronize()
        want, want_n = tracker.normal_equations(lv["vertices"], lv["normals"], maps.points, maps.normals,
                                                lv["intrinsics"], cam, mod)
        got, got_n = static_reduce(N, torch_gpu, lv["vertices"], lv["normals"], maps.points, maps.normals,
                                   lv["intrinsics"], cam, mod, opt.max_point_distance, opt.min_normal_dot)
        print(f"level {level}: {got_n} correspondences (context reduce {want_n})", flush=True)
        assert got_n == want_n and got_n > 1000
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), np.abs(got - want).max()
    p.close()


def test_static_reduce_rejects_oversized_maps(K, torch_gpu):
    from kinectfusion_amd import _native as N

    t = torch_gpu.zeros((4,), dtype=torch_gpu.float32, device="cuda")
    out = torch_gpu.zeros((32,), dtype=torch_gpu.float64, device="cuda")
    kk = (C.c_float * 4)(1.0, 1.0, 0.0, 0.0)
    eye = (C.c_float * 12)(1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0)
    big = 4097  # 4097 x 4097 pixels > KF_ICP_STATIC_MAX_PIXELS; nothing is read
    status = N.hip().kf_icp_reduce_static(t.data_ptr(), t.data_ptr(), big, big, t.data_ptr(), t.data_ptr(), 1, 1,
                                          kk, eye, eye, 0.05, 0.9, out.data_ptr(), out.data_ptr(), None)
    assert status == N.KF_ERR_INVALID_ARGUMENT
    assert b"KF_ICP_STATIC_MAX_PIXELS" in N.hip().kf_last_error(None)
