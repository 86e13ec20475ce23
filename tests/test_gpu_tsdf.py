"""GPU parity of TSDF fusion (tdm_integrate / tdm_integrate_frames) against the
oracle's restatement of Integrate (depth-fusion/demo.cu:18-68).

Bar: weights bit-identical, TSDF bit-identical (the kernel keeps the
reference's operation order with contraction off; 1e-5 absolute is the
contract's tolerance, the test asserts exact equality and reports the
maximum difference otherwise)."""
import glob
import os

import numpy as np
import pytest

import oracle
import tdm
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

TOL_TSDF = 1e-5


def _gpu_fuse(torch, K, poses, depths, dims, origin, vs, trunc, fused):
    dev = torch.device("cuda:0")
    n = dims[0] * dims[1] * dims[2]
    tsdf = torch.empty(n, dtype=torch.float32, device=dev)
    weight = torch.empty(n, dtype=torch.float32, device=dev)
    tdm.tsdf_init(tsdf, weight)
    dK = torch.from_numpy(np.ascontiguousarray(K, np.float32)).to(dev)
    dP = torch.from_numpy(np.ascontiguousarray(poses, np.float32).reshape(-1, 16)).to(dev)
    dD = torch.from_numpy(np.ascontiguousarray(depths, np.float32)).to(dev)
    if fused:
        tdm.integrate_frames(dK, dP, dD, dims, origin, vs, trunc, tsdf, weight)
    else:
        for f in range(len(depths)):
            tdm.integrate(dK, dP[f], dD[f], dims, origin, vs, trunc, tsdf, weight)
    torch.cuda.synchronize()
    shape = (dims[2], dims[1], dims[0])
    return tsdf.cpu().numpy().reshape(shape), weight.cpu().numpy().reshape(shape)


def _oracle_fuse(K, poses, depths, dims, origin, vs, trunc):
    t, w = oracle.new_grid(dims)
    for f in range(len(depths)):
        oracle.integrate(K, poses[f], depths[f], dims, origin, vs, trunc, t, w)
    return t, w


def _assert_same(got, ref):
    (gt, gw), (rt, rw) = got, ref
    assert np.array_equal(gw, rw), f"weights differ at {np.count_nonzero(gw != rw)} voxels"
    diff = np.abs(gt.astype(np.float64) - rt)
    assert diff.max() <= TOL_TSDF
    assert np.array_equal(gt.view(np.uint32), rt.view(np.uint32)), f"max |dTSDF| {diff.max():g}"


@pytest.mark.parametrize("fused", [False, True])
def test_synthetic_frames_bit_exact(torch_gpu, synth_frames, fused):
    depths, poses, K = synth_frames
    dims = (96, 80, 72)
    origin = np.array([-1.44, -1.2, 0.4], np.float32)
    vs, trunc = np.float32(0.03), np.float32(0.15)
    print(f"case synthetic 4 frames dims={dims} fused={fused}", flush=True)
    got = _gpu_fuse(torch_gpu, K, poses, depths, dims, origin, vs, trunc, fused)
    ref = _oracle_fuse(K, poses, depths, dims, origin, vs, trunc)
    assert ref[1].sum() > 1000  # the scene is actually seen
    _assert_same(got, ref)


@pytest.mark.parametrize("fused", [False, True])
def test_rounding_boundaries_bit_exact(torch_gpu, fused):
    """Voxels that project onto (near) half-pixel ties and voxels on both
    sides of the camera plane: 1 mm voxels seen head-on by fx = 500 put every
    other column on pixel coordinate k + 0.5 up to float rounding, so the
    reciprocal shortcut must hand them to the reference's division; a camera
    placed inside the grid gives tiles that straddle pc_z = 0."""
    dims = (256, 192, 64)
    origin = np.array([-0.32, -0.24, 1.0], np.float32)
    vs, trunc = np.float32(0.001), np.float32(0.005)
    K = np.array([[500, 0, 320], [0, 500, 240], [0, 0, 1]], np.float32)
    H, W = 480, 640
    poses = np.tile(np.eye(4, dtype=np.float32), (3, 1, 1))
    poses[1, 2, 3] = 1.03  # camera inside the grid: slices on both sides of it
    a = np.float32(0.002)  # a slight turn about y
    poses[2, :3, :3] = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]],
                                np.float32)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float32)
    depths = np.stack([
        np.full((H, W), 1.03, np.float32),
        np.float32(0.02) + np.float32(1e-5) * xx,
        np.float32(1.01) + np.float32(4e-5) * (xx + yy),
    ])
    depths[:, ::7, ::5] = 0.0  # dropouts
    print(f"case rounding boundaries dims={dims} fused={fused}", flush=True)
    got = _gpu_fuse(torch_gpu, K, poses, depths, dims, origin, vs, trunc, fused)
    ref = _oracle_fuse(K, poses, depths, dims, origin, vs, trunc)
    assert ref[1].sum() > 100000
    _assert_same(got, ref)


def test_full_size_grid_bit_exact(torch_gpu):
    """The benchmark's workload: all 50 frames into its 256^3 grid at 1 cm,
    per frame and fused."""
    from tdm import synth

    mm, c2w = synth.depth_frames(50, count=50)
    depths = synth.depth_to_metres(mm)
    poses = synth.relative_poses(c2w)
    K = synth.intrinsics()
    args = (synth.TSDF_DIMS, synth.TSDF_ORIGIN, synth.TSDF_VOXEL, synth.TSDF_TRUNC)
    print("case 256^3 x 50 frames", flush=True)
    ref = _oracle_fuse(K, poses, depths, *args)
    assert ref[1].sum() > 1.5e8
    for fused in (False, True):
        _assert_same(_gpu_fuse(torch_gpu, K, poses, depths, *args, fused), ref)


def test_sample_frames_bit_exact(torch_gpu):
    """The reference's own sample RGB-D frames (data/sample/depth-fusion-demo),
    5 of them committed as fixtures, fused with the demo's pose arithmetic."""
    d = os.path.join(GOLDEN, "sample_frames")
    K = tdm.load_matrix(os.path.join(d, "camera-intrinsics.txt"), 3, 3)
    files = sorted(glob.glob(os.path.join(d, "frame-*.depth.png")))
    assert len(files) >= 3
    base = tdm.load_matrix(files[0].replace(".depth.png", ".pose.txt"), 4, 4)
    base_inv = oracle.invert_matrix(base)
    poses, depths = [], []
    for f in files:
        c2w = tdm.load_matrix(f.replace(".depth.png", ".pose.txt"), 4, 4)
        poses.append(oracle.multiply_matrix(base_inv, c2w))
        depths.append(oracle.depth_from_mm(tdm.read_png_u16(f)))
    dims = (250, 250, 250)  # the demo's origin/truncation at 12 mm voxels
    origin = np.array([-1.5, -1.5, 0.5], np.float32)
    vs = np.float32(0.012)
    trunc = np.float32(vs * np.float32(5))
    print("case sample frames 250^3", flush=True)
    ref = _oracle_fuse(K, poses, depths, dims, origin, vs, trunc)
    for fused in (False, True):
        _assert_same(_gpu_fuse(torch_gpu, K, poses, depths, dims, origin, vs, trunc, fused), ref)


def test_odd_grid_shapes(torch_gpu, synth_frames):
    """dim_y > 1024 (beyond the reference's one-thread-per-y launch) and dims
    that are not multiples of the lane or z-run tiling."""
    depths, poses, K = synth_frames
    dims = (37, 1100, 19)
    origin = np.array([-0.5, -2.0, 1.0], np.float32)
    vs, trunc = np.float32(0.004), np.float32(0.02)
    print(f"case odd dims {dims}", flush=True)
    ref = _oracle_fuse(K, poses[:2], depths[:2], dims, origin, vs, trunc)
    for fused in (False, True):
        _assert_same(_gpu_fuse(torch_gpu, K, poses[:2], depths[:2], dims, origin, vs, trunc, fused), ref)


def test_invalid_depth_leaves_grid(torch_gpu, synth_frames):
    _, poses, K = synth_frames
    depths = np.zeros((2, 480, 640), np.float32)
    depths[1, :, :] = 7.0  # beyond the 6 m cut-off
    dims = (32, 32, 32)
    got_t, got_w = _gpu_fuse(torch_gpu, K, poses[:2], depths, dims, np.zeros(3, np.float32),
                             np.float32(0.05), np.float32(0.25), True)
    assert np.all(got_w == 0) and np.all(got_t == 1)


def test_plane_closed_form(torch_gpu):
    """Fronto-parallel wall at depth D seen by an identity camera: on the optical
    axis pc = (0, 0, z) exactly, so tsdf = min(1, (D - z)/trunc) and w = frames."""
    K = np.array([[500, 0, 32], [0, 500, 24], [0, 0, 1]], np.float32)
    pose = np.eye(4, dtype=np.float32)
    D = np.float32(2.0)
    depth = np.full((48, 64), D, np.float32)
    vs = np.float32(0.25)
    origin = np.array([-1.0, -1.0, 0.25], np.float32)  # voxel (4, 4, z) is on the axis
    dims = (9, 9, 12)
    trunc = np.float32(0.5)
    # two identical frames: (d*1 + d)/2 == d exactly, so the average is exact
    t, w = _gpu_fuse(torch_gpu, K, np.stack([pose] * 2), np.stack([depth] * 2), dims, origin, vs, trunc, True)
    for z in range(dims[2]):
        zc = np.float32(origin[2] + np.float32(z) * vs)
        diff = np.float32(D - zc)
        if diff <= -trunc:
            assert w[z, 4, 4] == 0
        else:
            assert w[z, 4, 4] == 2
            assert t[z, 4, 4] == np.float32(min(np.float32(1.0), np.float32(diff / trunc)))


def surface_gpu(torch, tsdf, weight, dims, origin, vs, capacity=None):
    dt = torch.from_numpy(np.ascontiguousarray(tsdf)).to("cuda:0")
    dw = torch.from_numpy(np.ascontiguousarray(weight)).to("cuda:0")
    n = tdm.surface_points(dt, dw, dims, origin, vs, None, 0)  # count only
    cap = n if capacity is None else capacity
    pts = torch.full((max(cap, 1), 3), -7.0, dtype=torch.float32, device="cuda:0")
    n2 = tdm.surface_points(dt, dw, dims, origin, vs, pts, cap)
    assert n2 == n
    return n, pts[:cap].cpu().numpy()


def test_surface_points_bit_exact(torch_gpu, synth_frames):
    """tsdf.ply's point set (depth-fusion/utils.hpp:14-56) from the GPU
    compaction equals the oracle's, bit for bit and in order; a smaller
    capacity gets the same leading points."""
    depths, poses, K = synth_frames
    dims = (160, 120, 136)
    origin = np.array([-0.8, -0.6, 0.9], np.float32)
    vs, trunc = np.float32(0.01), np.float32(0.05)
    t, w = _gpu_fuse(torch_gpu, K, poses, depths, dims, origin, vs, trunc, True)
    ref = oracle.surface_points(t, w, origin, vs)
    assert len(ref) > 1000
    n, pts = surface_gpu(torch_gpu, t, w, dims, origin, vs)
    assert n == len(ref)
    assert np.array_equal(pts.view(np.uint32), ref.view(np.uint32))
    n, pts = surface_gpu(torch_gpu, t, w, dims, origin, vs, capacity=len(ref) // 3)
    assert n == len(ref)
    assert np.array_equal(pts.view(np.uint32), ref[: len(ref) // 3].view(np.uint32))
    # nothing passes: empty set
    n, _ = surface_gpu(torch_gpu, np.ones_like(t), w, dims, origin, vs)
    assert n == 0


def test_integrate_scenes_vs_oracle(torch_gpu, synth_frames):
    """Three scenes (different poses/frames/origins) in one launch, each equal
    to the oracle's frame loop bit for bit."""
    torch = torch_gpu
    depths, poses, K = synth_frames
    dims = (72, 60, 56)
    vs, trunc = np.float32(0.01), np.float32(0.05)
    origins = np.array([[-0.4, -0.3, 1.0], [-0.3, -0.35, 1.2], [-0.5, -0.2, 0.9]], np.float32)
    order = [[0, 1, 2, 3], [3, 2, 1, 0], [1, 3, 0, 2]]
    D = np.stack([depths[o] for o in order])
    P = np.stack([poses[o] for o in order]).astype(np.float32)
    dev = "cuda:0"
    n = dims[0] * dims[1] * dims[2]
    t = torch.empty(3 * n, device=dev)
    w = torch.empty(3 * n, device=dev)
    tdm.tsdf_init(t, w)
    dK = torch.from_numpy(np.ascontiguousarray(K, np.float32)).to(dev)
    tdm.integrate_scenes(dK, torch.from_numpy(P.reshape(3, 4, 16)).to(dev),
                         torch.from_numpy(np.ascontiguousarray(D)).to(dev), dims,
                         torch.from_numpy(origins).to(dev), vs, trunc, t, w)
    torch.cuda.synchronize()
    for s in range(3):
        ref = _oracle_fuse(K, P[s], D[s], dims, origins[s], vs, trunc)
        assert ref[1].sum() > 1000
        got = (t[s * n:(s + 1) * n].cpu().numpy().reshape(ref[0].shape),
               w[s * n:(s + 1) * n].cpu().numpy().reshape(ref[0].shape))
        _assert_same(got, ref)


def test_fusion_pipeline_vs_oracle(torch_gpu, tmp_path):
    """Frames read from PNG by the overlapped pipeline give the oracle's
    grids bit for bit (the oracle decodes the same PNGs with PIL-checked
    code and runs the reference's frame loop)."""
    from tdm import synth

    torch = torch_gpu
    mm, c2w = synth.depth_frames(50, count=6)
    poses = synth.relative_poses(c2w)
    paths = []
    for i in range(len(mm)):
        p = str(tmp_path / f"frame-{i:06d}.depth.png")
        synth.write_png16(p, mm[i])
        paths.append(p)
    K = synth.intrinsics()
    dims = (96, 80, 72)
    origin = np.array([-0.5, -0.4, 1.0], np.float32)
    vs, trunc = np.float32(0.01), np.float32(0.05)
    fus = tdm.Fusion(K, 480, 640, dims, origin, vs, trunc, threads=3)
    for p, P in zip(paths, poses):
        fus.add_png(p, P)
    n = dims[0] * dims[1] * dims[2]
    t = torch.empty(n, device="cuda:0")
    w = torch.empty(n, device="cuda:0")
    fus.copy_grids(t, w)
    torch.cuda.synchronize()
    fus.close()
    depths = np.stack([oracle.depth_from_mm(tdm.read_png_u16(p)) for p in paths])
    assert np.array_equal(depths, synth.depth_to_metres(mm))
    ref = _oracle_fuse(K, poses, depths, dims, origin, vs, trunc)
    assert ref[1].sum() > 10000
    shape = (dims[2], dims[1], dims[0])
    _assert_same((t.cpu().numpy().reshape(shape), w.cpu().numpy().reshape(shape)), ref)


def test_fusion_create_failure_is_clean(torch_gpu):
    """A grid that cannot be allocated fails tdm_fusion_create with an error
    (the partly built pipeline is released) and a normal pipeline still works."""
    K = np.array([[585, 0, 320], [0, 585, 240], [0, 0, 1]], np.float32)
    origin = np.zeros(3, np.float32)
    print("case fusion create with an unallocatable grid", flush=True)
    with pytest.raises(tdm.TdmError):
        tdm.Fusion(K, 480, 640, (65536, 65536, 65536), origin, 0.01, 0.05, threads=1)
    fus = tdm.Fusion(K, 480, 640, (32, 32, 32), origin, 0.01, 0.05, threads=1)
    fus.wait()
    fus.close()


def test_large_image_pixel_indices(torch_gpu):
    """A 4100 x 4100 frame (more pixels than float32 counts exactly, 2^24):
    voxels that project past pixel 2^24 must read their own pixel.  Depth and
    dropouts change from pixel to pixel inside the truncation band, so a
    neighbouring pixel would show in the grid."""
    H = W = 4100
    K = np.array([[3000, 0, 2050], [0, 3000, 4050], [0, 0, 1]], np.float32)
    yy, xx = np.mgrid[0:H, 0:W]
    depth = (np.float32(0.955) + np.float32(0.002) * ((xx * 7 + yy * 13) % 7)).astype(np.float32)
    depth[(xx * 3 + yy) % 4 == 0] = 0.0
    depth = depth[None]
    poses = np.eye(4, dtype=np.float32)[None]
    dims = (256, 32, 40)
    origin = np.array([-0.064, 0.0, 0.95], np.float32)  # rows up to pixel row 4099
    vs, trunc = np.float32(0.0005), np.float32(0.01)
    print("case 4100x4100 frame", flush=True)
    got = _gpu_fuse(torch_gpu, K, poses, depth, dims, origin, vs, trunc, False)
    ref = _oracle_fuse(K, poses, depth, dims, origin, vs, trunc)
    assert ref[1].sum() > 10000
    _assert_same(got, ref)
