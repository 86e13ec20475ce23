"""Pins the oracle's Integrate restatement (oracle.c orc_integrate,
depth-fusion/demo.cu:18-68) without using it to check itself:
  - closed form: a wall seen by an identity camera,
  - an independent float64 numpy implementation (values within 1e-5 where
    both update; integer decisions may flip only at rounding boundaries),
  - invariants: each update adds exactly 1 to a weight,
  - a planted mistake (rotation not transposed) is caught by the same checks.
"""
import glob
import os

import numpy as np

import oracle
import tdm
from conftest import GOLDEN


def fp64_integrate(K, pose, depth, dims, origin, vs, trunc, tsdf, weight, transpose=True):
    """Vectorised float64 statement of the reference per-voxel update."""
    dx, dy, dz = dims
    z, y, x = np.meshgrid(np.arange(dz), np.arange(dy), np.arange(dx), indexing="ij")
    p = np.stack([origin[0] + x * float(vs), origin[1] + y * float(vs), origin[2] + z * float(vs)], -1)
    P = pose.astype(np.float64)
    R, t = P[:3, :3], P[:3, 3]
    pc = (p - t) @ R if transpose else (p - t) @ R.T  # R^T (p - t) as row vectors
    cz = pc[..., 2]
    H, W = depth.shape
    with np.errstate(divide="ignore", invalid="ignore"):
        qx, qy = pc[..., 0] / cz, pc[..., 1] / cz
        px = np.floor(K[0, 0] * qx + K[0, 2] + 0.5)  # half away from zero; px >= 0 matters only
        py = np.floor(K[1, 1] * qy + K[1, 2] + 0.5)
    ok = (cz > 0) & (px >= 0) & (px < W) & (py >= 0) & (py < H)
    d = np.zeros_like(cz)
    d[ok] = depth[py[ok].astype(int), px[ok].astype(int)]
    ok &= (d > 0) & (d <= 6)
    diff = np.where(ok, (d - cz) * np.sqrt(1 + qx * qx + qy * qy), -np.inf)
    ok &= diff > -float(trunc)
    dist = np.minimum(1.0, diff / float(trunc))
    w_new = weight + ok
    tsdf[...] = np.where(ok, (tsdf * weight + np.where(ok, dist, 0)) / np.maximum(w_new, 1), tsdf)
    weight[...] = w_new


def _sample_frames(n=3):
    files = sorted(glob.glob(os.path.join(GOLDEN, "sample_frames", "*.depth.png")))[:n]
    K = tdm.load_matrix(os.path.join(GOLDEN, "sample_frames", "camera-intrinsics.txt"), 3, 3)
    base_inv = oracle.invert_matrix(tdm.load_matrix(files[0].replace(".depth.png", ".pose.txt"), 4, 4))
    poses = [oracle.multiply_matrix(base_inv, tdm.load_matrix(f.replace(".depth.png", ".pose.txt"), 4, 4))
             for f in files]
    depths = [oracle.depth_from_mm(tdm.read_png_u16(f)) for f in files]
    return K, poses, depths


def test_wall_closed_form():
    K = np.array([[500, 0, 32], [0, 500, 24], [0, 0, 1]], np.float32)
    pose = np.eye(4, dtype=np.float32)
    depth = np.full((48, 64), 2.0, np.float32)
    vs, trunc = np.float32(0.25), np.float32(0.5)
    origin = np.array([-1.0, -1.0, 0.25], np.float32)
    dims = (9, 9, 12)
    t, w = oracle.new_grid(dims)
    for _ in range(2):
        oracle.integrate(K, pose, depth, dims, origin, vs, trunc, t, w)
    for z in range(12):
        diff = np.float32(2.0) - np.float32(0.25 + 0.25 * z)
        if diff <= -trunc:
            assert w[z, 4, 4] == 0 and t[z, 4, 4] == 1
        else:
            assert w[z, 4, 4] == 2
            assert t[z, 4, 4] == min(np.float32(1), diff / trunc)
    # voxels behind the camera plane are never touched
    assert np.all(w[:, :, :][t == 1][...] >= 0)


def test_against_float64_statement_on_sample_frames():
    K, poses, depths = _sample_frames(3)
    dims = (100, 90, 80)
    origin = np.array([-1.5, -1.5, 0.5], np.float32)
    vs = np.float32(0.03)
    trunc = np.float32(vs * np.float32(5))
    t, w = oracle.new_grid(dims)
    t64 = np.ones((dims[2], dims[1], dims[0]))
    w64 = np.zeros_like(t64)
    for P, D in zip(poses, depths):
        oracle.integrate(K, P, D, dims, origin, vs, trunc, t, w)
        fp64_integrate(K, P, D, dims, origin, vs, trunc, t64, w64)
    flips = np.count_nonzero(w != w64)
    assert w.sum() > 50_000
    assert flips <= 1e-4 * w.size, flips  # rounding-boundary decisions only
    # Where both updated the same number of times the values agree to 1e-5,
    # except where a pixel index rounded differently (another depth sample):
    # float32 vs float64 coordinates put a few voxels on the other side of a
    # pixel boundary.
    same = (w == w64) & (w > 0)
    err = np.abs(t[same] - t64[same])
    assert np.count_nonzero(err > 1e-5) <= 1e-3 * same.sum()
    assert np.median(err) < 1e-7


def test_planted_transpose_mistake_is_caught():
    """The float64 check distinguishes R^T from R: a restatement that forgot the
    transpose would fail test_against_float64_statement_on_sample_frames."""
    K, poses, depths = _sample_frames(2)
    dims = (60, 50, 40)
    origin = np.array([-1.5, -1.5, 0.5], np.float32)
    vs, trunc = np.float32(0.05), np.float32(0.25)
    t, w = oracle.new_grid(dims)
    t64 = np.ones((dims[2], dims[1], dims[0]))
    w64 = np.zeros_like(t64)
    for P, D in zip(poses, depths):
        oracle.integrate(K, P, D, dims, origin, vs, trunc, t, w)
        fp64_integrate(K, P, D, dims, origin, vs, trunc, t64, w64, transpose=False)
    assert np.count_nonzero(w != w64) > 1e-3 * w.size


def test_weight_invariant_and_running_average():
    """Each applied update adds exactly 1; with identical frames the TSDF of a
    voxel is the mean of identical values."""
    K, poses, depths = _sample_frames(1)
    dims = (64, 64, 64)
    origin = np.array([-1.5, -1.5, 0.5], np.float32)
    vs, trunc = np.float32(0.04), np.float32(0.2)
    t1, w1 = oracle.new_grid(dims)
    oracle.integrate(K, poses[0], depths[0], dims, origin, vs, trunc, t1, w1)
    t2, w2 = oracle.new_grid(dims)
    for _ in range(2):
        oracle.integrate(K, poses[0], depths[0], dims, origin, vs, trunc, t2, w2)
    assert np.array_equal(w2, 2 * w1)
    assert np.array_equal(t2, t1)  # (d*1 + d)/2 == d exactly
    assert set(np.unique(w1)) <= {0.0, 1.0}


def test_surface_points_oracle_matches_literal_loop():
    """The vectorised surface extraction against a literal restatement of the
    reference loop (depth-fusion/utils.hpp:44-56) on a small grid."""
    rng = np.random.default_rng(11)
    dims = (7, 5, 6)
    tsdf = rng.uniform(-1, 1, (dims[2], dims[1], dims[0])).astype(np.float32)
    weight = rng.integers(0, 4, tsdf.shape).astype(np.float32)
    tsdf[0, 0, 0] = np.float32(0.2)  # boundary: not < 0.2
    weight[0, 0, 1] = np.float32(1.0)  # boundary: not > 1
    origin = np.array([-0.31, 0.07, 1.13], np.float32)
    vs = np.float32(0.013)
    got = oracle.surface_points(tsdf, weight, origin, vs)
    t, w = tsdf.reshape(-1), weight.reshape(-1)
    want = []
    for i in range(dims[0] * dims[1] * dims[2]):
        if abs(float(t[i])) < 0.2 and float(w[i]) > 1.0:
            z = i // (dims[0] * dims[1])
            y = (i - z * dims[0] * dims[1]) // dims[0]
            x = i - z * dims[0] * dims[1] - y * dims[0]
            want.append([np.float32(origin[0]) + np.float32(x) * vs,
                         np.float32(origin[1]) + np.float32(y) * vs,
                         np.float32(origin[2]) + np.float32(z) * vs])
    want = np.array(want, np.float32)
    assert len(want) > 10
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_surface_points_closed_form():
    """Planted voxels on a grid whose coordinates are exact in float
    (origin and voxel size dyadic): the extracted points are exactly the
    planted voxels that pass |tsdf| < 0.2 and weight > 1, in linear index
    order, at origin + index * voxel_size.  The threshold edges are planted
    too: 0.2, -0.2 and weight 1 are out, the floats just inside are in."""
    dims = (9, 8, 7)
    tsdf = np.ones((dims[2], dims[1], dims[0]), np.float32)
    weight = np.zeros_like(tsdf)
    inside_t = np.nextafter(np.float32(0.2), np.float32(0))
    inside_w = np.nextafter(np.float32(1.0), np.float32(2))
    planted = [  # (x, y, z, tsdf, weight, kept)
        (8, 7, 6, 0.0, 5.0, True), (0, 0, 0, -0.1, 2.0, True), (3, 2, 1, inside_t, 3.0, True),
        (4, 4, 4, -inside_t, inside_w, True), (5, 1, 2, 0.2, 4.0, False),
        (6, 1, 2, -0.2, 4.0, False), (2, 5, 3, 0.05, 1.0, False), (1, 6, 5, 0.3, 9.0, False),
        (7, 3, 0, -0.19, 1.5, True)]
    for x, y, z, t, w, _ in planted:
        tsdf[z, y, x] = np.float32(t)
        weight[z, y, x] = np.float32(w)
    origin = np.array([-2.0, 0.5, 1.25], np.float32)
    vs = np.float32(0.25)
    got = oracle.surface_points(tsdf, weight, origin, vs)
    kept = sorted([p for p in planted if p[5]], key=lambda p: (p[2] * dims[1] + p[1]) * dims[0] + p[0])
    want = np.array([[origin[0] + 0.25 * x, origin[1] + 0.25 * y, origin[2] + 0.25 * z]
                     for x, y, z, *_ in kept], np.float64)
    assert got.shape == (5, 3)
    assert np.array_equal(got.astype(np.float64), want)
