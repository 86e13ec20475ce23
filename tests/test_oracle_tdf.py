"""Pins the oracle's TDF pipeline (oracle.c, core/demo.cu:22-57, 97-135,
164-218) against library routines and brute force:
  - exact-decode TDF == table[EDT^2] with scipy.ndimage.distance_transform_edt,
  - pure-Python brute force of the 11^3 search on a tiny grid,
  - the float32 decode against an independent numpy float32 emulation, with
    the mis-decode counts (0 at 256^3, 5 111 392 at 300^3),
  - occupancy / keypoint coordinates against numpy float32 rounding,
  - patch gather against numpy slicing with zero padding."""
import numpy as np
import pytest
from scipy import ndimage

import oracle
from tdm import synth


def lut(r):
    d2 = np.arange(3 * 8 * 8 + 1, dtype=np.float32)
    v = np.float32(1) - np.sqrt(d2) / np.float32(r)
    return np.where(v > 0, v, np.float32(0)).astype(np.float32)


def edt_tdf(occ, r):
    d = ndimage.distance_transform_edt(occ == 0)
    d2 = np.rint(d * d).astype(np.int64)
    out = np.zeros(occ.shape, np.float32)
    near = d2 < r * r
    out[near] = lut(r)[d2[near]]
    return out


@pytest.mark.parametrize("r", [2, 5])
def test_exact_tdf_equals_truncated_edt(r):
    rng = np.random.default_rng(r)
    occ = (rng.random((40, 50, 70)) < 0.003).astype(np.float32)
    vs = np.float32(0.01)
    trunc = np.float32(vs * np.float32(r))
    got = oracle.compute_tdf(occ, vs, trunc)
    assert np.array_equal(got, edt_tdf(occ, r))
    assert np.array_equal(got[occ > 0], np.ones(int(occ.sum()), np.float32))


def test_cloud_tdf_equals_truncated_edt():
    cloud = synth.fragment(20_000, seed=5)
    vs = np.float32(0.04)
    dims, origin = oracle.grid_params(cloud, vs, 15)
    occ = oracle.occupancy(cloud, dims, origin, vs)
    assert np.array_equal(oracle.compute_tdf(occ, vs, np.float32(0.2)), edt_tdf(occ, 5))


def test_brute_force_tiny_grid():
    rng = np.random.default_rng(9)
    occ = (rng.random((6, 7, 9)) < 0.05).astype(np.float32)
    r = 3
    dz, dy, dx = occ.shape
    want = np.zeros_like(occ)
    pts = np.argwhere(occ > 0)
    for z in range(dz):
        for y in range(dy):
            for x in range(dx):
                if occ[z, y, x] > 0:
                    want[z, y, x] = 1
                    continue
                best = np.float32(0)
                for (iz, iy, ix) in pts:
                    if abs(iz - z) <= r and abs(iy - y) <= r and abs(ix - x) <= r:
                        v = np.float32(1) - np.sqrt(np.float32((x - ix) ** 2 + (y - iy) ** 2 + (z - iz) ** 2)) / np.float32(r)
                        best = max(best, v)
                want[z, y, x] = best
    got = oracle.compute_tdf(occ, np.float32(0.01), np.float32(0.03))
    assert np.array_equal(got, want)
    assert np.array_equal(oracle.compute_tdf(occ, np.float32(0.01), np.float32(0.03), "reference"), want)


def test_float_decode_counts():
    for dims, bad_want in (((256, 256, 256), 0), ((300, 300, 300), 5_111_392)):
        n = dims[0] * dims[1] * dims[2]
        x, y, z = oracle.float_decode_all(n, dims[0], dims[1])
        v = np.arange(n)
        bad = (x != v % dims[0]) | (y != (v // dims[0]) % dims[1]) | (z != v // (dims[0] * dims[1]))
        assert int(bad.sum()) == bad_want
        # the C restatement agrees with the numpy emulation on a sample
        idx = np.random.default_rng(0).choice(n, 2000, replace=False)
        if bad_want:
            idx = np.concatenate([idx, np.nonzero(bad)[0][:2000]])
        for i in idx:
            assert oracle.float_decode(int(i), dims[0], dims[1]) == (x[i], y[i], z[i])


def test_float_decode_exact_up_to_2_24_voxels():
    """The batched plan skips the reference-decode pass on grids of at most
    2^24 voxels: there the float32 decode of every index equals the integer
    one.  Shapes at the limit with one huge, one flat and one cubic slice,
    and a grid just above the limit where the decode does move voxels."""
    for dims in ((4095, 4097, 1), (1, 1, 1 << 24), (255, 257, 256), (1024, 1024, 16)):
        n = dims[0] * dims[1] * dims[2]
        assert n <= 1 << 24
        x, y, z = oracle.float_decode_all(n, dims[0], dims[1])
        v = np.arange(n)
        assert np.array_equal(x, v % dims[0]), dims
        assert np.array_equal(y, (v // dims[0]) % dims[1]), dims
        assert np.array_equal(z, v // (dims[0] * dims[1])), dims
    dims = (419, 321, 125)  # 16.8 M voxels, just above 2^24
    n = dims[0] * dims[1] * dims[2]
    x, y, z = oracle.float_decode_all(n, dims[0], dims[1])
    v = np.arange(n)
    assert np.any((x != v % dims[0]) | (z != v // (dims[0] * dims[1])))


def test_occupancy_and_keypoint_rounding():
    cloud = synth.fragment(5000, seed=6)
    vs = np.float32(0.01)
    dims, origin = oracle.grid_params(cloud, vs, 15)
    lo = cloud.min(0)
    hi = cloud.max(0)
    want_dims = [int(np.round(np.float32(hi[a] - lo[a]) / vs)) + 1 + 30 for a in range(3)]
    assert dims == want_dims
    assert np.array_equal(origin, (lo - np.float32(15) * vs + vs / np.float32(2)).astype(np.float32))
    # round() of the reference is half away from zero; coordinates here are >= 0
    q = (cloud - origin) / vs
    g = np.where(q - np.floor(q) == np.float32(0.5), np.floor(q) + 1, np.rint(q)).astype(np.int64)
    occ = oracle.occupancy(cloud, dims, origin, vs)
    want = np.zeros_like(occ)
    want[g[:, 2], g[:, 1], g[:, 0]] = 1
    assert np.array_equal(occ, want)
    assert np.array_equal(oracle.keypoint_grid(cloud[:100], origin, vs), g[:100])


def test_patch_gather_zero_padded():
    rng = np.random.default_rng(8)
    tdf = rng.random((40, 35, 50), dtype=np.float32)
    kg = np.array([[25, 17, 20], [0, 0, 0], [49, 34, 39], [14, 15, 16]], np.int32)
    got = oracle.extract_patches(tdf, kg)
    pad = np.pad(tdf, 30)
    for k, p in zip(kg, got):
        x, y, z = k + 30
        assert np.array_equal(p, pad[z - 15:z + 15, y - 15:y + 15, x - 15:x + 15])


def test_accurate_tdf_oracle_vs_brute_force():
    """The KD-tree-assisted accurate TDF oracle against a plain brute force
    over every point (float32, the reference's summation order) on a tiny
    grid, and basic closed forms: a voxel holding a point at its centre reads
    1, voxels farther than the margin from every point read 0."""
    rng = np.random.default_rng(12)
    dims = (14, 11, 9)
    origin = np.array([0.1, -0.2, 0.3], np.float32)
    vs, tr = np.float32(0.01), np.float32(0.05)
    # points up to 4 voxels beyond the grid on every side count too
    lo, hi = -4.5, np.array(dims) + 3.5
    pts = (origin + rng.uniform(lo, hi, (60, 3)) * vs).astype(np.float32)
    pts[0] = origin + np.float32(3) * vs  # exactly on the centre of voxel (3, 3, 3)
    pts[1] = origin + np.array([-2.2, 5, 4], np.float32) * vs  # outside, 2.2 voxels from x = 0
    got = oracle.tdf_accurate(pts, dims, origin, vs, tr)
    z, y, x = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    c = np.stack([origin[0] + x.astype(np.float32) * vs, origin[1] + y.astype(np.float32) * vs,
                  origin[2] + z.astype(np.float32) * vs], -1).reshape(-1, 3)
    e = c[:, None, :] - pts[None, :, :]
    d2 = (e[..., 0] * e[..., 0] + e[..., 1] * e[..., 1]) + e[..., 2] * e[..., 2]
    d = np.sqrt(d2.min(1))
    want = (np.float32(1) - np.minimum(d, tr) / tr).reshape(dims[2], dims[1], dims[0])
    assert np.array_equal(got, want)
    assert got[3, 3, 3] == 1.0
    # voxel (0, 5, 4) is 2.2 voxels from the outside point: 1 - 2.2/5 or more
    assert got[4, 5, 0] >= np.float32(1) - np.float32(0.022) / tr - np.float32(1e-6)
    g = np.round((pts - origin) / vs)
    outside = ~np.all((g >= 0) & (g < np.array(dims)), axis=1)
    assert outside.sum() > 10
    far = d > tr
    assert np.all(got.reshape(-1)[far] == 0.0)
