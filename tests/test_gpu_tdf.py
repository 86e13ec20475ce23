"""GPU parity of point cloud -> TDF (core/demo.cu:22-158) and the patch gather
(core/demo.cu:195-218) against the oracle.  Integer work (grid dims, voxel
coordinates, keypoint coordinates) and the TDF values are bit-exact."""
import numpy as np
import pytest

import oracle
import tdm
from tdm import synth

pytestmark = pytest.mark.gpu


def _dev(torch, a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda:0")


@pytest.fixture(scope="module")
def cloud():
    return synth.fragment(30_000, seed=11)


def test_grid_params_match(torch_gpu, cloud):
    for vs in (np.float32(0.01), np.float32(0.04), np.float32(0.013)):
        dims, origin = tdm.grid_params(_dev(torch_gpu, cloud), vs, 15)
        rdims, rorigin = oracle.grid_params(cloud, vs, 15)
        assert dims == rdims
        assert np.array_equal(origin.view(np.uint32), rorigin.view(np.uint32))


@pytest.mark.parametrize("decode", [tdm.DECODE_EXACT, tdm.DECODE_REFERENCE])
def test_tdf_from_points_bit_exact(torch_gpu, cloud, decode):
    vs = np.float32(0.04)
    trunc = np.float32(vs * np.float32(5))
    pts = _dev(torch_gpu, cloud)
    dims, origin = tdm.grid_params(pts, vs, 15)
    print(f"case tdf_from_points dims={dims} decode={decode}", flush=True)
    tdf = torch_gpu.empty(dims[2], dims[1], dims[0], dtype=torch_gpu.float32, device="cuda:0")
    tdm.tdf_from_points(pts, dims, origin, vs, trunc, tdf, decode)
    occ = oracle.occupancy(cloud, dims, origin, vs)
    ref = oracle.compute_tdf(occ, vs, trunc, "reference" if decode else "exact")
    got = tdf.cpu().numpy()
    assert np.count_nonzero(got == 1.0) == np.count_nonzero(occ)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), \
        f"{np.count_nonzero(got != ref)} voxels differ"


def test_occupancy_and_compute_tdf(torch_gpu, cloud):
    vs = np.float32(0.05)
    trunc = np.float32(vs * np.float32(5))
    dims, origin = oracle.grid_params(cloud, vs, 15)
    occ_d = torch_gpu.empty(dims[2], dims[1], dims[0], dtype=torch_gpu.float32, device="cuda:0")
    tdm.occupancy(_dev(torch_gpu, cloud), dims, origin, vs, occ_d)
    occ = oracle.occupancy(cloud, dims, origin, vs)
    assert np.array_equal(occ_d.cpu().numpy(), occ)
    tdf = torch_gpu.empty_like(occ_d)
    tdm.compute_tdf(occ_d, tdf, dims, vs, trunc)
    ref = oracle.compute_tdf(occ, vs, trunc)
    assert np.array_equal(tdf.cpu().numpy().view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("r", [1, 2, 3, 4, 6, 8])
def test_search_radius_variants(torch_gpu, r):
    rng = np.random.default_rng(100 + r)
    dims = (131, 45, 38)  # x spans several 64-voxel tiles and a partial word
    occ = (rng.random((dims[2], dims[1], dims[0])) < 0.002).astype(np.float32)
    vs = np.float32(0.01)
    trunc = np.float32(vs * np.float32(r))
    assert int(round(float(trunc / vs))) == r
    occ_d = _dev(torch_gpu, occ)
    tdf = torch_gpu.empty_like(occ_d)
    tdm.compute_tdf(occ_d, tdf, dims, vs, trunc)
    ref = oracle.compute_tdf(occ, vs, trunc)
    assert np.array_equal(tdf.cpu().numpy().view(np.uint32), ref.view(np.uint32))


def test_empty_and_single_point(torch_gpu):
    dims = (40, 30, 20)
    occ = np.zeros((20, 30, 40), np.float32)
    vs, trunc = np.float32(0.01), np.float32(0.05)
    occ_d = _dev(torch_gpu, occ)
    tdf = torch_gpu.full_like(occ_d, 7.0)
    tdm.compute_tdf(occ_d, tdf, dims, vs, trunc)
    assert np.all(tdf.cpu().numpy() == 0)
    occ[0, 0, 0] = 1  # a corner point: the clipped cube
    occ[19, 29, 39] = 1
    tdm.compute_tdf(_dev(torch_gpu, occ), tdf, dims, vs, trunc)
    assert np.array_equal(tdf.cpu().numpy(), oracle.compute_tdf(occ, vs, trunc))


def test_tdf_from_points_edges(torch_gpu):
    """Clustered points (thousands per bit word, more than a scatter
    workgroup's 4096), points outside the grid, an empty cloud and an output
    grid that starts 4 bytes past a 16-byte boundary, each on a grid that
    first held a larger cloud's result; every voxel is compared."""
    rng = np.random.default_rng(7)
    vs, trunc = np.float32(0.01), np.float32(0.05)
    dims, origin = (133, 37, 29), np.array([-0.3, -0.1, 0.2], np.float32)
    n = dims[0] * dims[1] * dims[2]
    lo = origin + np.float32(3) * vs
    span = np.array([0.9, 0.2, 0.15], np.float32)
    sparse = (lo + rng.random((9000, 3)) * span).astype(np.float32)
    cluster = (lo + np.float32(0.5) * span + rng.random((20000, 3)) * np.float32(0.05)).astype(np.float32)
    outside = np.array([[5.0, 5.0, 5.0], [-5.0, 0.0, 0.3], [0.0, 0.0, 9.0]], np.float32)
    cases = {
        "mixed": np.concatenate([cluster[:7000], sparse, outside, cluster[7000:]]),
        "cluster": cluster,
        "outside": outside,
        "empty": np.zeros((0, 3), np.float32),
    }
    backing = torch_gpu.empty(n + 1, dtype=torch_gpu.float32, device="cuda:0")
    tdf = backing[1:]  # 4 bytes past the allocation's 256-byte alignment
    for name, pts in cases.items():
        print(f"case tdf_from_points edges: {name} ({len(pts)} points)", flush=True)
        backing.fill_(7.0)
        tdm.tdf_from_points(_dev(torch_gpu, pts), dims, origin, vs, trunc, tdf)
        got = tdf.cpu().numpy().reshape(dims[2], dims[1], dims[0])
        occ = oracle.occupancy(pts, dims, origin, vs)
        ref = oracle.compute_tdf(occ, vs, trunc)
        assert backing[0].item() == 7.0, "write before the grid"
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), \
            f"{name}: {np.count_nonzero(got != ref)} voxels differ"


def test_compute_tdf_in_place(torch_gpu):
    """occ and tdf may be one buffer: the bits are packed before the output
    is written."""
    rng = np.random.default_rng(3)
    dims = (100, 40, 33)
    occ = (rng.random((dims[2], dims[1], dims[0])) < 0.003).astype(np.float32)
    vs, trunc = np.float32(0.01), np.float32(0.05)
    buf = _dev(torch_gpu, occ)
    tdm.compute_tdf(buf, buf, dims, vs, trunc)
    ref = oracle.compute_tdf(occ, vs, trunc)
    assert np.array_equal(buf.cpu().numpy().view(np.uint32), ref.view(np.uint32))


def test_reference_decode_beyond_2_24(torch_gpu):
    """On a grid of more than 2^24 voxels the reference's float32 decode of
    voxel_idx (core/demo.cu:32-34) is wrong for part of the voxels; the
    reference decode mode reproduces its output there."""
    dims = (300, 300, 190)  # 17.1 M voxels
    n = dims[0] * dims[1] * dims[2]
    rng = np.random.default_rng(7)
    occ = (rng.random((dims[2], dims[1], dims[0])) < 0.0005).astype(np.float32)
    vs, trunc = np.float32(0.01), np.float32(0.05)
    occ_d = _dev(torch_gpu, occ)
    ref_mode = torch_gpu.empty_like(occ_d)
    exact_mode = torch_gpu.empty_like(occ_d)
    print("case refdecode 300x300x190", flush=True)
    tdm.compute_tdf(occ_d, ref_mode, dims, vs, trunc, tdm.DECODE_REFERENCE)
    tdm.compute_tdf(occ_d, exact_mode, dims, vs, trunc, tdm.DECODE_EXACT)
    got_ref = ref_mode.cpu().numpy().reshape(-1)
    got_exact = exact_mode.cpu().numpy().reshape(-1)
    x, y, z = oracle.float_decode_all(n, dims[0], dims[1])
    v = np.arange(n)
    bad = (x != v % dims[0]) | (y != (v // dims[0]) % dims[1]) | (z != v // (dims[0] * dims[1]))
    assert bad.sum() > 0
    # where the decode is right both modes agree
    assert np.array_equal(got_ref[~bad], got_exact[~bad])
    # every mis-decoded voxel plus a random sample, against the oracle
    idx = np.concatenate([np.nonzero(bad)[0], rng.choice(n, 20000, replace=False)])
    ref = oracle.compute_tdf_ref_at(occ, idx, vs, trunc)
    assert np.array_equal(got_ref[idx].view(np.uint32), ref.view(np.uint32))


def test_extract_patches_bit_exact(torch_gpu, cloud):
    vs = np.float32(0.02)
    trunc = np.float32(vs * np.float32(5))
    dims, origin = oracle.grid_params(cloud, vs, 15)
    occ = oracle.occupancy(cloud, dims, origin, vs)
    tdf = oracle.compute_tdf(occ, vs, trunc)
    idx = synth.keypoint_indices(len(cloud), 64, seed=12)
    kp = cloud[idx]
    kg = tdm.keypoints_to_grid(kp, origin, vs)
    assert np.array_equal(kg, oracle.keypoint_grid(kp, origin, vs))
    # keypoints near and beyond the border read zeros outside the grid
    kg = np.concatenate([kg, np.array([[0, 0, 0], [dims[0] - 1, dims[1] - 1, dims[2] - 1], [14, 15, 16]], np.int32)])
    out = torch_gpu.empty(len(kg), 30, 30, 30, dtype=torch_gpu.float32, device="cuda:0")
    tdm.extract_patches(_dev(torch_gpu, tdf), dims, _dev(torch_gpu, kg), out)
    assert np.array_equal(out.cpu().numpy(), oracle.extract_patches(tdf, kg))


def test_accurate_tdf_bit_exact(torch_gpu):
    """tdm_tdf_accurate equals the oracle's voxel-centre nearest-point TDF
    bit for bit (float32, the reference's summation order, correctly rounded
    sqrt) on a synthetic fragment grid."""
    torch = torch_gpu
    cloud = synth.fragment(8_000, view=3, seed=77)
    vs, tr = np.float32(0.01), np.float32(0.05)
    full, origin = oracle.grid_params(cloud, vs, 15)
    # a 90^3 sub-block around the cloud's median keeps the oracle fast; it
    # cuts through the cloud, so points just outside it are the nearest points
    # of its border voxels; shift the origin by whole voxels
    med = np.median(cloud, axis=0)
    off = np.clip(np.round((med - origin) / vs).astype(int) - 45, 0, None)
    origin = (origin + off.astype(np.float32) * vs).astype(np.float32)
    dims = [min(90, f - o) for f, o in zip(full, off)]
    want = oracle.tdf_accurate(cloud, dims, origin, vs, tr)
    pts = torch.from_numpy(cloud).to("cuda:0")
    got = torch.full((dims[2], dims[1], dims[0]), -1.0, device="cuda:0")
    tdm.tdf_accurate(pts, dims, origin, vs, tr, got)
    torch.cuda.synchronize()
    g = got.cpu().numpy()
    assert np.count_nonzero(want) > 1000
    assert np.array_equal(g.view(np.uint32), want.view(np.uint32)), np.abs(g - want).max()
    # the cut matters: dropping the outside points changes border voxels
    gi = np.round((cloud - origin) / vs)
    inside = np.all((gi >= 0) & (gi < np.array(dims)), axis=1)
    assert 0 < inside.sum() < len(cloud)
    assert not np.array_equal(oracle.tdf_accurate(cloud[inside], dims, origin, vs, tr), want)


@pytest.mark.parametrize("trunc_vox", [5, 12, 30])
def test_accurate_tdf_dense_and_wide_reach(torch_gpu, trunc_vox):
    """Dense clusters put thousands of candidates on a tile (several work
    items of 1024 candidates, merged in the grid) and wide truncations walk
    5^3 and 9^3 bins: still the oracle's values bit for bit, and voxels out of
    reach read 0."""
    torch = torch_gpu
    rng = np.random.default_rng(100 + trunc_vox)
    vs = np.float32(0.01)
    tr = np.float32(vs * np.float32(trunc_vox))
    wide = trunc_vox > 20  # keep the brute-force oracle to seconds
    dims = [26, 22, 20] if wide else [44, 38, 36]
    origin = np.array([0.1, -0.2, 0.3], np.float32)
    dense = origin + (np.array([10, 12, 9]) + rng.random((1500 if wide else 6000, 3)) * 6
                      ).astype(np.float32) * vs
    sparse = origin + (rng.random((400 if wide else 1500, 3)) * np.array([60, 50, 50]) - 8
                       ).astype(np.float32) * vs
    cloud = np.concatenate([dense, sparse]).astype(np.float32)
    print(f"case accurate dense trunc={trunc_vox} voxels", flush=True)
    want = oracle.tdf_accurate(cloud, dims, origin, vs, tr)
    got = torch.full((dims[2], dims[1], dims[0]), -1.0, device="cuda:0")
    tdm.tdf_accurate(torch.from_numpy(cloud).to("cuda:0"), dims, origin, vs, tr, got)
    torch.cuda.synchronize()
    g = got.cpu().numpy()
    assert np.array_equal(g.view(np.uint32), want.view(np.uint32)), np.abs(g - want).max()
    if trunc_vox == 5:
        assert np.count_nonzero(want == 0) > 0 and np.count_nonzero(want) > 1000
