"""File formats and host pose arithmetic of the C ABI (CPU only), against
hand-built byte layouts, PIL, numpy and the oracle."""
import glob
import os
import struct

import numpy as np
import pytest

import oracle
import tdm
from conftest import GOLDEN, REFERENCE

FRAMES = os.path.join(GOLDEN, "sample_frames")


def test_ply_layout_and_round_trip(tmp_path):
    p = str(tmp_path / "c.ply")
    pts = np.random.default_rng(0).random((17, 3), dtype=np.float32)
    tdm.write_ply_points(p, pts)
    raw = open(p, "rb").read()
    header = (b"ply\nformat binary_little_endian 1.0\nelement vertex 17\nproperty float x\n"
              b"property float y\nproperty float z\nend_header\n")
    assert raw == header + pts.tobytes()
    assert np.array_equal(tdm.read_ply_points(p), pts)


def test_ply_bad_header(tmp_path):
    p = str(tmp_path / "bad.ply")
    open(p, "wb").write(b"ply\nformat\nelement vertex zero\n\n\n\n\n")
    with pytest.raises(tdm.TdmError) as e:
        tdm.read_ply_points(p)
    assert e.value.status == tdm.ERR_FORMAT
    open(p, "wb").write(b"ply\nformat\nelement vertex 5\n\n\n\n\n" + b"\0" * 12)
    with pytest.raises(tdm.TdmError):
        tdm.read_ply_points(p)  # truncated


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="reference tree not mounted")
def test_sample_clouds_and_grid_dims():
    """The reference's sample clouds (data/sample/3dmatch-demo) and the grid
    sizes core/demo.cu:116-118 gives them."""
    for name, n, dims in (("single-depth-1.ply", 252179, [370, 253, 324]),
                          ("single-depth-2.ply", 284874, [299, 255, 263])):
        pts = tdm.read_ply_points(os.path.join(REFERENCE, "data", "sample", "3dmatch-demo", name))
        assert len(pts) == n
        d, _ = oracle.grid_params(pts, np.float32(0.01), 15)
        assert d == dims


def test_sample_cloud_fixture():
    """The committed copy of data/sample/3dmatch-demo/single-depth-2.ply reads
    as the reference's demo would (7 header lines, N on line 3)."""
    pts = tdm.read_ply_points(os.path.join(GOLDEN, "sample_clouds", "single-depth-2.ply"))
    assert pts.shape == (284874, 3) and pts.dtype == np.float32
    d, _ = oracle.grid_params(pts, np.float32(0.01), 15)
    assert d == [299, 255, 263]


def test_keypts_desc_feat_grid_layouts(tmp_path):
    rng = np.random.default_rng(1)
    kp = rng.random((5, 3), dtype=np.float32)
    p = str(tmp_path / "a.keypts.bin")
    tdm.write_keypts_bin(p, kp)
    assert open(p, "rb").read() == struct.pack("<f", 5.0) + kp.tobytes()
    assert np.array_equal(tdm.read_keypts_bin(p), kp)

    desc = rng.random((4, 512), dtype=np.float32)
    p = str(tmp_path / "a.desc.3dmatch.bin")
    tdm.write_desc_bin(p, desc)
    assert open(p, "rb").read() == struct.pack("<ff", 4.0, 512.0) + desc.tobytes()
    assert np.array_equal(tdm.read_desc_bin(p), desc)

    p = str(tmp_path / "feat.bin")
    tdm.write_feat_bin(p, desc)
    assert open(p, "rb").read() == struct.pack("<5f", 4, 512, 1, 1, 1) + desc.tobytes()

    grid = rng.random((3, 4, 5), dtype=np.float32)  # dz, dy, dx
    p = str(tmp_path / "tsdf.bin")
    tdm.write_grid_bin(p, grid, [-1.5, -1.5, 0.5], 0.006, 0.03)
    raw = open(p, "rb").read()
    assert raw[:32] == struct.pack("<8f", 5, 4, 3, -1.5, -1.5, 0.5, 0.006, 0.03)
    assert raw[32:] == grid.tobytes()
    g, o, vs, tr = tdm.read_grid_bin(p)
    assert np.array_equal(g, grid) and vs == np.float32(0.006) and tr == np.float32(0.03)

    vols = rng.random((3, 30, 30, 30), dtype=np.float32)
    p = str(tmp_path / "data.bin")
    open(p, "wb").write(struct.pack("<4f", 3, 30, 30, 30) + vols.tobytes())
    assert np.array_equal(tdm.read_patch_volumes(p), vols)


def test_png_decoder_matches_pil():
    from PIL import Image

    files = sorted(glob.glob(os.path.join(FRAMES, "*.depth.png")))
    assert files
    for f in files:
        ours = tdm.read_png_u16(f)
        pil = np.array(Image.open(f)).astype(np.uint16)
        assert ours.shape == (480, 640)
        assert np.array_equal(ours, pil)
        m = tdm.read_depth_png(f)
        assert np.array_equal(m, oracle.depth_from_mm(ours))


def test_depth_conversion_is_ieee_division():
    mm = np.arange(0, 65536, dtype=np.uint16)
    d = oracle.depth_from_mm(mm)
    want = mm.astype(np.float32) / np.float32(1000.0)
    want[want > np.float32(6.0)] = 0
    assert np.array_equal(d, want)
    assert d[6000] == np.float32(6.0) and d[6001] == 0


def test_matrix_io_and_pose_arithmetic():
    K = tdm.load_matrix(os.path.join(FRAMES, "camera-intrinsics.txt"), 3, 3)
    assert np.array_equal(K, np.array([[585, 0, 320], [0, 585, 240], [0, 0, 1]], np.float32))
    poses = [tdm.load_matrix(f, 4, 4) for f in sorted(glob.glob(os.path.join(FRAMES, "*.pose.txt")))]
    rng = np.random.default_rng(2)
    mats = poses + [rng.normal(size=(4, 4)).astype(np.float32) for _ in range(50)]
    for m in mats:
        inv = tdm.invert_matrix(m)
        # the product's table-driven cofactors equal the oracle's restatement bit for bit
        assert np.array_equal(inv.view(np.uint32), oracle.invert_matrix(m).view(np.uint32))
        ref = np.linalg.inv(m.astype(np.float64))
        assert np.allclose(inv, ref, rtol=1e-3, atol=1e-4 * np.abs(ref).max())
    for a, b in zip(mats[:-1], mats[1:]):
        prod = tdm.multiply_matrix(a, b)
        assert np.array_equal(prod.view(np.uint32), oracle.multiply_matrix(a, b).view(np.uint32))
        assert np.allclose(prod, a.astype(np.float64) @ b.astype(np.float64), rtol=1e-5, atol=1e-5)
    with pytest.raises(tdm.TdmError):
        tdm.invert_matrix(np.zeros((4, 4), np.float32))


def test_surface_ply(tmp_path):
    rng = np.random.default_rng(3)
    t = rng.uniform(-1, 1, (6, 7, 8)).astype(np.float32)
    w = rng.integers(0, 4, (6, 7, 8)).astype(np.float32)
    p = str(tmp_path / "tsdf.ply")
    origin = np.array([-1.5, -1.5, 0.5], np.float32)
    vs = np.float32(0.006)
    n = tdm.save_surface_ply(p, t, w, origin, vs)
    z, y, x = np.nonzero((np.abs(t) < 0.2) & (w > 1))
    assert n == len(z)
    pts = tdm.read_ply_points(p)
    want = np.stack([origin[0] + x.astype(np.float32) * vs, origin[1] + y.astype(np.float32) * vs,
                     origin[2] + z.astype(np.float32) * vs], axis=1)
    assert np.array_equal(pts, want)
