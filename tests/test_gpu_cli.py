"""The drop-in executables (the reference's boundary: depth-fusion/demo,
core/demo, evaluation/*/get3DMatchDescriptors) run on the GPU and compared
with the oracle through the files they write."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

import oracle
import tdm
from conftest import GOLDEN, rel_l2
from tdm import netdef, synth

pytestmark = pytest.mark.gpu

BIN = tdm.BIN_DIR


def run(args, cwd):
    print("run:", " ".join(args), flush=True)
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.fixture(scope="module")
def net_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    netdef.write_net_json(str(d / "3dmatch-net-test.json"))
    shutil.copy(d / "3dmatch-net-test.json", d / "tmp.json")
    w = netdef.he_normal_weights(5)
    tdm.marvin_write(str(d / "3dmatch-weights-snapshot-137000.marvin"), w)
    return d, dict(w)


def test_fusion_demo_matches_oracle(torch_gpu, tmp_path):
    frames = os.path.join(GOLDEN, "sample_frames")
    K = tdm.load_matrix(os.path.join(frames, "camera-intrinsics.txt"), 3, 3)
    vs = np.float32(0.012)
    trunc = np.float32(vs * np.float32(5))
    dims = (200, 200, 200)
    origin = np.array([-1.5, -1.5, 0.5], np.float32)
    out = run([os.path.join(BIN, "fusion_demo"), os.path.join(frames, "camera-intrinsics.txt"), frames,
               "8", "8", "5", "-1.5", "-1.5", "0.5", "0.012", str(float(trunc)), "--dims", "200", "200",
               "200", "--tsdf-bin", "tsdf.bin", "--weight-bin", "weight.bin"], str(tmp_path))
    assert out.count("Fusing:") == 5
    # oracle: the demo's host prelude + Integrate per frame
    base_inv = oracle.invert_matrix(tdm.load_matrix(os.path.join(frames, "frame-000008.pose.txt"), 4, 4))
    t, w = oracle.new_grid(dims)
    for i in range(8, 13):
        pre = os.path.join(frames, f"frame-{i:06d}")
        P = oracle.multiply_matrix(base_inv, tdm.load_matrix(pre + ".pose.txt", 4, 4))
        D = oracle.depth_from_mm(tdm.read_png_u16(pre + ".depth.png"))
        oracle.integrate(K, P, D, dims, origin, vs, trunc, t, w)
    gt, go, gvs, gtr = tdm.read_grid_bin(str(tmp_path / "tsdf.bin"))
    gw, _, _, _ = tdm.read_grid_bin(str(tmp_path / "weight.bin"))
    assert np.array_equal(gw, w)
    assert np.array_equal(gt.view(np.uint32), t.view(np.uint32))
    assert gvs == vs and gtr == trunc and np.array_equal(go, origin)
    # tsdf.ply: voxels with |tsdf| < 0.2 and weight > 1, at origin + idx*vs
    pts = tdm.read_ply_points(str(tmp_path / "tsdf.ply"))
    z, y, x = np.nonzero((np.abs(t) < 0.2) & (w > 1))
    want = np.stack([origin[0] + x.astype(np.float32) * vs, origin[1] + y.astype(np.float32) * vs,
                     origin[2] + z.astype(np.float32) * vs], 1)
    assert len(pts) > 1000 and np.array_equal(pts, want)
    # the serial loop and the fused pass write the same files as the default
    # overlapped pipeline
    for k, mode in enumerate(("--fused", "--serial"), start=2):
        run([os.path.join(BIN, "fusion_demo"), os.path.join(frames, "camera-intrinsics.txt"), frames,
             "8", "8", "5", "-1.5", "-1.5", "0.5", "0.012", str(float(trunc)), "--dims", "200", "200",
             "200", "--tsdf-bin", f"tsdf{k}.bin", mode, "--out", f"tsdf{k}.ply"], str(tmp_path))
        assert open(tmp_path / f"tsdf{k}.bin", "rb").read() == open(tmp_path / "tsdf.bin", "rb").read()
        assert open(tmp_path / f"tsdf{k}.ply", "rb").read() == open(tmp_path / "tsdf.ply", "rb").read()


def test_descriptor_demo_matches_oracle(torch_gpu, net_dir, tmp_path):
    d, w = net_dir
    cloud = synth.fragment(40_000, seed=41)
    tdm.write_ply_points(str(tmp_path / "cloud.ply"), cloud)
    idx = synth.keypoint_indices(len(cloud), 7, seed=42)
    (tmp_path / "kp.txt").write_text("\n".join(str(i) for i in idx))
    for f in ("3dmatch-net-test.json", "3dmatch-weights-snapshot-137000.marvin"):
        shutil.copy(d / f, tmp_path / f)
    run([os.path.join(BIN, "descriptor_demo"), "cloud.ply", "frag", "--keypoint-indices", "kp.txt",
         "--tdf-bin", "tdf.bin"], str(tmp_path))
    kp = tdm.read_keypts_bin(str(tmp_path / "frag.keypts.bin"))
    assert np.array_equal(kp, cloud[idx])
    vs = np.float32(0.01)
    dims, origin = oracle.grid_params(cloud, vs, 15)
    occ = oracle.occupancy(cloud, dims, origin, vs)
    tdf = oracle.compute_tdf(occ, vs, np.float32(0.05))
    g, go, _, _ = tdm.read_grid_bin(str(tmp_path / "tdf.bin"))
    assert np.array_equal(go, origin) and np.array_equal(g, tdf)
    patches = oracle.extract_patches(tdf, oracle.keypoint_grid(kp, origin, vs))
    ref = oracle.net_forward_fp64(patches, w, netdef.layer_shapes())[-1].reshape(len(kp), 512)
    desc = tdm.read_desc_bin(str(tmp_path / "frag.desc.3dmatch.bin"))
    assert desc.shape == (7, 512)
    assert rel_l2(desc, ref) <= 1e-5
    # --precision fp16 (the reference build's half storage) gives what the
    # library's fp16 net gives on the same grid and keypoints, bit for bit
    run([os.path.join(BIN, "descriptor_demo"), "cloud.ply", "h", "--keypoint-indices", "kp.txt",
         "--precision", "fp16"], str(tmp_path))
    half = tdm.read_desc_bin(str(tmp_path / "h.desc.3dmatch.bin"))
    net = tdm.Net(str(d / "3dmatch-net-test.json"), tdm.PREC_FP16)
    net.load_weights(str(d / "3dmatch-weights-snapshot-137000.marvin"))
    feat = torch_gpu.empty(7, 512, device="cuda:0")
    kpg = oracle.keypoint_grid(kp, origin, vs).astype(np.int32)
    net.forward_grid(torch_gpu.from_numpy(tdf).to("cuda:0"), dims,
                     torch_gpu.from_numpy(kpg).to("cuda:0"), feat)
    assert np.array_equal(half.view(np.uint32), feat.cpu().numpy().view(np.uint32))
    assert rel_l2(half, ref) <= 1e-2
    net.close()
    # seeded draws are reproducible
    run([os.path.join(BIN, "descriptor_demo"), "cloud.ply", "s1", "--seed", "7", "--num-keypoints", "50"], str(tmp_path))
    run([os.path.join(BIN, "descriptor_demo"), "cloud.ply", "s2", "--seed", "7", "--num-keypoints", "50"], str(tmp_path))
    assert open(tmp_path / "s1.desc.3dmatch.bin", "rb").read() == open(tmp_path / "s2.desc.3dmatch.bin", "rb").read()


def test_get3dmatch_descriptors_both_forms(torch_gpu, net_dir, tmp_path):
    d, w = net_dir
    for f in ("tmp.json", "3dmatch-weights-snapshot-137000.marvin"):
        shutil.copy(d / f, tmp_path / f)
    rng = np.random.default_rng(43)
    vols = (rng.random((53, 30, 30, 30)) < 0.05).astype(np.float32)  # 53: not a multiple of 50
    open(tmp_path / "data.bin", "wb").write(struct.pack("<4f", 53, 30, 30, 30) + vols.tobytes())
    run([os.path.join(BIN, "get3dmatch_descriptors"), "data.bin"], str(tmp_path))
    raw = open(tmp_path / "feat.bin", "rb").read()
    assert struct.unpack("<5f", raw[:20]) == (53, 512, 1, 1, 1)
    feat = np.frombuffer(raw[20:], np.float32).reshape(53, 512)
    ref = oracle.net_forward_fp64(vols[:4], w, netdef.layer_shapes())[-1].reshape(4, 512)
    assert rel_l2(feat[:4], ref) <= 1e-5

    # shape2pose form: TDF grid with a 3-float header + grid-coordinate keypoints
    grid = rng.random((40, 45, 50), dtype=np.float32)
    open(tmp_path / "TDF.bin", "wb").write(struct.pack("<3f", 50, 45, 40) + grid.tobytes())
    # 20.5 and 24.5 round away from zero (C round, as the reference's
    # get3DMatchDescriptors.cu:49-51), where half-to-even would give 20 and 24
    kpg = np.array([[20.2, 22.7, 19.5], [30, 25, 20], [15, 15, 15], [20.5, 24.5, 17.5]], np.float32)
    open(tmp_path / "kp.bin", "wb").write(struct.pack("<f", len(kpg)) + kpg.tobytes())
    run([os.path.join(BIN, "get3dmatch_descriptors"), "TDF.bin", "kp.bin"], str(tmp_path))
    desc = tdm.read_desc_bin(str(tmp_path / "kp.3dmatch.descriptors.bin"))
    kpi = (np.sign(kpg) * np.floor(np.abs(kpg) + np.float32(0.5))).astype(np.int32)
    assert kpi[3].tolist() == [21, 25, 18]
    patches = oracle.extract_patches(grid, kpi)
    ref = oracle.net_forward_fp64(patches, w, netdef.layer_shapes())[-1].reshape(len(kpg), 512)
    assert rel_l2(desc, ref) <= 1e-5
    # a keypoint rounded half to even instead would read a patch one voxel off
    off = oracle.net_forward_fp64(oracle.extract_patches(grid, np.round(kpg[3:]).astype(np.int32)), w,
                                  netdef.layer_shapes())[-1].reshape(1, 512)
    assert rel_l2(desc[3:], off) > 1e-3

    # fp16 and fp8 are selectable; fp16 equals the library's fp16 net bit for bit
    for prec, tol in (("fp16", 1e-2), ("fp8", 0.15)):
        run([os.path.join(BIN, "get3dmatch_descriptors"), "data.bin", "--precision", prec],
            str(tmp_path))
        raw = open(tmp_path / "feat.bin", "rb").read()
        got = np.frombuffer(raw[20:], np.float32).reshape(53, 512)
        assert rel_l2(got[:4], oracle.net_forward_fp64(vols[:4], w, netdef.layer_shapes())[-1]
                      .reshape(4, 512)) <= tol
        if prec == "fp16":
            net = tdm.Net(str(tmp_path / "tmp.json"), tdm.PREC_FP16)
            net.load_weights(str(tmp_path / "3dmatch-weights-snapshot-137000.marvin"))
            f16 = torch_gpu.empty(53, 512, device="cuda:0")
            net.forward(torch_gpu.from_numpy(vols).to("cuda:0"), f16)
            assert np.array_equal(got.view(np.uint32), f16.cpu().numpy().view(np.uint32))
            net.close()


def test_descriptor_demo_on_reference_sample_cloud(torch_gpu, net_dir, tmp_path):
    """data/sample/3dmatch-demo/single-depth-2.ply (284 874 points, a
    299 x 255 x 263 = 20.05 M-voxel grid, above 2^24, so core/demo.cu:32-34's
    float32 decode mis-places voxels): descriptor_demo --decode reference
    writes the reference's TDF bit for bit, the keypoints it was given and
    their descriptors (core/demo.cu:69-135, 236-274)."""
    d, w = net_dir
    for f in ("3dmatch-net-test.json", "3dmatch-weights-snapshot-137000.marvin"):
        shutil.copy(d / f, tmp_path / f)
    ply = os.path.join(GOLDEN, "sample_clouds", "single-depth-2.ply")
    cloud = tdm.read_ply_points(ply)
    assert len(cloud) == 284_874
    idx = synth.keypoint_indices(len(cloud), 6, seed=44)
    (tmp_path / "kp.txt").write_text("\n".join(str(i) for i in idx))
    run([os.path.join(BIN, "descriptor_demo"), ply, "test", "--keypoint-indices", "kp.txt",
         "--tdf-bin", "tdf.bin", "--decode", "reference"], str(tmp_path))
    vs = np.float32(0.01)
    trunc = vs * np.float32(5)  # core/demo.cu: trunc_margin = voxel_size * 5 in float
    dims, origin = oracle.grid_params(cloud, vs, 15)
    assert list(dims) == [299, 255, 263]
    occ = oracle.occupancy(cloud, dims, origin, vs)
    ref = oracle.compute_tdf(occ, vs, trunc, "reference")
    exact = oracle.compute_tdf(occ, vs, trunc, "exact")
    assert np.count_nonzero(ref != exact) > 1000  # the decodes differ on this grid
    g, go, gvs, gtr = tdm.read_grid_bin(str(tmp_path / "tdf.bin"))
    assert np.array_equal(go, origin) and gvs == vs and gtr == trunc
    assert np.array_equal(g.view(np.uint32), ref.view(np.uint32))
    # keypts.bin: float32 N, then N x 3 world coordinates
    raw = open(tmp_path / "test.keypts.bin", "rb").read()
    assert struct.unpack("<f", raw[:4])[0] == 6.0 and len(raw) == 4 + 6 * 12
    kp = tdm.read_keypts_bin(str(tmp_path / "test.keypts.bin"))
    assert np.array_equal(kp, cloud[idx])
    # desc.3dmatch.bin: float32 N, float32 512, then N x 512
    raw = open(tmp_path / "test.desc.3dmatch.bin", "rb").read()
    assert struct.unpack("<2f", raw[:8]) == (6.0, 512.0) and len(raw) == 8 + 6 * 512 * 4
    desc = tdm.read_desc_bin(str(tmp_path / "test.desc.3dmatch.bin"))
    patches = oracle.extract_patches(ref, oracle.keypoint_grid(kp, origin, vs))
    want = oracle.net_forward_fp64(patches, w, netdef.layer_shapes())[-1].reshape(6, 512)
    assert rel_l2(desc, want) <= 1e-5
