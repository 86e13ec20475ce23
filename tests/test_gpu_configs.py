"""Parity at the full sizes of the benchmark configurations, against the
oracle (not against other HIP entry points).

C2  one 200k-point fragment (seed 3): grid, TDF in both decodes on its
    419 x 321 x 347 grid (46.7 M voxels, above 2^24, so the float32 decode of
    core/demo.cu:32-34 mis-places voxels), 5000 keypoints, all 5000 patches,
    and descriptors from a max_batch 1024 autotuned net (the benchmarked
    configuration) per layer and at the output against FP64.
C3  see bench.py's precision study (a measurement, not a parity case).
C4  64 fragments of 20k-500k points (extents 1-4 m, seed 40) in one TDF plan,
    the largest, the smallest and two others against the oracle in exact
    decode and three in reference decode; BF16 descriptors of one multi-grid
    pass over all their keypoints against FP64; 16 scenes x 50 frames into
    256^3 grids in one launch, two scenes against the oracle.
C5  the registration server at 5000 keypoints per fragment and a 50k-trial
    RANSAC cap, with a pair of 300k+ points: alone == in a mix, and the whole
    chain (TDF, descriptors, exact mutual nearest neighbours, ransac.m's loop)
    against the oracle.
C1 (50 frames into 256^3) is in test_gpu_tsdf.py.
"""
import numpy as np
import pytest

import oracle
import registration_oracle as ro
import tdm
from conftest import rel_l2
from tdm import netdef, synth

pytestmark = pytest.mark.gpu

VS = np.float32(0.01)
TRUNC = np.float32(0.05)
FP32_TOL, FP16_TOL, BF16_TOL = 1e-5, 1e-2, 3e-2


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


@pytest.fixture(scope="module")
def bench_net(tmp_path_factory):
    """The benchmark's net: He-normal weights (seed 5), zero biases."""
    d = tmp_path_factory.mktemp("cfg")
    js, wt = str(d / "net.json"), str(d / "w.marvin")
    netdef.write_net_json(js)
    w = netdef.he_normal_weights(5)
    tdm.marvin_write(wt, w)
    return js, wt, dict(w)


# ---- C2 ----------------------------------------------------------------------------


@pytest.fixture(scope="module")
def c2():
    cloud = synth.fragment(200_000, seed=3)
    idx = synth.keypoint_indices(len(cloud), 5000, seed=4)
    dims, origin = oracle.grid_params(cloud, VS, 15)
    occ = oracle.occupancy(cloud, dims, origin, VS)
    exact = oracle.compute_tdf(occ, VS, TRUNC, "exact")
    return cloud, idx, dims, origin, occ, exact


def test_c2_grid_and_tdf_both_decodes(torch_gpu, c2):
    torch = torch_gpu
    cloud, _, dims, origin, occ, exact = c2
    assert dims[0] * dims[1] * dims[2] > 1 << 24
    pts = _dev(torch, cloud)
    gd, go = tdm.grid_params(pts, VS, 15)
    assert gd == dims and _bits_equal(go, origin)
    ref = oracle.compute_tdf(occ, VS, TRUNC, "reference")
    assert np.count_nonzero(ref != exact) > 100_000  # the decodes do differ here
    for decode, want in ((tdm.DECODE_EXACT, exact), (tdm.DECODE_REFERENCE, ref)):
        print(f"case C2 TDF dims={dims} decode={decode}", flush=True)
        tdf = torch.full((dims[2], dims[1], dims[0]), -1.0, device="cuda:0")
        tdm.tdf_from_points(pts, dims, origin, VS, TRUNC, tdf, decode)
        got = tdf.cpu().numpy()
        assert _bits_equal(got, want), f"{np.count_nonzero(got != want)} voxels differ"


def test_c2_keypoints_and_patches(torch_gpu, c2):
    torch = torch_gpu
    cloud, idx, dims, origin, _, exact = c2
    kp = cloud[idx]
    kg = tdm.keypoints_to_grid(kp, origin, VS)
    rk = oracle.keypoint_grid(kp, origin, VS)
    assert np.array_equal(kg, rk)
    print("case C2 5000 patches", flush=True)
    out = torch.empty(len(kg), 30, 30, 30, device="cuda:0")
    tdm.extract_patches(_dev(torch, exact), dims, _dev(torch, kg), out)
    got = out.cpu().numpy()
    for c0 in range(0, len(kg), 1000):  # bounded host memory
        want = oracle.extract_patches(exact, rk[c0:c0 + 1000])
        assert _bits_equal(got[c0:c0 + 1000], want)


@pytest.mark.parametrize("prec,tol", [(tdm.PREC_FP32, FP32_TOL), (tdm.PREC_FP16, FP16_TOL),
                                      (tdm.PREC_BF16, BF16_TOL)])
def test_c2_descriptors_benchmarked_net(torch_gpu, c2, bench_net, prec, tol):
    """max_batch 1024 with the tuned tile variants (split-K counts follow
    max_batch): every layer of 12 patches against FP64, the fused gather of
    all 5000 keypoints equal to the patch path bit for bit."""
    torch = torch_gpu
    cloud, idx, dims, origin, _, exact = c2
    js, wt, w = bench_net
    net = tdm.Net(js, prec, max_batch=1024)
    net.load_weights(wt)
    variants = net.autotune(1024)
    print(f"case C2 descriptors prec={prec} variants={variants}", flush=True)
    rk = oracle.keypoint_grid(cloud[idx], origin, VS)
    tdf = _dev(torch, exact)
    kgd = _dev(torch, rk)
    fused = torch.empty(len(rk), 512, device="cuda:0")
    net.forward_grid(tdf, dims, kgd, fused)
    patches = torch.empty(len(rk), 30, 30, 30, device="cuda:0")
    tdm.extract_patches(tdf, dims, kgd, patches)
    viap = torch.empty_like(fused)
    net.forward(patches, viap)
    assert torch.equal(fused, viap)
    pick = np.linspace(0, len(rk) - 1, 12).astype(int)
    p = oracle.extract_patches(exact, rk[pick])
    ref = oracle.net_forward_fp64(p, w, netdef.layer_shapes())
    data = _dev(torch, p)
    errs = []
    for i, (_n, _ci, co, _k, _d, od, pooled) in enumerate(netdef.layer_shapes()):
        e = pooled if pooled is not None else od
        out = torch.empty(len(p), co, e, e, e, device="cuda:0")
        net.forward_layer(data, i, out)
        errs.append(rel_l2(out.cpu().numpy(), ref[i]))
    print("per-layer rel-L2:", ["%.2e" % e for e in errs], flush=True)
    assert max(errs) <= tol
    d = rel_l2(fused.cpu().numpy()[pick], ref[-1].reshape(len(p), 512))
    print(f"descriptor rel-L2 {d:.2e}", flush=True)
    assert d <= tol
    net.close()


# ---- C4 ----------------------------------------------------------------------------


@pytest.fixture(scope="module")
def c4_clouds():
    """The batched workload's 64 fragments and keypoints (bench.py's draws)."""
    rng = np.random.default_rng(40)
    out = []
    for f in range(64):
        n = int(rng.integers(20_000, 500_001))
        cloud = synth.cropped_fragment(40 * 1000 + f, n, float(rng.uniform(1.0, 4.0)))
        k = int(rng.integers(1000, 10_001))
        idx = rng.integers(0, n, k)
        out.append((cloud, idx))
    return out


def test_c4_tdf_plan_and_multigrid_descriptors(torch_gpu, c4_clouds, bench_net):
    torch = torch_gpu
    clouds = []
    for cloud, _ in c4_clouds:
        pts = _dev(torch, cloud)
        dims, origin = tdm.grid_params(pts, VS, 15)
        clouds.append((pts, torch.full((dims[2], dims[1], dims[0]), -1.0, device="cuda:0"),
                       dims, origin))
    nvox = [d[0] * d[1] * d[2] for _, _, d, _ in clouds]
    order = np.argsort(nvox)
    checked = [int(order[-1]), int(order[0]), int(order[21]), int(order[42])]
    print(f"case C4 plan of 64 grids ({sum(nvox)} voxels), checking {checked}", flush=True)
    plan = tdm.TdfPlan(clouds, VS, TRUNC, tdm.DECODE_EXACT)
    plan.run()
    torch.cuda.synchronize()
    exact = {}
    for f in checked:
        cloud = c4_clouds[f][0]
        pts, tdf, dims, origin = clouds[f]
        rd, ro_ = oracle.grid_params(cloud, VS, 15)
        assert rd == dims and _bits_equal(ro_, origin)
        exact[f] = oracle.compute_tdf(oracle.occupancy(cloud, dims, origin, VS), VS, TRUNC)
        assert _bits_equal(tdf.cpu().numpy(), exact[f]), f"fragment {f} {dims}"

    # BF16 descriptors of every keypoint of the batch in one multi-grid pass;
    # 3 keypoints of each checked grid against FP64
    js, wt, w = bench_net
    kgs, gids = [], []
    for f, (cloud, idx) in enumerate(c4_clouds):
        kgs.append(tdm.keypoints_to_grid(cloud[idx], clouds[f][3], VS))
        gids.append(np.full(len(idx), f, np.int32))
    kg, gid = np.concatenate(kgs), np.concatenate(gids)
    net = tdm.Net(js, tdm.PREC_BF16, max_batch=1024)
    net.load_weights(wt)
    net.autotune(1024)
    table = tdm.grid_table([(tdf, dims) for _, tdf, dims, _ in clouds])
    feat = torch.empty(len(kg), 512, device="cuda:0")
    print(f"case C4 multi-grid bf16 pass of {len(kg)} keypoints", flush=True)
    net.forward_grids(table, _dev(torch, kg), _dev(torch, gid), feat)
    torch.cuda.synchronize()
    got, want = [], []
    for f in checked:
        rows = np.nonzero(gid == f)[0][:3]
        p = oracle.extract_patches(exact[f], kg[rows])
        want.append(oracle.net_forward_fp64(p, w, netdef.layer_shapes())[-1].reshape(len(p), 512))
        got.append(feat[_dev(torch, rows)].cpu().numpy())
    e = rel_l2(np.concatenate(got), np.concatenate(want))
    print(f"bf16 descriptor rel-L2 {e:.2e}", flush=True)
    assert e <= BF16_TOL
    net.close()
    plan.close()

    # reference decode: the largest grid (above 2^24 voxels) and two small ones
    sub = [checked[0], checked[1], checked[2]]
    assert nvox[sub[0]] > 1 << 24
    refc = []
    for f in sub:
        pts, _, dims, origin = clouds[f]
        refc.append((pts, torch.full((dims[2], dims[1], dims[0]), -1.0, device="cuda:0"), dims,
                     origin))
    plan = tdm.TdfPlan(refc, VS, TRUNC, tdm.DECODE_REFERENCE)
    plan.run()
    torch.cuda.synchronize()
    for f, (pts, tdf, dims, origin) in zip(sub, refc):
        occ = oracle.occupancy(c4_clouds[f][0], dims, origin, VS)
        want = oracle.compute_tdf(occ, VS, TRUNC, "reference")
        assert _bits_equal(tdf.cpu().numpy(), want), f"fragment {f} reference decode"
    if nvox[sub[-1]] <= 1 << 24:  # below 2^24 voxels the decodes agree
        assert _bits_equal(want, exact[sub[-1]])
    plan.close()


def test_c4_sixteen_scenes_fifty_frames(torch_gpu):
    torch = torch_gpu
    S, F = 16, 50
    depths = []
    for sc in range(S):
        d, c2w = synth.depth_frames_torch(torch, F, scene_seed=50 + sc, noise_seed=100 + sc)
        depths.append(d)
    D = torch.stack(depths)
    poses = synth.relative_poses(c2w)
    P = _dev(torch, np.broadcast_to(poses.reshape(1, F, 16), (S, F, 16)))
    origins = _dev(torch, np.tile(synth.TSDF_ORIGIN, (S, 1)))
    K = synth.intrinsics()
    dims = synth.TSDF_DIMS
    n = dims[0] * dims[1] * dims[2]
    t = torch.empty(S * n, device="cuda:0")
    w = torch.empty(S * n, device="cuda:0")
    tdm.tsdf_init(t, w)
    print("case C4 16 scenes x 50 frames into 256^3", flush=True)
    tdm.integrate_scenes(_dev(torch, K), P, D, dims, origins, synth.TSDF_VOXEL, synth.TSDF_TRUNC,
                         t, w)
    torch.cuda.synchronize()
    for sc in (0, S - 1):
        frames = D[sc].cpu().numpy()
        rt, rw = oracle.new_grid(dims)
        for f in range(F):
            oracle.integrate(K, poses[f], frames[f], dims, synth.TSDF_ORIGIN, synth.TSDF_VOXEL,
                             synth.TSDF_TRUNC, rt, rw)
        assert rw.sum() > 1e8
        assert _bits_equal(w[sc * n:(sc + 1) * n].cpu().numpy(), rw.reshape(-1))
        assert _bits_equal(t[sc * n:(sc + 1) * n].cpu().numpy(), rt.reshape(-1)), f"scene {sc}"


# ---- C5 ----------------------------------------------------------------------------

NK, TRIALS = 5000, 50_000


def test_c5_server_full_size_against_oracle(torch_gpu, bench_net):
    torch = torch_gpu
    js, wt, w = bench_net
    net = tdm.Net(js, tdm.PREC_BF16, max_batch=1024)
    net.load_weights(wt)
    net.autotune(1024)
    reqs = []
    for k in range(3):
        n1, n2 = (320_000, 300_000) if k == 0 else (None, None)
        p1, p2, R, t = synth.fragment_pair(70 + k, n1=n1, n2=n2)
        rng = np.random.default_rng(170 + k)
        k1 = rng.integers(0, len(p1), NK).astype(np.int32)
        k2 = rng.integers(0, len(p2), NK).astype(np.int32)
        reqs.append((p1, p2, k1, k2, 700 + k))
    cfg = dict(keypoints=NK, max_pass_keypoints=4 * NK, ransac_max_trials=TRIALS,
               memory_budget=6 << 30)
    print(f"case C5 sizes {[(len(r[0]), len(r[1])) for r in reqs]}", flush=True)
    srv = tdm.Server(net, max_inflight=1, **cfg)
    alone = [srv.wait(srv.submit(p1, p2, k1, k2, seed=s)) for p1, p2, k1, k2, s in reqs]
    srv.close()
    srv = tdm.Server(net, max_inflight=3, **cfg)
    tickets = [srv.submit(p1, p2, k1, k2, seed=s) for p1, p2, k1, k2, s in reqs]
    mixed = [srv.wait(tk) for tk in tickets]
    st = srv.stats()
    srv.close()
    print("mix stats", st, flush=True)
    assert st["max_pass_requests"] > 1
    for a, b in zip(alone, mixed):
        assert a["status"] == 0 and b["status"] == 0
        assert np.array_equal(a["rt"], b["rt"])
        assert (a["num_matches"], a["num_inliers"], a["trials"]) == \
            (b["num_matches"], b["num_inliers"], b["trials"])

    # the oracle chain for the 300k+ pair
    p1, p2, k1, k2, seed = reqs[0]
    descs, kps = [], []
    for pts, kidx in ((p1, k1), (p2, k2)):
        dims, origin = oracle.grid_params(pts, VS, 15)
        tdf = oracle.compute_tdf(oracle.occupancy(pts, dims, origin, VS), VS, TRUNC)
        g = torch.empty(dims[2], dims[1], dims[0], device="cuda:0")
        tdm.tdf_from_points(_dev(torch, pts), dims, origin, VS, TRUNC, g)
        assert _bits_equal(g.cpu().numpy(), tdf)
        kg = oracle.keypoint_grid(pts[kidx], origin, VS)
        f = torch.empty(NK, 512, device="cuda:0")
        net.forward_grid(g, dims, _dev(torch, kg), f)
        f = f.cpu().numpy()
        pp = oracle.extract_patches(tdf, kg[:2])
        ref = oracle.net_forward_fp64(pp, w, netdef.layer_shapes())[-1].reshape(2, 512)
        assert rel_l2(f[:2], ref) <= BF16_TOL
        descs.append(f)
        kps.append(pts[kidx])
    nn_ab, _ = ro.nn_exact(descs[0], descs[1])
    nn_ba, _ = ro.nn_exact(descs[1], descs[0])
    pairs = ro.mutual_pairs(nn_ab, nn_ba)
    x, y = kps[0][pairs[:, 0]], kps[1][pairs[:, 1]]
    ort, oinl, _best, orun, _ = ro.ransac_fit_rt(x, y, 0.05, TRIALS, seed)
    res = alone[0]
    print(f"C5 matches {len(pairs)} inliers {len(oinl)} trials {orun}", flush=True)
    assert res["num_matches"] == len(pairs)
    assert res["num_inliers"] == len(oinl) and res["trials"] == orun
    assert np.abs(res["rt"] - ort).max() < 1e-9
    net.close()
