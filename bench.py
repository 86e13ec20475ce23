"""Throughput of the 3DMatch hot path on one MI355X.

Headline (value): descriptors/s of the plain fp32 path on the descriptor
workload -- a 200k-point synthetic fragment -> TDF grid (0.01 m voxels,
truncation 5 voxels, padding 15) -> 5000 keypoints -> 30^3 patches -> the
3DMatch net (He-normal weights) -> 5000 x 512 descriptors.  One step is that
whole pass with the points already in HBM: grid size (GPU bounding box),
TDF, keypoint grid coordinates, and the fused-gather descriptor forward.

Also reported: TSDF fusion (50 synthetic 640x480 frames into a 256^3 grid,
fused single pass and per-frame launches) in voxel-updates/s with its HBM
roofline, the bf16 descriptor path, per-layer times, and the CPU baseline
(the oracle's C loop nests on a bounded sample, scaled).

Also: the fp16 path (the reference's default half storage), a precision
study (rel-L2 vs FP64 per layer, exact mutual-NN agreement), matching and
RANSAC kernel times, and the serving form (p50/p99 registration latency,
Poisson arrival sweep) of the registration server.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
                       [--no-cpu-baseline] [--no-extras] [--no-serving]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "3dmatch-toolbox_amd"))

METRIC = "3DMatch descriptors/s + TSDF voxel-updates/s on 1 GPU; p50/p99 reg. latency"
HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"fp32": 157.3, "bf16": 2500.0, "fp16": 2500.0, "fp8": 5000.0}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def descriptor_workload(n_points=200_000, n_kp=5000):
    import numpy as np

    from tdm import netdef, synth

    cloud = synth.fragment(n_points, seed=3)
    idx = synth.keypoint_indices(len(cloud), n_kp, seed=4)
    weights = netdef.he_normal_weights(5)
    return cloud, cloud[idx].astype(np.float32), weights


class DescriptorPipeline:
    """TDF + descriptors for one fragment, everything on one stream."""

    def __init__(self, torch, cloud, keypts, js, wt, precision, batch):
        import tdm

        self.torch = torch
        self.tdm = tdm
        self.dev = torch.device("cuda:0")
        self.pts = torch.from_numpy(cloud).to(self.dev)
        self.keypts = keypts
        self.net = tdm.Net(js, precision, max_batch=batch)
        self.net.load_weights(wt)
        if precision == tdm.PREC_FP8:
            calibrate_fp8(torch, self.net, js, wt, self.pts, keypts[:256])
        # tile autotuner: bit-identical results, fastest tile per layer
        self.variants = self.net.autotune(batch)
        self.vs = 0.01
        self.trunc = 0.05
        dims, _ = tdm.grid_params(self.pts, self.vs, 15)
        self.tdf = torch.empty(dims[2] * dims[1] * dims[0], dtype=torch.float32, device=self.dev)
        self.feat = torch.empty(len(keypts), self.net.output_dim, dtype=torch.float32, device=self.dev)
        self.kp_dev = torch.empty(len(keypts), 3, dtype=torch.int32, device=self.dev)
        self.kp_host = torch.empty(len(keypts), 3, dtype=torch.int32).pin_memory()
        self.dims = dims

    def step(self):
        tdm, torch = self.tdm, self.torch
        import numpy as np

        dims, origin = tdm.grid_params(self.pts, np.float32(self.vs), 15)
        assert dims == self.dims
        tdm.tdf_from_points(self.pts, dims, origin, np.float32(self.vs), np.float32(self.trunc), self.tdf)
        kg = tdm.keypoints_to_grid(self.keypts, origin, np.float32(self.vs))
        self.kp_host.copy_(torch.from_numpy(kg))
        self.kp_dev.copy_(self.kp_host, non_blocking=True)
        self.net.forward_grid(self.tdf, dims, self.kp_dev, self.feat)


def measured_traffic(precision, patches):
    """conv2's HBM bytes per launch of `patches` patches, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload (measured on a
    launch of its stated size and scaled by patch count; None if there is no
    summary for the precision)."""
    import re

    path = os.path.join(REPO, "profiles", f"hbm_traffic_descriptor_pipeline_{precision}.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"conv2 \(`[^`]*`\): (\d+) MB read and (\d+) MB written per (\d+)-patch launch",
                  open(path).read())
    if not m:
        return None
    return (float(m.group(1)) + float(m.group(2))) * 1e6 * patches / float(m.group(3))


def layer_tflops(net):
    from tdm import netdef

    out = {}
    for i, (name, cin, cout, k, _d, od, _p) in enumerate(netdef.layer_shapes()):
        ms, passes, patches = net.layer_stats(i)
        flops = 2 * od ** 3 * cout * cin * k ** 3
        out[name] = flops * patches / (ms / 1e3) / 1e12 if ms > 0 else None
    return out


def calibrate_fp8(torch, net, js, wt, pts, keypts):
    """Activation scales of an fp8 net from an fp32 net on the patches around
    `keypts` of the cloud `pts` (device)."""
    import numpy as np

    import tdm

    dims, origin = tdm.grid_params(pts, np.float32(0.01), 15)
    tdf = torch.empty(dims[2], dims[1], dims[0], device=pts.device)
    tdm.tdf_from_points(pts, dims, origin, np.float32(0.01), np.float32(0.05), tdf)
    kg = torch.from_numpy(tdm.keypoints_to_grid(keypts, origin, np.float32(0.01))).to(pts.device)
    patches = torch.empty(len(kg), 30, 30, 30, device=pts.device)
    tdm.extract_patches(tdf, dims, kg, patches)
    ref = tdm.Net(js, tdm.PREC_FP32, max_batch=256)
    ref.load_weights(wt)
    scales = net.calibrate(ref, patches)
    ref.close()
    return scales


def net_passes(torch, precision, batch, reps=3):
    """Descriptor forward only: `reps` passes of `batch` keypoints of the
    200k-point fragment's TDF through the autotuned net (for kernel traces and
    counters of the convolutions); per-layer ms per pass by HIP events."""
    import numpy as np

    import tdm
    from tdm import netdef

    cloud, keypts, weights = descriptor_workload(200_000, batch)
    tmp = tempfile.mkdtemp(prefix="tdm_net_")
    js, wt = os.path.join(tmp, "net.json"), os.path.join(tmp, "w.marvin")
    netdef.write_net_json(js)
    tdm.marvin_write(wt, weights)
    prec = {"fp32": tdm.PREC_FP32, "bf16": tdm.PREC_BF16, "fp16": tdm.PREC_FP16}[precision]
    pipe = DescriptorPipeline(torch, cloud, keypts, js, wt, prec, batch)
    pipe.step()  # TDF, keypoints and one warm pass
    torch.cuda.synchronize()
    pipe.net.set_profiling(True)
    for _ in range(reps):
        pipe.net.forward_grid(pipe.tdf, pipe.dims, pipe.kp_dev, pipe.feat)
    torch.cuda.synchronize()
    layers = {}
    for i, name in enumerate(pipe.net.layer_names):
        ms, passes, _ = pipe.net.layer_stats(i)
        layers[name] = ms / max(passes, 1)
    out = {"precision": precision, "batch": batch, "tile_variants": pipe.variants,
           "ms_per_pass": layers}
    pipe.net.close()
    return out


def time_steps(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def tsdf_bench(torch, reps=3):
    """TSDF fusion of 50 frames into 256^3: fused single pass and per-frame."""
    import numpy as np

    import tdm
    from tdm import synth

    mm, c2w = synth.depth_frames(50)
    depth = synth.depth_to_metres(mm)
    poses = synth.relative_poses(c2w)
    dev = torch.device("cuda:0")
    dims = synth.TSDF_DIMS
    n = dims[0] * dims[1] * dims[2]
    F, H, W = depth.shape
    K = torch.from_numpy(synth.intrinsics()).to(dev)
    P = torch.from_numpy(poses.reshape(F, 16)).to(dev)
    D = torch.from_numpy(depth).to(dev)
    tsdf = torch.empty(n, dtype=torch.float32, device=dev)
    weight = torch.empty(n, dtype=torch.float32, device=dev)
    args = (dims, synth.TSDF_ORIGIN, synth.TSDF_VOXEL, synth.TSDF_TRUNC, tsdf, weight)

    def fused():
        tdm.integrate_frames(K, P, D, *args)

    def per_frame():
        for f in range(F):
            tdm.integrate(K, P[f], D[f], *args)

    out = {}
    for name, fn in (("fused", fused), ("per_frame", per_frame)):
        times = []
        for _ in range(reps + 1):
            tdm.tsdf_init(tsdf, weight)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1))
        ms = float(np.median(times[1:]))
        w = weight.cpu().numpy()
        updates = float(w.sum())  # each update adds exactly 1
        touched = float(np.count_nonzero(w))
        if name == "fused":
            algo_bytes = 16.0 * touched + F * 4.0 * H * W
        else:
            algo_bytes = 16.0 * updates + F * 4.0 * H * W
        out[name] = {
            "ms": ms,
            "voxel_updates_per_s": n * F / (ms / 1e3),
            "applied_updates": updates,
            "hbm_gbs": algo_bytes / (ms / 1e3) / 1e9,
            "hbm_frac": algo_bytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
            "algorithmic_bytes": algo_bytes,
        }
    return out


def tdf_bench(torch, pts, reps=5):
    """TDF of the descriptor workload's fragment (occupancy scatter + the
    truncated distance transform, one call of tdm_tdf_from_points) in both
    decode modes, and the accurate point-distance TDF.  Algorithmic bytes:
    12 per point read + 4 per TDF voxel written."""
    import numpy as np

    import tdm

    vs, tr = np.float32(0.01), np.float32(0.05)
    dims, origin = tdm.grid_params(pts, vs, 15)
    nvox = dims[0] * dims[1] * dims[2]
    tdf = torch.empty(nvox, dtype=torch.float32, device=pts.device)
    algo = 12.0 * pts.shape[0] + 4.0 * nvox

    def timed(fn):
        times = []
        for _ in range(reps + 1):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1))
        return float(np.median(times[1:]))

    out = {"dims": dims, "points": int(pts.shape[0]), "voxels": nvox, "algorithmic_bytes": algo}
    for name, dec in (("exact_decode", tdm.DECODE_EXACT), ("reference_decode", tdm.DECODE_REFERENCE)):
        ms = timed(lambda: tdm.tdf_from_points(pts, dims, origin, vs, tr, tdf, decode=dec))
        out[name] = {"ms": ms, "gvoxels_per_s": nvox / (ms / 1e3) / 1e9,
                     "hbm_gbs": algo / (ms / 1e3) / 1e9,
                     "hbm_frac": algo / (ms / 1e3) / 1e9 / HBM_PEAK_GBS}
    ms = timed(lambda: tdm.tdf_accurate(pts, dims, origin, vs, tr, tdf))
    out["accurate"] = {"ms": ms, "gvoxels_per_s": nvox / (ms / 1e3) / 1e9,
                       "hbm_frac": algo / (ms / 1e3) / 1e9 / HBM_PEAK_GBS}
    return out


def fusion_io_bench(torch, n_frames=50):
    """The frame loop of depth-fusion/demo.cu:156-183 from PNG files on disk
    (50 synthetic frames into 256^3): the overlapped pipeline (tdm_fusion_*:
    decoder threads, pinned ring, u16 upload, GPU conversion, Integrate)
    against the serial loop (host decode, copy, Integrate, frame by frame).
    Frames/s of each, wall clock from the first frame to the grids ready."""
    import numpy as np

    import tdm
    from tdm import synth

    mm, c2w = synth.depth_frames(50, count=n_frames)
    poses = synth.relative_poses(c2w)
    K = synth.intrinsics()
    d = tempfile.mkdtemp(prefix="tdm_frames_")
    paths = []
    for i in range(n_frames):
        p = os.path.join(d, f"frame-{i:06d}.depth.png")
        synth.write_png16(p, mm[i])
        paths.append(p)
    dev = torch.device("cuda:0")
    dims = synth.TSDF_DIMS
    n = dims[0] * dims[1] * dims[2]
    args = (dims, synth.TSDF_ORIGIN, synth.TSDF_VOXEL, synth.TSDF_TRUNC)
    out = {}
    # serial: decode on the host, copy, integrate, one frame at a time
    t = torch.empty(n, device=dev)
    w = torch.empty(n, device=dev)
    dK = torch.from_numpy(K).to(dev)
    dP = torch.from_numpy(poses.reshape(-1, 16)).to(dev)
    tdm.tsdf_init(t, w)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, p in enumerate(paths):
        dD = torch.from_numpy(tdm.read_depth_png(p)).to(dev)
        tdm.integrate(dK, dP[i], dD, *args, t, w)
    torch.cuda.synchronize()
    serial = time.perf_counter() - t0
    # pipeline
    fus = tdm.Fusion(K, 480, 640, dims, synth.TSDF_ORIGIN, synth.TSDF_VOXEL, synth.TSDF_TRUNC,
                     threads=4)
    t0 = time.perf_counter()
    for p, P in zip(paths, poses):
        fus.add_png(p, P)
    fus.wait()
    pipe = time.perf_counter() - t0
    t2 = torch.empty(n, device=dev)
    w2 = torch.empty(n, device=dev)
    fus.copy_grids(t2, w2)
    torch.cuda.synchronize()
    fus.close()
    same = bool(torch.equal(t, t2) and torch.equal(w, w2))
    for p in paths:
        os.remove(p)
    os.rmdir(d)
    out = {"frames": n_frames, "serial_frames_per_s": n_frames / serial,
           "pipeline_frames_per_s": n_frames / pipe, "speedup": serial / pipe,
           "grids_identical": same, "decoder_threads": 4,
           "config": "50 synthetic 640x480 16-bit PNG depth frames from disk into 256^3"}
    return out


def tsdf_scenes_bench(torch, n_scenes=16, reps=3):
    """Batched TSDF: n_scenes synthetic scenes (sphere seeds 50..), 50 frames
    each, fused into n_scenes 256^3 grids in one launch."""
    import numpy as np

    import tdm
    from tdm import synth

    dev = torch.device("cuda:0")
    F = 50
    depths = []
    for sc in range(n_scenes):
        d, c2w = synth.depth_frames_torch(torch, F, scene_seed=50 + sc, noise_seed=100 + sc)
        depths.append(d)
    D = torch.stack(depths)
    poses = synth.relative_poses(c2w)  # the camera circle is the same in every scene
    P = torch.from_numpy(np.ascontiguousarray(np.broadcast_to(
        poses.reshape(1, F, 16), (n_scenes, F, 16)))).to(dev)
    origins = torch.from_numpy(np.tile(synth.TSDF_ORIGIN, (n_scenes, 1))).to(dev)
    K = torch.from_numpy(synth.intrinsics()).to(dev)
    dims = synth.TSDF_DIMS
    n = dims[0] * dims[1] * dims[2]
    tsdf = torch.empty(n_scenes * n, device=dev)
    weight = torch.empty(n_scenes * n, device=dev)
    times = []
    for _ in range(reps + 1):
        tdm.tsdf_init(tsdf, weight)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        tdm.integrate_scenes(K, P, D, dims, origins, synth.TSDF_VOXEL, synth.TSDF_TRUNC, tsdf,
                             weight)
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    ms = float(np.median(times[1:]))
    touched = float((weight > 0).sum().item())
    H, W = D.shape[-2:]
    algo = 16.0 * touched + n_scenes * F * 4.0 * H * W
    return {"scenes": n_scenes, "frames": F, "ms": ms,
            "voxel_updates_per_s": n_scenes * n * F / (ms / 1e3),
            "applied_updates": float(weight.sum().item()),
            "hbm_gbs": algo / (ms / 1e3) / 1e9,
            "config": "16 scenes x 50 frames, 640x480, into 16 grids of 256^3 in one launch"}


def mutual_nn(torch, a, b):
    """Exact nearest neighbours of a's rows in b and the mutual mask, with the
    library's matcher (the knnsearch pair of core/demo.m:49-55)."""
    import tdm

    ab = torch.empty(len(a), dtype=torch.int32, device=a.device)
    ba = torch.empty(len(b), dtype=torch.int32, device=a.device)
    tdm.match_nn(a.contiguous(), b.contiguous(), ab, ba)
    ab, ba = ab.long(), ba.long()
    idx = torch.arange(len(a), device=a.device)
    mutual = ba[ab] == idx
    return ab, mutual


def _nn_margins(torch, a, b):
    """Per row of a: nearest and second-nearest Euclidean distances to b's rows
    (fp64 on the device, exact enough to rank the fp32 descriptors)."""
    a64, b64 = a.double(), b.double()
    d2 = (a64 * a64).sum(1, keepdim=True) + (b64 * b64).sum(1)[None, :] - 2.0 * a64 @ b64.T
    top = torch.topk(d2, 2, dim=1, largest=False).values.clamp_min(0).sqrt()
    return top[:, 0], top[:, 1]


def precision_study(torch, js, wt, weights, n_frag=20, n_kp=5000, layer_patches=8):
    """Precision form (20 fragments x 5000 keypoints = 100k patches): for each
    precision, descriptor rel-L2 vs the fp32 path and mutual-NN agreement with
    fp32 on the fragment pairs (0,1), (2,3), ..., (18,19); the fp32 floor (the
    same agreement between two fp32 nets whose split-K summation orders
    differ); the NN-margin distribution of pair (0,1) with each precision's
    disagreement per margin bucket; per-layer rel-L2 vs the FP64 oracle on a
    few patches."""
    import numpy as np

    import tdm
    from tdm import netdef, synth

    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    dev = torch.device("cuda:0")
    names = {tdm.PREC_FP32: "fp32", tdm.PREC_BF16: "bf16", tdm.PREC_FP16: "fp16",
             tdm.PREC_FP8: "fp8"}
    nets = {p: tdm.Net(js, p, max_batch=1024) for p in names}
    # fp32 with another summation order: split-K counts follow max_batch
    nets["fp32_alt"] = tdm.Net(js, tdm.PREC_FP32, max_batch=64)
    for n in nets.values():
        n.load_weights(wt)
    c0 = synth.fragment(200_000, view=0, seed=10)
    calibrate_fp8(torch, nets[tdm.PREC_FP8], js, wt, torch.from_numpy(c0).to(dev),
                  c0[synth.keypoint_indices(len(c0), 256, seed=999)])
    desc = {p: [] for p in nets}
    first_patches = None
    for f in range(n_frag):
        cloud = synth.fragment(200_000, view=f, seed=10 + f)
        pts = torch.from_numpy(cloud).to(dev)
        dims, origin = tdm.grid_params(pts, 0.01, 15)
        tdf = torch.empty(dims[2], dims[1], dims[0], device=dev)
        tdm.tdf_from_points(pts, dims, origin, np.float32(0.01), np.float32(0.05), tdf)
        kp = cloud[synth.keypoint_indices(len(cloud), n_kp, seed=100 + f)]
        kg = tdm.keypoints_to_grid(kp, origin, np.float32(0.01))
        kgd = torch.from_numpy(kg).to(dev)
        for p, net in nets.items():
            out = torch.empty(n_kp, 512, device=dev)
            net.forward_grid(tdf, dims, kgd, out)
            desc[p].append(out)
        if f == 0:
            patches = torch.empty(layer_patches, 30, 30, 30, device=dev)
            tdm.extract_patches(tdf, dims, kgd[:layer_patches], patches)
            first_patches = patches
        del tdf
    torch.cuda.synchronize()
    d32 = torch.cat(desc[tdm.PREC_FP32])
    pairs = [(2 * k, 2 * k + 1) for k in range(n_frag // 2)]
    ref_nn = [mutual_nn(torch, desc[tdm.PREC_FP32][i], desc[tdm.PREC_FP32][j]) for i, j in pairs]
    # margins of pair (0, 1) in fp32: relative gap (d2 - d1) / d1 per query
    d1, d2 = _nn_margins(torch, desc[tdm.PREC_FP32][0], desc[tdm.PREC_FP32][1])
    gap = ((d2 - d1) / d1.clamp_min(1e-30)).cpu().numpy()
    edges = [0.0, 1e-4, 1e-3, 3e-3, 1e-2, 3e-2, 1e-1, np.inf]
    bucket = np.digitize(gap, edges[1:-1])
    vs_fp32 = {}
    for p in list(names)[1:] + ["fp32_alt"]:
        name = names.get(p, p)
        dp = torch.cat(desc[p])
        agree, nn_agree = [], []
        for (i, j), (ab32, m32) in zip(pairs, ref_nn):
            abp, mp = mutual_nn(torch, desc[p][i], desc[p][j])
            agree.append(float((m32 & mp & (ab32 == abp)).sum()) / max(int(m32.sum()), 1))
            nn_agree.append(float((ab32 == abp).float().mean()))
        same = (ref_nn[0][0] == mutual_nn(torch, desc[p][0], desc[p][1])[0]).cpu().numpy()
        pert = (torch.linalg.norm(desc[p][0] - desc[tdm.PREC_FP32][0], dim=1) /
                torch.linalg.norm(desc[tdm.PREC_FP32][0], dim=1)).cpu().numpy()
        vs_fp32[name] = {
            "descriptor_rel_l2": float(torch.linalg.norm(dp - d32) / torch.linalg.norm(d32)),
            "median_per_descriptor_rel_err": float(np.median(pert)),
            "mutual_nn_agreement_mean": float(np.mean(agree)),
            "mutual_nn_agreement_min": float(np.min(agree)),
            "mutual_nn_agreement_pair01": agree[0],
            "nn_agreement_mean": float(np.mean(nn_agree)),
            "nn_disagreement_by_margin_pair01": [
                float(1.0 - same[bucket == b].mean()) if np.any(bucket == b) else None
                for b in range(len(edges) - 1)],
        }
    margins = {
        "relative_gap_bucket_edges": [e if np.isfinite(e) else "inf" for e in edges],
        "queries_per_bucket": [int(np.sum(bucket == b)) for b in range(len(edges) - 1)],
        "median_relative_gap": float(np.median(gap)),
    }
    shapes = netdef.layer_shapes()
    p_host = first_patches.cpu().numpy()
    ref = oracle.net_forward_fp64(p_host, dict(weights), shapes)
    per_layer = {}
    for p, name in names.items():
        errs = []
        for i, (_n, _ci, co, _k, _d, od, pooled) in enumerate(shapes):
            e = pooled if pooled is not None else od
            out = torch.empty(layer_patches, co, e, e, e, device=dev)
            nets[p].forward_layer(first_patches, i, out)
            got = out.cpu().numpy().astype(np.float64)
            errs.append(float(np.linalg.norm(got - ref[i]) / np.linalg.norm(ref[i])))
        per_layer[name] = dict(zip([s[0] for s in shapes], errs))
    for n in nets.values():
        n.close()
    return {
        "fragments": n_frag, "keypoints_per_fragment": n_kp, "patches": n_frag * n_kp,
        "pairs": pairs,
        "vs_fp32": vs_fp32,
        "fp32_floor": "fp32_alt: the fp32 net at max_batch 64 (other split-K summation order)",
        "mutual_nn_pairs_fp32_pair01": int(ref_nn[0][1].sum()),
        "nn_margins_fp32_pair01": margins,
        "per_layer_rel_l2_vs_fp64": per_layer,
        "per_layer_patches": layer_patches,
    }


def _rot_err_deg(Ra, Rb):
    import numpy as np

    c = (np.trace(Ra.T @ Rb) - 1.0) / 2.0
    return float(np.degrees(np.arccos(np.clip(c, -1.0, 1.0))))


def registration_kernels(torch, js, wt, n_kp=5000, trials=50000):
    """Matching and RANSAC on one request's real descriptors: mutual NN of
    two 5000 x 512 sets (one 25.6 GFLOP fp32 distance GEMM on MFMA plus the
    top-2 epilogue) and RANSAC with 50k hypotheses, timed with HIP events."""
    import numpy as np

    import tdm
    from tdm import synth

    dev = torch.device("cuda:0")
    net = tdm.Net(js, tdm.PREC_FP16, max_batch=1024)
    net.load_weights(wt)
    p1, p2, R, t = synth.fragment_pair(7, 200_000, 200_000)
    feats, kps = [], []
    rng = np.random.default_rng(8)
    for pts in (p1, p2):
        dp = torch.from_numpy(pts).to(dev)
        dims, origin = tdm.grid_params(dp, 0.01, 15)
        tdf = torch.empty(dims[2], dims[1], dims[0], device=dev)
        tdm.tdf_from_points(dp, dims, origin, np.float32(0.01), np.float32(0.05), tdf)
        kidx = rng.integers(0, len(pts), n_kp)
        kg = tdm.keypoints_to_grid(pts[kidx], origin, np.float32(0.01))
        f = torch.empty(n_kp, 512, device=dev)
        net.forward_grid(tdf, dims, torch.from_numpy(kg).to(dev), f)
        feats.append(f)
        kps.append(torch.from_numpy(np.ascontiguousarray(pts[kidx])).to(dev))
    net.close()
    ab = torch.empty(n_kp, dtype=torch.int32, device=dev)
    ba = torch.empty(n_kp, dtype=torch.int32, device=dev)
    refined = torch.zeros(2, dtype=torch.int32, device=dev)
    pairs = torch.empty((n_kp, 2), dtype=torch.int32, device=dev)
    npairs = torch.zeros(1, dtype=torch.int32, device=dev)
    rt = torch.zeros(12, dtype=torch.float64, device=dev)
    inl = torch.empty(n_kp, dtype=torch.int32, device=dev)
    stats = torch.zeros(4, dtype=torch.int32, device=dev)

    def match():
        tdm.match_nn(feats[0], feats[1], ab, ba, refined=refined)
        tdm.mutual_pairs(ab, ba, pairs, npairs)

    def ransac():
        tdm.ransac_rt(kps[0], kps[1], rt, inl, stats, pairs=pairs, num_pairs=npairs,
                      max_pairs=n_kp, threshold=0.05, max_trials=trials, seed=1)

    out = {}
    for name, fn in (("match_ms", match), ("ransac_ms", ransac)):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 10
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        out[name] = e0.elapsed_time(e1) / reps
    flops = 2.0 * n_kp * n_kp * 512
    st = stats.cpu().numpy()
    out.update({
        "keypoints": n_kp, "dim": 512,
        "distance_gemm_tflops": flops / (out["match_ms"] / 1e3) / 1e12,
        "refined_queries": refined.cpu().numpy().tolist(),
        "mutual_pairs": int(npairs.item()),
        "ransac_trials_cap": trials, "ransac_trials_run": int(st[2]),
        "ransac_inliers": int(st[0]),
    })
    return out


def serving_bench(torch, js, wt, rates=(2.0, 4.0, 8.0), per_rate=16, unloaded=8, prec="bf16",
                  slots=3):
    """Serving form: fragment-pair registration requests (points U[20k, 500k],
    50-80 % overlap, rigid motion <= 60 deg / 1 m, 5000 keypoints each, RANSAC
    50k trials at 5 cm) through the continuous-batching server with a `prec`
    net and `slots` request streams (3 + the net's stream = the 4 hardware
    queues HIP gives a process by default).  Unloaded latency: one request at
    a time; then Poisson arrivals at each rate (no sweep when rates is
    empty).  Motion errors are against the generating motion; they are not a
    metric here: the weights are random He-normal, not the trained ones, so
    recovery is not expected."""
    import threading

    import numpy as np

    import tdm
    from tdm import synth

    pid = {"bf16": tdm.PREC_BF16, "fp16": tdm.PREC_FP16}[prec]
    net = tdm.Net(js, pid, max_batch=1024)
    net.load_weights(wt)
    net.autotune(1024)
    n_req = unloaded + per_rate * len(rates)
    reqs = [synth.fragment_pair(70 + k) for k in range(n_req)]
    srv = tdm.Server(net, memory_budget=24 << 30, max_inflight=slots, keypoints=5000,
                     max_pass_keypoints=20000, ransac_max_trials=50000)

    def summarize(results, wall):
        lat = np.array([r["latency_ms"] for r in results])
        return {"requests": len(results), "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
                "requests_per_s": len(results) / wall,
                "mean_queue_ms": float(np.mean([r["queue_ms"] for r in results]))}

    # warm-up (first-use allocations)
    p1, p2, _, _ = reqs[0]
    srv.wait(srv.submit(p1, p2, seed=1))
    out = {}
    res, errs = [], []
    t0 = time.perf_counter()
    for k in range(unloaded):
        p1, p2, R, t = reqs[k]
        r = srv.wait(srv.submit(p1, p2, seed=100 + k))
        res.append(r)
        errs.append((_rot_err_deg(r["rt"][:, :3], R), float(np.linalg.norm(r["rt"][:, 3] - t))))
    out["unloaded"] = summarize(res, time.perf_counter() - t0)
    sweep = []
    rng = np.random.default_rng(9)
    base = unloaded
    for rate in rates:
        batch = reqs[base:base + per_rate]
        base += per_rate
        gaps = rng.exponential(1.0 / rate, len(batch))
        tickets = [None] * len(batch)

        def submitter():
            start = time.perf_counter()
            at = 0.0
            for i, (p1, p2, _, _) in enumerate(batch):
                at += gaps[i]
                delay = start + at - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                tickets[i] = srv.submit(p1, p2, seed=1000 + i)

        t0 = time.perf_counter()
        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        results = [srv.wait(tk) for tk in tickets]
        wall = time.perf_counter() - t0
        s = summarize(results, wall)
        s["arrival_rate"] = rate
        s["mean_pass_keypoints"] = float(np.mean([r["pass_keypoints"] for r in results]))
        for r, (_, _, R, t) in zip(results, batch):
            errs.append((_rot_err_deg(r["rt"][:, :3], R), float(np.linalg.norm(r["rt"][:, 3] - t))))
        sweep.append(s)
        log(f"serving rate {rate}/s: p50 {s['p50_ms']:.1f} ms p99 {s['p99_ms']:.1f} ms")
    out["poisson"] = sweep
    st = srv.stats()
    srv.close()
    net.close()
    e = np.array(errs)
    out["server_stats"] = st
    out["motion_recovery_not_a_metric"] = {
        "note": ("random He-normal weights, not the trained 3DMatch weights (which cannot be "
                 "downloaded here): recovery only shows the pipeline runs end to end"),
        "rotation_error_deg_median": float(np.median(e[:, 0])),
        "translation_error_m_median": float(np.median(e[:, 1])),
        "recovered_fraction": float(np.mean((e[:, 0] < 5.0) & (e[:, 1] < 0.1))),
    }
    out["config"] = ("fragment pairs of the synthetic scene, points U[20k,500k], overlap "
                     "U[0.5,0.8], rotation <= 60 deg, shift <= 1 m, 5000 keypoints per "
                     f"fragment, {prec} net, RANSAC 50k trials / 0.05 m, {slots} slots, "
                     "24 GiB budget")
    return out


def batched_bench(torch, js, wt, n_frag=64, seed=40, replays=2, prec="bf16"):
    """Batched form: n_frag fragments (points U[20k, 500k], extents U[1, 4] m,
    keypoints U[1000, 10000]) -> one TDF plan (four launches for all grids)
    -> one multi-grid `prec` descriptor pass over every keypoint, the whole
    batch captured in a hipGraph and replayed.  Reports descriptors/s of a
    replay and of the same work issued eagerly."""
    import numpy as np

    import tdm
    from tdm import synth

    dev = torch.device("cuda:0")
    rng = np.random.default_rng(seed)
    clouds, kps, gids = [], [], []
    total_pts = 0
    for f in range(n_frag):
        n = int(rng.integers(20_000, 500_001))
        cloud = synth.cropped_fragment(seed * 1000 + f, n, float(rng.uniform(1.0, 4.0)))
        pts = torch.from_numpy(cloud).to(dev)
        dims, origin = tdm.grid_params(pts, 0.01, 15)
        tdf = torch.empty(dims[2], dims[1], dims[0], device=dev)
        clouds.append((pts, tdf, dims, origin))
        k = int(rng.integers(1000, 10_001))
        idx = rng.integers(0, n, k)
        kps.append(tdm.keypoints_to_grid(cloud[idx], origin, np.float32(0.01)))
        gids.append(np.full(k, f, np.int32))
        total_pts += n
    kp = torch.from_numpy(np.concatenate(kps)).to(dev)
    gid = torch.from_numpy(np.concatenate(gids)).to(dev)
    n_kp = len(kp)
    plan = tdm.TdfPlan(clouds)
    table = tdm.grid_table([(tdf, dims) for _, tdf, dims, _ in clouds])
    net = tdm.Net(js, {"bf16": tdm.PREC_BF16, "fp16": tdm.PREC_FP16}[prec], max_batch=1024)
    net.load_weights(wt)
    variants = net.autotune(1024)
    feat = torch.empty(n_kp, 512, device=dev)

    def step():
        plan.run()
        net.forward_grids(table, kp, gid, feat)

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step()
    torch.cuda.synchronize()
    eager = time.perf_counter() - t0
    want = feat.clone()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            step()
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    same = bool(torch.equal(feat, want))
    t0 = time.perf_counter()
    for _ in range(replays):
        g.replay()
    torch.cuda.synchronize()
    graph = (time.perf_counter() - t0) / replays
    # TDF part alone (plan run) for its own rate
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    plan.run()
    e1.record()
    torch.cuda.synchronize()
    tdf_ms = e0.elapsed_time(e1)
    voxels = sum(d[0] * d[1] * d[2] for _, _, d, _ in clouds)
    plan.close()
    # the same plan with the reference's float32 voxel decode
    plan = tdm.TdfPlan(clouds, 0.01, 0.05, tdm.DECODE_REFERENCE)
    plan.run()
    e0.record()
    plan.run()
    e1.record()
    torch.cuda.synchronize()
    tdf_ref_ms = e0.elapsed_time(e1)
    plan.close()
    net.close()
    return {"fragments": n_frag, "points": total_pts, "keypoints": n_kp, "voxels": voxels,
            "eager_s": eager, "graph_s": graph,
            "descriptors_per_s_graph": n_kp / graph, "descriptors_per_s_eager": n_kp / eager,
            "graph_equals_eager": same, "tdf_plan_ms": tdf_ms,
            "tdf_plan_reference_decode_ms": tdf_ref_ms,
            "tdf_gvoxels_per_s": voxels / (tdf_ms / 1e3) / 1e9, "tile_variants": variants,
            "config": ("64 fragments, points U[20k,500k], extents U[1,4] m, keypoints "
                       f"U[1000,10000], {prec} net, TDF plan + multi-grid pass in one hipGraph")}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _omp_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_baseline(cloud, keypts, weights, n_kp, conv_patches=250):
    """Oracle (C loop nests, OpenMP) on the descriptor workload: occupancy and
    the reference's ComputeTDF body (float32 decode, core/demo.cu:22-57) over
    the whole grid, and `conv_patches` patches (5 batches of the reference's
    50) through the fp32 net, scaled to the 5000 keypoints."""
    import numpy as np

    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    from tdm import netdef

    vs, trunc = np.float32(0.01), np.float32(0.05)
    dims, origin = oracle.grid_params(cloud, vs, 15)
    t0 = time.perf_counter()
    occ = oracle.occupancy(cloud, dims, origin, vs)
    t_occ = time.perf_counter() - t0
    nvox = dims[0] * dims[1] * dims[2]
    t0 = time.perf_counter()
    tdf = oracle.compute_tdf(occ, vs, trunc, "reference")
    t_tdf = time.perf_counter() - t0
    wdict = dict(weights)
    shapes = netdef.layer_shapes()
    patches = oracle.extract_patches(tdf, oracle.keypoint_grid(keypts[:conv_patches], origin, vs))
    t0 = time.perf_counter()
    oracle.net_forward_fp32(patches, wdict, shapes)
    t_conv = (time.perf_counter() - t0) / conv_patches * n_kp
    total = t_occ + t_tdf + t_conv
    threads = _omp_threads()
    return {
        "value": n_kp / total,
        "unit": "descriptors/s",
        "cores": threads,
        "kind": "port",
        "cpu": cpu_model(),
        "sample": (f"oracle C loop nests (gcc -O3 -fopenmp -ffp-contract=off, {threads} threads, "
                   f"{cpu_model()}): occupancy of all {len(cloud)} points and the reference's "
                   f"ComputeTDF body over all {nvox} voxels, measured; fp32 net on {conv_patches} "
                   f"patches (x{n_kp / conv_patches:.0f}); seconds: occupancy {t_occ:.2f}, "
                   f"tdf {t_tdf:.1f}, net {t_conv:.1f} (scaled)"),
        "tdf_voxels_per_s": nvox / t_tdf,
    }


def cpu_baseline_tsdf():
    """The oracle's Integrate (the reference kernel body, OpenMP over (z, y)
    columns) on the TSDF workload: 50 frames into 256^3, whole, measured."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    from tdm import synth

    mm, c2w = synth.depth_frames(50)
    depth = synth.depth_to_metres(mm)
    poses = synth.relative_poses(c2w)
    K = synth.intrinsics()
    dims = synth.TSDF_DIMS
    t, w = oracle.new_grid(dims)
    t0 = time.perf_counter()
    for f in range(len(depth)):
        oracle.integrate(K, poses[f], depth[f], dims, synth.TSDF_ORIGIN, synth.TSDF_VOXEL,
                         synth.TSDF_TRUNC, t, w)
    sec = time.perf_counter() - t0
    n = dims[0] * dims[1] * dims[2]
    threads = _omp_threads()
    return {"value": n * len(depth) / sec, "unit": "voxel-updates/s", "cores": threads,
            "kind": "port", "cpu": cpu_model(),
            "sample": (f"oracle orc_integrate (gcc -O3 -fopenmp -ffp-contract=off, {threads} "
                       f"threads): all 50 frames into 256^3, {sec:.2f} s measured")}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--precision", choices=["fp32", "bf16", "fp16"], default="fp32",
                    help="precision of the headline pipeline (default fp32, the plain path)")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the serving sweep (e.g. under a profiler)")
    ap.add_argument("--only", nargs="+", choices=["tsdf", "tdf", "net", "batched"],
                    help="run only these workloads and print their figures (for counter runs); "
                         "net: three descriptor passes of --batch keypoints at --precision; "
                         "batched: the 64-fragment batched form in bf16 (hipGraph)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy 3 times (reference for HBM counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 4, "copies": 3}))
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(the workloads are sized for one MI355X); run with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import tdm
    from tdm import netdef

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    tdm.lib()
    if args.only:
        out = {}
        if "tdf" in args.only:
            log("tdf")
            cloud, _, _ = descriptor_workload(200_000, 5000)
            out["tdf"] = tdf_bench(torch, torch.from_numpy(cloud).to("cuda:0"))
        if "tsdf" in args.only:
            log("tsdf fusion")
            out["tsdf"] = tsdf_bench(torch)
        if "net" in args.only:
            log(f"descriptor passes, {args.precision}")
            out["net"] = net_passes(torch, args.precision, args.batch)
        if "batched" in args.only:
            log("batched form (hipGraph), bf16")
            tmp = tempfile.mkdtemp(prefix="tdm_bench_")
            js, wt = os.path.join(tmp, "net.json"), os.path.join(tmp, "w.marvin")
            netdef.write_net_json(js)
            tdm.marvin_write(wt, netdef.he_normal_weights(5))
            out["batched"] = batched_bench(torch, js, wt, replays=1)
        print(json.dumps(out), flush=True)
        return 0
    log("building inputs")
    n_kp = 5000
    cloud, keypts, weights = descriptor_workload(200_000, n_kp)
    tmp = tempfile.mkdtemp(prefix="tdm_bench_")
    js, wt = os.path.join(tmp, "net.json"), os.path.join(tmp, "w.marvin")
    netdef.write_net_json(js)
    tdm.marvin_write(wt, weights)

    prec_id = {"fp32": tdm.PREC_FP32, "bf16": tdm.PREC_BF16, "fp16": tdm.PREC_FP16}[args.precision]
    pipe = DescriptorPipeline(torch, cloud, keypts, js, wt, prec_id, args.batch)
    log(f"{args.precision} descriptor pipeline: grid {pipe.dims}, batch {args.batch}")
    pipe.net.set_profiling(True)
    sec = time_steps(torch, pipe.step, args.steps, args.warmup)
    value = n_kp / sec
    shapes = netdef.layer_shapes()
    layers = {}
    for i, (name, cin, cout, k, _d, od, _p) in enumerate(shapes):
        ms, passes, patches = pipe.net.layer_stats(i)
        flops = 2 * od ** 3 * cout * cin * k ** 3
        layers[name] = {"ms_per_pass": ms / max(passes, 1), "passes": passes,
                        "tflops": flops * patches / (ms / 1e3) / 1e12 if ms > 0 else None}
    ms2, passes2, patches2 = pipe.net.layer_stats(1)
    conv2_flops = 2 * 26 ** 3 * 64 * 64 * 27
    achieved = conv2_flops * patches2 / (ms2 / 1e3) / 1e12
    pipe.net.set_profiling(False)
    log(f"{args.precision}: {value:.1f} descriptors/s, conv2 {achieved:.1f} TFLOP/s")

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "descriptors/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": {
            "workload": (f"point cloud -> TDF -> 3DMatch descriptors, {args.precision}: 200k-point synthetic "
                         "fragment, 0.01 m voxels, truncation 5 voxels, TDF with the exact "
                         "integer voxel decode (the reference's float32 decode misplaces voxels "
                         "on this 46.7 M-voxel grid), 5000 keypoints, 30^3 patches, He-normal "
                         f"weights, passes of {args.batch} patches"),
            "decode": "exact",
            "points": len(cloud), "keypoints": n_kp, "grid": pipe.dims, "batch": args.batch,
            "tile_variants": pipe.variants,
        },
        "roofline": {
            "kernel": (f"conv2: {args.precision} MFMA implicit GEMM 64->64 ch, 28^3->26^3, "
                       "fused bias+ReLU+2^3 max pool"),
            "bound": "mfma",
            "achieved": achieved,
            "peak": MFMA_PEAK_TFLOPS[args.precision],
            "unit": "TFLOP/s",
            "frac": achieved / MFMA_PEAK_TFLOPS[args.precision],
            "traffic": measured_traffic(args.precision, patches2 / max(passes2, 1)),
            "traffic_source": "profiles/hbm_traffic_descriptor_pipeline_<precision>.md, scaled to the average patches per launch",
            "flop_per_launch": conv2_flops * (patches2 / max(passes2, 1)),
            "avg_launch_ms": ms2 / max(passes2, 1),
        },
        "layers": layers,
    }
    pipe.net.close()
    del pipe
    torch.cuda.empty_cache()

    if not args.no_extras:
        # the reference's own batch (core/demo.cu:190,195: 50 patches per forward)
        log("fp32 descriptor pipeline, batch 50")
        p50 = DescriptorPipeline(torch, cloud, keypts, js, wt, tdm.PREC_FP32, 50)
        s50 = time_steps(torch, p50.step, args.steps, args.warmup)
        result["fp32_batch50"] = {"descriptors_per_s": n_kp / s50, "ms_per_step": s50 * 1e3,
                                  "tile_variants": p50.variants,
                                  "config": "the headline workload in passes of 50 patches"}
        p50.net.close()
        del p50
        for prec, name in ((tdm.PREC_BF16, "bf16"), (tdm.PREC_FP16, "fp16"),
                           (tdm.PREC_FP8, "fp8")):
            log(f"{name} descriptor pipeline")
            pb = DescriptorPipeline(torch, cloud, keypts, js, wt, prec, args.batch)
            pb.net.set_profiling(True)
            sb = time_steps(torch, pb.step, args.steps, args.warmup)
            ms2b, passes2b, patches2b = pb.net.layer_stats(1)
            tf = conv2_flops * patches2b / (ms2b / 1e3) / 1e12
            result[name] = {
                "descriptors_per_s": n_kp / sb,
                "ms_per_step": sb * 1e3,
                "speedup_vs_fp32": (n_kp / sb) / value,
                "conv2_tflops": tf,
                "conv2_frac_of_peak": tf / MFMA_PEAK_TFLOPS[name],
                "tile_variants": pb.variants,
                "layer_tflops": layer_tflops(pb.net),
            }
            pb.net.close()
            del pb
            torch.cuda.empty_cache()
        log("precision study")
        result["precision"] = precision_study(torch, js, wt, weights)
        torch.cuda.empty_cache()
        log("tdf")
        result["tdf"] = tdf_bench(torch, torch.from_numpy(cloud).to("cuda:0"))
        result["tdf"]["config"] = "the 200k-point fragment of the descriptor workload, 0.01 m voxels, 5-voxel truncation, 15-voxel padding"
        log("tsdf fusion")
        result["tsdf"] = tsdf_bench(torch)
        result["tsdf"]["config"] = "50 synthetic 640x480 depth frames into a 256^3 grid, 0.01 m voxels, truncation 0.05 m"
        log("tsdf scenes batch")
        result["tsdf"]["scenes_batch"] = tsdf_scenes_bench(torch)
        torch.cuda.empty_cache()
        log("fusion from PNG files: pipeline vs serial loop")
        result["tsdf"]["fusion_io"] = fusion_io_bench(torch)
        torch.cuda.empty_cache()
        log("registration kernels")
        result["registration_kernels"] = registration_kernels(torch, js, wt)
        for prec in ("bf16", "fp16"):
            log(f"batched form (hipGraph), {prec}")
            result["batched" if prec == "bf16" else "batched_fp16"] = batched_bench(
                torch, js, wt, prec=prec)
            torch.cuda.empty_cache()
        if not args.no_serving:
            log("serving, bf16")
            result["serving"] = serving_bench(torch, js, wt, prec="bf16")
            result["registration_latency_ms"] = {
                "p50": result["serving"]["unloaded"]["p50_ms"],
                "p99": result["serving"]["unloaded"]["p99_ms"], "precision": "bf16"}
            log("serving, fp16 (unloaded only)")
            result["serving_fp16"] = serving_bench(torch, js, wt, rates=(), prec="fp16")

    if not args.no_cpu_baseline:
        log("cpu baseline (oracle loop nests)")
        result["cpu_baseline"] = cpu_baseline(cloud, keypts, weights, n_kp)
        log("cpu baseline, TSDF (oracle Integrate)")
        result["cpu_baseline_tsdf"] = cpu_baseline_tsdf()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
