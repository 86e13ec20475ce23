"""GPU parity of the descriptor network (Net::forward, core/marvin.hpp:7653-7679)
against the FP64 oracle (torch CPU float64 conv3d/relu/max_pool3d).

Tolerances (relative L2 over each layer's whole output, NCDHW):
  fp32 path  <= 1e-5 per layer and for the descriptors;
  bf16 path  <= 3e-2 per layer, <= 3e-2 for the descriptors;
  fp16 path  <= 1e-2 per layer and for the descriptors;
  fp8 path   <= 1.5e-1 per layer and for the descriptors (e4m3 with per-channel
             weight scales and per-layer activation scales calibrated from the
             fp32 net on the same patches).
"""
import os

import numpy as np
import pytest

import oracle
import tdm
from conftest import rel_l2
from tdm import netdef, synth

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5
BF16_TOL = 3e-2
FP16_TOL = 1e-2
FP8_TOL = 1.5e-1  # e4m3: 3 mantissa bits, per-channel weight / per-layer activation scales


@pytest.fixture(scope="module")
def net_files(tmp_path_factory):
    d = tmp_path_factory.mktemp("net")
    js = str(d / "net.json")
    wt = str(d / "weights.marvin")
    netdef.write_net_json(js)
    tensors = netdef.he_normal_weights(5)
    # non-zero biases exercise the bias epilogue
    rng = np.random.default_rng(9)
    tensors = [(n, a if n.endswith(".weight") else rng.normal(0, 0.05, a.shape).astype(np.float32))
               for n, a in tensors]
    tdm.marvin_write(wt, tensors)
    return js, wt, dict(tensors)


@pytest.fixture(scope="module")
def patches():
    """TDF patches around surface keypoints of a synthetic fragment."""
    cloud = synth.fragment(60_000, seed=21)
    vs = np.float32(0.01)
    dims, origin = oracle.grid_params(cloud, vs, 15)
    occ = oracle.occupancy(cloud, dims, origin, vs)
    tdf = oracle.compute_tdf(occ, vs, np.float32(0.05))
    idx = synth.keypoint_indices(len(cloud), 6, seed=22)
    kg = oracle.keypoint_grid(cloud[idx], origin, vs)
    return oracle.extract_patches(tdf, kg), tdf, dims, kg


@pytest.fixture(scope="module")
def fp64(net_files, patches):
    _, _, w = net_files
    return oracle.net_forward_fp64(patches[0], w, netdef.layer_shapes())


def _run_layers(torch, net, p):
    shapes = netdef.layer_shapes()
    data = torch.from_numpy(p).to("cuda:0")
    outs = []
    for i, (_n, _ci, co, _k, _d, od, pooled) in enumerate(shapes):
        e = pooled if pooled is not None else od
        out = torch.empty(len(p), co, e, e, e, dtype=torch.float32, device="cuda:0")
        net.forward_layer(data, i, out)
        outs.append(out.cpu().numpy())
    return outs


PRECISIONS = [(tdm.PREC_FP32, FP32_TOL), (tdm.PREC_BF16, BF16_TOL), (tdm.PREC_FP16, FP16_TOL),
              (tdm.PREC_FP8, FP8_TOL)]


def _make_net(torch, js, wt, prec, calib, max_batch=0):
    """A net of `prec`; fp8 nets are calibrated from an fp32 net on `calib`."""
    net = tdm.Net(js, prec, max_batch=max_batch)
    net.load_weights(wt)
    if prec == tdm.PREC_FP8:
        ref = tdm.Net(js, tdm.PREC_FP32, max_batch=max_batch)
        ref.load_weights(wt)
        net.calibrate(ref, torch.from_numpy(np.ascontiguousarray(calib)).to("cuda:0"))
        ref.close()
    return net


@pytest.mark.parametrize("prec,tol", PRECISIONS)
def test_per_layer_vs_fp64(torch_gpu, net_files, patches, fp64, prec, tol):
    js, wt, _ = net_files
    net = _make_net(torch_gpu, js, wt, prec, patches[0], max_batch=4)
    outs = _run_layers(torch_gpu, net, patches[0])
    errs = [rel_l2(o, r) for o, r in zip(outs, fp64)]
    print("per-layer rel-L2:", ["%.2e" % e for e in errs], flush=True)
    for name, e in zip(net.layer_names, errs):
        assert e <= tol, f"{name}: rel-L2 {e:.3e} > {tol}"


@pytest.mark.parametrize("prec,tol", PRECISIONS)
def test_descriptors_vs_fp64(torch_gpu, net_files, patches, fp64, prec, tol):
    js, wt, _ = net_files
    net = _make_net(torch_gpu, js, wt, prec, patches[0])
    p = patches[0]
    data = torch_gpu.from_numpy(p).to("cuda:0")
    feat = torch_gpu.empty(len(p), 512, dtype=torch_gpu.float32, device="cuda:0")
    net.forward(data, feat)
    got = feat.cpu().numpy()
    e = rel_l2(got, fp64[-1].reshape(len(p), 512))
    print(f"descriptor rel-L2 {e:.3e}", flush=True)
    assert e <= tol


def test_grid_path_equals_patch_path(torch_gpu, net_files, patches):
    """Patch gather fused into conv1 gives bit-identical descriptors, also for
    keypoints whose patch reaches beyond the grid (zeros outside, the gather's
    bounds-checked path)."""
    js, wt, _ = net_files
    p0, tdf, dims, kg0 = patches
    border = np.array([[0, 0, 0], [dims[0] - 1, dims[1] - 1, dims[2] - 1], [14, 15, 16]], np.int32)
    kg = np.ascontiguousarray(np.concatenate([kg0, border]))
    p = oracle.extract_patches(tdf, kg)
    for prec in (tdm.PREC_FP32, tdm.PREC_BF16, tdm.PREC_FP16, tdm.PREC_FP8):
        net = _make_net(torch_gpu, js, wt, prec, p0)
        a = torch_gpu.empty(len(p), 512, dtype=torch_gpu.float32, device="cuda:0")
        b = torch_gpu.empty_like(a)
        net.forward(torch_gpu.from_numpy(p).to("cuda:0"), a)
        net.forward_grid(torch_gpu.from_numpy(tdf).to("cuda:0"), dims,
                         torch_gpu.from_numpy(kg).to("cuda:0"), b)
        assert torch_gpu.equal(a, b), prec


def test_batch_invariance(torch_gpu, net_files, patches):
    """A patch's descriptor does not depend on the other patches of its pass."""
    js, wt, _ = net_files
    p = patches[0]
    net = tdm.Net(js, tdm.PREC_FP32, max_batch=4)
    net.load_weights(wt)
    full = torch_gpu.empty(len(p), 512, dtype=torch_gpu.float32, device="cuda:0")
    net.forward(torch_gpu.from_numpy(p).to("cuda:0"), full)
    one = torch_gpu.empty(1, 512, dtype=torch_gpu.float32, device="cuda:0")
    for i in range(len(p)):
        net.forward(torch_gpu.from_numpy(p[i:i + 1]).to("cuda:0"), one)
        assert torch_gpu.equal(one[0], full[i])


def test_fp32_vs_oracle_fp32_loop(torch_gpu, net_files, patches):
    """Against the float32 C loop nest (a different summation order): the gap is
    summation-order noise, far below the fp64 tolerance."""
    js, wt, w = net_files
    p = patches[0][:2]
    ref = oracle.net_forward_fp32(p, w, netdef.layer_shapes()).reshape(2, 512)
    net = tdm.Net(js, tdm.PREC_FP32)
    net.load_weights(wt)
    feat = torch_gpu.empty(2, 512, dtype=torch_gpu.float32, device="cuda:0")
    net.forward(torch_gpu.from_numpy(p).to("cuda:0"), feat)
    assert rel_l2(feat.cpu().numpy(), ref) <= FP32_TOL


def test_multi_grid_pass_equals_per_grid(torch_gpu, net_files):
    """Keypoints of three fragments interleaved in one multi-grid pass give
    the same descriptors, bit for bit, as one pass per grid."""
    torch = torch_gpu
    js, wt, _ = net_files
    vs = np.float32(0.01)
    grids, kps, host_tdfs = [], [], []
    for k, (n, view) in enumerate([(30_000, 0), (45_000, 10), (20_000, 25)]):
        cloud = synth.fragment(n, view=view, seed=40 + k)
        dims, origin = oracle.grid_params(cloud, vs, 15)
        occ = oracle.occupancy(cloud, dims, origin, vs)
        tdf = oracle.compute_tdf(occ, vs, np.float32(0.05))
        idx = synth.keypoint_indices(len(cloud), 5 + 2 * k, seed=50 + k)
        kg = oracle.keypoint_grid(cloud[idx], origin, vs)
        if k == 1:  # a patch reaching beyond the grid's corner
            kg = np.ascontiguousarray(np.concatenate([kg, np.array([[1, 2, 0]], np.int32)]))
        grids.append((torch.from_numpy(tdf).to("cuda:0"), dims))
        kps.append(kg)
        host_tdfs.append(tdf)
    order = np.concatenate([np.full(len(k), i) for i, k in enumerate(kps)])
    kp_all = np.concatenate(kps)
    perm = np.random.default_rng(3).permutation(len(order))
    table = tdm.grid_table(grids)
    calib = oracle.extract_patches(host_tdfs[0], kps[0][:4])
    for prec in (tdm.PREC_FP32, tdm.PREC_BF16, tdm.PREC_FP16, tdm.PREC_FP8):
        net = _make_net(torch, js, wt, prec, calib, max_batch=8)
        per = []
        for (t, dims), kg in zip(grids, kps):
            f = torch.empty(len(kg), 512, dtype=torch.float32, device="cuda:0")
            net.forward_grid(t, dims, torch.from_numpy(kg).to("cuda:0"), f)
            per.append(f)
        want = torch.cat(per)[torch.from_numpy(perm).to("cuda:0")]
        got = torch.empty_like(want)
        net.forward_grids(table, torch.from_numpy(np.ascontiguousarray(kp_all[perm])).to("cuda:0"),
                          torch.from_numpy(order[perm].astype(np.int32)).to("cuda:0"), got)
        assert torch.equal(got, want), prec


NUM_VARIANTS = 17  # conv tile variants (tdm_net_set_layer_variant)


def test_tile_variants_bit_identical_and_autotune(torch_gpu, net_files, patches):
    """Every tile variant of every convolution gives the same descriptors bit
    for bit, so the autotuner changes speed only."""
    torch = torch_gpu
    js, wt, _ = net_files
    p = torch.from_numpy(patches[0]).to("cuda:0")
    for prec in (tdm.PREC_FP32, tdm.PREC_BF16, tdm.PREC_FP16, tdm.PREC_FP8):
        net = _make_net(torch, js, wt, prec, patches[0], max_batch=4)
        ref = torch.empty(len(p), 512, dtype=torch.float32, device="cuda:0")
        net.forward(p, ref)
        for v in range(NUM_VARIANTS):
            for i in range(1, len(net.layer_names)):
                net.set_layer_variant(i, v)  # variants that do not fit fall back
            got = torch.empty_like(ref)
            net.forward(p, got)
            assert torch.equal(got, ref), (prec, v)
        chosen = net.autotune(4)
        print("autotuned variants", prec, chosen, flush=True)
        assert all(-1 <= c < NUM_VARIANTS for c in chosen)
        got = torch.empty_like(ref)
        net.forward(p, got)
        assert torch.equal(got, ref)
        net.close()


def test_tile_variants_bit_identical_without_split(torch_gpu, net_files):
    """64 patches at max_batch 64: the early layers run without split-K (conv2
    with its fused pool, conv3, conv4), so every variant's direct epilogue is
    compared too."""
    torch = torch_gpu
    js, wt, _ = net_files
    rng = np.random.default_rng(17)
    p = torch.from_numpy(rng.random((64, 30, 30, 30), dtype=np.float32)).to("cuda:0")
    for prec in (tdm.PREC_FP32, tdm.PREC_BF16, tdm.PREC_FP16):
        net = tdm.Net(js, prec, max_batch=64)
        net.load_weights(wt)
        ref = torch.empty(64, 512, dtype=torch.float32, device="cuda:0")
        net.forward(p, ref)
        for v in range(NUM_VARIANTS):
            for i in range(1, len(net.layer_names)):
                net.set_layer_variant(i, v)
            got = torch.empty_like(ref)
            net.forward(p, got)
            assert torch.equal(got, ref), (prec, v)
        net.close()

