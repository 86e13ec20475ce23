"""Batched form on the GPU: the TDF of many clouds of different sizes in
three launches (bit-identical to one call per cloud), and the whole batch --
TDF plan run + one multi-grid descriptor pass -- captured in a hipGraph and
replayed (every replay equal to the eager result, with eager work and input
changes between replays)."""
import numpy as np
import pytest

import oracle
import tdm
from tdm import netdef, synth

pytestmark = pytest.mark.gpu


def _clouds(torch, specs, vs=np.float32(0.01)):
    out = []
    for k, (n, view, shrink) in enumerate(specs):
        cloud = synth.fragment(max(n, 1), view=view, seed=60 + k)[:n]
        dims, origin = (oracle.grid_params(cloud, vs, 15) if n else ([40, 33, 20],
                        np.zeros(3, np.float32)))
        dims = [max(1, d - shrink) for d in dims]  # shrink: points fall outside
        pts = torch.from_numpy(np.ascontiguousarray(cloud)).to("cuda:0")
        tdf = torch.full((dims[2], dims[1], dims[0]), -1.0, device="cuda:0")
        out.append((pts, tdf, dims, origin))
    return out


SPECS = [(20_000, 0, 0), (0, 0, 0), (55_000, 7, 0), (9_000, 13, 40), (31_000, 29, 0)]


@pytest.mark.parametrize("decode", [tdm.DECODE_EXACT, tdm.DECODE_REFERENCE])
def test_batch_tdf_equals_per_cloud(torch_gpu, decode):
    """Every grid of the plan equals the oracle's occupancy + ComputeTDF of
    its cloud bit for bit (points outside a shrunk grid dropped, an empty
    cloud all zeros), and the per-cloud entry point gives the same bits."""
    torch = torch_gpu
    clouds = _clouds(torch, SPECS)
    plan = tdm.TdfPlan(clouds, 0.01, 0.05, decode)
    plan.run()
    torch.cuda.synchronize()
    mode = "exact" if decode == tdm.DECODE_EXACT else "reference"
    for pts, tdf, dims, origin in clouds:
        print(f"case plan cloud dims={dims} decode={mode}", flush=True)
        occ = oracle.occupancy(pts.cpu().numpy(), dims, origin, np.float32(0.01))
        want = oracle.compute_tdf(occ, np.float32(0.01), np.float32(0.05), mode)
        got = tdf.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), dims
        ref = torch.empty_like(tdf)
        tdm.tdf_from_points(pts, dims, origin, np.float32(0.01), np.float32(0.05), ref, decode)
        torch.cuda.synchronize()
        assert torch.equal(tdf, ref), dims
    plan.close()


def test_batch_reference_decode_two_large_grids(torch_gpu):
    """Reference decode in a plan with two grids above 2^24 voxels (each
    gathers from its exact TDF through the same plan-owned scratch, one
    after the other) and a small one between them (no pass needed): every
    grid equals the oracle's float32-decode ComputeTDF bit for bit."""
    torch = torch_gpu
    vs, tr = np.float32(0.01), np.float32(0.05)
    clouds, hosts = [], []
    # views of a 4 m cube (grids of 23 M and 19 M voxels) around a 1 m one
    for seed, n, extent in ((900, 20_000, 4.0), (1900, 8_000, 1.0), (903, 20_000, 4.0)):
        cloud = synth.cropped_fragment(seed, n, extent)
        dims, origin = oracle.grid_params(cloud, vs, 15)
        pts = torch.from_numpy(np.ascontiguousarray(cloud)).to("cuda:0")
        clouds.append((pts, torch.full((dims[2], dims[1], dims[0]), -1.0, device="cuda:0"), dims,
                       origin))
        hosts.append(cloud)
    nvox = [d[0] * d[1] * d[2] for _, _, d, _ in clouds]
    print(f"case plan reference decode, grids {nvox}", flush=True)
    assert nvox[0] > 1 << 24 and nvox[2] > 1 << 24 and nvox[1] <= 1 << 24
    plan = tdm.TdfPlan(clouds, vs, tr, tdm.DECODE_REFERENCE)
    plan.run()
    torch.cuda.synchronize()
    for cloud, (pts, tdf, dims, origin) in zip(hosts, clouds):
        occ = oracle.occupancy(cloud, dims, origin, vs)
        want = oracle.compute_tdf(occ, vs, tr, "reference")
        assert np.array_equal(tdf.cpu().numpy().view(np.uint32), want.view(np.uint32)), dims
    plan.close()


def test_batch_graph_replay(torch_gpu, tmp_path):
    torch = torch_gpu
    js, wt = str(tmp_path / "n.json"), str(tmp_path / "w.marvin")
    netdef.write_net_json(js)
    tdm.marvin_write(wt, netdef.he_normal_weights(5))
    net = tdm.Net(js, tdm.PREC_BF16, max_batch=64)
    net.load_weights(wt)
    clouds = _clouds(torch, [(20_000, 0, 0), (35_000, 9, 0), (12_000, 21, 0)])
    plan = tdm.TdfPlan(clouds)
    rng = np.random.default_rng(1)
    kps, gids = [], []
    for g, (pts, _, dims, origin) in enumerate(clouds):
        p = pts.cpu().numpy()
        idx = rng.integers(0, len(p), 40 + 10 * g)
        kps.append(tdm.keypoints_to_grid(p[idx], origin, np.float32(0.01)))
        gids.append(np.full(len(idx), g, np.int32))
    kp = torch.from_numpy(np.concatenate(kps)).to("cuda:0")
    gid = torch.from_numpy(np.concatenate(gids)).to("cuda:0")
    table = tdm.grid_table([(tdf, dims) for _, tdf, dims, _ in clouds])
    feat = torch.empty(len(kp), 512, device="cuda:0")

    def step():
        plan.run()
        net.forward_grids(table, kp, gid, feat)

    step()  # warm-up: buffers and uploads happen outside the capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            step()
    torch.cuda.synchronize()
    for rep in range(3):
        # change the inputs in place, then compare a replay with the eager run
        for pts, _, _, _ in clouds:
            pts.add_(torch.randn_like(pts) * 0.002)
        torch.cuda.synchronize()
        step()
        torch.cuda.synchronize()
        want = feat.clone()
        tdfs = [t.clone() for _, t, _, _ in clouds]
        feat.zero_()
        for _, t, _, _ in clouds:
            t.fill_(-3.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(feat, want), rep
        for (_, t, _, _), w in zip(clouds, tdfs):
            assert torch.equal(t, w), rep
    plan.close()
    net.close()
